"""Writes tests/golden/defects.npz: input/output vectors of the reference's own
defect-detector functions (RISA/risaLib/src/DetectorInterpolation/
interpolationFunctions.h:28-97, called from Attenuation.cu:191-196).

The vectors come from oracle/_ref/ref_defects, a driver of ours compiled
against that header where it lies in the reference checkout
(`make -C oracle ref`).  Run this script in a container that has the
checkout; the fixture it writes is data only and travels without it.

Cases cover what the reference tests leave open: dead (constant) columns,
flickering columns whose +-2 neighbours wrap through unsigned modulo at
detector 0 and D-1, runs of flagged detectors, a run across the D-1 -> 0
seam, the (det0 - 1) % D wrap when detector 0 is flagged (a non-power-of-two
D makes that wrap land on an unrelated detector), and threshold sums that
wrap at the segment start ((i - j) % (D/2) with unsigned i).
"""
from __future__ import annotations

import os
import subprocess
import sys
import tempfile
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ROOT = HERE.parents[1]
DRIVER = ROOT / "oracle" / "_ref" / "ref_defects"


def run_reference(data: np.ndarray, tmin: float, tmax: float):
    P, D = data.shape
    with tempfile.TemporaryDirectory() as td:
        fin, fout = os.path.join(td, "in.bin"), os.path.join(td, "out.bin")
        with open(fin, "wb") as f:
            f.write(np.array([D, P], np.uint32).tobytes())
            f.write(np.array([tmin, tmax], np.float64).tobytes())
            f.write(np.ascontiguousarray(data, np.uint16).tobytes())
        subprocess.run([str(DRIVER), fin, fout], check=True)
        raw = open(fout, "rb").read()
    flags = np.frombuffer(raw[:4 * D], np.int32).copy()
    rep = np.frombuffer(raw[4 * D:], np.uint16).reshape(P, D).copy()
    return flags, rep


def base(rng, P, D, level=30000.0, noise=30.0):
    """Smooth object-like structure along the projections (similar variation
    on neighbouring detectors) plus a little noise."""
    p = np.arange(P)[:, None]
    phase = np.linspace(0, 2 * np.pi, D)[None, :]
    x = level + 8000.0 * np.sin(2 * np.pi * p / P + phase) + noise * rng.standard_normal((P, D))
    return np.clip(np.rint(x), 0, 65535).astype(np.uint16)


def cases():
    rng = np.random.default_rng(20240611)
    out = []
    # 1. dead columns in the middle of both segments (low variance)
    d = base(rng, 24, 64)
    d[:, 10] = 200
    d[:, 45] = 201
    out.append(("dead_mid", d, 0.67, 1.5))
    # 2. flickering column (high variance) -> +-2 neighbours flagged
    d = base(rng, 24, 64)
    d[::2, 20] = 60000
    d[1::2, 20] = 100
    out.append(("flicker_mid", d, 0.67, 1.5))
    # 3. flicker at detector 0 and D-1 with non-power-of-two D: (detInd+offset) % D wraps unsigned
    d = base(rng, 20, 72)
    d[::2, 0] = 60000
    d[1::2, 0] = 50
    d[::3, 71] = 65000
    out.append(("flicker_edges_D72", d, 0.67, 1.5))
    # 4. dead detector 0 (interpolation reads (0 - 1) % D) and a run across the seam
    d = base(rng, 16, 72)
    d[:, 0] = 210
    d[:, 70] = 205
    d[:, 71] = 207
    out.append(("dead_seam_D72", d, 0.67, 1.5))
    # 5. runs of adjacent dead detectors and one at the segment boundary D/2
    d = base(rng, 16, 96)
    d[:, 30:34] = 220
    d[:, 48] = 230
    d[:, 47] = 231
    out.append(("dead_runs", d, 0.67, 1.5))
    # 6. the ROFEX width with random dead columns (as the 1 % dead-detector workload)
    d = base(rng, 40, 432)
    dead = rng.choice(432, 5, replace=False)
    d[:, dead] = np.rint(200 + 5 * rng.standard_normal((40, 5))).astype(np.uint16)
    out.append(("rofex432", d, 0.67, 1.5))
    # 7. the alternative thresholds kept in the reference config (0.33, 2.0)
    d = base(rng, 30, 128, noise=200.0)
    d[:, 5] = 300
    d[::2, 100] = 64000
    out.append(("alt_thresholds", d, 0.33, 2.0))
    # 8. noise-dominated data: many detectors flagged, long runs to interpolate
    x = 30000.0 + 300.0 * rng.standard_normal((20, 64))
    out.append(("noise_only", np.rint(x).astype(np.uint16), 0.67, 1.5))
    return out


def main():
    if not DRIVER.exists():
        sys.exit(f"{DRIVER} missing: run `make -C oracle ref` in a container with the reference checkout")
    arrays = {}
    names = []
    for name, data, tmin, tmax in cases():
        flags, rep = run_reference(data, tmin, tmax)
        names.append(name)
        arrays[f"{name}__data"] = data
        arrays[f"{name}__thresh"] = np.array([tmin, tmax], np.float64)
        arrays[f"{name}__flags"] = flags
        arrays[f"{name}__repaired"] = rep
        print(f"{name}: D={data.shape[1]} P={data.shape[0]} flagged={int(flags.sum())} "
              f"changed={int((rep != data).sum())}")
    arrays["names"] = np.array(names)
    np.savez_compressed(HERE / "defects.npz", **arrays)


if __name__ == "__main__":
    main()
