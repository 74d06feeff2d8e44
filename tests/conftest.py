"""Shared fixtures.  `-m gpu` tests need the MI355X; everything else runs on
the CPU.  Inputs are seeded synthetic frames from the library's generator."""
from __future__ import annotations

import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")
    config.addinivalue_line("markers", "slow: takes more than a few seconds on the CPU")


@pytest.fixture(scope="session")
def risa():
    import risa_amd

    risa_amd.load()
    return risa_amd


@pytest.fixture(scope="session")
def orc():
    from oracle import oracle

    oracle.lib()
    return oracle


@pytest.fixture(scope="session")
def geom_c0(risa):
    return risa.geom_from_config(risa.config_path("c0"))


@pytest.fixture(scope="session")
def geom_c1(risa):
    return risa.geom_from_config(risa.config_path("c1"))


class FrameSet:
    """Seeded raw frames + flat-field reference for one geometry."""

    def __init__(self, g, config_id, frames, R=4, dead_fraction=0.0):
        from risa_amd import synth

        self.g = g
        planes = g.numberOfPlanes
        self.dead = synth.dead_columns(g.numberOfFanDetectors, dead_fraction, 7000 + config_id) if dead_fraction else None
        ref = synth.reference_frames(g, R, seed=900 + config_id, dead=self.dead)
        self.avg_ref, self.avg_dark = synth.prepare_reference(g, ref)
        self.raw = np.stack([synth.fan_frame(g, f % planes, 1000 * config_id + f // planes, 55 + f, dead=self.dead)
                             for f in range(frames)])
        self.planes = np.array([f % planes for f in range(frames)], np.int32)


@pytest.fixture(scope="session")
def frames_c0(geom_c0):
    return FrameSet(geom_c0, 0, 3)


def gpu_available():
    try:
        import torch

        return torch.cuda.is_available()
    except Exception:
        return False


@pytest.fixture(scope="session")
def cuda():
    import torch

    if not torch.cuda.is_available():
        pytest.fail("GPU test selected but no HIP device is visible")
    return torch.device("cuda", 0)


def rel_rmse(a, ref, mask=None):
    a = np.asarray(a, np.float64)
    ref = np.asarray(ref, np.float64)
    if mask is not None:
        a, ref = a[mask], ref[mask]
    den = np.sqrt(np.mean(ref * ref))
    return float(np.sqrt(np.mean((a - ref) ** 2)) / (den if den > 0 else 1.0))


def dot_gap(pairs):
    """Relative gap of the dot-product test <A u, v> = <u, A^T v> summed over
    random pairs: sum |lhs - rhs| / sum |lhs|.  One pair's ratio is
    ill-conditioned -- with independent zero-mean vectors <A u, v> lands near
    0 now and then, and float32 summation noise (about 1e-6 of the inner
    product's typical size) then reads as a large ratio; summing over pairs
    keeps the denominator at its typical size.  pairs: [(lhs, rhs), ...]."""
    num = sum(abs(a - b) for a, b in pairs)
    den = sum(abs(a) for a, _ in pairs)
    return num / den


def disk_mask(N):
    c = (N - 1) * 0.5
    y, x = np.mgrid[0:N, 0:N].astype(np.float32)
    d = (x - np.float32(c)) ** 2 + (y - np.float32(c)) ** 2
    return d.astype(np.float64) <= N * N * 0.25
