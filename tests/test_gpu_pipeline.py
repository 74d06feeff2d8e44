"""The C++ glados-style stage pipeline on the GPU (RISA/example/main.cpp's
wiring): frames streamed host -> device -> host through H2D, optionally
Reordering (module-major ROFEX frames), Attenuation, Fan2Para, Filter,
Backprojection (mask fused), optionally Masking, and D2H -- or, with
--fused, the Reconstruction stage that runs the same span in batches through
the chain.  The slices it saves are checked against the FP64 oracle chain."""
from __future__ import annotations

import json
import subprocess

import numpy as np
import pytest

from conftest import ROOT, disk_mask, rel_rmse

pytestmark = pytest.mark.gpu

EXE = ROOT / "risa_amd" / "bin" / "risa_pipeline"


def run_pipeline(cfg, frames, save, out, extra=()):
    cmd = [str(EXE), str(cfg), "--frames", str(frames), "--unique", "4", "--save", str(save), str(out), *extra]
    print("case:", " ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=240)
    print(r.stdout, r.stderr[-2000:], flush=True)
    assert r.returncode == 0, r.stderr
    return json.loads(r.stdout.strip().splitlines()[-1])


def read_saved(path, N):
    raw = np.fromfile(path, dtype=np.uint8)
    rec = 12 + 4 * N * N
    assert raw.size % rec == 0
    out = []
    for i in range(raw.size // rec):
        hdr = raw[i * rec:i * rec + 12].view(np.int32)
        img = raw[i * rec + 12:(i + 1) * rec].view(np.float32).reshape(N, N)
        out.append((int(hdr[0]), int(hdr[1]), int(hdr[2]), img))
    return out


@pytest.mark.parametrize("cfg,extra,precision", [("c0", (), "fp32"), ("c1", ("--masking",), "fp32"),
                                                 ("c0", ("--module-major",), "fp32"), ("c0", (), "mixed"),
                                                 ("c1", (), "mixed"), ("c1", ("--fused",), "fp32"),
                                                 ("c0", ("--fused", "--module-major"), "fp32"),
                                                 ("c1", ("--fused",), "mixed")])
def test_stage_pipeline_matches_fp64_oracle(risa, orc, cuda, tmp_path, cfg, extra, precision):
    """precision = "mixed" wires FilterMixed (DFT-GEMM on MFMA, FP16 sinogram)
    and BackprojectionMixed (FP16 sinogram, FP32 accumulation): within the
    mixed-precision bound 2e-4; the FP32 stages within 1e-5."""
    from risa_amd import synth

    path = risa.config_path(cfg)
    if precision != "fp32":
        path = tmp_path / f"{cfg}_{precision}.cfg"
        path.write_text(risa.config_path(cfg).read_text() + f'\nprecision = "{precision}"\n')
    g = risa.geom_from_config(path)
    N, planes = g.numberOfPixels, g.numberOfPlanes
    frames = 24 if cfg == "c0" else 16
    out = tmp_path / "slices.bin"
    res = run_pipeline(path, frames, 6, out, extra)
    assert res["frames"] == frames
    assert res["precision"] == precision
    assert res["stages"] == ("fused" if "--fused" in extra else "staged")
    assert res["lost_frames"] == 0
    tol = 1e-5 if precision == "fp32" else 2e-4
    saved = read_saved(out, N)
    assert len(saved) == 6
    # the stage's reference average: R synthetic flat fields, seed 901 (Attenuation(configFile))
    R = 8
    avg_ref, avg_dark = synth.prepare_reference(g, synth.reference_frames(g, R, seed=901))
    oc = orc.Chain(g, avg_ref, avg_dark)
    m = disk_mask(N)
    for idx, u, k, img in saved:
        raw = synth.fan_frame(g, k, 1000 + u, 55 + u * planes + k)
        ref = oc.f64(raw, k)
        err = rel_rmse(img, ref, m)
        print(f"  {cfg} {precision} slice {idx} (frame {u}, plane {k}): rel RMSE {err:.2e}", flush=True)
        assert err < tol
        assert (img[~m] == g.maskingValue).all()
    oc.close()


@pytest.mark.parametrize("extra", [(), ("--fused",)])
def test_stage_pipeline_counts_lost_frames(risa, cuda, tmp_path, extra):
    """A source that leaves out every 5th index (frames lost before the GPU):
    H2D (or the fused Reconstruction stage) counts the missing indices and
    the loss against the last index (H2D.cu:102-107), D2H (or
    Reconstruction) reports its rate (D2H.cu:85-93); every emitted frame is
    reconstructed."""
    frames, k = 40, 5
    res = run_pipeline(risa.config_path("c0"), frames, 0, tmp_path / "unused.bin", ("--skip-every", str(k), *extra))
    # emitted indices: 0, 1, 2, 3, 5, ... (i % 5 == 4 left out)
    idx, i = [], 0
    while len(idx) < frames:
        if i % k != k - 1:
            idx.append(i)
        i += 1
    lost = idx[-1] + 1 - len(idx)
    assert res["frames"] == frames
    assert res["lost_frames"] == lost
    assert abs(res["loss_percent"] - lost / idx[-1] * 100.0) < 1e-3
    assert res["d2h_images_per_s"] > 0


@pytest.mark.parametrize("mode", ["offline", "online"])
def test_stage_pipeline_tiff_output(risa, orc, cuda, tmp_path, mode):
    """The pipeline's sink writing TIFF through risa::OfflineSaver
    (OfflineSaver.cu:66-178): offline, BigTIFF sequences of outputBufferSize
    pages per plane (the remainder at the end); online, the latest slice per
    plane.  Every page read back (Pillow) is one of the replayed frames'
    reconstructions (FP64 oracle, rel RMSE < 1e-5)."""
    from PIL import Image, ImageSequence
    from risa_amd import synth

    src = risa.config_path("c0").read_text()
    src = src.replace('outputPath = "Reco/"', f'outputPath = "{tmp_path}/"')
    src = src.replace('mode = "offline"', f'mode = "{mode}"').replace("outputBufferSize = 100", "outputBufferSize = 5")
    cfg = tmp_path / "out.cfg"
    cfg.write_text(src)
    g = risa.geom_from_config(cfg)
    # online mode writes at most every 20 ms: enough frames to outlast that
    N, unique, frames = g.numberOfPixels, 4, (12 if mode == "offline" else 2000)
    res = run_pipeline(cfg, frames, 0, tmp_path / "unused.bin", ("--tiff",))
    assert res["frames"] == frames
    avg_ref, avg_dark = synth.prepare_reference(g, synth.reference_frames(g, 8, seed=901))
    oc = orc.Chain(g, avg_ref, avg_dark)
    m = disk_mask(N)
    refs = [oc.f64(synth.fan_frame(g, 0, 1000 + u, 55 + u), 0) for u in range(unique)]
    oc.close()
    if mode == "offline":
        names = [f"reco_plane0_{i}seq.tif" for i in range(3)]
        counts = [5, 5, 2]
    else:
        names, counts = ["reco_plane0_0.tif"], [1]
    assert sorted(p.name for p in tmp_path.glob("reco_*.tif")) == sorted(names)
    for name, cnt in zip(names, counts):
        pages = [np.array(pg, dtype=np.float32) for pg in ImageSequence.Iterator(Image.open(tmp_path / name))]
        assert len(pages) == cnt
        for img in pages:
            assert img.shape == (N, N)
            err = min(rel_rmse(img, r, m) for r in refs)
            print(f"  {name}: rel RMSE {err:.2e}", flush=True)
            assert err < 1e-5

