"""GPU parity tests of the workloads beyond the single-stage checks: the C1
production geometry with dead detectors and every window in both precisions
(plain / mixed), uneven batched backprojection in one persistent launch, and
the hand-written adjoint (dot-product tests, the FP64 oracle's gradient and
central finite differences).  Everything is called through the C-ABI."""
from __future__ import annotations

import math

import numpy as np
import pytest

from conftest import FrameSet, disk_mask, rel_rmse

pytestmark = pytest.mark.gpu


def _t(a, dev):
    import torch

    return torch.from_numpy(np.ascontiguousarray(a)).to(dev)


def _sync():
    import torch

    torch.cuda.synchronize()


@pytest.fixture(scope="module")
def frames_c1(geom_c1):
    # one frame per plane, 1 % dead detector columns (also dead in the reference frames)
    return FrameSet(geom_c1, 1, 2, R=4, dead_fraction=0.01)


@pytest.mark.parametrize("window", [("sheppLogan", 1), ("cosine", 2), ("hamming", 3)], ids=lambda w: w[0])
def test_c1_chains_vs_fp64_oracle(risa, orc, geom_c1, frames_c1, cuda, window):
    """Plain FP32 within 1e-5 and mixed precision (DFT-GEMM on MFMA, FP16
    sinogram) within 2e-4 relative RMSE of the FP64 oracle, per frame; the
    max abs error is reported beside it."""
    import torch
    from risa_amd import device

    name, ft = window
    g = geom_c1.replace(filterType=ft)
    fs = frames_c1
    oc = orc.Chain(g, fs.avg_ref, fs.avg_dark)
    refs = [oc.f64(fs.raw[f], int(fs.planes[f])) for f in range(len(fs.planes))]
    oc.close()
    m = disk_mask(g.numberOfPixels)
    for mode, tol in ((risa.CHAIN_PLAIN, 1e-5), (risa.CHAIN_MIXED, 2e-4)):
        print(f"case: C1 chain mode {mode} window {name}", flush=True)
        ch = device.Chain(g, fs.avg_ref, fs.avg_dark, mode=mode, max_frames=len(fs.planes))
        img = torch.empty((len(fs.planes),) + g.img_shape, dtype=torch.float32, device=cuda)
        ch(_t(fs.raw, cuda), fs.planes, img)
        _sync()
        got = img.cpu().numpy()
        ch.close()
        for f, ref in enumerate(refs):
            err = rel_rmse(got[f], ref, m)
            mx = float(np.abs(got[f][m] - ref[m]).max())
            print(f"  {name} mode {mode} frame {f} (plane {fs.planes[f]}): rel RMSE {err:.3e}, max abs {mx:.3e}, "
                  f"ref rms {np.sqrt(np.mean(ref[m] ** 2)):.3e}", flush=True)
            assert err < tol
            assert (got[f][~m] == g.maskingValue).all()


def _uneven_jobs(n=96):
    Ns, Ps = (256, 384, 512, 768), (360, 500, 720, 1000)
    rng = np.random.default_rng(96)
    jobs = []
    for i in range(n):
        jobs.append(dict(N=Ns[i % 4], P=Ps[(i // 4) % 4], D=int(256 + (i * 37) % 513),
                         dtype=(i // 16) % 2, rot=float(rng.uniform(-45, 45)), mask=int(i % 3 != 0)))
    return jobs


def test_c3_batched_uneven_geometries(risa, orc, geom_c0, cuda):
    """96 reconstructions mixing 256^2..768^2 grids, 360..1000 angles and
    256..768 detectors, FP32 and FP16 sinograms, in one persistent launch;
    each against the oracle's backprojector for its own geometry.  The FP16
    jobs are compared with the oracle on the dequantised sinogram."""
    import torch
    from risa_amd import device, synth

    specs = _uneven_jobs()
    scale16 = 256.0
    jobs, host = [], []
    for i, s in enumerate(specs):
        sino = synth.parallel_sinogram(3000 + i, 190.0, s["P"], s["D"], rotation=s["rot"])
        if s["dtype"] == 1:
            q = (sino * scale16).astype(np.float16)
            dev_sino = _t(q.view(np.uint16), cuda)
            ref_in = q.astype(np.float32) / np.float32(scale16)
        else:
            dev_sino = _t(sino, cuda)
            ref_in = sino
        img = torch.full((s["N"], s["N"]), float("nan"), dtype=torch.float32, device=cuda)
        jobs.append(dict(sino=dev_sino, img=img, N=s["N"], P=s["P"], D=s["D"], dtype=s["dtype"],
                         in_scale=scale16 if s["dtype"] else 1.0, rotationOffset=s["rot"], mask=s["mask"],
                         maskingValue=-1.0))
        host.append(ref_in)
    print(f"case: C3 {len(jobs)} uneven jobs", flush=True)
    bb = device.BatchedBackprojection(jobs)
    bb.autotune()
    for rep in range(2):                       # the tile counter is reset by the kernel itself
        for j in jobs:
            j["img"].fill_(float("nan"))
        bb()
        _sync()
        worst = {0: 0.0, 1: 0.0}
        for i, (s, j) in enumerate(zip(specs, jobs)):
            g = geom_c0.replace(numberOfPixels=s["N"], numberOfParallelProjections=s["P"],
                                numberOfParallelDetectors=s["D"], rotationOffset=s["rot"],
                                fusedMasking=s["mask"], maskingValue=-1.0)
            ref = orc.bp_linear_f32(g, host[i])
            got = j["img"].cpu().numpy()
            assert np.isfinite(got).all(), i
            m = disk_mask(s["N"]) if s["mask"] else np.ones((s["N"], s["N"]), bool)
            if s["mask"]:
                ref = orc.masking_f32(g.replace(normalization=0), ref)
                assert (got[~m] == -1.0).all(), i
            err = rel_rmse(got, ref, m)
            worst[s["dtype"]] = max(worst[s["dtype"]], err)
            assert err < (2e-6 if s["dtype"] == 0 else 5e-4), (i, s, err)
        print(f"  pass {rep}: worst rel RMSE FP32 jobs {worst[0]:.2e}, FP16 jobs {worst[1]:.2e}", flush=True)
    bb.close()


# ------------------------------------------------------------------ C4 --
@pytest.fixture(scope="module")
def c4(risa, geom_c0):
    """C4: 16 slices, 512 detectors x 720 angles -> 512^2, ramp taps."""
    g = geom_c0.replace(numberOfParallelDetectors=512, numberOfParallelProjections=720, numberOfPixels=512,
                        filterType=0, fusedMasking=1)
    from risa_amd import synth

    S = 16
    # attenuation sinograms from noisy uint16 counts, seeded as C0
    sino = np.stack([synth.parallel_sinogram(4000 + s, 190.0, 720, 512, noise_seed=77 + s) for s in range(S)])
    target = np.stack([synth.phantom_image(4000 + s, 190.0, 512) for s in range(S)]) * np.float32(190.0 / 512)
    return g, sino.astype(np.float32), target.astype(np.float32)


def test_c4_adjoint_dot_products(risa, c4, cuda):
    """<BP s, x> = <s, BP^T x> and <F s, y> = <s, F^T y> on the GPU kernels,
    accumulated in FP64 on the host: relative gap over 8 random pairs
    (sum |lhs - rhs| / sum |lhs|) <= 1e-5 for each operator."""
    import torch
    from conftest import dot_gap
    from risa_amd import device

    g, sino, _ = c4
    P, D, N = g.numberOfParallelProjections, g.numberOfParallelDetectors, g.numberOfPixels
    rng = np.random.default_rng(44)
    bp = device.Backprojection(g)
    filt = device.Filter(g)
    bpairs, fpairs = [], []
    for trial in range(8):
        s = rng.standard_normal((P, D)).astype(np.float32)
        x = rng.standard_normal((N, N)).astype(np.float32)
        ds, dx = _t(s, cuda), _t(x, cuda)
        bs = torch.empty((N, N), device=cuda)
        btx = torch.empty((P, D), device=cuda)
        bp(ds, bs)
        bp.adjoint(dx, btx)
        _sync()
        bpairs.append((float(np.sum(bs.cpu().numpy().astype(np.float64) * x)),
                       float(np.sum(s.astype(np.float64) * btx.cpu().numpy()))))
        y = rng.standard_normal((P, D)).astype(np.float32)
        fs_, fty = _t(s, cuda), _t(y, cuda)
        filt(fs_, 1)
        filt.adjoint(fty, 1)
        _sync()
        fpairs.append((float(np.sum(fs_.cpu().numpy().astype(np.float64) * y)),
                       float(np.sum(s.astype(np.float64) * fty.cpu().numpy()))))
        print(f"  pair {trial}: BP {bpairs[-1][0]:.9e} vs {bpairs[-1][1]:.9e}, "
              f"filter {fpairs[-1][0]:.9e} vs {fpairs[-1][1]:.9e}", flush=True)
    gb, gf = dot_gap(bpairs), dot_gap(fpairs)
    print(f"  BP dot gap {gb:.2e}, filter dot gap {gf:.2e}", flush=True)
    assert gb <= 1e-5
    assert gf <= 1e-5
    bp.close()
    filt.close()


@pytest.mark.parametrize("D,P,S,gscale", [(512, 720, 3, 1e-7), (500, 9, 1, 1.0), (256, 33, 2, 1e3)])
def test_filter_grad_taps_vs_fp64_oracle(risa, orc, geom_c0, cuda, D, P, S, gscale):
    """d<F_h x, g>/dh on the GPU (one complex FFT per row) against the FP64
    oracle, also when g is many orders of magnitude below x, with odd row counts."""
    import torch
    from risa_amd import device

    g = geom_c0.replace(numberOfParallelDetectors=D, numberOfParallelProjections=P)
    rng = np.random.default_rng(D + P)
    x = (rng.standard_normal((S, P, D)) * 5.0).astype(np.float32)
    gr = (rng.standard_normal((S, P, D)) * gscale).astype(np.float32)
    filt = device.Filter(g)
    dt = torch.empty(D // 2 + 1, dtype=torch.float32, device=cuda)
    filt.grad_taps(_t(x, cuda), _t(gr, cuda), S, dt)
    _sync()
    ref = orc.grad_taps_f64(x.astype(np.float64), gr.astype(np.float64))
    err = float(np.linalg.norm(dt.cpu().numpy() - ref) / np.linalg.norm(ref))
    print(f"  grad taps D={D} rows={S * P} g scale {gscale:g}: rel L2 {err:.2e}", flush=True)
    assert err < 1e-5
    filt.close()


def test_c4_gradients_vs_fp64_oracle_and_finite_differences(risa, orc, c4, cuda):
    """Loss = MSE(FBP(sino; taps), phantom) over 16 slices.  The autograd
    gradients (through the ctypes shim) w.r.t. the attenuation sinogram and
    the 257 filter taps match the FP64 oracle's gradient, and central finite
    differences (h = 1e-3) of the FP64 oracle loss along random directions."""
    import torch
    from autograd_shim import fbp
    from risa_amd import device

    g, sino, target = c4
    S, P, D = sino.shape
    N = g.numberOfPixels
    taps0 = orc.taps_f32(g)
    assert taps0.shape == (257,)
    bp = device.Backprojection(g)
    filt = device.Filter(g)
    x = _t(sino, cuda).requires_grad_(True)
    taps = torch.from_numpy(taps0.copy()).requires_grad_(True)
    tgt = _t(target, cuda)
    img = fbp(x, taps, filt, bp)
    loss = torch.mean((img - tgt) ** 2)
    loss.backward()
    _sync()
    gx = x.grad.cpu().numpy().astype(np.float64)
    gt = taps.grad.numpy().astype(np.float64)
    n = S * N * N

    # FP64 oracle: img = BP(F_h s); dL/dimg = 2 (img - t) / n; ds = F^T BP^T dimg; dh = grad_taps(s, BP^T dimg)
    m = disk_mask(N)
    h64 = taps0.astype(np.float64)

    def forward64(s_, h_, ks=range(S)):
        return np.stack([orc.masking_f64(g, orc.bp_linear_f64(g, orc.filter_rows_f64(h_, s_[k]))) for k in ks])

    def loss64(s_, h_, ks=range(S)):
        ks = list(ks)
        return float(np.sum((forward64(s_, h_, ks) - target[ks].astype(np.float64)) ** 2) / n)

    img64 = forward64(sino.astype(np.float64), h64)
    dimg = 2.0 * (img64 - target) / n
    dimg[:, ~m] = 0.0                                   # masked pixels are constants of the input
    bt = np.stack([orc.bp_adjoint_f64(g, dimg[k]) for k in range(S)])
    gx64 = np.stack([orc.filter_rows_f64(h64, bt[k]) for k in range(S)])
    gt64 = orc.grad_taps_f64(sino.astype(np.float64), bt)
    ex = rel_rmse(gx, gx64)
    et = float(np.linalg.norm(gt - gt64) / np.linalg.norm(gt64))
    print(f"  grad sino vs FP64 oracle: rel RMSE {ex:.2e}; grad taps rel L2 {et:.2e}", flush=True)
    assert ex < 1e-4
    assert et < 1e-4

    # central differences of the FP64 loss, h = 1e-3 (the loss is quadratic:
    # the central difference is exact up to rounding)
    rng = np.random.default_rng(7)
    hstep = 1e-3
    for trial in range(2):
        # directions with a large projection on the gradient (a white-noise
        # direction's projection cancels down to the FP32 gradient's noise):
        # the analytic gradient itself on one slice, then a smooth random field
        v = np.zeros_like(sino, dtype=np.float64)
        k = trial * 5
        if trial == 0:
            v[k] = gx[k] / np.abs(gx[k]).max()
        else:
            v[k] = np.cumsum(np.cumsum(rng.standard_normal((P, D)), axis=0), axis=1)
            v[k] /= np.abs(v[k]).max()
        # v touches slice k only, so only that slice's share of the loss changes
        fd = (loss64(sino + hstep * v, h64, [k]) - loss64(sino - hstep * v, h64, [k])) / (2 * hstep)
        an = float(np.sum(gx * v))
        print(f"  FD sino dir {trial}: fd {fd:.6e} analytic {an:.6e}", flush=True)
        assert abs(fd - an) <= 1e-4 * abs(fd) + 1e-12
        u = rng.standard_normal(257) * 1e-3
        fd = (loss64(sino, h64 + hstep * u) - loss64(sino, h64 - hstep * u)) / (2 * hstep)
        an = float(np.sum(gt * u))
        print(f"  FD taps dir {trial}: fd {fd:.6e} analytic {an:.6e}", flush=True)
        assert abs(fd - an) <= 1e-4 * abs(fd) + 1e-12
    bp.close()
    filt.close()
