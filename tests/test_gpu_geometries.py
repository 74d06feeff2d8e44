"""GPU parity of the paths every detector count takes: the filter at lengths
whose factors are not all 2, 3 and 5 (zero-padded transform of the same
circular convolution), odd counts in the rebinning fold, C3's 96 uneven
geometries filtered *and* backprojected, and run calls recorded into a
hipGraph (they allocate nothing and do not synchronise).  Everything goes
through the C-ABI and is compared with the oracle on the same inputs."""
from __future__ import annotations

import numpy as np
import pytest

from conftest import disk_mask, dot_gap, rel_rmse

pytestmark = pytest.mark.gpu


def _t(a, dev):
    import torch

    return torch.from_numpy(np.ascontiguousarray(a)).to(dev)


def _sync():
    import torch

    torch.cuda.synchronize()


# 7, 257 and 293 are prime; 261 = 9 * 29; 303 = 3 * 101; 1021 is prime;
# 2047 = 23 * 89; 2048 and 1000 take the direct transform.
ANY_D = [7, 257, 261, 293, 303, 511, 1000, 1021, 2047, 2048]


@pytest.mark.parametrize("D", ANY_D)
def test_filter_fft32_any_length(risa, orc, geom_c0, cuda, D):
    """FFT32 filtering of any D against the FP64 circulant (the reference's
    unpadded length-D cuFFT convolution, Filter.cu:155-163)."""
    from risa_amd import device

    print(f"case: FFT32 filter D={D}", flush=True)
    g = geom_c0.replace(numberOfParallelDetectors=D, numberOfParallelProjections=37, filterType=1)
    x = (np.random.default_rng(D).standard_normal((2,) + g.par_shape) * 3.0).astype(np.float32)
    flt = device.Filter(g, 0)
    t = _t(x, cuda)
    flt(t, 2)
    _sync()
    ref = orc.filter_rows_f64(orc.taps_f32(g).astype(np.float64), x.astype(np.float64))
    err = rel_rmse(t.cpu().numpy(), ref)
    print(f"  rel RMSE vs FP64 {err:.2e}", flush=True)
    assert err < 2e-6
    flt.close()


# Transforms longer than 4096 run pass by pass through global memory:
# 2049 = 3 * 683 (padded to 4320), 4099 (prime, padded to 8640), 10007
# (prime, padded to 20250), 12345 = 3 * 5 * 823 (padded to 25000), 5000 and
# 16384 (direct); 3000 and 4096 stay in LDS.  Three slices of 37 rows leave
# the last row pair half empty and pairs straddling sinograms.
LONG_D = [(2049, 37), (3000, 37), (4096, 37), (4099, 37), (5000, 37), (16384, 5), (10007, 5), (12345, 3)]


@pytest.mark.parametrize("D,P", LONG_D)
def test_filter_fft32_long_rows(risa, orc, geom_c0, cuda, D, P):
    """FFT32 filtering of long rows against the FP64 circulant; the frame
    list, the FP16 output and the in-place call give the same values; the
    adjoint passes the dot test; the tap gradient refuses lengths its LDS
    kernels do not hold."""
    import torch
    from risa_amd import _lib, device

    print(f"case: FFT32 filter long rows D={D} P={P}", flush=True)
    S = 3
    g = geom_c0.replace(numberOfParallelDetectors=D, numberOfParallelProjections=P, filterType=3)
    rng = np.random.default_rng(D + P)
    x = (rng.standard_normal((S,) + g.par_shape) * 3.0).astype(np.float32)
    flt = device.Filter(g, 0)
    t = _t(x, cuda)
    flt(t, S)
    _sync()
    got = t.cpu().numpy()
    ref = orc.filter_rows_f64(orc.taps_f32(g).astype(np.float64), x.astype(np.float64))
    err = rel_rmse(got, ref)
    print(f"  rel RMSE vs FP64 {err:.2e}", flush=True)
    assert err < 2e-6
    sinos = [_t(x[k], cuda) for k in range(S)]
    flt.run_list(sinos)
    _sync()
    for k in range(S):
        assert torch.equal(sinos[k], t[k]), f"list sinogram {k}"
    out16 = torch.empty((S,) + g.par_shape, dtype=torch.float16, device=cuda)
    flt.f16(_t(x, cuda), out16, S, 256.0)
    _sync()
    assert torch.equal(out16, (t * 256.0).half())
    y = rng.standard_normal((S,) + g.par_shape).astype(np.float32)
    fty = _t(y, cuda)
    flt.adjoint(fty, S)
    _sync()
    lhs = float(np.sum(got.astype(np.float64) * y))
    rhs = float(np.sum(x.astype(np.float64) * fty.cpu().numpy()))
    gap = abs(lhs - rhs) / max(abs(lhs), abs(rhs))
    print(f"  dot test rel gap {gap:.2e}", flush=True)
    assert gap <= 1e-5
    if D > 2048:
        dt = torch.empty(D // 2 + 1, dtype=torch.float32, device=cuda)
        with pytest.raises(_lib.RisaError, match="RISA_ERR_UNSUPPORTED"):
            flt.grad_taps(t, t, S, dt)
    flt.close()


def test_filter_fft32_long_rows_many_chunks(risa, orc, geom_c0, cuda):
    """1100 rows of 16384 detectors are 550 row pairs, filtered in two chunks
    of the global-memory transform (512 pairs each at this length); rows on
    both sides of the chunk boundary and the last row against the FP64
    circulant."""
    from risa_amd import device

    D, P = 16384, 1100
    print(f"case: FFT32 filter D={D}, {P} rows in two chunks", flush=True)
    g = geom_c0.replace(numberOfParallelDetectors=D, numberOfParallelProjections=P, filterType=1)
    x = (np.random.default_rng(11).standard_normal((P, D)) * 2.0).astype(np.float32)
    flt = device.Filter(g, 0)
    t = _t(x, cuda)
    flt(t, 1)
    _sync()
    rows = [0, 1, 1022, 1023, 1024, 1025, P - 1]
    ref = orc.filter_rows_f64(orc.taps_f32(g).astype(np.float64), x[rows].astype(np.float64))
    err = rel_rmse(t.cpu().numpy()[rows], ref)
    print(f"  rel RMSE vs FP64 {err:.2e}", flush=True)
    assert err < 2e-6
    flt.close()


@pytest.mark.parametrize("D", [7, 257, 261, 293, 303, 1021])
def test_filter_gemm_and_adjoint_any_length(risa, orc, geom_c0, cuda, D):
    """GEMM mode (no FFT length condition) in place against the FP64
    circulant, and the adjoint (the same symmetric operator, through the
    padded FFT) by the dot-product test."""
    import torch
    from risa_amd import device

    print(f"case: GEMM filter and adjoint D={D}", flush=True)
    g = geom_c0.replace(numberOfParallelDetectors=D, numberOfParallelProjections=29, filterType=2)
    rng = np.random.default_rng(D + 5)
    x = (rng.standard_normal((3,) + g.par_shape) * 2.0).astype(np.float32)
    flt = device.Filter(g, mode=1)
    t = _t(x, cuda)
    flt(t, 3)
    _sync()
    ref = orc.filter_rows_f64(orc.taps_f64(g), x.astype(np.float64))
    err = rel_rmse(t.cpu().numpy(), ref)
    print(f"  GEMM rel RMSE vs FP64 {err:.2e}", flush=True)
    assert err < 5e-6
    y = rng.standard_normal((3,) + g.par_shape).astype(np.float32)
    fx, fty = _t(x, cuda), _t(y, cuda)
    flt(fx, 3)
    flt.adjoint(fty, 3)
    _sync()
    lhs = float(np.sum(fx.cpu().numpy().astype(np.float64) * y))
    rhs = float(np.sum(x.astype(np.float64) * fty.cpu().numpy()))
    gap = abs(lhs - rhs) / max(abs(lhs), abs(rhs))
    print(f"  dot test <Fx, y> vs <x, F^T y>: rel gap {gap:.2e}", flush=True)
    assert gap <= 1e-5
    flt.close()


@pytest.mark.parametrize("D,S,gscale", [(7, 1, 1.0), (257, 2, 1e-6), (293, 1, 1e3), (1021, 1, 1.0), (2047, 1, 1.0)])
def test_filter_grad_taps_any_length(risa, orc, geom_c0, cuda, D, S, gscale):
    """Tap gradient for padded lengths (summed cross-correlation, one inverse
    transform, wrap, cosine transform) against the FP64 oracle's
    sum_rows Re(conj(G_k) X_k) w_k."""
    import torch
    from risa_amd import device

    print(f"case: grad taps D={D} slices={S}", flush=True)
    g = geom_c0.replace(numberOfParallelDetectors=D, numberOfParallelProjections=41)
    rng = np.random.default_rng(D * 3 + S)
    x = (rng.standard_normal((S,) + g.par_shape) * 5.0).astype(np.float32)
    gr = (rng.standard_normal((S,) + g.par_shape) * gscale).astype(np.float32)
    flt = device.Filter(g)
    dt = torch.empty(D // 2 + 1, dtype=torch.float32, device=cuda)
    flt.grad_taps(_t(x, cuda), _t(gr, cuda), S, dt)
    _sync()
    ref = orc.grad_taps_f64(x.astype(np.float64), gr.astype(np.float64))
    err = float(np.linalg.norm(dt.cpu().numpy() - ref) / np.linalg.norm(ref))
    print(f"  rel L2 vs FP64 oracle {err:.2e}", flush=True)
    assert err < 1e-5
    flt.close()


@pytest.mark.parametrize("D", [255, 257])
def test_fan2para_odd_detectors(risa, orc, geom_c0, frames_c0, cuda, D):
    """Odd numberOfParallelDetectors: the reference's mirror for odd D
    (cuda_kernels_fan2para.h:164-169) leaves output D/2 with one addend and
    gives output D/2 - 1 three; reference-arithmetic weights agree with the
    faithful FP32 oracle, FP64 weights with the FP64 oracle."""
    import torch
    from risa_amd import device

    print(f"case: fan2para odd D={D}", flush=True)
    g = geom_c0.replace(numberOfParallelDetectors=D)
    fs = frames_c0
    mask = orc.area_mask(g)
    tables = orc.Tables(g, 0)
    att = orc.attenuation_f32(g, fs.raw[1], 0, fs.avg_ref, fs.avg_dark, mask)
    ref32 = orc.gather_f32(g, tables, 0, att)
    ref64 = orc.gather_f64(g, tables, 0, att.astype(np.float64))
    out = torch.empty(g.par_shape, dtype=torch.float32, device=cuda)
    device.Fan2Para(g, weights=1)(_t(att, cuda), 0, out)
    _sync()
    got = out.cpu().numpy()
    assert np.array_equal(got == 0, ref32 == 0)
    assert np.max(np.abs(got - ref32)) <= 2e-6 * np.max(np.abs(ref32))
    h = D // 2
    assert np.max(np.abs(got[:, h - 1] - ref32[:, h - 1])) <= 2e-6 * np.max(np.abs(ref32))
    f2p = device.Fan2Para(g)
    f2p(_t(att, cuda), 0, out)
    _sync()
    got = out.cpu().numpy()
    err = rel_rmse(got, ref64)
    print(f"  FP64-weight gather rel RMSE vs FP64 oracle {err:.3e}", flush=True)
    assert err < 5e-7
    # a batch over both orders of frames equals the single-frame launches
    fan = _t(np.stack([att, att[::-1].copy()]), cuda)
    outb = torch.empty((2,) + g.par_shape, dtype=torch.float32, device=cuda)
    f2p.batch(fan, [0, 0], outb)
    for f in range(2):
        one = torch.empty(g.par_shape, dtype=torch.float32, device=cuda)
        f2p(fan[f], 0, one)
        assert torch.equal(one, outb[f])


def test_fan2para_c1_both_planes(risa, orc, geom_c1, cuda):
    """The production rebinning (768 x 1000 fan -> 512 x 1000, two tilted
    planes, 1 % dead detectors) compared directly with the oracle's gather on
    each plane: reference-arithmetic weights within 2e-6 of the maximum of the
    faithful FP32 oracle (same zero pattern), FP64 weights within 5e-7 rel
    RMSE of the FP64 oracle."""
    import torch
    from conftest import FrameSet
    from risa_amd import device

    g = geom_c1
    fs = FrameSet(g, 1, 2, R=4, dead_fraction=0.01)
    mask = orc.area_mask(g)
    tables = orc.Tables(g, 0)
    f32w, f64w = device.Fan2Para(g, weights=1), device.Fan2Para(g)
    out = torch.empty(g.par_shape, dtype=torch.float32, device=cuda)
    for f in range(2):
        k = int(fs.planes[f])
        print(f"case: fan2para C1 plane {k}", flush=True)
        att = orc.attenuation_f32(g, fs.raw[f], k, fs.avg_ref, fs.avg_dark, mask)
        ref32 = orc.gather_f32(g, tables, k, att)
        ref64 = orc.gather_f64(g, tables, k, att.astype(np.float64))
        f32w(_t(att, cuda), k, out)
        _sync()
        got = out.cpu().numpy()
        assert np.array_equal(got == 0, ref32 == 0)
        assert np.max(np.abs(got - ref32)) <= 2e-6 * np.max(np.abs(ref32))
        f64w(_t(att, cuda), k, out)
        _sync()
        err = rel_rmse(out.cpu().numpy(), ref64)
        print(f"  plane {k}: FP64-weight gather rel RMSE vs FP64 oracle {err:.3e}", flush=True)
        assert err < 5e-7
    tables.close()


@pytest.mark.parametrize("N,P,D,rot,mask", [(512, 720, 512, 0.0, 1), (257, 181, 200, 13.0, 1),
                                             (128, 90, 128, -30.0, 0), (96, 50, 160, 5.0, 1)])
def test_adjoint_vs_fp64_oracle_batch_and_dot(risa, orc, geom_c0, cuda, N, P, D, rot, mask):
    """The adjoint projector (sheared paths for D <= N; the gather kernel for
    D > N, last case) against the FP64 oracle's exact transpose of the
    backprojector on masked images (rel RMSE <= 2e-5: float angle table
    against the oracle's exact angles), the batched entry bit-equal to
    per-slice calls, and the dot-product test <BP s, x> = <s, BP^T x> against
    the GPU's own forward (relative gap over 8 random pairs <= 1e-5)."""
    import torch
    from risa_amd import device

    print(f"case: adjoint N={N} P={P} D={D} rot={rot} mask={mask}", flush=True)
    g = geom_c0.replace(numberOfPixels=N, numberOfParallelProjections=P, numberOfParallelDetectors=D,
                        rotationOffset=rot, fusedMasking=mask, filterType=0)
    S = 3
    rng = np.random.default_rng(N + P + D)
    x = rng.standard_normal((S, N, N)).astype(np.float32)
    bp = device.Backprojection(g)
    dx = _t(x, cuda)
    single = torch.empty((S, P, D), dtype=torch.float32, device=cuda)
    for i in range(S):
        bp.adjoint(dx[i], single[i])
    batch = torch.full((S, P, D), float("nan"), dtype=torch.float32, device=cuda)
    bp.adjoint_batch(dx, S, batch)
    _sync()
    assert torch.equal(single, batch)
    m = disk_mask(N) if mask else np.ones((N, N), bool)
    for i in range(S):
        ref = orc.bp_adjoint_f64(g, np.where(m, x[i], 0.0))
        err = rel_rmse(batch[i].cpu().numpy(), ref)
        print(f"  slice {i}: rel RMSE vs FP64 transpose {err:.2e}", flush=True)
        assert err < 2e-5
    pairs = []
    for trial in range(8):
        s = rng.standard_normal((P, D)).astype(np.float32)
        xi = rng.standard_normal((N, N)).astype(np.float32)
        bs = torch.empty((N, N), dtype=torch.float32, device=cuda)
        btx = torch.empty((P, D), dtype=torch.float32, device=cuda)
        bp(_t(s, cuda), bs)
        bp.adjoint(_t(xi, cuda), btx)
        _sync()
        pairs.append((float(np.sum(np.where(m, bs.cpu().numpy().astype(np.float64), 0.0) * xi)),
                      float(np.sum(s.astype(np.float64) * btx.cpu().numpy()))))
    gap = dot_gap(pairs)
    print(f"  dot test over 8 pairs: rel gap {gap:.2e} (" +
          ", ".join(f"{a:.6e}/{b:.6e}" for a, b in pairs) + ")", flush=True)
    assert gap <= 1e-5
    bp.close()


def _lut(g):
    from risa_amd import _lib

    P = g.numberOfParallelProjections
    s, c, k = np.empty(P, np.float32), np.empty(P, np.float32), np.empty(3, np.float32)
    _lib.check(_lib.load().risa_bp_lut(g, _lib.ptr(s), _lib.ptr(c), _lib.ptr(k)))
    return s, c, k


_WEIGHT_CASES = [(128, 90, 128, -30.0, 0), (97, 41, 80, 13.0, 1), (64, 36, 64, 0.0, 1)]


@pytest.mark.parametrize("N,P,D,rot,mask", _WEIGHT_CASES)
def test_forward_weights_match_rounding_model(risa, geom_c0, cuda, N, P, D, rot, mask):
    """The FP32 linear backprojector's weights, read off one-hot sinograms
    (one sample per slice, so each output pixel is one weight times norm),
    equal the float32 model of its coordinates bit for bit (to one ulp of
    the product with norm), for every tile variant."""
    import torch
    import bp_rounding as br
    from risa_amd import device

    print(f"case: forward weights N={N} P={P} D={D} rot={rot} mask={mask}", flush=True)
    g = geom_c0.replace(numberOfPixels=N, numberOfParallelProjections=P, numberOfParallelDetectors=D,
                        rotationOffset=rot, fusedMasking=mask, filterType=0, maskingValue=0.0)
    sT, cT, k3 = _lut(g)
    scale, norm, center = k3
    cI = int(D * 0.5)
    y, x = np.mgrid[0:N, 0:N]
    inside = ~br.outside(x, y, N, center) if mask else np.ones((N, N), bool)
    rng = np.random.default_rng(N + P)
    S = 24
    angles = rng.choice(P, S, replace=False)
    bins = rng.integers(D // 4, 3 * D // 4, S)
    sino = np.zeros((S, P, D), np.float32)
    sino[np.arange(S), angles, bins] = 1.0
    bp = device.Backprojection(g)
    worst = 0.0
    for v in range(device.bp_num_variants()):
        TX, TY = device.bp_variant_shape(False, v)[:2]
        bp.variant = v
        img = torch.empty((S, N, N), dtype=torch.float32, device=cuda)
        bp.batch(_t(sino, cuda), S, img)
        _sync()
        got = img.cpu().numpy()
        for i in range(S):
            a, kk = angles[i], bins[i]
            k, f = br.weights(x, y, cT[a], sT[a], TX, TY, center, scale, cI)
            w = np.where(k == kk, (1.0 - f).astype(np.float32), np.where(k + 1 == kk, f.astype(np.float32), 0.0))
            want = (w.astype(np.float32) * norm).astype(np.float32)
            err = np.abs(np.where(inside, got[i] - want, 0.0)).max()
            worst = max(worst, err / norm)
            assert err <= np.spacing(norm), f"variant {v} ({TX}x{TY}) angle {a} bin {kk}: max |diff| {err:.3e}"
    print(f"  worst |GPU - model| over {device.bp_num_variants()} variants x {S} angles: {worst:.2e} of norm",
          flush=True)
    bp.close()


@pytest.mark.parametrize("N,S", [(200, 5), (256, 16), (256, 4), (256, 3)])
def test_batch_equals_single_every_variant(risa, geom_c0, cuda, N, S):
    """Every FP32 tile variant, two-slices-per-workgroup ones included,
    reconstructs each slice of a batch bit for bit as a one-slice launch
    does: odd batches (the last workgroup row holds one slice), and batches
    whose slice groups are dealt to XCDs whole (1, 2, 8 or 16 groups on a
    tile grid that is a multiple of 8) or not (3 or 5 groups)."""
    import torch
    from risa_amd import device

    g = geom_c0.replace(numberOfPixels=N, numberOfParallelProjections=120, numberOfParallelDetectors=160)
    N, P, D = g.numberOfPixels, g.numberOfParallelProjections, g.numberOfParallelDetectors
    rng = np.random.default_rng(5 + S)
    sino = _t(rng.standard_normal((S, P, D)).astype(np.float32), cuda)
    bp = device.Backprojection(g)
    for v in range(device.bp_num_variants()):
        shape = device.bp_variant_shape(False, v)
        print(f"case: batch of {S} vs single, variant {v} {shape}", flush=True)
        bp.variant = v
        img = torch.full((S, N, N), float("nan"), dtype=torch.float32, device=cuda)
        bp.batch(sino, S, img)
        one = torch.full((N, N), float("nan"), dtype=torch.float32, device=cuda)
        for i in range(S):
            bp(sino[i], one)
            _sync()
            assert torch.equal(img[i], one), f"variant {v} {shape} slice {i}"
    bp.close()


@pytest.mark.parametrize("N,P,D,rot,mask", _WEIGHT_CASES)
def test_adjoint_weights_match_rounding_model(risa, geom_c0, cuda, N, P, D, rot, mask):
    """The adjoint projector's weights, read off one-hot images (one pixel per
    slice, so each output sample is one weight times norm), equal the same
    float32 model of the forward's coordinates (a few ulps of norm), for
    every tile variant: the adjoint is the exact transpose of the forward
    variant in use."""
    import torch
    import bp_rounding as br
    from risa_amd import device

    print(f"case: adjoint weights N={N} P={P} D={D} rot={rot} mask={mask}", flush=True)
    g = geom_c0.replace(numberOfPixels=N, numberOfParallelProjections=P, numberOfParallelDetectors=D,
                        rotationOffset=rot, fusedMasking=mask, filterType=0, maskingValue=0.0)
    sT, cT, k3 = _lut(g)
    scale, norm, center = k3
    cI = int(D * 0.5)
    rng = np.random.default_rng(3 * N + P)
    S = 16
    px = rng.integers(0, N, 4 * S)
    py = rng.integers(0, N, 4 * S)
    keep = ~br.outside(px, py, N, center) if mask else np.ones(px.shape, bool)
    px, py = px[keep][:S], py[keep][:S]
    assert len(px) == S
    img = np.zeros((S, N, N), np.float32)
    img[np.arange(S), py, px] = 1.0
    bp = device.Backprojection(g)
    worst = 0.0
    for v in range(device.bp_num_variants()):
        TX, TY = device.bp_variant_shape(False, v)[:2]
        bp.variant = v
        sino = torch.empty((S, P, D), dtype=torch.float32, device=cuda)
        bp.adjoint_batch(_t(img, cuda), S, sino)
        _sync()
        got = sino.cpu().numpy()
        for i in range(S):
            want = np.zeros((P, D))
            k, f = br.weights(px[i], py[i], cT, sT, TX, TY, center, scale, cI)
            for kk, ww in ((k, 1.0 - f), (k + 1, f)):
                ok = (kk >= 0) & (kk < D)
                want[np.arange(P)[ok], kk[ok]] += ww[ok] * float(norm)
            err = np.abs(got[i] - want).max()
            worst = max(worst, err / norm)
            assert err <= 4 * np.spacing(norm), \
                f"variant {v} ({TX}x{TY}) pixel ({px[i]}, {py[i]}): max |diff| {err:.3e}"
    print(f"  worst |GPU - model| over 10 variants x {S} pixels: {worst:.2e} of norm", flush=True)
    bp.close()


@pytest.mark.parametrize("N,P,D,rot,mask", [(257, 181, 200, 13.0, 1), (512, 120, 512, 0.0, 1)])
def test_adjoint_methods_agree(risa, geom_c0, cuda, N, P, D, rot, mask):
    """The sheared-path adjoint and the gather adjoint (method 1) are the same
    operator -- both recompute the forward's coordinates -- so they agree to
    float summation order (rel RMSE <= 1e-6)."""
    import torch
    from risa_amd import device

    print(f"case: adjoint methods N={N} P={P} D={D} rot={rot} mask={mask}", flush=True)
    g = geom_c0.replace(numberOfPixels=N, numberOfParallelProjections=P, numberOfParallelDetectors=D,
                        rotationOffset=rot, fusedMasking=mask, filterType=0)
    rng = np.random.default_rng(N * P)
    S = 3
    x = _t(rng.standard_normal((S, N, N)).astype(np.float32), cuda)
    bp = device.Backprojection(g)
    path = torch.empty((S, P, D), dtype=torch.float32, device=cuda)
    gath = torch.empty_like(path)
    bp.adjoint_batch(x, S, path)
    bp.adjoint_method(1)
    bp.adjoint_batch(x, S, gath)
    _sync()
    err = rel_rmse(gath.cpu().numpy(), path.cpu().numpy())
    print(f"  rel RMSE gather vs sheared paths {err:.2e}", flush=True)
    assert err <= 1e-6
    bp.close()


@pytest.mark.parametrize("N,P,D,rot,mask", [(257, 181, 200, 13.0, 1), (128, 90, 128, -30.0, 0)])
def test_adjoint_dot_every_tile_variant(risa, orc, geom_c0, cuda, N, P, D, rot, mask):
    """The adjoint is the transpose of the forward tile variant in use (its
    weights recompute the forward's float coordinates, tile origins
    included): <BP s, x> = <s, BP^T x>, relative gap over 8 random pairs
    <= 1e-5 for every FP32 variant."""
    import torch
    from risa_amd import device

    print(f"case: adjoint dot, every variant, N={N} P={P} D={D} rot={rot} mask={mask}", flush=True)
    g = geom_c0.replace(numberOfPixels=N, numberOfParallelProjections=P, numberOfParallelDetectors=D,
                        rotationOffset=rot, fusedMasking=mask, filterType=0)
    m = disk_mask(N) if mask else np.ones((N, N), bool)
    bp = device.Backprojection(g)
    rng = np.random.default_rng(7 * N + P)
    worst = 0.0
    for v in range(device.bp_num_variants()):
        bp.variant = v
        pairs = []
        for trial in range(8):
            s = rng.standard_normal((P, D)).astype(np.float32)
            x = rng.standard_normal((N, N)).astype(np.float32)
            bs = torch.empty((N, N), dtype=torch.float32, device=cuda)
            btx = torch.empty((P, D), dtype=torch.float32, device=cuda)
            bp(_t(s, cuda), bs)
            bp.adjoint(_t(x, cuda), btx)
            _sync()
            pairs.append((float(np.sum(np.where(m, bs.cpu().numpy().astype(np.float64), 0.0) * x)),
                          float(np.sum(s.astype(np.float64) * btx.cpu().numpy()))))
        gap = dot_gap(pairs)
        worst = max(worst, gap)
        print(f"  variant {v}: dot gap over 8 pairs {gap:.2e}", flush=True)
        assert gap <= 1e-5, f"variant {v}: " + ", ".join(f"{a:.9e}/{b:.9e}" for a, b in pairs)
    print(f"  worst dot gap over {device.bp_num_variants()} variants: {worst:.2e}", flush=True)
    bp.close()


def _uneven_jobs(n=96):
    Ns, Ps = (256, 384, 512, 768), (360, 500, 720, 1000)
    rng = np.random.default_rng(96)
    jobs = []
    for i in range(n):
        jobs.append(dict(N=Ns[i % 4], P=Ps[(i // 4) % 4], D=int(256 + (i * 37) % 513),
                         dtype=(i // 16) % 2, rot=float(rng.uniform(-45, 45)), mask=int(i % 3 != 0),
                         window=(i // 3) % 4))
    return jobs


def test_c3_filtered_uneven_geometries(risa, orc, geom_c0, cuda):
    """C3 as filtered backprojection: each of the 96 geometries (256..768
    detectors, most of them not 2/3/5-smooth) is filtered on the GPU -- FP32
    jobs by the FFT, FP16 jobs by the DFT-GEMM writing FP16 -- and all 96 are
    backprojected in one persistent launch.  FP32 jobs are compared with the
    FP64 oracle's filter + backprojection (rel RMSE <= 2e-5: the kernels take
    the reference's float angle table, the FP64 oracle exact angles, which
    moves t by up to ~1e-4 detector bins at 768^2), FP16 jobs within the
    mixed-precision bound 2e-4."""
    import torch
    from risa_amd import device, synth

    specs = _uneven_jobs()
    scale16 = 256.0
    jobs, refs = [], []
    print(f"case: C3 filtered, {len(specs)} jobs", flush=True)
    for i, s in enumerate(specs):
        g = geom_c0.replace(numberOfPixels=s["N"], numberOfParallelProjections=s["P"],
                            numberOfParallelDetectors=s["D"], rotationOffset=s["rot"], fusedMasking=s["mask"],
                            maskingValue=-1.0, filterType=s["window"])
        sino = synth.parallel_sinogram(5000 + i, 190.0, s["P"], s["D"], rotation=s["rot"])
        flt = device.Filter(g, mode=1 if s["dtype"] else 0)
        if s["dtype"] == 1:
            out16 = torch.empty((s["P"], s["D"]), dtype=torch.float16, device=cuda)
            flt.f16(_t(sino, cuda), out16, 1, scale16)
            dev_sino = out16
        else:
            dev_sino = _t(sino, cuda)
            flt(dev_sino, 1)
        flt.close()
        img = torch.full((s["N"], s["N"]), float("nan"), dtype=torch.float32, device=cuda)
        jobs.append(dict(sino=dev_sino, img=img, N=s["N"], P=s["P"], D=s["D"], dtype=s["dtype"],
                         in_scale=scale16 if s["dtype"] else 1.0, rotationOffset=s["rot"], mask=s["mask"],
                         maskingValue=-1.0))
        filt64 = orc.filter_rows_f64(orc.taps_f64(g), sino.astype(np.float64))
        ref = orc.bp_linear_f64(g, filt64)
        if s["mask"]:
            ref = orc.masking_f64(g.replace(normalization=0), ref)
        refs.append((g, ref))
    _sync()
    bb = device.BatchedBackprojection(jobs)
    bb.autotune()
    for j in jobs:
        j["img"].fill_(float("nan"))
    bb()
    _sync()
    worst = {0: 0.0, 1: 0.0}
    for i, (s, j, (g, ref)) in enumerate(zip(specs, jobs, refs)):
        got = j["img"].cpu().numpy()
        assert np.isfinite(got).all(), i
        m = disk_mask(s["N"]) if s["mask"] else np.ones((s["N"], s["N"]), bool)
        if s["mask"]:
            assert (got[~m] == -1.0).all(), i
        err = rel_rmse(got, ref, m)
        worst[s["dtype"]] = max(worst[s["dtype"]], err)
        assert err < (2e-5 if s["dtype"] == 0 else 2e-4), (i, s, err)
    print(f"  worst rel RMSE vs FP64 filter+BP: FP32 jobs {worst[0]:.2e}, FP16 jobs {worst[1]:.2e}", flush=True)
    bb.close()


def test_run_calls_capture_in_graph(risa, orc, geom_c0, cuda):
    """The run calls that use scratch buffers (GEMM-mode in-place filter, tap
    gradient, FP16 plain-layout backprojection, the gather adjoint (D > N),
    the sheared-path adjoint in two 8-slice passes (D <= N), the long-row
    filter's global-memory passes) and the persistent uneven-batch launch
    (its tile counter reset by a kernel) record into a hipGraph: they
    allocate nothing and do not synchronise.  Each replay -- with eager work
    and synchronisations between replays -- equals the eager result bit for
    bit."""
    import torch
    from risa_amd import device

    g = geom_c0.replace(numberOfParallelDetectors=261, numberOfParallelProjections=120, numberOfPixels=192)
    P, D, N = g.numberOfParallelProjections, g.numberOfParallelDetectors, g.numberOfPixels
    rng = np.random.default_rng(123)
    S = 5
    x0 = _t((rng.standard_normal((S, P, D)) * 2).astype(np.float32), cuda)
    gr = _t(rng.standard_normal((S, P, D)).astype(np.float32), cuda)
    img0 = _t(rng.standard_normal((N, N)).astype(np.float32), cuda)
    S16 = 18                                    # two passes through the 16-slice staging buffer
    h16 = _t((rng.standard_normal((S16, P, D)) * 100).astype(np.float16), cuda)
    fg = device.Filter(g, mode=1)
    ff = device.Filter(g, mode=0)
    bp = device.Backprojection(g)
    gp = g.replace(numberOfParallelDetectors=160)   # scale <= 1: sheared paths
    bpp = device.Backprojection(gp)
    SA = 10
    imgs = _t(rng.standard_normal((SA, N, N)).astype(np.float32), cuda)
    gl = g.replace(numberOfParallelDetectors=4099, numberOfParallelProjections=7)   # transform length 8640
    fl = device.Filter(gl, mode=0)
    xl0 = _t(rng.standard_normal((3,) + gl.par_shape).astype(np.float32), cuda)
    jobs = []
    for i, (Nj, Pj, Dj) in enumerate([(64, 90, 70), (96, 120, 130), (128, 60, 97), (80, 150, 101)]):
        f16 = i % 2
        sino = rng.standard_normal((Pj, Dj)).astype(np.float32)
        jobs.append(dict(sino=_t(sino.astype(np.float16) if f16 else sino, cuda),
                         img=torch.empty((Nj, Nj), dtype=torch.float32, device=cuda), N=Nj, P=Pj, D=Dj,
                         dtype=f16, in_scale=1.0, rotationOffset=7.0 * i, mask=i % 2, maskingValue=-1.0))
    bb = device.BatchedBackprojection(jobs)

    def work(o):
        o["x"].copy_(x0)
        fg(o["x"], S)                           # GEMM mode, in place through the staging buffer
        ff.grad_taps(x0, gr, S, o["dt"])        # padded-length tap gradient
        bp.f16(h16, 256.0, S16, o["img16"])     # plain-layout FP16 through the staging buffer
        bp.adjoint(img0, o["adj"])              # transposed image buffer
        bpp.adjoint_batch(imgs, SA, o["adjp"])  # padded copies, tile origins, partials
        o["xl"].copy_(xl0)
        fl(o["xl"], 3)                          # global-memory transform passes, in place
        bb()                                    # persistent launch over the uneven jobs
        for i, j in enumerate(jobs):
            o[f"bb{i}"].copy_(j["img"])

    def outs():
        o = dict(x=torch.empty_like(x0), dt=torch.empty(D // 2 + 1, device=cuda),
                 img16=torch.empty((S16, N, N), device=cuda), adj=torch.empty((P, D), device=cuda),
                 adjp=torch.empty((SA, P, 160), device=cuda), xl=torch.empty_like(xl0))
        for i, j in enumerate(jobs):
            o[f"bb{i}"] = torch.empty_like(j["img"])
        return o

    print("case: run calls under graph capture (GEMM filter, grad taps, FP16 BP, adjoints, long-row filter, "
          "persistent batch)", flush=True)
    eager = outs()
    work(eager)
    _sync()
    cap = outs()
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph, stream=s):
        work(cap)
    for rep in range(3):
        for v in cap.values():
            v.fill_(float("nan"))
        _sync()
        graph.replay()
        _sync()
        for k in eager:
            assert torch.equal(cap[k], eager[k]), (rep, k)
        work(outs())                            # eager work between replays
        _sync()
    del graph
    for h in (fg, ff, bp, bpp, fl, bb):
        h.close()


def test_chain_graph_cache_bound_and_recapture(risa, orc, geom_c0, frames_c0, cuda):
    """The chain keeps at most 128 graph executables (least recently replayed
    released after a device synchronisation); changing the variant or the
    input layout after capture drains the device, releases the graphs and the
    next graph run captures anew, with results equal to eager runs."""
    import torch
    from risa_amd import device

    g, fs = geom_c0, frames_c0
    raw = _t(fs.raw[:1], cuda)
    planes = fs.planes[:1]
    ch = device.Chain(g, fs.avg_ref, fs.avg_dark, mode=risa.CHAIN_PLAIN, max_frames=1)
    want = torch.empty((1,) + g.img_shape, dtype=torch.float32, device=cuda)
    ch(raw, planes, want)
    _sync()
    n_img = 136
    imgs = torch.full((n_img,) + g.img_shape, float("nan"), dtype=torch.float32, device=cuda)
    s = torch.cuda.Stream()
    print(f"case: chain graph cache, {n_img} buffer sets", flush=True)
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        for k in range(n_img):
            ch(raw, planes, imgs[k:k + 1], use_graph=True)
        for k in range(4):                      # replays of evicted and kept shapes
            ch(raw, planes, imgs[k:k + 1], use_graph=True)
    s.synchronize()
    assert ch.graph_count() <= 128
    for k in range(n_img):
        assert torch.equal(imgs[k], want[0]), k
    print("case: chain variant change after capture", flush=True)
    v0 = ch.variant
    others = [v for v in (0, 1, 2, 3) if v != (v0 & 0xFF)]
    for v in others[:3]:
        ch.variant = (v0 & ~0xFF) | v
        assert ch.graph_count() == 0
        eager = torch.empty_like(want)
        ch(raw, planes, eager)
        out = torch.full_like(want, float("nan"))
        s.wait_stream(torch.cuda.current_stream())   # the fill and the eager run come first
        with torch.cuda.stream(s):
            ch(raw, planes, out, use_graph=True)
        s.synchronize()
        _sync()
        assert ch.graph_count() == 1
        assert torch.equal(out, eager), v
    ch.close()
