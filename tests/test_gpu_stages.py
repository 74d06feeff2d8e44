"""GPU parity of each stage against the oracle, through the C-ABI.

Tolerances: integer/geometric decisions bit-exact (mask pixels); floating
point as relative RMSE against the FP64 oracle (plain FP32 path: 1e-5).
"""
from __future__ import annotations

import numpy as np
import pytest

from conftest import ROOT, disk_mask, rel_rmse

pytestmark = pytest.mark.gpu


def _t(a, dev):
    import torch

    return torch.from_numpy(np.ascontiguousarray(a)).to(dev)


def _sync():
    import torch

    torch.cuda.synchronize()


def test_attenuation_matches_oracle(risa, orc, geom_c0, frames_c0, cuda):
    import torch
    from risa_amd import device

    print("case: attenuation C0", flush=True)
    g, fs = geom_c0, frames_c0
    mask = orc.area_mask(g)
    att = device.Attenuation(g, fs.avg_ref, fs.avg_dark)
    out = torch.empty(g.fan_shape, dtype=torch.float32, device=cuda)
    for f in range(len(fs.planes)):
        att(_t(fs.raw[f], cuda), int(fs.planes[f]), out)
        _sync()
        got = out.cpu().numpy()
        ref32 = orc.attenuation_f32(g, fs.raw[f], fs.planes[f], fs.avg_ref, fs.avg_dark, mask)
        ref64 = orc.attenuation_f64(g, fs.raw[f], fs.planes[f], fs.avg_ref, fs.avg_dark, mask)
        assert np.array_equal(got == 0, ref32 == 0)           # the area mask, bit-exact
        assert np.max(np.abs(got - ref64)) <= 2e-6 * max(1.0, np.max(np.abs(ref64)))
        assert rel_rmse(got, ref64) < 1e-6


def test_attenuation_batch_equals_single(risa, geom_c0, frames_c0, cuda):
    import torch
    from risa_amd import device

    print("case: attenuation batch C0", flush=True)
    g, fs = geom_c0, frames_c0
    att = device.Attenuation(g, fs.avg_ref, fs.avg_dark)
    raw = _t(fs.raw, cuda)
    outb = torch.empty((len(fs.planes),) + g.fan_shape, dtype=torch.float32, device=cuda)
    att.batch(raw, fs.planes, outb)
    for f in range(len(fs.planes)):
        one = torch.empty(g.fan_shape, dtype=torch.float32, device=cuda)
        att(raw[f], int(fs.planes[f]), one)
        assert torch.equal(one, outb[f])


@pytest.mark.parametrize("cfg", ["c0", "c1"])
def test_reorder_and_module_major_attenuation(risa, orc, cfg, cuda):
    """Reordering on the device (Reordering.cu:127-150, 180-204) bit-exact vs
    the oracle, and the attenuation reading module-major frames through the
    same permutation bit-exact vs Reordering followed by Attenuation: batched
    and single-frame launches, 4-wide loads and (misaligned input) the
    one-sample kernel."""
    import torch
    from risa_amd import device, synth

    g = risa.geom_from_config(risa.config_path(cfg))
    P, D, planes = g.numberOfFanProjections, g.numberOfFanDetectors, g.numberOfPlanes
    print(f"case: reorder + module-major attenuation {cfg} ({P} x {D})", flush=True)
    F = 3
    raw_mm = np.random.default_rng(17).integers(100, 60000, size=(F, P * D), dtype=np.uint16)
    ref = orc.reorder_u16(raw_mm, P, D)
    mm = _t(raw_mm, cuda)
    ordered = _t(np.zeros((F, P, D), np.uint16), cuda)
    device.reorder(mm, F, P, D, ordered)
    _sync()
    np.testing.assert_array_equal(ordered.cpu().numpy(), ref)
    # a misaligned output takes the one-sample kernel
    obuf = _t(np.zeros(F * P * D + 1, np.uint16), cuda)
    device.reorder(mm, F, P, D, obuf[1:])
    _sync()
    np.testing.assert_array_equal(obuf[1:].cpu().numpy().reshape(F, P, D), ref)

    avg_ref, avg_dark = synth.prepare_reference(g, synth.reference_frames(g, 4, seed=917))
    pl = np.array([f % planes for f in range(F)], np.int32)
    plain = device.Attenuation(g, avg_ref, avg_dark)
    fused = device.Attenuation(g, avg_ref, avg_dark)
    fused.input_layout = risa.LAYOUT_MODULE_MAJOR
    assert fused.input_layout == risa.LAYOUT_MODULE_MAJOR and plain.input_layout == risa.LAYOUT_DETECTOR_MAJOR
    want = torch.empty((F,) + g.fan_shape, dtype=torch.float32, device=cuda)
    got = torch.empty_like(want)
    plain.batch(ordered, pl, want)
    fused.batch(mm, pl, got)
    assert torch.equal(got, want)
    # misaligned module-major input takes the one-sample kernel
    buf = torch.zeros(F * P * D + 1, dtype=mm.dtype, device=cuda)
    buf[1:] = mm.reshape(-1)
    for f in range(F):
        one = torch.empty(g.fan_shape, dtype=torch.float32, device=cuda)
        fused(mm[f], int(pl[f]), one)
        assert torch.equal(one, want[f])
        fused(buf[1 + f * P * D:1 + (f + 1) * P * D], int(pl[f]), one)
        assert torch.equal(one, want[f])


def test_chain_module_major_input(risa, orc, geom_c0, frames_c0, cuda):
    """The chain reading ROFEX module-major frames (the reference's H2D ->
    Reordering -> Attenuation, the reordering fused) gives the images of the
    detector-major chain bit for bit: plain and mixed, eager and graph;
    switching the layout back after capture releases the graphs and the next
    graph run reads detector-major frames again."""
    import torch
    from risa_amd import _lib, device

    g, fs = geom_c0, frames_c0
    P, D = g.numberOfFanProjections, g.numberOfFanDetectors
    # module-major frames whose reordering is fs.raw (D = 27 * 16 here)
    mm = np.stack([f.reshape(P, 27, 16).transpose(1, 0, 2).reshape(-1) for f in fs.raw])
    np.testing.assert_array_equal(orc.reorder_u16(mm, P, D), fs.raw)
    raw, raw_mm = _t(fs.raw, cuda), _t(mm, cuda)
    F = len(fs.planes)
    for mode in (risa.CHAIN_PLAIN, risa.CHAIN_MIXED):
        print(f"case: chain module-major input, mode {mode}", flush=True)
        a = device.Chain(g, fs.avg_ref, fs.avg_dark, mode=mode, max_frames=F)
        b = device.Chain(g, fs.avg_ref, fs.avg_dark, mode=mode, max_frames=F)
        b.variant = a.variant
        b.set_input_layout(risa.LAYOUT_MODULE_MAJOR)
        s = torch.cuda.Stream()   # graph replays need a non-default stream
        for use_graph in (False, True):
            ia = torch.empty((F,) + g.img_shape, dtype=torch.float32, device=cuda)
            ib = torch.empty_like(ia)
            _sync()
            with torch.cuda.stream(s):
                a(raw, fs.planes, ia, use_graph=use_graph)
                b(raw_mm, fs.planes, ib, use_graph=use_graph)
            _sync()
            assert torch.equal(ia, ib)
        assert b.graph_count() == 1
        b.set_input_layout(risa.LAYOUT_DETECTOR_MAJOR)
        assert b.graph_count() == 0
        ic = torch.full_like(ia, float("nan"))
        with torch.cuda.stream(s):
            b(raw, fs.planes, ic, use_graph=True)
        _sync()
        assert torch.equal(ia, ic)
        a.close()
        b.close()


def test_defect_repair_matches_reference_vectors(risa, cuda):
    """Defect detection and repair on the device against the vectors written
    by the reference's own interpolationFunctions.h:28-97 (tests/golden/),
    bit-exact, three frames per call; the cases include the unsigned-wrap
    neighbours at detectors 0 and D-1 and a run wrapping across the seam."""
    from risa_amd import device

    z = np.load(ROOT / "tests" / "golden" / "defects.npz", allow_pickle=False)
    for name in z["names"]:
        data, thr = z[f"{name}__data"], z[f"{name}__thresh"]
        P, D = data.shape
        print(f"case: defect repair {name} ({P} x {D})", flush=True)
        frames = _t(np.stack([data] * 3), cuda)
        flags = _t(np.full((3, D), 7, np.int32), cuda)
        device.repair_defects(frames, 3, P, D, float(thr[0]), float(thr[1]), flags)
        _sync()
        for f in range(3):
            np.testing.assert_array_equal(flags[f].cpu().numpy(), z[f"{name}__flags"])
            np.testing.assert_array_equal(frames[f].cpu().numpy(), z[f"{name}__repaired"])


@pytest.mark.parametrize("cfg", ["c0", "c1"])
def test_prepare_reference_on_device_matches_host(risa, cfg, cuda):
    """Flat-field preparation on the device (repair of every reference frame,
    the Q2 block average, the zero dark average) equals the host path bit for
    bit on seeded flat fields with dead columns, and an attenuation handle
    built from the device averages gives the host-built handle's output."""
    import torch
    from risa_amd import device, synth

    g = risa.geom_from_config(risa.config_path(cfg))
    P, D, planes, R = g.numberOfFanProjections, g.numberOfFanDetectors, g.numberOfPlanes, 4
    print(f"case: device flat-field preparation {cfg}", flush=True)
    dead = synth.dead_columns(D, 0.02, 31)
    ref = synth.reference_frames(g, R, seed=77, dead=dead)
    avg_ref_h, avg_dark_h = synth.prepare_reference(g, ref)
    # the host path's repaired frames (its outputs held in named arrays: the
    # call writes through their addresses)
    import ctypes
    from risa_amd import _lib
    rep_h = ref.copy()
    out_ref, out_dark = np.empty_like(avg_ref_h), np.empty_like(avg_dark_h)
    _lib.check(_lib.load().risa_attenuation_prepare_reference(ctypes.byref(g), rep_h.ctypes.data, R,
                                                               out_ref.ctypes.data, out_dark.ctypes.data))
    np.testing.assert_array_equal(out_ref, avg_ref_h)
    frames = _t(ref, cuda)
    avg_ref = torch.empty(avg_ref_h.shape, dtype=torch.float32, device=cuda)
    avg_dark = torch.full(avg_dark_h.shape, 5.0, dtype=torch.float32, device=cuda)
    device.prepare_reference(g, frames, R, avg_ref, avg_dark)
    _sync()
    assert (rep_h != ref).any()   # the dead columns were repaired
    np.testing.assert_array_equal(frames.cpu().numpy(), rep_h)
    np.testing.assert_array_equal(avg_ref.cpu().numpy(), avg_ref_h)
    np.testing.assert_array_equal(avg_dark.cpu().numpy(), avg_dark_h)
    raw = _t(synth.fan_frame(g, planes - 1, 3, 5, dead=dead), cuda)
    a = device.Attenuation(g, avg_ref_h, avg_dark_h)
    b = device.Attenuation(g, avg_ref, avg_dark)
    oa = torch.empty(g.fan_shape, dtype=torch.float32, device=cuda)
    ob = torch.empty_like(oa)
    a(raw, planes - 1, oa)
    b(raw, planes - 1, ob)
    _sync()
    assert torch.equal(oa, ob)


def test_fan2para_matches_oracle(risa, orc, geom_c0, frames_c0, cuda):
    import torch
    from risa_amd import device

    print("case: fan2para C0", flush=True)
    g, fs = geom_c0, frames_c0
    mask = orc.area_mask(g)
    tables = orc.Tables(g, 0)
    att = orc.attenuation_f32(g, fs.raw[0], 0, fs.avg_ref, fs.avg_dark, mask)
    ref32 = orc.gather_f32(g, tables, 0, att)
    ref64 = orc.gather_f64(g, tables, 0, att.astype(np.float64))
    out = torch.empty(g.par_shape, dtype=torch.float32, device=cuda)
    # FP32 weights: the reference's own arithmetic, up to FMA contraction
    device.Fan2Para(g, weights=1)(_t(att, cuda), 0, out)
    _sync()
    got = out.cpu().numpy()
    assert np.array_equal(got == 0, ref32 == 0)
    assert np.max(np.abs(got - ref32)) <= 2e-6 * np.max(np.abs(ref32))
    # FP64 weights (default): agrees with the FP64 oracle to FP32 rounding
    device.Fan2Para(g)(_t(att, cuda), 0, out)
    _sync()
    got = out.cpu().numpy()
    assert np.array_equal(got == 0, ref32 == 0)
    print(f"  fp64-weight gather rel RMSE vs FP64 oracle {rel_rmse(got, ref64):.3e}; "
          f"reference-weight oracle {rel_rmse(ref32, ref64):.3e}", flush=True)
    assert rel_rmse(got, ref64) < 5e-7


def test_fan2para_batch_equals_single(risa, geom_c1, cuda):
    """A batched launch (frames grouped by plane, each workgroup looping over
    its group's frames) gives every frame exactly its single-frame result."""
    import torch
    from risa_amd import device, synth

    print("case: fan2para batch C1, 9 frames over 2 planes", flush=True)
    g = geom_c1
    planes = np.array([0, 1, 1, 0, 1, 0, 0, 1, 0], np.int32)
    rng = np.random.default_rng(17)
    fan = torch.from_numpy(rng.standard_normal((len(planes),) + g.fan_shape).astype(np.float32)).to(cuda)
    f2p = device.Fan2Para(g)
    outb = torch.empty((len(planes),) + g.par_shape, dtype=torch.float32, device=cuda)
    f2p.batch(fan, planes, outb)
    for f in range(len(planes)):
        one = torch.empty(g.par_shape, dtype=torch.float32, device=cuda)
        f2p(fan[f], int(planes[f]), one)
        assert torch.equal(one, outb[f]), f"frame {f}"


def test_filter_fft_matches_oracle(risa, orc, geom_c0, cuda):
    import torch
    from risa_amd import device

    for name, ftype in [("ramp", 0), ("sheppLogan", 1), ("cosine", 2), ("hamming", 3), ("hanning", 4)]:
        print(f"case: filter FFT C0 {name}", flush=True)
        g = geom_c0.replace(filterType=ftype)
        rng = np.random.default_rng(3 + ftype)
        x = rng.standard_normal((3,) + g.par_shape).astype(np.float32)
        flt = device.Filter(g, 0)
        assert np.array_equal(flt.taps(), orc.taps_f32(g))
        t = _t(x, cuda)
        flt(t, 3)
        _sync()
        ref = orc.filter_rows_f64(orc.taps_f32(g).astype(np.float64), x.astype(np.float64))
        assert rel_rmse(t.cpu().numpy(), ref) < 2e-6


@pytest.mark.parametrize("D", [256, 384, 500, 512, 640, 768])
def test_filter_fft_lengths(risa, orc, geom_c0, cuda, D):
    import torch
    from risa_amd import device

    print(f"case: filter FFT length {D}", flush=True)
    g = geom_c0.replace(numberOfParallelDetectors=D, numberOfParallelProjections=7)
    x = np.random.default_rng(D).standard_normal(g.par_shape).astype(np.float32)
    flt = device.Filter(g, 0)
    t = _t(x, cuda)
    flt(t, 1)
    _sync()
    ref = orc.filter_rows_f64(orc.taps_f32(g).astype(np.float64), x.astype(np.float64))
    assert rel_rmse(t.cpu().numpy(), ref) < 2e-6


def test_backprojection_variants_match_oracle(risa, orc, geom_c0, cuda):
    import torch
    from risa_amd import device, synth

    g = geom_c0
    sino = synth.parallel_sinogram(11, g.imageWidth, g.numberOfParallelProjections, g.numberOfParallelDetectors)
    filt = orc.filter_rows_f64(orc.taps_f32(g).astype(np.float64), sino.astype(np.float64)).astype(np.float32)
    ref = orc.masking_f64(g, orc.bp_linear_f64(g, filt.astype(np.float64)))
    m = disk_mask(g.numberOfPixels)
    bp = device.Backprojection(g)
    st = _t(filt, cuda)
    for v in range(64):
        try:
            bp.variant = v
        except risa.RisaError:
            assert v >= 4
            break
        print(f"case: backprojection C0 variant {v}", flush=True)
        img = torch.full(g.img_shape, float("nan"), dtype=torch.float32, device=cuda)
        bp(st, img)
        _sync()
        got = img.cpu().numpy()
        assert np.all(got[~m] == np.float32(g.maskingValue))      # fused mask, bit-exact
        assert rel_rmse(got, ref, m) < 2e-6, v


def test_backprojection_unmasked_and_rotated(risa, orc, geom_c0, cuda):
    import torch
    from risa_amd import device

    print("case: backprojection unmasked, rotationOffset 17.5, N != D", flush=True)
    g = geom_c0.replace(fusedMasking=0, rotationOffset=17.5, numberOfPixels=200, numberOfParallelProjections=180)
    sino = np.random.default_rng(5).standard_normal(g.par_shape).astype(np.float32)
    ref64 = orc.bp_linear_f64(g, sino.astype(np.float64))
    ref32 = orc.bp_linear_f32(g, sino)
    bp = device.Backprojection(g)
    img = torch.empty(g.img_shape, dtype=torch.float32, device=cuda)
    bp(_t(sino, cuda), img)
    _sync()
    got = img.cpu().numpy()
    # White noise is the worst case for rounding of t: the kernel forms
    # t = fma(yp, sin, xp*cos) where the faithful form rounds the product and
    # the sum separately (the reference's nvcc build contracts it in an order
    # we cannot see), and one ulp of t (|t| ~ 100) moves each weight by ~1e-5.
    # Smooth (filtered) sinograms agree to 2e-6 (test above).
    assert rel_rmse(got, ref32) < 1e-5
    assert rel_rmse(got, ref64) < 2e-5


def test_backprojection_nearest_and_tex_bit_exact(risa, orc, geom_c0, cuda):
    import torch
    from risa_amd import device

    sino = np.random.default_rng(9).standard_normal(geom_c0.par_shape).astype(np.float32)
    for label, kw, fn in [("nearest", dict(interpolationType=1), orc.bp_nearest_f32),
                          ("tex", dict(useTextureMemory=1), orc.bp_tex_f32)]:
        print(f"case: backprojection {label}", flush=True)
        g = geom_c0.replace(fusedMasking=0, **kw)
        bp = device.Backprojection(g)
        img = torch.empty(g.img_shape, dtype=torch.float32, device=cuda)
        bp(_t(sino, cuda), img)
        _sync()
        got, ref = img.cpu().numpy(), fn(g, sino)
        # The sample index comes from rounding/truncating t, so one ulp of t
        # at a boundary flips the pick: the kernels evaluate t with the
        # oracle's roundings (no contraction beyond the fused multiply-adds
        # both name) and the sums in the same order, so every pixel is equal.
        print(f"  {label}: {np.count_nonzero(got != ref)} of {got.size} pixels differ", flush=True)
        np.testing.assert_array_equal(got, ref)


def test_masking_stage(risa, orc, geom_c0, cuda):
    import torch
    from risa_amd import device

    for norm in (0, 1):
        print(f"case: masking normalization={norm}", flush=True)
        g = geom_c0.replace(normalization=norm, maskingValue=-1.5)
        img = np.random.default_rng(2).standard_normal(g.img_shape).astype(np.float32)
        t = _t(img, cuda)
        device.Masking(g)(t)
        _sync()
        ref = orc.masking_f32(g, img)
        got = t.cpu().numpy()
        if norm:
            assert np.max(np.abs(got - ref)) <= 1e-6
        else:
            assert np.array_equal(got, ref)


def test_chain_plain_lane_groups_equal_single_frames(risa, geom_c0, frames_c0, cuda):
    """A plain chain call of 37 frames (2 lanes x groups of up to 16 frames,
    the last lane step a partial group, backprojected with multi-slice
    tiles) gives every frame bit for bit what a one-frame chain with the same
    tile variant gives it, eager and as a replayed graph."""
    import torch
    from risa_amd import device

    g, fs = geom_c0, frames_c0
    F = 37
    idx = np.arange(F) % len(fs.planes)
    raw = _t(fs.raw[idx], cuda)
    planes = fs.planes[idx].astype(np.int32)
    a = device.Chain(g, fs.avg_ref, fs.avg_dark, mode=risa.CHAIN_PLAIN, max_frames=F)
    b = device.Chain(g, fs.avg_ref, fs.avg_dark, mode=risa.CHAIN_PLAIN, max_frames=1)
    b.variant = a.variant
    print(f"case: chain plain C0, {F} frames in lane groups, variant {a.variant & 0xFF}", flush=True)
    img = torch.full((F,) + g.img_shape, float("nan"), dtype=torch.float32, device=cuda)
    a(raw, planes, img)
    one = torch.empty((1,) + g.img_shape, dtype=torch.float32, device=cuda)
    for f in range(F):
        b(raw[f:f + 1], planes[f:f + 1], one)
        _sync()
        assert torch.equal(one[0], img[f]), f"frame {f}"
    img2 = torch.full_like(img, float("nan"))
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        a(raw, planes, img2, use_graph=True)
    s.synchronize()
    assert torch.equal(img2, img)
    a.close()
    b.close()


def test_chain_mixed_lane_groups_equal_library_calls(risa, geom_c0, frames_c0, cuda):
    """A mixed chain call of 70 frames (2 lanes, groups of 36 and 34 frames,
    the last x4 slice group half empty) gives every frame bit for bit what
    the four library calls give it over the whole batch (attenuation,
    rebinning, DFT-GEMM filter with x4 FP16 output, FP16 backprojection with
    the chain's tile variant), eager and as a replayed graph."""
    import torch
    from risa_amd import device

    g, fs = geom_c0, frames_c0
    F = 70
    idx = np.arange(F) % len(fs.planes)
    raw = _t(fs.raw[idx], cuda)
    planes = fs.planes[idx].astype(np.int32)
    a = device.Chain(g, fs.avg_ref, fs.avg_dark, mode=risa.CHAIN_MIXED, max_frames=F)
    print(f"case: chain mixed C0, {F} frames in lane groups, variant {a.variant >> 8}", flush=True)
    img = torch.full((F,) + g.img_shape, float("nan"), dtype=torch.float32, device=cuda)
    a(raw, planes, img)
    att = device.Attenuation(g, fs.avg_ref, fs.avg_dark)
    f2p = device.Fan2Para(g)
    flt = device.Filter(g, mode=1)
    bp = device.Backprojection(g)
    bp.variant = a.variant
    fan = torch.empty((F,) + g.fan_shape, dtype=torch.float32, device=cuda)
    att.batch(raw, planes, fan)
    par = torch.empty((F,) + g.par_shape, dtype=torch.float32, device=cuda)
    f2p.batch(fan, planes, par)
    x4 = torch.zeros(((F + 3) // 4 * 4,) + g.par_shape, dtype=torch.float16, device=cuda)
    flt.f16x4(par, x4, F, 256.0)
    ref = torch.full_like(img, float("nan"))
    bp.f16x4(x4, 256.0, F, ref)
    _sync()
    for f in range(F):
        assert torch.equal(img[f], ref[f]), f"frame {f}"
    img2 = torch.full_like(img, float("nan"))
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        a(raw, planes, img2, use_graph=True)
    s.synchronize()
    assert torch.equal(img2, img)
    for h in (a, att, f2p, flt, bp):
        h.close()


def test_chain_plain_matches_fp64_oracle(risa, orc, geom_c0, frames_c0, cuda):
    import torch
    from risa_amd import device

    print("case: chain plain C0", flush=True)
    g, fs = geom_c0, frames_c0
    ch = device.Chain(g, fs.avg_ref, fs.avg_dark, mode=risa.CHAIN_PLAIN, max_frames=len(fs.planes))
    raw = _t(fs.raw, cuda)
    img = torch.empty((len(fs.planes),) + g.img_shape, dtype=torch.float32, device=cuda)
    ch(raw, fs.planes, img)
    _sync()
    got = img.cpu().numpy()
    oc = orc.Chain(g, fs.avg_ref, fs.avg_dark)
    m = disk_mask(g.numberOfPixels)
    for f in range(len(fs.planes)):
        ref = oc.f64(fs.raw[f], int(fs.planes[f]))
        err = rel_rmse(got[f], ref, m)
        print(f"  frame {f}: rel RMSE vs FP64 oracle {err:.3e}", flush=True)
        assert err < 1e-5
        assert np.all(got[f][~m] == 0.0)
    # graph replay (captured on a side stream) gives the same images, replay
    # after replay, with eager work in between
    s = torch.cuda.Stream()
    for rep in range(3):
        img2 = torch.full_like(img, float("nan"))
        with torch.cuda.stream(s):
            ch(raw, fs.planes, img2, use_graph=True)
        s.synchronize()
        assert torch.equal(img, img2), rep
        ch(raw, fs.planes, img)
        _sync()


def test_chain_variant_override(risa, orc, geom_c0, frames_c0, cuda):
    """The chain's backprojector variant can replace the autotuner's pick
    (each tile shape stays within the plain chain's tolerance), also after
    graphs were captured: the change releases them and the next graph run
    captures the new shape."""
    import torch
    from risa_amd import _lib, device

    print("case: chain variant override C0", flush=True)
    g, fs = geom_c0, frames_c0
    oc = orc.Chain(g, fs.avg_ref, fs.avg_dark)
    m = disk_mask(g.numberOfPixels)
    raw = _t(fs.raw[:2], cuda)
    planes = fs.planes[:2]
    refs = [oc.f64(fs.raw[f], int(planes[f])) for f in range(2)]
    for v in (0, 2):
        ch = device.Chain(g, fs.avg_ref, fs.avg_dark, mode=risa.CHAIN_PLAIN, max_frames=2)
        ch.variant = (ch.variant & ~0xFF) | v
        assert ch.variant & 0xFF == v
        img = torch.empty((2,) + g.img_shape, dtype=torch.float32, device=cuda)
        s = torch.cuda.Stream()
        with torch.cuda.stream(s):
            ch(raw, planes, img, use_graph=True)
        s.synchronize()
        got = img.cpu().numpy()
        for f in range(2):
            assert rel_rmse(got[f], refs[f], m) < 1e-5, (v, f)
        w = 2 - v
        ch.variant = (ch.variant & ~0xFF) | w
        assert ch.graph_count() == 0
        img2 = torch.full_like(img, float("nan"))
        with torch.cuda.stream(s):
            ch(raw, planes, img2, use_graph=True)
        s.synchronize()
        got = img2.cpu().numpy()
        for f in range(2):
            assert rel_rmse(got[f], refs[f], m) < 1e-5, (w, f)
        ch.close()


def test_chain_mixed_within_2e4(risa, orc, geom_c0, frames_c0, cuda):
    import torch
    from risa_amd import device

    print("case: chain mixed C0", flush=True)
    g, fs = geom_c0, frames_c0
    ch = device.Chain(g, fs.avg_ref, fs.avg_dark, mode=risa.CHAIN_MIXED, max_frames=len(fs.planes))
    img = torch.empty((len(fs.planes),) + g.img_shape, dtype=torch.float32, device=cuda)
    ch(_t(fs.raw, cuda), fs.planes, img)
    _sync()
    got = img.cpu().numpy()
    oc = orc.Chain(g, fs.avg_ref, fs.avg_dark)
    m = disk_mask(g.numberOfPixels)
    for f in range(len(fs.planes)):
        ref = oc.f64(fs.raw[f], int(fs.planes[f]))
        err = rel_rmse(got[f], ref, m)
        print(f"  frame {f}: mixed rel RMSE vs FP64 oracle {err:.3e}", flush=True)
        assert err < 2e-4


def test_chain_mixed_large_call(risa, orc, geom_c0, frames_c0, cuda):
    """A mixed call of 53 frames (13 whole 4-slice groups and a partial one):
    eager and graph-replayed runs give identical images, each frame equal to
    the same frame run alone."""
    import torch
    from risa_amd import device

    print("case: chain mixed C0, 53 frames per call", flush=True)
    g, fs = geom_c0, frames_c0
    n = 53
    idx = np.arange(n) % len(fs.planes)
    raw = _t(np.ascontiguousarray(fs.raw[idx]), cuda)
    planes = fs.planes[idx]
    ch = device.Chain(g, fs.avg_ref, fs.avg_dark, mode=risa.CHAIN_MIXED, max_frames=n)
    img = torch.empty((n,) + g.img_shape, dtype=torch.float32, device=cuda)
    ch(raw, planes, img)
    _sync()
    s = torch.cuda.Stream()
    for rep in range(2):
        img2 = torch.full_like(img, float("nan"))
        with torch.cuda.stream(s):
            ch(raw, planes, img2, use_graph=True)
        s.synchronize()
        assert torch.equal(img, img2), rep
    one = torch.empty((1,) + g.img_shape, dtype=torch.float32, device=cuda)
    for f in (0, 17, 52):
        ch(raw[f:f + 1], planes[f:f + 1], one)
        _sync()
        assert torch.equal(one[0], img[f]), f


@pytest.mark.gpu
@pytest.mark.parametrize("D,P,slices", [(512, 1000, 3), (256, 500, 5), (500, 37, 2), (768, 130, 4)])
def test_filter_gemm_split_matches_fp64(risa, orc, geom_c0, cuda, D, P, slices):
    """DFT-GEMM on MFMA with the FP16 hi+lo split (FP32 output) against the
    FP64 circulant; the FP16 and x4 FP16 outputs are the same numbers rounded
    once, bit for bit, with the x4 layout's padding slices zero."""
    import torch
    from risa_amd import device

    print(f"case: GEMM filter D={D} P={P} slices={slices}", flush=True)
    g = geom_c0.replace(numberOfParallelDetectors=D, numberOfParallelProjections=P, filterType=1)
    rng = np.random.default_rng(D * 7 + P)
    x = (rng.standard_normal((slices, P, D)) * 2.0).astype(np.float32)
    filt = device.Filter(g, mode=1)
    y = _t(x, cuda)
    filt(y, slices)
    _sync()
    ref = orc.filter_rows_f64(orc.taps_f64(g), x.astype(np.float64))
    err = rel_rmse(y.cpu().numpy(), ref)
    print(f"  GEMM FP32 output rel RMSE vs FP64 {err:.2e}", flush=True)
    assert err < 5e-6
    scale = 256.0
    h16 = torch.empty((slices, P, D), dtype=torch.float16, device=cuda)
    filt.f16(_t(x, cuda), h16, slices, scale)
    groups = (slices + 3) // 4
    x4 = torch.full((groups, P, D, 4), float("nan"), dtype=torch.float16, device=cuda)
    filt.f16x4(_t(x, cuda), x4, slices, scale)
    _sync()
    a = h16.cpu().numpy().view(np.uint16)
    b = x4.cpu().numpy().view(np.uint16)
    for s in range(groups * 4):
        lane = b[s // 4, :, :, s % 4]
        if s < slices:
            np.testing.assert_array_equal(lane, a[s])
        else:
            assert (lane == 0).all()
    err16 = rel_rmse(a.view(np.float16).astype(np.float64) / scale, ref)
    print(f"  FP16 output rel RMSE vs FP64 {err16:.2e}", flush=True)
    assert err16 < 5e-4
    filt.close()


@pytest.mark.gpu
@pytest.mark.parametrize("slices", [1, 4, 6, 18])
def test_backprojection_f16_matches_oracle(risa, orc, geom_c1, cuda, slices):
    """FP16-sinogram / FP32-accumulate backprojection: the plain-layout and
    x4-layout entry points give identical images, within the FP16 weight
    rounding of the oracle on the dequantised sinogram, for every tile
    variant; padding slices of the last group are not written."""
    import torch
    from risa_amd import device, synth

    g = geom_c1.replace(numberOfParallelProjections=360)
    P, D, N = g.numberOfParallelProjections, g.numberOfParallelDetectors, g.numberOfPixels
    scale = 256.0
    sino = np.stack([synth.parallel_sinogram(70 + s, 190.0, P, D) for s in range(slices)])
    q = (sino * scale).astype(np.float16)
    deq = q.astype(np.float32) / np.float32(scale)
    groups = (slices + 3) // 4
    x4 = np.zeros((groups * 4, P, D), np.float16)
    x4[:slices] = q
    x4 = np.ascontiguousarray(x4.reshape(groups, 4, P, D).transpose(0, 2, 3, 1))
    bp = device.Backprojection(g)
    m = disk_mask(N)
    refs = [orc.bp_linear_f32(g, deq[s]) for s in range(slices)]
    v = 0
    while True:
        try:
            bp.variant = v << 8
        except Exception:
            break
        print(f"case: FP16 BP variant {v} slices {slices}", flush=True)
        img_a = torch.full((slices + 1, N, N), 7.0, dtype=torch.float32, device=cuda)
        img_b = torch.full((slices + 1, N, N), 7.0, dtype=torch.float32, device=cuda)
        bp.f16(_t(q.view(np.uint16), cuda), scale, slices, img_a)
        bp.f16x4(_t(x4.view(np.uint16), cuda), scale, slices, img_b)
        _sync()
        a, b = img_a.cpu().numpy(), img_b.cpu().numpy()
        np.testing.assert_array_equal(a, b)
        assert (a[slices] == 7.0).all()
        worst = max(rel_rmse(a[s], refs[s], m) for s in range(slices))
        print(f"  variant {v}: worst rel RMSE vs oracle on dequantised input {worst:.2e}", flush=True)
        assert worst < 2e-4
        assert all((a[s][~m] == g.maskingValue).all() for s in range(slices))
        v += 1
    assert v >= 4
    bp.close()
