"""The *_run_list entry points (separately allocated frames, one launch per
batch) give every frame bit for bit what the contiguous batch entry gives it,
through the C-ABI: attenuation, rebinning (grouped by plane, odd D with its
centre-column kernel), the FFT filter (D = 512 in registers, other lengths,
row pairs that straddle two frames when P is odd) and every FP32
backprojector tile variant, whole and partial slice groups."""
from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _frames(t):
    """Each frame of `t` copied into its own allocation."""
    return [t[i].clone() for i in range(t.shape[0])]


def _sync():
    import torch

    torch.cuda.synchronize()


@pytest.mark.parametrize("F", [7, 70])
def test_attenuation_list_equals_batch(risa, geom_c1, cuda, F):
    """(70 frames: two launches of the 64-frame table.)"""
    import torch
    from risa_amd import device

    print(f"case: attenuation list C1, {F} frames over 2 planes", flush=True)
    g = geom_c1
    rng = np.random.default_rng(3)
    planes = ((np.arange(F) * 7) // 3 % 2).astype(np.int32)
    ref = rng.uniform(20000, 60000, (g.numberOfPlanes,) + g.fan_shape).astype(np.float32)
    dark = rng.uniform(100, 300, (g.numberOfPlanes, g.numberOfFanDetectors)).astype(np.float32)
    att = device.Attenuation(g, ref, dark)
    raw = torch.from_numpy(rng.integers(0, 65535, (len(planes),) + g.fan_shape).astype(np.uint16)).to(cuda)
    outb = torch.empty((len(planes),) + g.fan_shape, dtype=torch.float32, device=cuda)
    att.batch(raw, planes, outb)
    raws = _frames(raw)
    outs = [torch.full(g.fan_shape, float("nan"), dtype=torch.float32, device=cuda) for _ in planes]
    att.run_list(raws, planes, outs)
    _sync()
    for f in range(len(planes)):
        assert torch.equal(outs[f], outb[f]), f"frame {f}"
    att.close()


@pytest.mark.parametrize("D,F", [(512, 9), (511, 9), (512, 70)])
def test_fan2para_list_equals_batch(risa, geom_c1, cuda, D, F):
    import torch
    from risa_amd import device

    print(f"case: fan2para list C1 D={D}, {F} frames over 2 planes", flush=True)
    g = geom_c1.replace(numberOfParallelDetectors=D)
    planes = ((np.arange(F) * 5) // 2 % 2).astype(np.int32)
    rng = np.random.default_rng(17)
    fan = torch.from_numpy(rng.standard_normal((len(planes),) + g.fan_shape).astype(np.float32)).to(cuda)
    f2p = device.Fan2Para(g)
    outb = torch.empty((len(planes),) + g.par_shape, dtype=torch.float32, device=cuda)
    f2p.batch(fan, planes, outb)
    outs = [torch.full(g.par_shape, float("nan"), dtype=torch.float32, device=cuda) for _ in planes]
    f2p.run_list(_frames(fan), planes, outs)
    _sync()
    for f in range(len(planes)):
        assert torch.equal(outs[f], outb[f]), f"frame {f}"
    f2p.close()


@pytest.mark.parametrize("P,D,S", [(1000, 512, 5), (999, 512, 5), (360, 500, 5), (361, 261, 5), (101, 512, 70)])
def test_filter_list_equals_batch(risa, geom_c0, cuda, P, D, S):
    import torch
    from risa_amd import device

    print(f"case: filter list P={P} D={D}, {S} sinograms", flush=True)
    g = geom_c0.replace(numberOfParallelProjections=P, numberOfParallelDetectors=D)
    rng = np.random.default_rng(P + D)
    x = torch.from_numpy(rng.standard_normal((S, P, D)).astype(np.float32)).to(cuda)
    filt = device.Filter(g)
    sinos = _frames(x)
    filt(x, S)
    filt.run_list(sinos)
    _sync()
    for f in range(S):
        assert torch.equal(sinos[f], x[f]), f"sinogram {f}"
    filt.close()


@pytest.mark.parametrize("S,N", [(7, 256), (16, 256), (70, 64)])
def test_backprojection_list_equals_batch_every_variant(risa, geom_c0, cuda, S, N):
    """(70 slices: two launches of the 64-slice table, the second a partial group.)"""
    import torch
    from risa_amd import device

    g = geom_c0.replace(numberOfPixels=N, numberOfParallelProjections=180, numberOfParallelDetectors=200)
    N, P, D = g.numberOfPixels, g.numberOfParallelProjections, g.numberOfParallelDetectors
    rng = np.random.default_rng(S)
    sino = torch.from_numpy(rng.standard_normal((S, P, D)).astype(np.float32)).to(cuda)
    sinos = _frames(sino)
    bp = device.Backprojection(g)
    for v in range(device.bp_num_variants()):
        shape = device.bp_variant_shape(False, v)
        print(f"case: backprojection list of {S}, variant {v} {shape}", flush=True)
        bp.variant = v
        imgb = torch.empty((S, N, N), dtype=torch.float32, device=cuda)
        bp.batch(sino, S, imgb)
        if S > 64:   # the contiguous batch against one-slice launches too
            one = torch.empty((N, N), dtype=torch.float32, device=cuda)
            for k in (0, 63, 64, S - 1):
                bp(sino[k], one)
                _sync()
                assert torch.equal(one, imgb[k]), f"variant {v} slice {k} (batch vs single)"
        imgs = [torch.full((N, N), float("nan"), dtype=torch.float32, device=cuda) for _ in range(S)]
        bp.run_list(sinos, imgs)
        _sync()
        for k in range(S):
            assert torch.equal(imgs[k], imgb[k]), f"variant {v} {shape} slice {k}"
    bp.close()
