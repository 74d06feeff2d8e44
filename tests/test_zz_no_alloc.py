"""The allocation-free tap gradient (risa_hip_filter_grad_taps_noalloc):
partial sums in a static device array instead of the handle's buffers.
Compared with risa_hip_filter_grad_taps through the C-ABI, bit for bit, on
two calls in a row -- the second with a smaller input -- for a 2/3/5-smooth
detector count (direct transform) and a padded one."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("D", [512, 261])
def test_grad_taps_noalloc_equals_allocating_entry(risa, geom_c0, cuda, D):
    import torch
    from risa_amd import _lib, device

    lib = _lib.load()
    g = geom_c0.replace(numberOfParallelDetectors=D, numberOfParallelProjections=97)
    P = g.numberOfParallelProjections
    flt = device.Filter(g)
    stream = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    rng = np.random.default_rng(D)
    for slices in (5, 2):                      # second call smaller: stale partials must not leak in
        print(f"case: grad taps noalloc D={D} slices={slices}", flush=True)
        x = torch.from_numpy((rng.standard_normal((slices, P, D)) * 3).astype(np.float32)).to(cuda)
        gr = torch.from_numpy(rng.standard_normal((slices, P, D)).astype(np.float32)).to(cuda)
        want = torch.full((D // 2 + 1,), float("nan"), device=cuda)
        got = torch.full((D // 2 + 1,), float("nan"), device=cuda)
        _lib.check(lib.risa_hip_filter_grad_taps(flt.h, _lib.ptr(x), _lib.ptr(gr), slices, _lib.ptr(want), stream))
        _lib.check(lib.risa_hip_filter_grad_taps_noalloc(flt.h, _lib.ptr(x), _lib.ptr(gr), slices, _lib.ptr(got),
                                                         stream))
        torch.cuda.synchronize()
        assert torch.isfinite(want).all()
        assert torch.equal(got, want)
    flt.close()
