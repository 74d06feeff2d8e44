"""Device wrappers over the C-ABI for torch tensors (test and bench harness).

Every call goes to librisa_hip.so on the tensor's device and torch's current
stream; there is no CPU path here.  Handles are released by close() or
when the wrapper is collected.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import BpJob, Geom, check, load, ptr


def _stream():
    import torch

    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def _dev(device) -> int:
    import torch

    d = torch.device(device) if not isinstance(device, int) else torch.device("cuda", device)
    return 0 if d.index is None else d.index


class _Handle:
    _destroy = None

    def __init__(self):
        self.h = C.c_void_p()

    def close(self):
        if self.h and self.h.value:
            check(getattr(load(), self._destroy)(self.h))
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def _ptr_list(tensors):
    """A host array of the device addresses of `tensors` (for the *_run_list entries)."""
    arr = (C.c_void_p * max(1, len(tensors)))(*[t.data_ptr() for t in tensors])
    return arr


def _require_cuda(*tensors):
    for t in tensors:
        if t is not None and not t.is_cuda:
            raise ValueError("device tensors are required")


class Attenuation(_Handle):
    _destroy = "risa_hip_attenuation_destroy"

    def __init__(self, g: Geom, avg_ref, avg_dark, device=0):
        """avg_ref / avg_dark: host arrays, or device tensors (the table is then built on the device)."""
        super().__init__()
        self.g = g
        if hasattr(avg_ref, "is_cuda") and avg_ref.is_cuda:
            _require_cuda(avg_dark)
            check(load().risa_hip_attenuation_create_device(C.byref(g), ptr(avg_ref), ptr(avg_dark), _dev(device),
                                                            _stream(), C.byref(self.h)))
            return
        avg_ref = np.ascontiguousarray(avg_ref, dtype=np.float32)
        avg_dark = np.ascontiguousarray(avg_dark, dtype=np.float32)
        check(load().risa_hip_attenuation_create(C.byref(g), ptr(avg_ref), ptr(avg_dark), _dev(device), C.byref(self.h)))

    def __call__(self, raw, plane: int, out):
        _require_cuda(raw, out)
        check(load().risa_hip_attenuation_run(self.h, ptr(raw), plane, ptr(out), _stream()))
        return out

    def batch(self, raw, planes, out):
        _require_cuda(raw, out)
        pl = np.ascontiguousarray(planes, dtype=np.int32)
        check(load().risa_hip_attenuation_run_batch(self.h, ptr(raw), ptr(pl), len(pl), ptr(out), _stream()))
        return out

    def run_list(self, raws, planes, outs):
        """Separately allocated frames raws[f] -> outs[f]."""
        _require_cuda(*raws, *outs)
        pl = np.ascontiguousarray(planes, dtype=np.int32)
        a, b = _ptr_list(raws), _ptr_list(outs)
        check(load().risa_hip_attenuation_run_list(self.h, C.addressof(a), ptr(pl), len(pl), C.addressof(b), _stream()))
        return outs

    @property
    def input_layout(self) -> int:
        v = C.c_int()
        check(load().risa_hip_attenuation_get_input_layout(self.h, C.byref(v)))
        return v.value

    @input_layout.setter
    def input_layout(self, layout: int):
        """RISA_LAYOUT_DETECTOR_MAJOR (0) or RISA_LAYOUT_MODULE_MAJOR (1)."""
        check(load().risa_hip_attenuation_set_input_layout(self.h, layout))


def repair_defects(frames, nframes: int, P: int, D: int, thresh_min: float, thresh_max: float, flags=None):
    """Defect detection and repair of device frames [nframes][P][D] in place."""
    _require_cuda(frames, flags)
    check(load().risa_hip_repair_defects(ptr(frames), nframes, P, D, thresh_min, thresh_max,
                                         ptr(flags) if flags is not None else None, _stream()))
    return frames


def prepare_reference(g: Geom, ref_frames, R: int, avg_ref, avg_dark):
    """Flat-field preparation of R*planes device reference frames (repaired in place)."""
    _require_cuda(ref_frames, avg_ref, avg_dark)
    check(load().risa_hip_attenuation_prepare_reference(C.byref(g), ptr(ref_frames), R, ptr(avg_ref), ptr(avg_dark),
                                                        _stream()))
    return avg_ref, avg_dark


def reorder(raw, frames: int, Pf: int, Df: int, out):
    """Module-major ROFEX frames -> detector-major (the Reordering stage)."""
    _require_cuda(raw, out)
    check(load().risa_hip_reorder_run(ptr(raw), frames, Pf, Df, ptr(out), _stream()))
    return out


class Fan2Para(_Handle):
    _destroy = "risa_hip_fan2para_destroy"

    def __init__(self, g: Geom, device=0, weights: int = 0):
        """weights: 0 = FP64 interpolation fractions (default), 1 = FP32 as the reference."""
        super().__init__()
        self.g = g
        check(load().risa_hip_fan2para_create_ex(C.byref(g), _dev(device), weights, C.byref(self.h)))

    def __call__(self, fan, plane: int, out):
        _require_cuda(fan, out)
        check(load().risa_hip_fan2para_run(self.h, ptr(fan), plane, ptr(out), _stream()))
        return out

    def batch(self, fan, planes, out):
        _require_cuda(fan, out)
        pl = np.ascontiguousarray(planes, dtype=np.int32)
        check(load().risa_hip_fan2para_run_batch(self.h, ptr(fan), ptr(pl), len(pl), ptr(out), _stream()))
        return out

    def run_list(self, fans, planes, outs):
        """Separately allocated frames fans[f] -> outs[f]."""
        _require_cuda(*fans, *outs)
        pl = np.ascontiguousarray(planes, dtype=np.int32)
        a, b = _ptr_list(fans), _ptr_list(outs)
        check(load().risa_hip_fan2para_run_list(self.h, C.addressof(a), ptr(pl), len(pl), C.addressof(b), _stream()))
        return outs

    def table_bytes(self) -> int:
        n = C.c_size_t()
        check(load().risa_hip_fan2para_table_bytes(self.h, C.byref(n)))
        return n.value


class Filter(_Handle):
    _destroy = "risa_hip_filter_destroy"

    def __init__(self, g: Geom, mode: int = 0, device=0):
        super().__init__()
        self.g = g
        self.mode = mode
        check(load().risa_hip_filter_create(C.byref(g), _dev(device), mode, C.byref(self.h)))

    def __call__(self, sino, slices: int = 1):
        _require_cuda(sino)
        check(load().risa_hip_filter_run(self.h, ptr(sino), slices, _stream()))
        return sino

    def run_list(self, sinos):
        """In place over separately allocated sinograms (FFT32 mode)."""
        _require_cuda(*sinos)
        a = _ptr_list(sinos)
        check(load().risa_hip_filter_run_list(self.h, C.addressof(a), len(sinos), _stream()))
        return sinos

    def f16(self, sino, out16, slices: int, out_scale: float):
        _require_cuda(sino, out16)
        check(load().risa_hip_filter_run_f16(self.h, ptr(sino), ptr(out16), slices, out_scale, _stream()))
        return out16

    def f16x4(self, sino, out_x4, slices: int, out_scale: float):
        _require_cuda(sino, out_x4)
        check(load().risa_hip_filter_run_f16x4(self.h, ptr(sino), ptr(out_x4), slices, out_scale, _stream()))
        return out_x4

    def adjoint(self, grad, slices: int = 1):
        _require_cuda(grad)
        check(load().risa_hip_filter_adjoint_run(self.h, ptr(grad), slices, _stream()))
        return grad

    def grad_taps(self, x, g, slices: int, dtaps):
        _require_cuda(x, g, dtaps)
        check(load().risa_hip_filter_grad_taps(self.h, ptr(x), ptr(g), slices, ptr(dtaps), _stream()))
        return dtaps

    def set_taps(self, taps: np.ndarray):
        t = np.ascontiguousarray(taps, dtype=np.float32)
        check(load().risa_hip_filter_set_taps(self.h, ptr(t)))

    def taps(self) -> np.ndarray:
        t = np.empty(self.g.numberOfParallelDetectors // 2 + 1, dtype=np.float32)
        check(load().risa_hip_filter_get_taps(self.h, ptr(t)))
        return t


def bp_num_variants(f16: bool = False) -> int:
    """Number of FP32 (or FP16-input) backprojection tile variants."""
    shape = (C.c_int * 7)()
    n = 0
    while load().risa_hip_bp_variant_shape(1 if f16 else 0, n, shape) == 0:
        n += 1
    return n


def bp_variant_shape(f16: bool, variant: int) -> tuple:
    """(tile width, tile height, pixels per thread, E, U, G, max angles) of a variant."""
    shape = (C.c_int * 7)()
    check(load().risa_hip_bp_variant_shape(1 if f16 else 0, variant, shape))
    return tuple(shape)


def bp_kernel_name(f16: bool, variant: int) -> str:
    """The backprojection kernel a variant launches, as profilers name it."""
    tx, ty, ppt, e, u, g, _ = bp_variant_shape(f16, variant)
    if f16:
        return f"risa::bp_linear_f16_kernel<{tx}, {ty}, {ppt}, {e}, {u}, {g}>"
    return f"risa::bp_linear_kernel<{tx}, {ty}, {ppt}, {e}, {u}, {g}>"


class Backprojection(_Handle):
    _destroy = "risa_hip_bp_destroy"

    def __init__(self, g: Geom, device=0):
        super().__init__()
        self.g = g
        check(load().risa_hip_bp_create(C.byref(g), _dev(device), C.byref(self.h)))

    def __call__(self, sino, img):
        _require_cuda(sino, img)
        check(load().risa_hip_bp_run(self.h, ptr(sino), ptr(img), _stream()))
        return img

    def batch(self, sino, slices: int, img):
        _require_cuda(sino, img)
        check(load().risa_hip_bp_run_batch(self.h, ptr(sino), slices, ptr(img), _stream()))
        return img

    def run_list(self, sinos, imgs):
        """Separately allocated slices sinos[k] -> imgs[k]."""
        _require_cuda(*sinos, *imgs)
        a, b = _ptr_list(sinos), _ptr_list(imgs)
        check(load().risa_hip_bp_run_list(self.h, C.addressof(a), len(sinos), C.addressof(b), _stream()))
        return imgs

    def f16(self, sino16, in_scale: float, slices: int, img):
        _require_cuda(sino16, img)
        check(load().risa_hip_bp_run_f16(self.h, ptr(sino16), in_scale, slices, ptr(img), _stream()))
        return img

    def f16x4(self, sino_x4, in_scale: float, slices: int, img):
        _require_cuda(sino_x4, img)
        check(load().risa_hip_bp_run_f16x4(self.h, ptr(sino_x4), in_scale, slices, ptr(img), _stream()))
        return img

    def adjoint(self, img, sino):
        _require_cuda(img, sino)
        check(load().risa_hip_bp_adjoint_run(self.h, ptr(img), ptr(sino), _stream()))
        return sino

    def adjoint_batch(self, img, slices: int, sino):
        _require_cuda(img, sino)
        check(load().risa_hip_bp_adjoint_run_batch(self.h, ptr(img), slices, ptr(sino), _stream()))
        return sino

    def adjoint_method(self, method: int):
        """0: sheared-path kernel where it applies (default); 1: gather kernel."""
        check(load().risa_hip_bp_adjoint_method(self.h, method))

    def autotune(self):
        check(load().risa_hip_bp_autotune(self.h, _stream()))

    def autotune_slices(self, slices: int):
        check(load().risa_hip_bp_autotune_slices(self.h, slices, _stream()))

    def autotune_batch(self, slices: int):
        """FP32 tile shape for batched launches of `slices` slices (two-slice tiles included)."""
        check(load().risa_hip_bp_autotune_batch(self.h, slices, _stream()))

    @property
    def slices_per_tile(self) -> int:
        n = C.c_int()
        check(load().risa_hip_bp_slices_per_tile(self.h, C.byref(n)))
        return n.value

    def autotune_f16(self, slices: int):
        check(load().risa_hip_bp_autotune_f16(self.h, slices, _stream()))

    @property
    def variant(self) -> int:
        v = C.c_int()
        check(load().risa_hip_bp_get_variant(self.h, C.byref(v)))
        return v.value

    @variant.setter
    def variant(self, v: int):
        check(load().risa_hip_bp_set_variant(self.h, v))


class Masking(_Handle):
    _destroy = "risa_hip_mask_destroy"

    def __init__(self, g: Geom, device=0):
        super().__init__()
        check(load().risa_hip_mask_create(C.byref(g), _dev(device), C.byref(self.h)))

    def __call__(self, img):
        _require_cuda(img)
        check(load().risa_hip_mask_run(self.h, ptr(img), _stream()))
        return img


class BatchedBackprojection(_Handle):
    """One persistent launch over jobs of different geometry."""

    _destroy = "risa_hip_bp_batched_destroy"

    def __init__(self, jobs, device=0):
        super().__init__()
        self.jobs = jobs
        arr = (BpJob * len(jobs))(*[self._job(**j) for j in jobs])
        check(load().risa_hip_bp_batched_create(arr, len(jobs), _dev(device), C.byref(self.h)))

    @staticmethod
    def _job(sino, img, N, P, D, dtype=0, in_scale=1.0, rotationOffset=0.0, mask=1, maskingValue=0.0):
        return BpJob(ptr(sino), ptr(img), N, P, D, dtype, in_scale, rotationOffset, mask, maskingValue)

    def __call__(self):
        check(load().risa_hip_bp_batched_run(self.h, _stream()))

    def autotune(self):
        check(load().risa_hip_bp_batched_autotune(self.h, _stream()))

    def set_variant(self, v: int):
        check(load().risa_hip_bp_batched_set_variant(self.h, v))


class Chain(_Handle):
    """The whole per-frame chain (plain FP32 or mixed precision)."""

    _destroy = "risa_hip_chain_destroy"

    def __init__(self, g: Geom, avg_ref: np.ndarray, avg_dark: np.ndarray, mode: int = 0, max_frames: int = 8,
                 device=0):
        super().__init__()
        self.g = g
        avg_ref = np.ascontiguousarray(avg_ref, dtype=np.float32)
        avg_dark = np.ascontiguousarray(avg_dark, dtype=np.float32)
        check(load().risa_hip_chain_create(C.byref(g), ptr(avg_ref), ptr(avg_dark), _dev(device), mode, max_frames,
                                           C.byref(self.h)))

    def __call__(self, raw, planes, img, use_graph: bool = False):
        _require_cuda(raw, img)
        pl = np.ascontiguousarray(planes, dtype=np.int32)
        check(load().risa_hip_chain_run(self.h, ptr(raw), ptr(pl), len(pl), ptr(img), 1 if use_graph else 0,
                                        _stream()))
        return img

    def intermediate(self, which: int) -> int:
        p = C.c_void_p()
        check(load().risa_hip_chain_intermediate(self.h, which, C.byref(p)))
        return p.value

    @property
    def variant(self) -> int:
        v = C.c_int()
        check(load().risa_hip_chain_get_variant(self.h, C.byref(v)))
        return v.value

    @variant.setter
    def variant(self, v: int):
        check(load().risa_hip_chain_set_variant(self.h, v))

    def set_input_layout(self, layout: int):
        check(load().risa_hip_chain_set_input_layout(self.h, layout))

    def graph_count(self) -> int:
        n = C.c_int()
        check(load().risa_hip_chain_graph_count(self.h, C.byref(n)))
        return n.value
