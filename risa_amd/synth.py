"""Seeded synthetic inputs (numpy front end of include/risa_synth.h)."""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import Geom, check, default_synth_opts, load


def _opts(dead=None, **kw):
    o = default_synth_opts(**kw)
    if dead is not None and len(dead):
        arr = np.ascontiguousarray(dead, dtype=np.int32)
        o.dead = arr.ctypes.data_as(C.POINTER(C.c_int32))
        o.n_dead = len(arr)
        o._keep = arr  # keep alive
    return o


def dead_columns(n_detectors: int, fraction: float, seed: int) -> np.ndarray:
    """Fixed-per-run set of dead detector columns (C1: 1 %)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    k = int(round(fraction * n_detectors))
    return np.sort(rng.choice(n_detectors, size=k, replace=False)).astype(np.int32)


def fan_frame(g: Geom, plane: int, seed: int, noise_seed: int, dead=None, **kw) -> np.ndarray:
    out = np.empty(g.fan_shape, dtype=np.uint16)
    o = _opts(dead, **kw)
    rc = load().risa_synth_fan_frame(C.byref(g), plane, seed, noise_seed, C.byref(o), out.ctypes.data)
    if rc != 0:
        raise RuntimeError("risa_synth_fan_frame failed")
    return out


def fan_integrals(g: Geom, plane: int, seed: int, **kw) -> np.ndarray:
    out = np.empty(g.fan_shape, dtype=np.float64)
    o = _opts(None, **kw)
    if load().risa_synth_fan_integrals(C.byref(g), plane, seed, C.byref(o), out.ctypes.data) != 0:
        raise RuntimeError("risa_synth_fan_integrals failed")
    return out


def reference_frames(g: Geom, R: int, seed: int, dead=None, **kw) -> np.ndarray:
    """R*planes reference frames, ordered (frame i, plane k) -> block i*planes + k."""
    out = np.empty((R * g.numberOfPlanes,) + g.fan_shape, dtype=np.uint16)
    o = _opts(dead, **kw)
    for b in range(out.shape[0]):
        if load().risa_synth_reference_frame(C.byref(g), seed * 1000003 + b + 1, C.byref(o), out[b].ctypes.data) != 0:
            raise RuntimeError("risa_synth_reference_frame failed")
    return out


def prepare_reference(g: Geom, ref_frames: np.ndarray):
    """Flat-field preparation through the library (Attenuation::init)."""
    R = ref_frames.shape[0] // g.numberOfPlanes
    ref = np.ascontiguousarray(ref_frames.copy())
    avg_ref = np.empty((g.numberOfPlanes,) + g.fan_shape, dtype=np.float32)
    avg_dark = np.empty((g.numberOfPlanes, g.numberOfFanDetectors), dtype=np.float32)
    check(load().risa_attenuation_prepare_reference(C.byref(g), ref.ctypes.data, R, avg_ref.ctypes.data,
                                                    avg_dark.ctypes.data))
    return avg_ref, avg_dark


def parallel_sinogram(seed: int, width: float, P: int, D: int, rotation: float = 0.0, noise_seed: int = 0,
                      **kw) -> np.ndarray:
    out = np.empty((P, D), dtype=np.float32)
    o = _opts(None, **kw)
    if load().risa_synth_parallel_sinogram(seed, width, P, D, rotation, noise_seed, C.byref(o), out.ctypes.data) != 0:
        raise RuntimeError("risa_synth_parallel_sinogram failed")
    return out


def phantom_image(seed: int, width: float, N: int, **kw) -> np.ndarray:
    out = np.empty((N, N), dtype=np.float32)
    o = _opts(None, **kw)
    if load().risa_synth_phantom_image(seed, width, N, C.byref(o), out.ctypes.data) != 0:
        raise RuntimeError("risa_synth_phantom_image failed")
    return out


def phantom(seed: int, width: float, **kw) -> np.ndarray:
    buf = np.empty((64, 6), dtype=np.float64)
    o = _opts(None, **kw)
    n = load().risa_synth_phantom(seed, width, C.byref(o), buf.ctypes.data, 64)
    if n < 0:
        raise RuntimeError("risa_synth_phantom failed")
    return buf[:n].copy()
