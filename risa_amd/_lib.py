"""ctypes binding of librisa_hip.so (the C-ABI declared in include/risa_hip.h
and include/risa_synth.h).

Loading fails loudly when the library has not been built: there is no
CPU fallback behind these functions.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

PKG_DIR = Path(__file__).resolve().parent
LIB_PATH = PKG_DIR / "librisa_hip.so"

MAX_PLANES = 8

RISA_OK = 0
STATUS_NAMES = {0: "RISA_OK", 1: "RISA_ERR_ARG", 2: "RISA_ERR_CONFIG", 3: "RISA_ERR_HIP",
                4: "RISA_ERR_UNSUPPORTED", 5: "RISA_ERR_ALLOC"}

FILTER_TYPES = {"ramp": 0, "sheppLogan": 1, "cosine": 2, "hamming": 3, "hanning": 4}
FILTER_MODE_FFT32 = 0
FILTER_MODE_GEMM_SPLIT = 1
CHAIN_PLAIN = 0
CHAIN_MIXED = 1
F32 = 0
F16 = 1
LAYOUT_DETECTOR_MAJOR = 0
LAYOUT_MODULE_MAJOR = 1


class Geom(C.Structure):
    """Mirror of risa_geom (include/risa_hip.h)."""

    _fields_ = [
        ("numberOfFanDetectors", C.c_int32),
        ("numberOfFanProjections", C.c_int32),
        ("numberOfPlanes", C.c_int32),
        ("numberOfDetectorModules", C.c_int32),
        ("numberOfParallelProjections", C.c_int32),
        ("numberOfParallelDetectors", C.c_int32),
        ("numberOfPixels", C.c_int32),
        ("filterType", C.c_int32),
        ("interpolationType", C.c_int32),
        ("useTextureMemory", C.c_int32),
        ("normalization", C.c_int32),
        ("fusedMasking", C.c_int32),
        ("xa", C.c_int32), ("xb", C.c_int32), ("xc", C.c_int32),
        ("xd", C.c_int32), ("xe", C.c_int32), ("xf", C.c_int32),
        ("reserved0", C.c_int32),
        ("reserved1", C.c_int32),
        ("sourceOffset", C.c_double),
        ("detectorDiameter", C.c_double),
        ("imageCenterX", C.c_double),
        ("imageCenterY", C.c_double),
        ("imageWidth", C.c_double),
        ("rotationOffset", C.c_double),
        ("cutoffFraction", C.c_double),
        ("maskingValue", C.c_double),
        ("lowerLimOffset", C.c_double),
        ("upperLimOffset", C.c_double),
        ("threshMin", C.c_double),
        ("threshMax", C.c_double),
        ("sourceDiameter", C.c_double * MAX_PLANES),
        ("deltaX", C.c_double * MAX_PLANES),
        ("deltaZ", C.c_double * MAX_PLANES),
        ("sourceAngle", C.c_double * MAX_PLANES),
    ]

    def copy(self) -> "Geom":
        g = Geom()
        C.memmove(C.byref(g), C.byref(self), C.sizeof(Geom))
        return g

    def replace(self, **kw) -> "Geom":
        g = self.copy()
        for k, v in kw.items():
            if k in ("sourceDiameter", "deltaX", "deltaZ", "sourceAngle"):
                arr = getattr(g, k)
                for i, x in enumerate(v):
                    arr[i] = x
            else:
                setattr(g, k, v)
        return g

    @property
    def fan_shape(self):
        return (self.numberOfFanProjections, self.numberOfFanDetectors)

    @property
    def par_shape(self):
        return (self.numberOfParallelProjections, self.numberOfParallelDetectors)

    @property
    def img_shape(self):
        return (self.numberOfPixels, self.numberOfPixels)


class BpJob(C.Structure):
    """Mirror of risa_bp_job."""

    _fields_ = [
        ("sino", C.c_void_p), ("img", C.c_void_p),
        ("N", C.c_int32), ("P", C.c_int32), ("D", C.c_int32),
        ("dtype", C.c_int32),
        ("in_scale", C.c_float),
        ("rotationOffset", C.c_float),
        ("mask", C.c_int32),
        ("maskingValue", C.c_float),
    ]


class SynthOpts(C.Structure):
    """Mirror of risa_synth_opts (include/risa_synth.h)."""

    _fields_ = [
        ("I0", C.c_double), ("dark_mean", C.c_double), ("dark_sd", C.c_double),
        ("mu_max", C.c_double), ("n_min", C.c_int32), ("n_max", C.c_int32),
        ("dead", C.POINTER(C.c_int32)), ("n_dead", C.c_int32), ("threads", C.c_int32),
    ]


class RisaError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"{STATUS_NAMES.get(code, code)}: {msg}")
        self.code = code


_lib = None

P = C.c_void_p
I32 = C.c_int32
FP = C.POINTER(C.c_float)

# name -> argtypes (restype int unless noted)
_SIGS = {
    "risa_abi_version": [],
    "risa_geom_from_config": [C.c_char_p, C.POINTER(Geom)],
    "risa_config_lookup_int": [C.c_char_p, C.c_char_p, C.POINTER(C.c_int)],
    "risa_config_lookup_string": [C.c_char_p, C.c_char_p, P, C.c_size_t],
    "risa_attenuation_area_mask": [C.POINTER(Geom), P],
    "risa_attenuation_prepare_reference": [C.POINTER(Geom), P, C.c_int, P, P],
    "risa_find_defect_detectors": [P, C.c_int, C.c_int, C.c_double, C.c_double, P],
    "risa_interpolate_defect_detectors": [P, P, C.c_int, C.c_int],
    "risa_fan2para_tables": [C.POINTER(Geom), C.c_int, P, P],
    "risa_filter_taps": [C.POINTER(Geom), P],
    "risa_bp_lut": [C.POINTER(Geom), P, P, P],
    "risa_hip_attenuation_create": [C.POINTER(Geom), P, P, C.c_int, C.POINTER(P)],
    "risa_hip_attenuation_run": [P, P, C.c_int, P, P],
    "risa_hip_attenuation_run_batch": [P, P, P, C.c_int, P, P],
    "risa_hip_attenuation_run_list": [P, P, P, C.c_int, P, P],
    "risa_hip_attenuation_destroy": [P],
    "risa_hip_attenuation_set_input_layout": [P, C.c_int],
    "risa_hip_attenuation_get_input_layout": [P, C.POINTER(C.c_int)],
    "risa_hip_reorder_run": [P, C.c_int, C.c_int, C.c_int, P, P],
    "risa_tiff_write_f32": [C.c_char_p, P, C.c_int, C.c_int, C.c_int, C.c_int],
    "risa_hip_repair_defects": [P, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, P, P],
    "risa_hip_attenuation_prepare_reference": [C.POINTER(Geom), P, C.c_int, P, P, P],
    "risa_hip_attenuation_create_device": [C.POINTER(Geom), P, P, C.c_int, P, C.POINTER(P)],
    "risa_hip_fan2para_create": [C.POINTER(Geom), C.c_int, C.POINTER(P)],
    "risa_hip_fan2para_create_ex": [C.POINTER(Geom), C.c_int, C.c_int, C.POINTER(P)],
    "risa_hip_fan2para_run": [P, P, C.c_int, P, P],
    "risa_hip_fan2para_run_batch": [P, P, P, C.c_int, P, P],
    "risa_hip_fan2para_run_list": [P, P, P, C.c_int, P, P],
    "risa_hip_fan2para_table_bytes": [P, C.POINTER(C.c_size_t)],
    "risa_hip_fan2para_destroy": [P],
    "risa_hip_filter_create": [C.POINTER(Geom), C.c_int, C.c_int, C.POINTER(P)],
    "risa_hip_filter_set_taps": [P, P],
    "risa_hip_filter_get_taps": [P, P],
    "risa_hip_filter_run": [P, P, C.c_int, P],
    "risa_hip_filter_run_list": [P, P, C.c_int, P],
    "risa_hip_filter_run_f16": [P, P, P, C.c_int, C.c_float, P],
    "risa_hip_filter_run_f16x4": [P, P, P, C.c_int, C.c_float, P],
    "risa_hip_filter_adjoint_run": [P, P, C.c_int, P],
    "risa_hip_filter_grad_taps": [P, P, P, C.c_int, P, P],
    "risa_hip_filter_grad_taps_noalloc": [P, P, P, C.c_int, P, P],
    "risa_hip_filter_destroy": [P],
    "risa_hip_bp_create": [C.POINTER(Geom), C.c_int, C.POINTER(P)],
    "risa_hip_bp_run": [P, P, P, P],
    "risa_hip_bp_run_batch": [P, P, C.c_int, P, P],
    "risa_hip_bp_run_list": [P, P, C.c_int, P, P],
    "risa_hip_bp_run_f16": [P, P, C.c_float, C.c_int, P, P],
    "risa_hip_bp_run_f16x4": [P, P, C.c_float, C.c_int, P, P],
    "risa_hip_bp_adjoint_run": [P, P, P, P],
    "risa_hip_bp_adjoint_run_batch": [P, P, C.c_int, P, P],
    "risa_hip_bp_adjoint_method": [P, C.c_int],
    "risa_hip_bp_set_variant": [P, C.c_int],
    "risa_hip_bp_get_variant": [P, C.POINTER(C.c_int)],
    "risa_hip_bp_autotune": [P, P],
    "risa_hip_bp_autotune_slices": [P, C.c_int, P],
    "risa_hip_bp_autotune_batch": [P, C.c_int, P],
    "risa_hip_bp_slices_per_tile": [P, C.POINTER(C.c_int)],
    "risa_hip_bp_autotune_f16": [P, C.c_int, P],
    "risa_hip_bp_variant_shape": [C.c_int, C.c_int, C.POINTER(C.c_int)],
    "risa_hip_bp_destroy": [P],
    "risa_hip_mask_create": [C.POINTER(Geom), C.c_int, C.POINTER(P)],
    "risa_hip_mask_run": [P, P, P],
    "risa_hip_mask_destroy": [P],
    "risa_hip_bp_batched_create": [C.POINTER(BpJob), C.c_int, C.c_int, C.POINTER(P)],
    "risa_hip_bp_batched_set_buffers": [P, C.POINTER(BpJob), C.c_int],
    "risa_hip_bp_batched_run": [P, P],
    "risa_hip_bp_batched_set_variant": [P, C.c_int],
    "risa_hip_bp_batched_autotune": [P, P],
    "risa_hip_bp_batched_destroy": [P],
    "risa_hip_chain_create": [C.POINTER(Geom), P, P, C.c_int, C.c_int, C.c_int, C.POINTER(P)],
    "risa_hip_chain_run": [P, P, P, C.c_int, P, C.c_int, P],
    "risa_hip_chain_intermediate": [P, C.c_int, C.POINTER(P)],
    "risa_hip_chain_set_variant": [P, C.c_int],
    "risa_hip_chain_get_variant": [P, C.POINTER(C.c_int)],
    "risa_hip_chain_set_input_layout": [P, C.c_int],
    "risa_hip_chain_graph_count": [P, C.POINTER(C.c_int)],
    "risa_hip_chain_destroy": [P],
    # synthetic source (include/risa_synth.h)
    "risa_synth_phantom": [C.c_uint64, C.c_double, C.POINTER(SynthOpts), P, C.c_int],
    "risa_synth_phantom_image": [C.c_uint64, C.c_double, C.c_int, C.POINTER(SynthOpts), P],
    "risa_synth_fan_integrals": [C.POINTER(Geom), C.c_int, C.c_uint64, C.POINTER(SynthOpts), P],
    "risa_synth_fan_frame": [C.POINTER(Geom), C.c_int, C.c_uint64, C.c_uint64, C.POINTER(SynthOpts), P],
    "risa_synth_reference_frame": [C.POINTER(Geom), C.c_uint64, C.POINTER(SynthOpts), P],
    "risa_synth_parallel_sinogram": [C.c_uint64, C.c_double, C.c_int, C.c_int, C.c_double, C.c_uint64,
                                     C.POINTER(SynthOpts), P],
}
_VOID = {"risa_synth_default_opts": [C.POINTER(SynthOpts)]}
_STR = {"risa_last_error": []}
_SIZE = {"risa_geom_size": []}


def exported_symbols():
    """Every function name the binding expects the library to export."""
    return sorted(list(_SIGS) + list(_VOID) + list(_STR) + list(_SIZE))


def load():
    """Load librisa_hip.so (built by __graft_entry__.build() / `make -C risa_amd`)."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise ImportError(f"{LIB_PATH} is not built: run `make -C risa_amd` (or __graft_entry__.build())")
    lib = C.CDLL(str(LIB_PATH), mode=os.RTLD_NOW | getattr(os, "RTLD_GLOBAL", 0))
    for name, args in _SIGS.items():
        f = getattr(lib, name)
        f.argtypes = args
        f.restype = C.c_int
    for name, args in _VOID.items():
        f = getattr(lib, name)
        f.argtypes = args
        f.restype = None
    for name, args in _STR.items():
        f = getattr(lib, name)
        f.argtypes = args
        f.restype = C.c_char_p
    for name, args in _SIZE.items():
        f = getattr(lib, name)
        f.argtypes = args
        f.restype = C.c_size_t
    if lib.risa_geom_size() != C.sizeof(Geom):
        raise ImportError("risa_geom layout mismatch between the library and the binding")
    _lib = lib
    return lib


def check(rc: int):
    if rc != RISA_OK:
        raise RisaError(rc, load().risa_last_error().decode(errors="replace"))


def ptr(a) -> int | None:
    """Address of a numpy array or torch tensor (or None)."""
    if a is None:
        return None
    if isinstance(a, np.ndarray):
        if not a.flags["C_CONTIGUOUS"]:
            raise ValueError("array must be C-contiguous")
        return a.ctypes.data
    if hasattr(a, "data_ptr"):
        if not a.is_contiguous():
            raise ValueError("tensor must be contiguous")
        return a.data_ptr()
    return int(a)


def default_synth_opts(**kw) -> SynthOpts:
    o = SynthOpts()
    load().risa_synth_default_opts(C.byref(o))
    for k, v in kw.items():
        setattr(o, k, v)
    return o
