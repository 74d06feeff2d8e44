"""risa_amd -- MI355X-native reconstruction hot path of RISA.

The product is librisa_hip.so (HIP kernels for gfx950 behind the C-ABI in
include/risa_hip.h) and the C++ glados-compatible stages in
risa_amd/csrc/stages/.  This Python package is the harness around it: a
ctypes binding (_lib), geometry/config helpers, thin device wrappers over the
C-ABI (device), and the seeded synthetic source (synth).
"""
from __future__ import annotations

from pathlib import Path

from ._lib import (CHAIN_MIXED, CHAIN_PLAIN, F16, F32, FILTER_MODE_FFT32, FILTER_MODE_GEMM_SPLIT,
                   FILTER_TYPES, LAYOUT_DETECTOR_MAJOR, LAYOUT_MODULE_MAJOR, BpJob, Geom, RisaError, check, load,
                   ptr)

CONFIG_DIR = Path(__file__).resolve().parent / "configs"

__all__ = ["Geom", "BpJob", "RisaError", "load", "check", "ptr", "geom_from_config", "config_path",
           "FILTER_TYPES", "FILTER_MODE_FFT32", "FILTER_MODE_GEMM_SPLIT", "CHAIN_PLAIN", "CHAIN_MIXED",
           "F32", "F16", "LAYOUT_DETECTOR_MAJOR", "LAYOUT_MODULE_MAJOR"]


def config_path(name: str) -> Path:
    """Path of a bundled workload configuration (c0, c1)."""
    p = CONFIG_DIR / (name if name.endswith(".cfg") else name + ".cfg")
    if not p.exists():
        raise FileNotFoundError(p)
    return p


def geom_from_config(path) -> Geom:
    """Read a libconfig-syntax configuration file through the library's parser."""
    import ctypes as C

    g = Geom()
    check(load().risa_geom_from_config(str(path).encode(), C.byref(g)))
    return g
