"""ctypes binding of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg -- never by the product (risa_amd/).
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "liboracle.so"
MAX_PLANES = 8


class OrcGeom(C.Structure):
    """orc_geom (risa_oracle.h); same layout as the product's risa_geom."""

    _fields_ = [
        ("numberOfFanDetectors", C.c_int32), ("numberOfFanProjections", C.c_int32),
        ("numberOfPlanes", C.c_int32), ("numberOfDetectorModules", C.c_int32),
        ("numberOfParallelProjections", C.c_int32), ("numberOfParallelDetectors", C.c_int32),
        ("numberOfPixels", C.c_int32), ("filterType", C.c_int32), ("interpolationType", C.c_int32),
        ("useTextureMemory", C.c_int32), ("normalization", C.c_int32), ("fusedMasking", C.c_int32),
        ("xa", C.c_int32), ("xb", C.c_int32), ("xc", C.c_int32), ("xd", C.c_int32), ("xe", C.c_int32),
        ("xf", C.c_int32), ("reserved0", C.c_int32), ("reserved1", C.c_int32),
        ("sourceOffset", C.c_double), ("detectorDiameter", C.c_double), ("imageCenterX", C.c_double),
        ("imageCenterY", C.c_double), ("imageWidth", C.c_double), ("rotationOffset", C.c_double),
        ("cutoffFraction", C.c_double), ("maskingValue", C.c_double), ("lowerLimOffset", C.c_double),
        ("upperLimOffset", C.c_double), ("threshMin", C.c_double), ("threshMax", C.c_double),
        ("sourceDiameter", C.c_double * MAX_PLANES), ("deltaX", C.c_double * MAX_PLANES),
        ("deltaZ", C.c_double * MAX_PLANES), ("sourceAngle", C.c_double * MAX_PLANES),
    ]


class OrcTables(C.Structure):
    _fields_ = [("P2", C.c_int), ("D", C.c_int), ("Df", C.c_int), ("Pf", C.c_int), ("planes", C.c_int),
                ("theta", C.POINTER(C.c_float)), ("gamma", C.POINTER(C.c_float)),
                ("s", C.POINTER(C.c_float)), ("alpha", C.POINTER(C.c_float))] + [
        (n, C.POINTER(C.c_int32)) for n in ("thetaAfter1", "thetaAfter2", "thetaBefore1", "thetaBefore2",
                                            "gammaAfter1", "gammaAfter2", "gammaBefore1", "gammaBefore2")] + [
        (n, C.POINTER(C.c_float)) for n in ("gammaGoal1", "gammaGoal2", "thetaGoal1", "thetaGoal2")] + [
        ("ray1", C.POINTER(C.c_int32)), ("ray2", C.POINTER(C.c_int32))]


class OrcChain(C.Structure):
    _fields_ = [("g", OrcGeom), ("t", OrcTables), ("mask", C.POINTER(C.c_float)),
                ("avg_ref", C.POINTER(C.c_float)), ("avg_dark", C.POINTER(C.c_float)),
                ("taps", C.POINTER(C.c_float)), ("taps64", C.POINTER(C.c_double))]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not LIB.exists():
            raise ImportError(f"{LIB} not built: run `make -C oracle`")
        _lib = C.CDLL(str(LIB))
        _lib.orc_geom_size.restype = C.c_size_t
        _lib.orc_ellipse_kreis_f32.restype = C.c_float
        _lib.orc_ellipse_kreis_f32.argtypes = [C.c_float] * 4 + [C.c_int]
        _lib.orc_ellipse_kreis_f64.restype = C.c_double
        _lib.orc_ellipse_kreis_f64.argtypes = [C.c_double] * 4
        _lib.orc_filter_rows_f32.restype = C.c_int
        _lib.orc_f2p_tables_build.restype = C.c_int
        _lib.orc_chain_init.restype = C.c_int
    return _lib


def geom(g) -> OrcGeom:
    """Copy any object with the risa_geom field names into an OrcGeom."""
    o = OrcGeom()
    for name, _ in OrcGeom._fields_:
        v = getattr(g, name)
        if name in ("sourceDiameter", "deltaX", "deltaZ", "sourceAngle"):
            arr = getattr(o, name)
            for i in range(MAX_PLANES):
                arr[i] = v[i]
        else:
            setattr(o, name, v)
    return o


def _p(a):
    # data_as returns a pointer object that keeps `a` alive for as long as the
    # pointer is: a converted temporary (_p(_c(x, dtype))) must not be freed
    # before the C function has read it
    return a.ctypes.data_as(C.c_void_p)


def _c(a, dtype):
    return np.ascontiguousarray(a, dtype=dtype)


def area_mask(g):
    og = geom(g)
    out = np.empty((og.numberOfFanProjections, og.numberOfFanDetectors), np.float32)
    lib().orc_area_mask(C.byref(og), _p(out))
    return out


def defect_window():
    w = np.empty(17, np.float64)
    lib().orc_defect_window(_p(w))
    return w


def find_defects(data, tmin, tmax):
    data = _c(data, np.uint16)
    P, D = data.shape
    flags = np.zeros(P * D, np.int32)
    w = defect_window()
    lib().orc_find_defects(_p(data), _p(w), _p(flags), C.c_uint(D), C.c_uint(P), C.c_double(tmin), C.c_double(tmax))
    return flags[:D].copy()


def interp_defects(data, flags):
    data = _c(data, np.uint16).copy()
    P, D = data.shape
    flags = _c(flags, np.int32)
    lib().orc_interp_defects(_p(data), _p(flags), C.c_uint(D), C.c_uint(P))
    return data


def reference_average(g, ref_frames, R):
    og = geom(g)
    ref = _c(ref_frames, np.uint16).copy()
    avg_ref = np.empty((og.numberOfPlanes, og.numberOfFanProjections, og.numberOfFanDetectors), np.float32)
    avg_dark = np.empty((og.numberOfPlanes, og.numberOfFanDetectors), np.float32)
    lib().orc_reference_average(C.byref(og), _p(ref), C.c_int(R), _p(avg_ref), _p(avg_dark))
    return avg_ref, avg_dark, ref


def reorder_u16(raw, P, D):
    """Reordering (Reordering.cu:180-204): module-major frames [frames][P*D]
    -> detector-major [frames][P][D]."""
    raw = _c(raw, np.uint16).reshape(-1, P * D)
    out = np.empty((raw.shape[0], P, D), np.uint16)
    rc = lib().orc_reorder_u16(C.c_int(P), C.c_int(D), C.c_int(raw.shape[0]), _p(raw), _p(out))
    if rc:
        raise ValueError("the reference's reordering table indexes past frames with fewer than 432 detectors")
    return out


def attenuation_f32(g, raw, plane, avg_ref, avg_dark, mask):
    og = geom(g)
    out = np.empty((og.numberOfFanProjections, og.numberOfFanDetectors), np.float32)
    lib().orc_attenuation_f32(C.byref(og), _p(_c(raw, np.uint16)), C.c_int(plane), _p(_c(avg_ref, np.float32)),
                              _p(_c(avg_dark, np.float32)), _p(_c(mask, np.float32)), _p(out))
    return out


def attenuation_f64(g, raw, plane, avg_ref, avg_dark, mask):
    og = geom(g)
    out = np.empty((og.numberOfFanProjections, og.numberOfFanDetectors), np.float64)
    lib().orc_attenuation_f64(C.byref(og), _p(_c(raw, np.uint16)), C.c_int(plane), _p(_c(avg_ref, np.float32)),
                              _p(_c(avg_dark, np.float32)), _p(_c(mask, np.float32)), _p(out))
    return out


class Tables:
    """Rebinning tables (orc_f2p_tables) with numpy views; freed on close()."""

    INT_NAMES = ("thetaBefore1", "thetaAfter1", "gammaBefore1", "gammaAfter1", "ray1",
                 "thetaBefore2", "thetaAfter2", "gammaBefore2", "gammaAfter2", "ray2")
    FLOAT_NAMES = ("thetaGoal1", "thetaGoal2", "gammaGoal1", "gammaGoal2")

    def __init__(self, g, variant=0):
        self.og = geom(g)
        self.t = OrcTables()
        rc = lib().orc_f2p_tables_build(C.byref(self.og), C.c_int(variant), C.byref(self.t))
        if rc != 0:
            raise RuntimeError("orc_f2p_tables_build failed")
        n = self.t.P2 * self.t.D * self.t.planes
        self.n = n

        def view(ptr, count):
            return np.ctypeslib.as_array(ptr, shape=(count,))

        self.arrays = {k: view(getattr(self.t, k), n) for k in self.INT_NAMES + self.FLOAT_NAMES}
        self.theta = view(self.t.theta, self.t.Pf)
        self.gamma = view(self.t.gamma, self.t.Df)
        self.s = view(self.t.s, self.t.D)
        self.alpha = view(self.t.alpha, self.t.P2)

    def plane(self, name, k):
        m = self.t.P2 * self.t.D
        return self.arrays[name][k * m:(k + 1) * m]

    def close(self):
        if self.t.theta:
            lib().orc_f2p_tables_free(C.byref(self.t))

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def gather_f32(g, tables: Tables, plane, fan):
    og = geom(g)
    out = np.empty((og.numberOfParallelProjections, og.numberOfParallelDetectors), np.float32)
    lib().orc_f2p_gather_f32(C.byref(og), C.byref(tables.t), C.c_int(plane), _p(_c(fan, np.float32)), _p(out))
    return out


def gather_f64(g, tables: Tables, plane, fan):
    og = geom(g)
    out = np.empty((og.numberOfParallelProjections, og.numberOfParallelDetectors), np.float64)
    lib().orc_f2p_gather_f64(C.byref(og), C.byref(tables.t), C.c_int(plane), _p(_c(fan, np.float64)), _p(out))
    return out


def ellipse_f32(alpha, dx, dz, diam, variant=0):
    return lib().orc_ellipse_kreis_f32(alpha, dx, dz, diam, variant)


def ellipse_f64(alpha, dx, dz, diam):
    return lib().orc_ellipse_kreis_f64(alpha, dx, dz, diam)


def taps_f32(g):
    og = geom(g)
    h = np.empty(og.numberOfParallelDetectors // 2 + 1, np.float32)
    lib().orc_filter_taps_f32(C.byref(og), _p(h))
    return h


def taps_f64(g):
    og = geom(g)
    h = np.empty(og.numberOfParallelDetectors // 2 + 1, np.float64)
    lib().orc_filter_taps_f64(C.byref(og), _p(h))
    return h


def filter_kernel_f64(h, D):
    c = np.empty(D, np.float64)
    lib().orc_filter_kernel_f64(_p(_c(h, np.float64)), C.c_int(D), _p(c))
    return c


def filter_rows_f64(h, x):
    x = _c(x, np.float64)
    D = x.shape[-1]
    rows = x.size // D
    out = np.empty_like(x)
    lib().orc_filter_rows_f64(_p(_c(h, np.float64)), C.c_int(D), C.c_int(rows), _p(x), _p(out))
    return out


def filter_rows_f32(h, x):
    y = _c(x, np.float32).copy()
    D = y.shape[-1]
    rc = lib().orc_filter_rows_f32(_p(_c(h, np.float32)), C.c_int(D), C.c_int(y.size // D), _p(y))
    if rc != 0:
        raise ValueError("unsupported FFT length")
    return y


def grad_taps_f64(x, g):
    x = _c(x, np.float64)
    g = _c(g, np.float64)
    D = x.shape[-1]
    dh = np.empty(D // 2 + 1, np.float64)
    lib().orc_filter_grad_taps_f64(C.c_int(D), C.c_int(x.size // D), _p(x), _p(g), _p(dh))
    return dh


def bp_lut_f32(g):
    og = geom(g)
    P = og.numberOfParallelProjections
    s = np.empty(P, np.float32)
    c = np.empty(P, np.float32)
    k = np.empty(3, np.float32)
    lib().orc_bp_lut_f32(C.byref(og), _p(s), _p(c), _p(k))
    return s, c, k


def _img(g, dtype):
    return np.empty((g.numberOfPixels, g.numberOfPixels), dtype)


def bp_linear_f32(g, sino):
    og = geom(g)
    out = _img(og, np.float32)
    lib().orc_bp_linear_f32(C.byref(og), _p(_c(sino, np.float32)), _p(out))
    return out


def bp_linear_f64(g, sino):
    og = geom(g)
    out = _img(og, np.float64)
    lib().orc_bp_linear_f64(C.byref(og), _p(_c(sino, np.float64)), _p(out))
    return out


def bp_adjoint_f64(g, img):
    og = geom(g)
    out = np.empty((og.numberOfParallelProjections, og.numberOfParallelDetectors), np.float64)
    lib().orc_bp_linear_adjoint_f64(C.byref(og), _p(_c(img, np.float64)), _p(out))
    return out


def bp_nearest_f32(g, sino):
    og = geom(g)
    out = _img(og, np.float32)
    lib().orc_bp_nearest_f32(C.byref(og), _p(_c(sino, np.float32)), _p(out))
    return out


def bp_tex_f32(g, sino):
    og = geom(g)
    out = _img(og, np.float32)
    lib().orc_bp_tex_f32(C.byref(og), _p(_c(sino, np.float32)), _p(out))
    return out


def masking_f32(g, img):
    og = geom(g)
    out = _c(img, np.float32).copy()
    lib().orc_masking_f32(C.byref(og), _p(out))
    return out


def masking_f64(g, img):
    og = geom(g)
    out = _c(img, np.float64).copy()
    lib().orc_masking_f64(C.byref(og), _p(out))
    return out


class Chain:
    """Init-time state of the oracle chain for one geometry."""

    def __init__(self, g, avg_ref, avg_dark, variant=0):
        self.og = geom(g)
        self.c = OrcChain()
        self._ref = _c(avg_ref, np.float32)
        self._dark = _c(avg_dark, np.float32)
        rc = lib().orc_chain_init(C.byref(self.c), C.byref(self.og), _p(self._ref), _p(self._dark), C.c_int(variant))
        if rc != 0:
            raise RuntimeError("orc_chain_init failed")

    def f32(self, raw, planes, threads=0):
        raw = _c(raw, np.uint16)
        planes = _c(planes, np.int32)
        frames = len(planes)
        N = self.og.numberOfPixels
        out = np.empty((frames, N, N), np.float32)
        lib().orc_chain_f32(C.byref(self.c), _p(raw), _p(planes), C.c_int(frames), _p(out), C.c_int(threads))
        return out

    def f64(self, raw, plane, intermediates=False):
        raw = _c(raw, np.uint16)
        og = self.og
        N = og.numberOfPixels
        img = np.empty((N, N), np.float64)
        att = np.empty((og.numberOfFanProjections, og.numberOfFanDetectors), np.float64)
        par = np.empty((og.numberOfParallelProjections, og.numberOfParallelDetectors), np.float64)
        filt = np.empty_like(par)
        lib().orc_chain_f64(C.byref(self.c), _p(raw), C.c_int(plane), _p(img), _p(att), _p(par), _p(filt))
        if intermediates:
            return img, att, par, filt
        return img

    def close(self):
        if self.c.mask:
            lib().orc_chain_free(C.byref(self.c))

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def rel_rmse(a, ref, mask=None):
    a = np.asarray(a, np.float64)
    ref = np.asarray(ref, np.float64)
    if mask is not None:
        a = a[mask]
        ref = ref[mask]
    den = np.sqrt(np.mean(ref * ref))
    return float(np.sqrt(np.mean((a - ref) ** 2)) / (den if den > 0 else 1.0))
