"""Float32 model of the linear backprojector's detector coordinates, for tests.

For pixel (x, y) and angle a, bp_linear_kernel (tile TX x TY) computes
    xp = ((float)x - center) * scale,  yp likewise,
    T  = floor(min over the tile's corners of fma(yp, sin, xp * cos)),
    t' = fma(yp, sin, fma(xp, cos, -T)),
and weighs sample k = floor(t') + T + cI by 1 - f and k + 1 by f, with
f = t' - floor(t').  This module reproduces those float32 operations
exactly on the host (fused multiply-adds rounded once), so tests can read the
backprojector's and the adjoint projector's weights off one-hot inputs and
compare them with the model bit for bit.  It models this library's kernels,
not the reference.
"""
from __future__ import annotations

import numpy as np

f32, f64 = np.float32, np.float64


def fma32(a, b, c):
    """float32 fma(a, b, c) rounded once: the product is exact in float64,
    the sum's float64 rounding error comes from TwoSum, and the float32
    rounding is corrected where the float64 sum sits on a float32 midpoint."""
    a, b, c = (np.asarray(v, f32).astype(f64) for v in (a, b, c))
    p = a * b
    s = p + c
    bb = s - p
    e = (p - (s - bb)) + (c - bb)
    r = np.asarray(s.astype(f32))
    r64 = r.astype(f64)
    d = s - r64
    up = np.nextafter(r, f32(np.inf)).astype(f64) - r64
    dn = r64 - np.nextafter(r, f32(-np.inf)).astype(f64)
    mid = np.where(d > 0, 2 * d == up, np.where(d < 0, -2 * d == dn, False)) & (e != 0)
    fix = mid & (np.sign(e) == np.sign(d))
    if np.any(fix):
        r = r.copy()
        r[fix] = np.nextafter(r[fix], np.where(d[fix] > 0, f32(np.inf), f32(-np.inf)).astype(f32))
    return r


def coord(i, center, scale):
    return ((np.asarray(i).astype(f32) - f32(center)) * f32(scale)).astype(f32)


def bp_t(xp, yp, c, s):
    return fma32(yp, s, (np.asarray(xp, f32) * f32(c)).astype(f32))


def tile_origin(c, s, tx, ty, TX, TY, center, scale):
    x0, y0 = np.asarray(tx) * TX, np.asarray(ty) * TY
    xa, xb = coord(x0, center, scale), coord(x0 + TX - 1, center, scale)
    ya, yb = coord(y0, center, scale), coord(y0 + TY - 1, center, scale)
    tmin = np.minimum(np.minimum(bp_t(xa, ya, c, s), bp_t(xb, ya, c, s)),
                      np.minimum(bp_t(xa, yb, c, s), bp_t(xb, yb, c, s)))
    return np.floor(tmin).astype(f32)


def weights(x, y, c, s, TX, TY, center, scale, cI):
    """(k, f): sample index floor(t') + T + cI and fraction of pixels (x, y)."""
    x, y = np.asarray(x), np.asarray(y)
    T = tile_origin(c, s, x // TX, y // TY, TX, TY, center, scale)
    xc = fma32(coord(x, center, scale), c, -T)
    tp = fma32(coord(y, center, scale), s, xc)
    k = np.floor(tp.astype(f64))
    return (k + T.astype(f64)).astype(np.int64) + cI, tp.astype(f64) - k


def outside(x, y, N, center):
    """The fused mask's test (float distance against N*N/4 in double)."""
    dX = np.asarray(x).astype(f32) - f32(center)
    dY = np.asarray(y).astype(f32) - f32(center)
    dist = (dX * dX).astype(f32) + (dY * dY).astype(f32)
    return dist.astype(f64) > N * N * 0.25
