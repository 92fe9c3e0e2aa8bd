"""CPU oracle for the dexel morphology path -- TEST INFRASTRUCTURE ONLY.

Python access to ``libvo_oracle.so`` (built from ``vo_oracle.cpp`` by
``oracle/build.sh``), an FP64 CPU restatement of voroffset's ``ours`` and
``brute_force`` methods.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import this package, and only as the
checker; the product path never calls it.  See vo_oracle.cpp's header for
which reference files each function restates and how each is pinned.
"""

from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

METHODS = {"ours": 0, "brute_force": 1}
OPS = {"noop": 0, "dilation": 1, "erosion": 2, "closing": 3, "opening": 4}

_i64p = ctypes.POINTER(ctypes.c_int64)
_f64p = ctypes.POINTER(ctypes.c_double)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "libvo_oracle.so")
        if not os.path.exists(path):
            raise RuntimeError(f"oracle library missing: {path} (run oracle/build.sh)")
        L = ctypes.CDLL(path)
        L.vo_oracle_morph.restype = ctypes.c_void_p
        L.vo_oracle_morph.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, _i64p,
                                      _f64p, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                      ctypes.c_int, _f64p, _f64p]
        L.vo_oracle_pass1.restype = ctypes.c_void_p
        L.vo_oracle_pass1.argtypes = [ctypes.c_int, ctypes.c_int, _i64p, _f64p, ctypes.c_double,
                                      ctypes.c_int]
        L.vo_oracle_count.restype = ctypes.c_int64
        L.vo_oracle_count.argtypes = [ctypes.c_void_p]
        L.vo_oracle_copy.argtypes = [ctypes.c_void_p, _i64p, _f64p, _f64p, _f64p]
        L.vo_oracle_free.argtypes = [ctypes.c_void_p]
        L.vo_oracle_last_error.restype = ctypes.c_char_p
        for name in ("vo_oracle_negate_ray", "vo_oracle_negate_ray_range"):
            f = getattr(L, name)
            f.restype = ctypes.c_int
            f.argtypes = [_f64p, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double]
        L.vo_oracle_rasterize.restype = ctypes.c_void_p
        L.vo_oracle_rasterize.argtypes = [_f64p, ctypes.c_int64, ctypes.POINTER(ctypes.c_int32), ctypes.c_int64,
                                          ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                          ctypes.c_double]
        L.vo_oracle_xor.restype = ctypes.c_void_p
        L.vo_oracle_xor.argtypes = [ctypes.c_int, ctypes.c_int, _i64p, _f64p, _i64p, _f64p, ctypes.c_double,
                                    ctypes.c_double]
        L.vo_oracle_union.restype = ctypes.c_int
        L.vo_oracle_union.argtypes = [_f64p, ctypes.c_int, _f64p, ctypes.c_int, _f64p, ctypes.c_int]
        _LIB = L
    return _LIB


def _p64(a):
    return a.ctypes.data_as(_i64p)


def _pf(a):
    return a.ctypes.data_as(_f64p)


def _csr(vol):
    off = np.ascontiguousarray(vol.off, np.int64)
    z = np.ascontiguousarray(vol.z, np.float64)
    return off, z


def morph(vol, op="dilation", radius=8.0, method="ours", threads=1, timings=None):
    """Apply a morphology op; returns (off, z) of the result (float64)."""
    L = lib()
    off, z = _csr(vol)
    zlo, zhi = vol.z_range()
    t1, t2 = ctypes.c_double(0), ctypes.c_double(0)
    h = L.vo_oracle_morph(METHODS[method], OPS[op], vol.nx, vol.ny, _p64(off), _pf(z),
                          float(radius), zlo, zhi, int(threads), ctypes.byref(t1), ctypes.byref(t2))
    if not h:
        raise RuntimeError(L.vo_oracle_last_error().decode())
    try:
        n = L.vo_oracle_count(h)
        o = np.empty(vol.ncols + 1, np.int64)
        out = np.empty(n, np.float64)
        L.vo_oracle_copy(h, _p64(o), _pf(out), None, None)
    finally:
        L.vo_oracle_free(h)
    if timings is not None:
        timings["time_1"], timings["time_2"] = t1.value, t2.value
    return o, out


def pass1(vol, radius, threads=1):
    """Pass-1 intermediate of `ours`: (off, y1, y2, rho) per column."""
    L = lib()
    off, z = _csr(vol)
    h = L.vo_oracle_pass1(vol.nx, vol.ny, _p64(off), _pf(z), float(radius), int(threads))
    if not h:
        raise RuntimeError(L.vo_oracle_last_error().decode())
    try:
        n = L.vo_oracle_count(h)
        o = np.empty(vol.ncols + 1, np.int64)
        y1, y2, r = (np.empty(n, np.float64) for _ in range(3))
        L.vo_oracle_copy(h, _p64(o), _pf(y1), _pf(y2), _pf(r))
    finally:
        L.vo_oracle_free(h)
    return o, y1, y2, r


def xor(a, b):
    """calculateXor per column (Voronoi.cpp:91-111) within a's z-range: (off, z)."""
    L = lib()
    oa, za = _csr(a)
    ob, zb = _csr(b)
    lo, hi = a.z_range()
    h = L.vo_oracle_xor(a.nx, a.ny, _p64(oa), _pf(za), _p64(ob), _pf(zb), lo, hi)
    if not h:
        raise RuntimeError(L.vo_oracle_last_error().decode())
    try:
        n = L.vo_oracle_count(h)
        o = np.empty(a.ncols + 1, np.int64)
        out = np.empty(n, np.float64)
        L.vo_oracle_copy(h, _p64(o), _pf(out), None, None)
    finally:
        L.vo_oracle_free(h)
    return o, out


def measure(off, z, spacing):
    """get_volume / numSegments (CompressedVolume.cpp:61-83) in numpy."""
    lens = np.diff(off)
    total, segs = 0.0, 0
    for c in np.nonzero(lens >= 2)[0]:
        col = z[off[c]:off[c] + (lens[c] // 2) * 2]
        total += float(np.sum(col[1::2] - col[0::2]))
        segs += int(lens[c] // 2)
    return total * spacing ** 3, segs


def rasterize(vertices, triangles, nx, ny, origin_x, origin_y, spacing):
    """Mesh -> dexel columns (Dexelize.cpp:56-225): (off, z) with z / spacing
    of the hits per column, sorted."""
    L = lib()
    V = np.ascontiguousarray(vertices, np.float64).reshape(-1, 3)
    F = np.ascontiguousarray(triangles, np.int32).reshape(-1, 3)
    h = L.vo_oracle_rasterize(_pf(V), V.shape[0], F.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), F.shape[0],
                              int(nx), int(ny), float(origin_x), float(origin_y), float(spacing))
    if not h:
        raise RuntimeError(L.vo_oracle_last_error().decode())
    try:
        n = L.vo_oracle_count(h)
        o = np.empty(int(nx) * int(ny) + 1, np.int64)
        out = np.empty(n, np.float64)
        L.vo_oracle_copy(h, _p64(o), _pf(out), None, None)
    finally:
        L.vo_oracle_free(h)
    return o, out


def negate_ray(col, lo, hi):
    c = np.zeros(len(col) + 2)
    c[:len(col)] = col
    n = lib().vo_oracle_negate_ray(_pf(c), len(col), len(c), lo, hi)
    return c[:n].copy()


def negate_ray_range(col, lo, hi):
    c = np.zeros(len(col) + 2)
    c[:len(col)] = col
    n = lib().vo_oracle_negate_ray_range(_pf(c), len(col), len(c), lo, hi)
    return c[:n].copy()


def union(a, b):
    a = np.ascontiguousarray(a, np.float64)
    b = np.ascontiguousarray(b, np.float64)
    out = np.zeros(len(a) + len(b))
    n = lib().vo_oracle_union(_pf(a), len(a), _pf(b), len(b), _pf(out), len(out))
    return out[:n].copy()
