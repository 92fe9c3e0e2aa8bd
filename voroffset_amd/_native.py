"""ctypes binding of libvoroffset_hip.so (the C ABI in include/voroffset_hip.h).

The library is loaded from this package's directory; a missing or unloadable
library raises immediately -- there is no CPU fallback.
"""

from __future__ import annotations

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libvoroffset_hip.so")

OK, ERR_ARGUMENT, ERR_DEVICE, ERR_CAPACITY, ERR_RANGE = 0, 1, 2, 3, 4
OP_CODES = {"noop": 0, "dilation": 1, "erosion": 2, "closing": 3, "opening": 4}
PRECISIONS = {"fp32": 0, "fp64": 1}


def storage_shift(z_lo, z_hi):
    """The integer z-shift FP32 endpoints are stored relative to: the centre
    of [z_lo, z_hi] rounded half away from zero, as the host entry point
    rounds it (std::round), so that every entry point quantizes a volume the
    same way."""
    c = 0.5 * (float(z_lo) + float(z_hi))
    return float(np.copysign(np.floor(abs(c) + 0.5), c))

# every symbol the header declares
EXPORTS = ("vo_create", "vo_destroy", "vo_last_error", "vo_morph", "vo_result_shape", "vo_result_copy",
           "vo_result_free", "vo_morph_device", "vo_last_stats", "vo_rasterize", "vo_morph_batch_device",
           "vo_last_batch_alone", "vo_measure_device", "vo_xor_device",
           "vo_dexel_mesh", "vo_mesh_shape", "vo_mesh_copy", "vo_mesh_free", "vo_autotune", "vo_tuning_save",
           "vo_tuning_load", "vo_tuning_get", "vo_pass1", "vo_triples_shape", "vo_triples_copy", "vo_triples_free", "vo_copy_device", "vo_morph_device_noalloc")


class VoError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"[vo status {code}] {msg}")
        self.code = code


class VoVolume(ctypes.Structure):
    _fields_ = [("nx", ctypes.c_int32), ("ny", ctypes.c_int32), ("offsets", ctypes.c_void_p),
                ("z", ctypes.c_void_p), ("z_lo", ctypes.c_double), ("z_hi", ctypes.c_double)]


class VoJob(ctypes.Structure):
    _fields_ = [("op", ctypes.c_int), ("radius", ctypes.c_double), ("vol", VoVolume)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"native library missing: {LIB_PATH} (build with python -m voroffset_amd.build)")
        L = ctypes.CDLL(LIB_PATH)
        vp, i32, i64, dbl = ctypes.c_void_p, ctypes.c_int, ctypes.c_int64, ctypes.c_double
        L.vo_create.argtypes = [ctypes.POINTER(vp), i32]
        L.vo_create.restype = i32
        L.vo_destroy.argtypes = [vp]
        L.vo_destroy.restype = None
        L.vo_last_error.argtypes = [vp]
        L.vo_last_error.restype = ctypes.c_char_p
        L.vo_morph.argtypes = [vp, i32, i32, dbl, ctypes.POINTER(VoVolume), ctypes.POINTER(vp),
                               ctypes.POINTER(dbl), ctypes.POINTER(dbl)]
        L.vo_morph.restype = i32
        L.vo_result_shape.argtypes = [vp, ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_int32),
                                      ctypes.POINTER(i64)]
        L.vo_result_shape.restype = i32
        L.vo_result_copy.argtypes = [vp, vp, vp]
        L.vo_result_copy.restype = i32
        L.vo_result_free.argtypes = [vp]
        L.vo_result_free.restype = None
        L.vo_morph_device.argtypes = [vp, i32, i32, dbl, ctypes.POINTER(VoVolume), ctypes.POINTER(VoVolume),
                                      ctypes.POINTER(i64), ctypes.POINTER(dbl), ctypes.POINTER(dbl)]
        L.vo_morph_device.restype = i32
        L.vo_last_stats.argtypes = [vp, ctypes.POINTER(i64), ctypes.POINTER(i64)]
        L.vo_last_stats.restype = i32
        L.vo_rasterize.argtypes = [vp, vp, i64, vp, i64, i32, i32, dbl, dbl, dbl, ctypes.POINTER(vp)]
        L.vo_rasterize.restype = i32
        L.vo_morph_batch_device.argtypes = [vp, i32, i32, ctypes.POINTER(VoJob), ctypes.POINTER(VoVolume),
                                            ctypes.POINTER(i64), ctypes.POINTER(dbl), ctypes.POINTER(dbl)]
        L.vo_morph_batch_device.restype = i32
        L.vo_measure_device.argtypes = [vp, i32, ctypes.POINTER(VoVolume), dbl, ctypes.POINTER(dbl),
                                        ctypes.POINTER(i64)]
        L.vo_measure_device.restype = i32
        L.vo_xor_device.argtypes = [vp, i32, ctypes.POINTER(VoVolume), ctypes.POINTER(VoVolume), dbl,
                                    ctypes.POINTER(VoVolume), ctypes.POINTER(i64), ctypes.POINTER(dbl)]
        L.vo_xor_device.restype = i32
        L.vo_dexel_mesh.argtypes = [vp, ctypes.POINTER(VoVolume), dbl, dbl, dbl, ctypes.POINTER(vp)]
        L.vo_dexel_mesh.restype = i32
        L.vo_mesh_shape.argtypes = [vp, ctypes.POINTER(i64), ctypes.POINTER(i64)]
        L.vo_mesh_shape.restype = i32
        L.vo_mesh_copy.argtypes = [vp, vp, vp]
        L.vo_mesh_copy.restype = i32
        L.vo_mesh_free.argtypes = [vp]
        L.vo_mesh_free.restype = None
        L.vo_autotune.argtypes = [vp, i32, ctypes.POINTER(VoVolume), i32, ctypes.POINTER(dbl), i32]
        L.vo_autotune.restype = i32
        L.vo_tuning_save.argtypes = [vp, ctypes.c_char_p]
        L.vo_tuning_save.restype = i32
        L.vo_tuning_load.argtypes = [vp, ctypes.c_char_p]
        L.vo_tuning_load.restype = i32
        L.vo_tuning_get.argtypes = [vp, i32, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)]
        L.vo_tuning_get.restype = i32
        L.vo_pass1.argtypes = [vp, i32, dbl, ctypes.POINTER(VoVolume), ctypes.POINTER(vp)]
        L.vo_pass1.restype = i32
        L.vo_triples_shape.argtypes = [vp, ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_int32),
                                       ctypes.POINTER(i64)]
        L.vo_triples_shape.restype = i32
        L.vo_triples_copy.argtypes = [vp, vp, vp, vp, vp]
        L.vo_triples_copy.restype = i32
        L.vo_triples_free.argtypes = [vp]
        L.vo_triples_free.restype = None
        L.vo_morph_device_noalloc.argtypes = [vp, i32, i32, dbl, ctypes.POINTER(VoVolume), vp, vp, i64,
                                              ctypes.POINTER(i64), ctypes.POINTER(dbl), ctypes.POINTER(dbl)]
        L.vo_morph_device_noalloc.restype = i32
        L.vo_copy_device.argtypes = [vp, i32, ctypes.POINTER(vp), ctypes.POINTER(vp), ctypes.POINTER(i64)]
        L.vo_copy_device.restype = i32
        L.vo_last_batch_alone.argtypes = [vp, ctypes.POINTER(ctypes.c_int)]
        L.vo_last_batch_alone.restype = i32
        _lib = L
    return _lib


class Context:
    """A device context (stream + workspace pool) of the native library."""

    def __init__(self, device=0):
        self._h = ctypes.c_void_p()
        rc = lib().vo_create(ctypes.byref(self._h), int(device))
        if rc != OK:
            raise VoError(rc, "vo_create failed: " + lib().vo_last_error(None).decode())

    def close(self):
        if self._h:
            lib().vo_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _raise(self, rc):
        raise VoError(rc, lib().vo_last_error(self._h).decode())

    def morph_host(self, op, precision, radius, nx, ny, off, z, z_lo, z_hi):
        """Host CSR (float64) in -> host CSR (float64) out, plus (time_1, time_2) in ms."""
        off = np.ascontiguousarray(off, np.int64)
        z = np.ascontiguousarray(z, np.float64)
        vol = VoVolume(nx, ny, off.ctypes.data, z.ctypes.data if z.size else None, z_lo, z_hi)
        res = ctypes.c_void_p()
        t1, t2 = ctypes.c_double(0), ctypes.c_double(0)
        rc = lib().vo_morph(self._h, OP_CODES[op], PRECISIONS[precision], float(radius), ctypes.byref(vol),
                            ctypes.byref(res), ctypes.byref(t1), ctypes.byref(t2))
        if rc != OK:
            self._raise(rc)
        try:
            rnx, rny, cnt = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int64()
            lib().vo_result_shape(res, ctypes.byref(rnx), ctypes.byref(rny), ctypes.byref(cnt))
            o = np.empty(rnx.value * rny.value + 1, np.int64)
            zz = np.empty(cnt.value, np.float64)
            lib().vo_result_copy(res, o.ctypes.data, zz.ctypes.data if zz.size else None)
        finally:
            lib().vo_result_free(res)
        return o, zz, (t1.value, t2.value)

    def pass1_host(self, precision, radius, nx, ny, off, z, z_lo, z_hi):
        """The pass-1 intermediate of a dilation (vo_pass1): (off, y1, y2, r),
        CSR over triples per column x + nx*y."""
        off = np.ascontiguousarray(off, np.int64)
        z = np.ascontiguousarray(z, np.float64)
        vol = VoVolume(nx, ny, off.ctypes.data, z.ctypes.data if z.size else None, z_lo, z_hi)
        res = ctypes.c_void_p()
        rc = lib().vo_pass1(self._h, PRECISIONS[precision], float(radius), ctypes.byref(vol), ctypes.byref(res))
        if rc != OK:
            self._raise(rc)
        try:
            rnx, rny, cnt = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int64()
            lib().vo_triples_shape(res, ctypes.byref(rnx), ctypes.byref(rny), ctypes.byref(cnt))
            o = np.empty(rnx.value * rny.value + 1, np.int64)
            y1, y2, r = (np.empty(cnt.value, np.float64) for _ in range(3))
            p = (lambda a: a.ctypes.data if a.size else None)
            lib().vo_triples_copy(res, o.ctypes.data, p(y1), p(y2), p(r))
        finally:
            lib().vo_triples_free(res)
        return o, y1, y2, r

    def morph_device(self, op, precision, radius, nx, ny, off_ptr, z_ptr, z_lo, z_hi):
        """Device pointers in -> (out VoVolume, count, (time_1, time_2)); the
        output lives in the context's workspace until the next call."""
        vin = VoVolume(nx, ny, off_ptr, z_ptr, z_lo, z_hi)
        vout = VoVolume()
        cnt = ctypes.c_int64()
        t1, t2 = ctypes.c_double(0), ctypes.c_double(0)
        rc = lib().vo_morph_device(self._h, OP_CODES[op], PRECISIONS[precision], float(radius), ctypes.byref(vin),
                                   ctypes.byref(vout), ctypes.byref(cnt), ctypes.byref(t1), ctypes.byref(t2))
        if rc != OK:
            self._raise(rc)
        return vout, cnt.value, (t1.value, t2.value)

    def morph_batch_device(self, precision, jobs):
        """Batched form.  jobs: sequence of (op, radius, nx, ny, off_ptr, z_ptr,
        z_lo, z_hi) with device pointers.  Returns ([(VoVolume, count)],
        (time_1, time_2)); outputs live in the workspace until the next call."""
        n = len(jobs)
        arr = (VoJob * max(n, 1))()
        for i, (op, r, nx, ny, off, z, zlo, zhi) in enumerate(jobs):
            arr[i] = VoJob(OP_CODES[op], float(r), VoVolume(nx, ny, off, z, zlo, zhi))
        outs = (VoVolume * max(n, 1))()
        cnt = (ctypes.c_int64 * max(n, 1))()
        t1, t2 = ctypes.c_double(0), ctypes.c_double(0)
        rc = lib().vo_morph_batch_device(self._h, PRECISIONS[precision], n, arr, outs, cnt, ctypes.byref(t1),
                                         ctypes.byref(t2))
        if rc != OK:
            self._raise(rc)
        return [(outs[i], cnt[i]) for i in range(n)], (t1.value, t2.value)

    def morph_device_noalloc(self, op, precision, radius, nx, ny, off_ptr, z_ptr, z_lo, z_hi, out_off_ptr,
                             out_z_ptr, capacity):
        """vo_morph_device_noalloc: the result goes to the caller's device
        buffers; returns (endpoint count, (time_1, time_2)).  On a too small
        out_z the VoError carries code ERR_CAPACITY and `.needed`."""
        vin = VoVolume(nx, ny, off_ptr, z_ptr, z_lo, z_hi)
        cnt = ctypes.c_int64(0)
        t1, t2 = ctypes.c_double(0), ctypes.c_double(0)
        rc = lib().vo_morph_device_noalloc(self._h, OP_CODES[op], PRECISIONS[precision], float(radius),
                                           ctypes.byref(vin), out_off_ptr, out_z_ptr, int(capacity),
                                           ctypes.byref(cnt), ctypes.byref(t1), ctypes.byref(t2))
        if rc != OK:
            err = VoError(rc, lib().vo_last_error(self._h).decode())
            err.needed = cnt.value
            raise err
        return cnt.value, (t1.value, t2.value)

    def copy_device(self, copies):
        """Device-to-device copies [(dst, src, nbytes)] on this context's
        stream, then one synchronisation (vo_copy_device)."""
        n = len(copies)
        if n == 0:
            return
        dst = (ctypes.c_void_p * n)(*[c[0] for c in copies])
        src = (ctypes.c_void_p * n)(*[c[1] for c in copies])
        nb = (ctypes.c_int64 * n)(*[int(c[2]) for c in copies])
        rc = lib().vo_copy_device(self._h, n, dst, src, nb)
        if rc != OK:
            self._raise(rc)

    def rasterize(self, vertices, triangles, nx, ny, origin_x, origin_y, spacing):
        """Mesh -> dexel columns (CSR offsets int64, z float64 in units of spacing)."""
        V = np.ascontiguousarray(vertices, np.float64).reshape(-1, 3)
        F = np.ascontiguousarray(triangles, np.int32).reshape(-1, 3)
        res = ctypes.c_void_p()
        rc = lib().vo_rasterize(self._h, V.ctypes.data if V.size else None, V.shape[0],
                                F.ctypes.data if F.size else None, F.shape[0], int(nx), int(ny),
                                float(origin_x), float(origin_y), float(spacing), ctypes.byref(res))
        if rc != OK:
            self._raise(rc)
        return self._take(res)

    def _take(self, res):
        try:
            rnx, rny, cnt = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int64()
            lib().vo_result_shape(res, ctypes.byref(rnx), ctypes.byref(rny), ctypes.byref(cnt))
            o = np.empty(rnx.value * rny.value + 1, np.int64)
            zz = np.empty(cnt.value, np.float64)
            lib().vo_result_copy(res, o.ctypes.data, zz.ctypes.data if zz.size else None)
        finally:
            lib().vo_result_free(res)
        return o, zz

    def measure_device(self, precision, nx, ny, off_ptr, z_ptr, spacing):
        """(spacing^3 * total segment length, number of segments) of a device volume."""
        v = VoVolume(nx, ny, off_ptr, z_ptr, 0.0, 0.0)
        vol, segs = ctypes.c_double(0), ctypes.c_int64(0)
        rc = lib().vo_measure_device(self._h, PRECISIONS[precision], ctypes.byref(v), float(spacing),
                                     ctypes.byref(vol), ctypes.byref(segs))
        if rc != OK:
            self._raise(rc)
        return vol.value, segs.value

    def xor_device(self, precision, a, b, spacing):
        """calculateXor of two device volumes given as VoVolume; returns
        (VoVolume in the workspace, endpoint count, measure)."""
        out, cnt, vol = VoVolume(), ctypes.c_int64(0), ctypes.c_double(0)
        rc = lib().vo_xor_device(self._h, PRECISIONS[precision], ctypes.byref(a), ctypes.byref(b), float(spacing),
                                 ctypes.byref(out), ctypes.byref(cnt), ctypes.byref(vol))
        if rc != OK:
            self._raise(rc)
        return out, cnt.value, vol.value

    def dexel_mesh(self, nx, ny, off, z, origin_x, origin_y, spacing):
        """dexel_dump's hexahedra: (vertices (n, 3) float64, quads (m, 4) int64)."""
        off = np.ascontiguousarray(off, np.int64)
        z = np.ascontiguousarray(z, np.float64)
        v = VoVolume(nx, ny, off.ctypes.data, z.ctypes.data if z.size else None, 0.0, 0.0)
        res = ctypes.c_void_p()
        rc = lib().vo_dexel_mesh(self._h, ctypes.byref(v), float(origin_x), float(origin_y), float(spacing),
                                 ctypes.byref(res))
        if rc != OK:
            self._raise(rc)
        try:
            nv, nq = ctypes.c_int64(), ctypes.c_int64()
            lib().vo_mesh_shape(res, ctypes.byref(nv), ctypes.byref(nq))
            vert = np.empty((nv.value, 3), np.float64)
            quad = np.empty((nq.value, 4), np.int64)
            lib().vo_mesh_copy(res, vert.ctypes.data if vert.size else None, quad.ctypes.data if quad.size else None)
        finally:
            lib().vo_mesh_free(res)
        return vert, quad

    TUNE_BUCKETS = 5

    def autotune(self, precision, nx, ny, off_ptr, z_ptr, radii, reps=2):
        """Measures and keeps the fastest shapes per radius bucket (device sample)."""
        v = VoVolume(nx, ny, off_ptr, z_ptr, 0.0, 0.0)
        r = (ctypes.c_double * max(len(radii), 1))(*map(float, radii))
        rc = lib().vo_autotune(self._h, PRECISIONS[precision], ctypes.byref(v), len(radii), r, int(reps))
        if rc != OK:
            self._raise(rc)
        return self.tuning()

    def tuning(self):
        """[(events, list capacity)] per radius bucket."""
        out = []
        for b in range(self.TUNE_BUCKETS):
            e, o = ctypes.c_int(), ctypes.c_int()
            lib().vo_tuning_get(self._h, b, ctypes.byref(e), ctypes.byref(o))
            out.append((e.value, o.value))
        return out

    def tuning_save(self, path):
        if lib().vo_tuning_save(self._h, str(path).encode()) != OK:
            raise VoError(ERR_ARGUMENT, f"cannot write {path}")

    def tuning_load(self, path):
        rc = lib().vo_tuning_load(self._h, str(path).encode())
        if rc != OK:
            self._raise(rc)

    def last_batch_alone(self):
        """Jobs of the last batch that were redone by the single-volume path."""
        n = ctypes.c_int(0)
        lib().vo_last_batch_alone(self._h, ctypes.byref(n))
        return n.value

    def last_stats(self):
        """(output endpoints, pass-2 tiles recomputed by the sweep form) of the last dilation."""
        a, b = ctypes.c_int64(), ctypes.c_int64()
        lib().vo_last_stats(self._h, ctypes.byref(a), ctypes.byref(b))
        return a.value, b.value


# -- device buffers -------------------------------------------------------
# Plain hipMalloc'd buffers for callers that keep volumes resident in HBM.
# They go through the HIP runtime this library is linked against (resolved by
# soname to the copy already loaded with the library), so one process never
# mixes two HIP runtimes.

_hip = None


def hip():
    global _hip
    if _hip is None:
        lib()
        H = ctypes.CDLL("libamdhip64.so.7")
        H.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        H.hipMalloc.restype = ctypes.c_int
        H.hipFree.argtypes = [ctypes.c_void_p]
        H.hipFree.restype = ctypes.c_int
        H.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        H.hipMemcpy.restype = ctypes.c_int
        H.hipDeviceSynchronize.restype = ctypes.c_int
        H.hipSetDevice.argtypes = [ctypes.c_int]
        H.hipSetDevice.restype = ctypes.c_int
        H.hipGetDeviceCount.argtypes = [ctypes.POINTER(ctypes.c_int)]
        H.hipGetDeviceCount.restype = ctypes.c_int
        _hip = H
    return _hip


def _hip_check(rc, what):
    if rc != 0:
        raise VoError(ERR_DEVICE, f"{what} failed with hipError {rc}")


def device_count():
    n = ctypes.c_int(0)
    rc = hip().hipGetDeviceCount(ctypes.byref(n))
    return n.value if rc == 0 else 0


def synchronize():
    _hip_check(hip().hipDeviceSynchronize(), "hipDeviceSynchronize")


class DeviceBuffer:
    """`nbytes` of uninitialised device memory (freed with free() or on GC)."""

    def __init__(self, nbytes):
        self.nbytes = int(nbytes)
        self.ptr = ctypes.c_void_p()
        _hip_check(hip().hipMalloc(ctypes.byref(self.ptr), max(self.nbytes, 16)), "hipMalloc")

    @property
    def address(self):
        return self.ptr.value

    def free(self):
        if self.ptr:
            hip().hipFree(self.ptr)
            self.ptr = ctypes.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class DeviceArray:
    """A device copy of a host numpy array (freed with free() or on GC)."""

    def __init__(self, host):
        host = np.ascontiguousarray(host)
        self.dtype, self.size = host.dtype, host.size
        self.ptr = ctypes.c_void_p()
        _hip_check(hip().hipMalloc(ctypes.byref(self.ptr), max(host.nbytes, 16)), "hipMalloc")
        if host.nbytes:
            _hip_check(hip().hipMemcpy(self.ptr, host.ctypes.data, host.nbytes, 1), "hipMemcpy H2D")

    @property
    def address(self):
        return self.ptr.value

    def free(self):
        if self.ptr:
            hip().hipFree(self.ptr)
            self.ptr = ctypes.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def copy_device(dst, src, nbytes):
    """Device-to-device copy (synchronous)."""
    if nbytes:
        _hip_check(hip().hipMemcpy(ctypes.c_void_p(dst), ctypes.c_void_p(src), int(nbytes), 3), "hipMemcpy D2D")


def to_host(address, dtype, count):
    out = np.empty(count, dtype)
    if count:
        _hip_check(hip().hipMemcpy(out.ctypes.data, ctypes.c_void_p(address), out.nbytes, 2), "hipMemcpy D2H")
    return out
