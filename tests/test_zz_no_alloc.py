"""vo_morph_device_noalloc (workspace in a static per-GPU scratch, result in
the caller's buffers) against vo_morph_device on the same inputs: bit-exact,
on two calls in a row with the second smaller, in both storage precisions and
for all four operations; a too small output buffer reports the capacity it
needs, a volume too large for the scratch is rejected."""

import numpy as np
import pytest

from voroffset_amd import _native, synth

pytestmark = pytest.mark.gpu


def _device_result(ctx, prec, op, r, vol, d_off, d_z, zlo, zhi):
    v, cnt, _ = ctx.morph_device(op, prec, r, vol.nx, vol.ny, d_off.address, d_z.address, zlo, zhi)
    dt = np.float32 if prec == "fp32" else np.float64
    return _native.to_host(v.offsets, np.int64, vol.ncols + 1), _native.to_host(v.z, dt, cnt)


@pytest.mark.parametrize("prec", ["fp32", "fp64"])
@pytest.mark.parametrize("op", ["dilation", "erosion", "closing", "opening"])
def test_noalloc_equals_device_entry_point(prec, op):
    dt = np.float32 if prec == "fp32" else np.float64
    ctx = _native.Context(0)
    keep = []
    try:
        for k, (nx, ny, r) in enumerate([(96, 80, 7.5), (41, 37, 3.0)]):   # the second call is smaller
            vol = synth.poisson_columns(nx, ny, lam=5, zmax=300, seed=50 + k, z_margin=20)
            zlo, zhi = vol.z_range()
            shift = _native.storage_shift(zlo, zhi) if prec == "fp32" else 0.0
            d_off, d_z = _native.DeviceArray(vol.off), _native.DeviceArray((vol.z - shift).astype(dt))
            keep += [d_off, d_z]
            ref_off, ref_z = _device_result(ctx, prec, op, r, vol, d_off, d_z, zlo - shift, zhi - shift)
            cap = ref_z.size + 100
            out_off = _native.DeviceBuffer((vol.ncols + 1) * 8)
            out_z = _native.DeviceBuffer(cap * np.dtype(dt).itemsize)
            keep += [out_off, out_z]
            cnt, _ = ctx.morph_device_noalloc(op, prec, r, vol.nx, vol.ny, d_off.address, d_z.address,
                                              zlo - shift, zhi - shift, out_off.address, out_z.address, cap)
            assert cnt == ref_z.size
            assert np.array_equal(_native.to_host(out_off.address, np.int64, vol.ncols + 1), ref_off)
            assert np.array_equal(_native.to_host(out_z.address, dt, cnt), ref_z)
            if cnt > 1:  # a too small output buffer: the capacity needed comes back
                with pytest.raises(_native.VoError) as e:
                    ctx.morph_device_noalloc(op, prec, r, vol.nx, vol.ny, d_off.address, d_z.address,
                                             zlo - shift, zhi - shift, out_off.address, out_z.address, cnt - 1)
                assert e.value.code == _native.ERR_CAPACITY and e.value.needed >= cnt
    finally:
        for b in keep:
            b.free()
        ctx.close()


def test_noalloc_rejects_volume_beyond_scratch():
    vol = synth.poisson_columns(512, 512, lam=6, zmax=1024, seed=3)
    zlo, zhi = vol.z_range()
    ctx = _native.Context(0)
    d_off, d_z = _native.DeviceArray(vol.off), _native.DeviceArray(vol.z)
    out_off, out_z = _native.DeviceBuffer((vol.ncols + 1) * 8), _native.DeviceBuffer(1 << 20)
    try:
        with pytest.raises(_native.VoError) as e:
            ctx.morph_device_noalloc("dilation", "fp64", 16.0, vol.nx, vol.ny, d_off.address, d_z.address,
                                     zlo, zhi, out_off.address, out_z.address, (1 << 20) // 8)
        assert e.value.code == _native.ERR_CAPACITY and "scratch" in str(e.value)
        # the context still works afterwards, with its own workspace
        _, cnt, _ = ctx.morph_device("dilation", "fp64", 16.0, vol.nx, vol.ny, d_off.address, d_z.address, zlo, zhi)
        assert cnt > 0
    finally:
        for b in (d_off, d_z, out_off, out_z):
            b.free()
        ctx.close()
