"""Autotuner (vo_autotune): choices come from the allowed sets, survive a
save/load round trip, are rejected when malformed, and never change results."""

import numpy as np
import pytest

from voroffset_amd import _native, synth

pytestmark = pytest.mark.gpu


def test_autotune_roundtrip_and_results_unchanged(tmp_path):
    vol = synth.poisson_columns(96, 80, lam=6, zmax=512, seed=21)
    ctx = _native.Context(0)
    d_off, d_z = _native.DeviceArray(vol.off), _native.DeviceArray(vol.z)
    try:
        zlo, zhi = vol.z_range()
        args = ("dilation", "fp64", 12.0, vol.nx, vol.ny, d_off.address, d_z.address, zlo, zhi)
        before_v, before_n, _ = ctx.morph_device(*args)
        before = (_native.to_host(before_v.offsets, np.int64, vol.ncols + 1),
                  _native.to_host(before_v.z, np.float64, before_n))
        table = ctx.autotune("fp64", vol.nx, vol.ny, d_off.address, d_z.address, [3.0, 12.0], reps=1)
        for events, ocap in table:
            assert events in (512, 768, 1024, 1536) and ocap in (8, 16, 24, 32)
        after_v, after_n, _ = ctx.morph_device(*args)
        assert after_n == before_n
        assert np.array_equal(_native.to_host(after_v.offsets, np.int64, vol.ncols + 1), before[0])
        assert np.array_equal(_native.to_host(after_v.z, np.float64, after_n), before[1])
        path = tmp_path / "tuning.txt"
        ctx.tuning_save(path)
        other = _native.Context(0)
        try:
            other.tuning_load(path)
            assert other.tuning() == table
            bad = tmp_path / "bad.txt"
            bad.write_text("voroffset_amd tuning 1\n0 999 16\n")
            with pytest.raises(_native.VoError):
                other.tuning_load(bad)
            assert other.tuning() == table
        finally:
            other.close()
    finally:
        d_off.free()
        d_z.free()
        ctx.close()
