"""Error paths forced from the host: malformed resident volumes on the device
entry points (checked on the device as vo_morph checks host volumes), inputs
outside the z-range the operation needs, an injected HIP error status
(VO_INJECT_HIP_ERROR), and the reason a context could not be created.  The
reference reports these through vor_assert (src/vor3d/Common.h:228,
Common.cpp:8-18); here they are status codes, and the context stays usable."""

import os

import numpy as np
import pytest

from voroffset_amd import _native, serve, synth

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = _native.Context(0)
    yield c
    c.close()


def _device(off, z, dtype=np.float64):
    return _native.DeviceArray(np.asarray(off, np.int64)), _native.DeviceArray(np.asarray(z, dtype))


def _run(ctx, op, r, off, z, zlo=0.0, zhi=100.0, prec="fp64"):
    d_off, d_z = _device(off, z, np.float32 if prec == "fp32" else np.float64)
    try:
        return ctx.morph_device(op, prec, r, 2, 1, d_off.address, d_z.address, zlo, zhi)
    finally:
        d_off.free()
        d_z.free()


GOOD = ([0, 2, 4], [10.0, 20.0, 30.0, 40.0])
BAD = {
    "odd_column": ([0, 3, 4], [10.0, 20.0, 30.0, 40.0]),
    "unsorted": ([0, 2, 4], [20.0, 10.0, 30.0, 40.0]),
    "nan": ([0, 2, 4], [10.0, np.nan, 30.0, 40.0]),
    "inf": ([0, 2, 4], [10.0, 20.0, 30.0, np.inf]),
    "decreasing_offsets": ([0, 4, 2], [10.0, 20.0, 30.0, 40.0]),
}


@pytest.mark.parametrize("prec", ["fp32", "fp64"])
@pytest.mark.parametrize("name", sorted(BAD))
@pytest.mark.parametrize("op", ["dilation", "erosion", "closing"])
def test_malformed_device_input(ctx, prec, name, op):
    off, z = BAD[name]
    with pytest.raises(_native.VoError) as e:
        _run(ctx, op, 2.0, off, z, prec=prec)
    assert e.value.code == _native.ERR_ARGUMENT, str(e.value)
    _, cnt, _ = _run(ctx, op, 0.5, *GOOD, prec=prec)  # the context still works
    assert cnt == 4  # each column keeps one segment (its neighbour is farther than 0.5)


@pytest.mark.parametrize("prec", ["fp32", "fp64"])
def test_out_of_range_device_input(ctx, prec):
    # erosion / opening: endpoints within [z_lo - 1, z_hi + 1]
    for op in ("erosion", "opening"):
        with pytest.raises(_native.VoError) as e:
            _run(ctx, op, 1.0, [0, 2, 2], [-1.5, 5.0, 0, 0][:2], zlo=0.0, zhi=100.0, prec=prec)
        assert e.value.code == _native.ERR_RANGE
        _run(ctx, op, 1.0, [0, 2, 2], [-1.0, 101.0], zlo=0.0, zhi=100.0, prec=prec)  # the edge is allowed
    # closing: the dilation by r must stay within the range
    with pytest.raises(_native.VoError) as e:
        _run(ctx, "closing", 5.0, [0, 2, 2], [3.0, 50.0], zlo=0.0, zhi=100.0, prec=prec)
    assert e.value.code == _native.ERR_RANGE
    _run(ctx, "closing", 4.0, [0, 2, 2], [3.0, 50.0], zlo=0.0, zhi=100.0, prec=prec)
    # dilation has no range precondition
    _run(ctx, "dilation", 50.0, [0, 2, 2], [-500.0, 500.0], zlo=0.0, zhi=100.0, prec=prec)


def test_batch_names_the_bad_job(ctx):
    arrays, jobs = [], []
    for k in range(5):
        off, z = GOOD if k != 3 else BAD["unsorted"]
        d_off, d_z = _device(off, z)
        arrays += [d_off, d_z]
        jobs.append(("dilation" if k % 2 else "erosion", 1.5, 2, 1, d_off.address, d_z.address, 0.0, 100.0))
    try:
        with pytest.raises(_native.VoError) as e:
            ctx.morph_batch_device("fp64", jobs)
        assert e.value.code == _native.ERR_ARGUMENT and "job 3" in str(e.value)
        jobs[3] = ("noop",) + jobs[3][1:]  # a pass-through job is not checked
        outs, _ = ctx.morph_batch_device("fp64", jobs)
        assert len(outs) == 5
    finally:
        for a in arrays:
            a.free()


def test_served_closing_out_of_range_raises():
    vol = synth.poisson_columns(20, 20, lam=6, zmax=300, seed=5, z_margin=10)
    srv = serve.MorphoServer("fp32", hbm_budget=1 << 30)
    try:
        h = srv.put(vol)
        srv.submit(h, "closing", 30.0)   # the dilation leaves [z_lo - 1, z_hi + 1]
        with pytest.raises(_native.VoError) as e:
            srv.step()
        assert e.value.code == _native.ERR_RANGE
        req = srv.submit(h, "closing", 5.0)
        srv.step()
        assert req.result >= 0
    finally:
        srv.close()


def test_injected_hip_error():
    os.environ["VO_INJECT_HIP_ERROR"] = "pass1 launch"
    try:
        bad = _native.Context(0)
    finally:
        del os.environ["VO_INJECT_HIP_ERROR"]
    good = _native.Context(0)
    try:
        vol = synth.poisson_columns(16, 16, lam=4, zmax=200, seed=3)
        zlo, zhi = vol.z_range()
        with pytest.raises(_native.VoError) as e:
            bad.morph_host("dilation", "fp64", 3.0, vol.nx, vol.ny, vol.off, vol.z, zlo, zhi)
        assert e.value.code == _native.ERR_DEVICE and "injected" in str(e.value)
        off, z, _ = good.morph_host("dilation", "fp64", 3.0, vol.nx, vol.ny, vol.off, vol.z, zlo, zhi)
        assert off[-1] == z.size > 0
    finally:
        bad.close()
        good.close()


def test_create_failure_reason():
    with pytest.raises(_native.VoError) as e:
        _native.Context(4096)
    assert "device" in str(e.value).split("vo_create failed:")[1]
    # the failed call leaves no error behind for the next context
    c = _native.Context(0)
    try:
        vol = synth.poisson_columns(8, 8, lam=3, zmax=100, seed=1)
        zlo, zhi = vol.z_range()
        c.morph_host("dilation", "fp64", 2.0, vol.nx, vol.ny, vol.off, vol.z, zlo, zhi)
    finally:
        c.close()
