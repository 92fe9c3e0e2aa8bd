"""Batched form (vo_morph_batch_device): many volumes of uneven size, radius
and operation in one batch, checked against the single-volume path (the same
kernels, so bit-identical) and against the oracle."""

import numpy as np
import pytest

import oracle
from voroffset_amd import synth
from voroffset_amd.volume import CompressedVolume

from .helpers import compare_csr, narrow_gap_mismatches, quantize_fp32, random_case

pytestmark = pytest.mark.gpu


def mixed_items(seed, n=10, side=(12, 60), r=(1.0, 9.0)):
    rng = np.random.default_rng(seed)
    items = []
    for k in range(n):
        nx, ny = int(rng.integers(*side)), int(rng.integers(*side))
        vol = synth.poisson_columns(nx, ny, lam=float(rng.uniform(2, 8)), zmax=300, seed=seed * 100 + k, z_margin=14)
        items.append((synth.OPS[k % 4], vol, float(rng.uniform(*r))))
    return items


def test_batch_equals_single_path_fp64(hip_fp64):
    items = mixed_items(1)
    res, t1, t2 = hip_fp64.apply_batch(items)
    assert t1 > 0 and t2 >= 0  # the program's device time; jobs redone alone add to t2
    for (op, vol, r), got in zip(items, res):
        one, _, _ = hip_fp64.apply(op, vol, r)
        assert (got.nx, got.ny) == (vol.nx, vol.ny)
        assert np.array_equal(got.off, one.off), op
        assert np.array_equal(got.z, one.z), op
        bad, err = compare_csr(got.off, got.z, *oracle.morph(vol, op, r))
        assert len(bad) == 0 and err <= 1e-9, (op, r, bad[:5], err)


def test_batch_equals_single_path_fp32(hip_fp32):
    """FP32 storage: the batch quantizes every volume relative to the same
    integer shift as the host entry point (the z-range centre rounded half
    away from zero; these volumes have odd z-extents, where rounding half to
    even would differ), so batched results equal the single-volume ones bit
    for bit."""
    items = [random_case(3000 + k, degenerate=False) for k in range(24)]
    assert any(float(v.extent[2]) % 2 == 1 for _, v, _ in items)
    res, _, _ = hip_fp32.apply_batch(items)
    for (op, vol, r), got in zip(items, res):
        one, _, _ = hip_fp32.apply(op, vol, r)
        assert np.array_equal(got.off, one.off), (op, r)
        assert np.array_equal(got.z, one.z), (op, r)


def test_batch_large_radii_fp64(hip_fp64):
    # radii beyond one coverage word per lane (R > 31 and R > 63) in one batch
    items = mixed_items(3, n=6, side=(8, 30), r=(20.0, 80.0))
    items = [(op, synth.poisson_columns(vol.nx, vol.ny, lam=4, zmax=300, seed=40 + k, z_margin=100), r)
             for k, (op, vol, r) in enumerate(items)]
    res, _, _ = hip_fp64.apply_batch(items)
    for (op, vol, r), got in zip(items, res):
        bad, err = compare_csr(got.off, got.z, *oracle.morph(vol, op, r))
        assert len(bad) == 0 and err <= 1e-9, (op, r, bad[:5], err)


def test_batch_fp32_against_oracle(hip_fp32):
    items = mixed_items(2, n=12)
    res, _, _ = hip_fp32.apply_batch(items)
    for (op, vol, r), got in zip(items, res):
        ref = oracle.morph(quantize_fp32(vol), op, r)
        bad, err = compare_csr(got.off, got.z, *ref)
        assert narrow_gap_mismatches(got.off, got.z, ref[0], ref[1], bad) and err <= 1e-4, (op, r, bad[:5], err)


def _vol(nx, ny, cols, zext):
    return CompressedVolume.from_columns(nx, ny, cols, origin=(0.0, 0.0, 0.0),
                                         extent=(float(nx), float(ny), zext), spacing=1.0, padding=0)


def test_batch_jobs_redone_alone(hip_fp64):
    # 200 segments in one column at r = 0.5 overflow the batched gather lists;
    # r = 130 is beyond the batched pass-1 form: both go through the
    # single-volume fallbacks inside the same call
    dense = _vol(5, 5, [[float(k) for k in range(0, 400, 2)] if i == 12 else [] for i in range(25)], 400.0)
    wide = synth.poisson_columns(9, 7, lam=3, zmax=200, seed=5, z_margin=140)
    plain = synth.poisson_columns(30, 20, lam=5, zmax=200, seed=6)
    items = [("dilation", dense, 0.5), ("dilation", wide, 130.0), ("dilation", plain, 4.0), ("closing", plain, 2.0)]
    res, _, _ = hip_fp64.apply_batch(items)
    for (op, vol, r), got in zip(items, res):
        bad, err = compare_csr(got.off, got.z, *oracle.morph(vol, op, r))
        assert len(bad) == 0 and err <= 1e-9, (op, r, bad[:5], err)
    assert hip_fp64.ctx.last_batch_alone() >= 2


def test_empty_batch_and_empty_grid(hip_fp64):
    res, t1, t2 = hip_fp64.apply_batch([])
    assert res == []
    empty = _vol(0, 0, [], 10.0)
    blank = _vol(6, 4, [[] for _ in range(24)], 10.0)
    res, _, _ = hip_fp64.apply_batch([("dilation", empty, 2.0), ("erosion", blank, 1.0), ("noop", blank, 1.0)])
    assert res[0].z.size == 0 and res[1].z.size == 0 and res[2].z.size == 0


def test_batch_uneven_specs_fp64(hip_fp64):
    """The uneven-batch generator (sides log-uniform, lambda U[2,10], r U[2,32],
    ops uniform), at reduced sides so the oracle finishes quickly."""
    specs = synth.batch_specs(16, seed=77, side_range=(24, 72), r_range=(2.0, 12.0))
    items = [(s["op"], synth.poisson_columns(s["side"], s["side"], lam=s["lam"], zmax=256, seed=s["seed"],
                                             z_margin=int(s["radius"]) + 4), s["radius"]) for s in specs]
    res, _, _ = hip_fp64.apply_batch(items)
    for (op, vol, r), got in zip(items, res):
        bad, err = compare_csr(got.off, got.z, *oracle.morph(vol, op, r, threads=4))
        assert len(bad) == 0 and err <= 1e-9, (op, r, vol.nx, bad[:5], err)


def test_batch_program_replays_from_a_graph():
    """The same batch (same jobs, pointers and sizes) runs eagerly, is
    captured on its second call and replayed after that: every call's results
    equal the first's, with other work on the context between the calls."""
    from voroffset_amd import _native
    items = mixed_items(5, n=8)
    ctx = _native.Context(0)
    keep, jobs = [], []
    try:
        for op, vol, r in items:
            zlo, zhi = vol.z_range()
            d_off, d_z = _native.DeviceArray(vol.off), _native.DeviceArray(vol.z)
            keep += [d_off, d_z]
            jobs.append((op, r, vol.nx, vol.ny, d_off.address, d_z.address, zlo, zhi))

        def results():
            outs, _ = ctx.morph_batch_device("fp64", jobs)
            return [(_native.to_host(v.offsets, np.int64, v.nx * v.ny + 1), _native.to_host(v.z, np.float64, c))
                    for v, c in outs]
        first = results()
        for k in range(4):
            got = results()
            for (o0, z0), (o1, z1) in zip(first, got):
                assert np.array_equal(o0, o1) and np.array_equal(z0, z1), k
            # eager work between the calls
            op, vol, r = items[k]
            zlo, zhi = vol.z_range()
            ctx.morph_device("dilation", "fp64", 2.0, vol.nx, vol.ny, keep[2 * k].address, keep[2 * k + 1].address,
                             zlo, zhi)
        for (op, vol, r), (o, z) in zip(items, first):
            bad, err = compare_csr(o, z, *oracle.morph(vol, op, r))
            assert len(bad) == 0 and err <= 1e-9, (op, r)
    finally:
        for d in keep:
            d.free()
        ctx.close()


@pytest.mark.parametrize("prec", ["fp32", "fp64"])
def test_batch_work_queue_equals_flat_grid(prec, monkeypatch):
    """The persistent work-queue form of the batched passes (VO_BATCH_QUEUE=1:
    a resident-sized grid taking tiles from per-XCD queues) gives the same
    results as one workgroup per tile, over more tiles than the grid holds,
    every radius class, and three calls in a row (eager, captured, replayed)."""
    from voroffset_amd import _native
    rng = np.random.default_rng(11)
    dt = np.float32 if prec == "fp32" else np.float64
    items = []
    for k in range(24):
        side = int(rng.integers(60, 200))
        r = float([rng.uniform(1.5, 31), rng.uniform(32, 63), rng.uniform(64, 100)][k % 3])
        vol = synth.poisson_columns(side, side, lam=float(rng.uniform(2, 8)), zmax=400, seed=900 + k,
                                    z_margin=int(r) + 4)
        items.append((synth.OPS[k % 4], vol, r))
    assert sum(v.nx * -(-v.ny // 64) for _, v, _ in items) > 8192  # more tiles than a resident grid
    keep, jobs = [], []
    flat = _native.Context(0)
    monkeypatch.setenv("VO_BATCH_QUEUE", "1")
    queued = _native.Context(0)
    monkeypatch.delenv("VO_BATCH_QUEUE")
    try:
        for op, vol, r in items:
            zlo, zhi = vol.z_range()
            d_off, d_z = _native.DeviceArray(vol.off), _native.DeviceArray(vol.z.astype(dt))
            keep += [d_off, d_z]
            jobs.append((op, r, vol.nx, vol.ny, d_off.address, d_z.address, zlo, zhi))

        def results(ctx):
            outs, _ = ctx.morph_batch_device(prec, jobs)
            return [(_native.to_host(v.offsets, np.int64, v.nx * v.ny + 1), _native.to_host(v.z, dt, c))
                    for v, c in outs]
        want = results(flat)
        assert flat.last_batch_alone() == 0
        for call in range(3):
            got = results(queued)
            assert queued.last_batch_alone() == 0
            for k, ((o0, z0), (o1, z1)) in enumerate(zip(want, got)):
                assert np.array_equal(o0, o1) and np.array_equal(z0, z1), (call, k, items[k][0], items[k][2])
    finally:
        for d in keep:
            d.free()
        flat.close()
        queued.close()
