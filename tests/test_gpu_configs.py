"""The workload configurations at their full sizes on the GPU, checked against
the oracle.

* Direct volume, 1024x1024, z-range [-80, 1104], Poisson(6) segments per
  column, Exp(40) gaps and lengths: dilation r=64, erosion r=16 and r=64,
  closing r=16 (dilation r=16 is in test_gpu_parity.py).
* Uneven batch: 256 volumes, side log-uniform in [64, 512], Poisson(U[2,10]),
  r U[2,32], ops uniform, in one vo_morph_batch_device call.
* Served stream: resident volumes of side up to 2048, Poisson(8), r up to 48,
  mixed operations through serve.MorphoServer.

A column's result depends only on input within the radius (twice the radius
for closing and opening), so the full-size result is compared with the
oracle's result for a cropped input on windows farther than that from the
crop's inner edges; windows at the grid's corner keep the true border.  Over
whole volumes the tests check size-independent properties: valid CSR
(even, sorted, disjoint, touching fused; float32 rounding may leave gaps of
width 0) and the inclusions D(A) ⊇ A, E(A) ⊆ A, O(A) ⊆ A, and C(A) ⊇ A
farther than r from the border (the erosion step treats the outside of the
grid as empty, Voronoi.cpp:18-55).  FP32 storage: the oracle runs on the same
float32-rounded input; bar as in test_gpu_parity.py (counts equal except
where only gaps < 1e-5 dexel differ, endpoints within 1e-4 dexel).  Closing
and opening go through two rounded intermediate volumes, so their inclusions
are checked with the segments shrunk by 2e-4 at both ends.
"""

import numpy as np
import pytest

import oracle
from voroffset_amd import serve, synth

from .helpers import (compare_csr, crop, csr_valid, inside, narrow_gap_mismatches, quantize_fp32, reach,
                      window_columns)

pytestmark = pytest.mark.gpu

W = 24  # window side
THREADS = 16


def check_window(res, vol_q, op, r, x0, y0, w=W):
    """The result on the w x w window at (x0, y0) against the oracle on a crop
    of the (float32-rounded) input; returns (count-mismatch columns, max error)."""
    m = reach(op, r)
    cx0, cy0 = max(0, x0 - m), max(0, y0 - m)
    cx1, cy1 = min(vol_q.nx, x0 + w + m), min(vol_q.ny, y0 + w + m)
    c = crop(vol_q, cx0, cy0, cx1 - cx0, cy1 - cy0)
    ro, rz = oracle.morph(c, op, r, threads=THREADS)
    ref = window_columns(ro, rz, c.nx, x0 - cx0, y0 - cy0, w, w)
    got = window_columns(res.off, res.z, res.nx, x0, y0, w, w)
    bad, err = compare_csr(got[0], got[1], ref[0], ref[1])
    assert narrow_gap_mismatches(got[0], got[1], ref[0], ref[1], bad), (op, r, x0, y0, bad[:8])
    assert err <= 1e-4, (op, r, x0, y0, err)
    return len(bad), err


def check_whole(res, vol_q, op, r):
    ro, rz = oracle.morph(vol_q, op, r, threads=THREADS)
    bad, err = compare_csr(res.off, res.z, ro, rz)
    assert narrow_gap_mismatches(res.off, res.z, ro, rz, bad), (op, r, bad[:8])
    assert err <= 1e-4, (op, r, err)


def check_properties(res, vol_q, op, r):
    assert (res.nx, res.ny) == (vol_q.nx, vol_q.ny)
    assert csr_valid(res.off, res.z, zero_gaps=True), op
    if op == "dilation":
        assert inside(vol_q.off, vol_q.z, res.off, res.z) == 0, op
    elif op == "closing":
        # the erosion treats the outside of the grid as empty, so a closing
        # contains its input only farther than r from the border
        m = int(r) + 1
        if vol_q.nx > 2 * m and vol_q.ny > 2 * m:
            a = crop(vol_q, m, m, vol_q.nx - 2 * m, vol_q.ny - 2 * m)
            c = crop(res, m, m, res.nx - 2 * m, res.ny - 2 * m)
            assert inside(a.off, a.z, c.off, c.z, 2e-4) == 0, op
    else:
        assert inside(res.off, res.z, vol_q.off, vol_q.z, 2e-4 if op == "opening" else 0.0) == 0, op


# -- direct volume, 1024 x 1024 -------------------------------------------

@pytest.fixture(scope="module")
def direct():
    vol = synth.poisson_columns(1024, 1024, lam=6, zmax=1024, seed=1234)
    return vol, quantize_fp32(vol)


@pytest.mark.parametrize("op,r", [("dilation", 64.0), ("erosion", 16.0), ("erosion", 64.0), ("closing", 16.0)])
def test_direct_volume_fp32(hip_fp32, direct, op, r):
    vol, vol_q = direct
    res, t1, t2 = hip_fp32.apply(op, vol, r)
    check_properties(res, vol_q, op, r)
    for x0, y0 in ((500, 300), (131, 877), (0, 0), (1024 - W, 1024 - W)):
        check_window(res, vol_q, op, r, x0, y0)


# -- uneven batch of 256 volumes ---------------------------------------------

@pytest.fixture(scope="module")
def batch256(hip_fp32):
    specs = synth.batch_specs(256, seed=4321)
    items = [(s["op"], synth.poisson_columns(s["side"], s["side"], lam=s["lam"], zmax=1024, seed=s["seed"]),
              s["radius"]) for s in specs]
    res, t1, t2 = hip_fp32.apply_batch(items)
    return items, res, hip_fp32.ctx.last_batch_alone()


def test_batch256_properties(batch256):
    items, res, alone = batch256
    assert len(res) == 256
    assert sum(v.ncols for _, v, _ in items) > 10_000_000
    for (op, vol, r), got in zip(items, res):
        check_properties(got, quantize_fp32(vol), op, r)


SAMPLE = sorted(np.random.default_rng(7).choice(256, 16, replace=False).tolist())


@pytest.mark.parametrize("k", SAMPLE)
def test_batch256_against_oracle(batch256, k):
    items, res, _ = batch256
    op, vol, r = items[k]
    check_whole(res[k], quantize_fp32(vol), op, r)


# -- served stream -------------------------------------------------------------

SERVED_POOL = [(2048, 11), (128, 12), (640, 13)]   # (side, seed)
SERVED_PLAN = [(0, "closing", 48.0), (1, "erosion", 48.0), (2, "opening", 17.3), (0, "dilation", 30.5),
               (1, "closing", 9.7), (2, "dilation", 2.2), (0, "erosion", 11.0), (1, "opening", 40.0),
               (2, "closing", 33.3), (0, "opening", 5.5)]


@pytest.fixture(scope="module")
def served():
    vols = [synth.poisson_columns(s, s, lam=8, zmax=1024, seed=sd) for s, sd in SERVED_POOL]
    srv = serve.MorphoServer("fp32", hbm_budget=24 << 30, max_batch=4)
    try:
        handles = [srv.put(v) for v in vols]
        reqs = [srv.submit(handles[i], op, r) for i, op, r in SERVED_PLAN]
        while srv.queue:
            srv.step()
        results = [srv.get(req.result, vols[i]) for (i, _, _), req in zip(SERVED_PLAN, reqs)]
        nbatches = len(srv.batches)
    finally:
        srv.close()
    return vols, results, nbatches


@pytest.mark.parametrize("k", range(len(SERVED_PLAN)))
def test_served_against_oracle(served, k):
    vols, results, nbatches = served
    assert nbatches >= 2  # the budget splits the stream into several batches
    i, op, r = SERVED_PLAN[k]
    vol_q = quantize_fp32(vols[i])
    res = results[k]
    check_properties(res, vol_q, op, r)
    if vol_q.nx <= 160:
        check_whole(res, vol_q, op, r)
    else:
        s = vol_q.nx
        for x0, y0 in ((s // 2, s // 3), (s - W, 0)):
            check_window(res, vol_q, op, r, x0, y0)
