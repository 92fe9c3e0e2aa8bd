"""Served form (voroffset_amd.serve): FIFO admission under an HBM budget on
the CPU; on the GPU, served results equal the single-volume path."""

import numpy as np
import pytest

from voroffset_amd import serve, synth


class _StubCtx:
    def close(self):
        pass


def _server(budget, max_batch=8):
    s = serve.MorphoServer("fp32", hbm_budget=budget, max_batch=max_batch, ctx=_StubCtx())
    for side in (64, 128, 256):
        s._register(serve.Resident(side, side, None, None, 0.0, 1.0, 0.0, 1.0, side * side * 12))
    return s


def test_footprint_grows_with_stages_and_size():
    a = serve.footprint_bytes(1000, 12000, "dilation")
    assert serve.footprint_bytes(1000, 12000, "erosion") > a
    assert serve.footprint_bytes(1000, 12000, "closing") > serve.footprint_bytes(1000, 12000, "erosion")
    assert serve.footprint_bytes(4000, 48000, "dilation") > 3.9 * a
    assert serve.footprint_bytes(1000, 12000, "noop") == 0
    with pytest.raises(ValueError):
        serve.footprint_bytes(1, 1, "melt")


def test_admission_is_fifo_within_budget():
    s = _server(budget=serve.footprint_bytes(128 * 128, 128 * 128 * 12, "dilation") * 2)
    reqs = [s.submit(h, "dilation", 3.0, arrival=float(i)) for i, h in enumerate([0, 1, 2, 0, 1])]
    b1 = s.admit()
    assert [r.rid for r in b1] == [0, 1]            # 64^2 + 128^2 fit, 256^2 does not
    b2 = s.admit()
    assert [r.rid for r in b2] == [2]               # the head always runs, alone if too big
    b3 = s.admit(now=3.5)
    assert [r.rid for r in b3] == [3]               # request 4 has not arrived yet
    assert len(s.queue) == 1 and reqs[4] is s.queue[0]


def test_max_batch_and_unknown_requests():
    s = _server(budget=1 << 40, max_batch=2)
    for h in (0, 0, 0):
        s.submit(h, "erosion", 1.0, arrival=0.0)
    assert len(s.admit()) == 2
    with pytest.raises(KeyError):
        s.submit(99, "dilation", 1.0)
    with pytest.raises(ValueError):
        s.submit(0, "melt", 1.0)


@pytest.mark.gpu
@pytest.mark.parametrize("prec", ["fp32", "fp64"])
def test_served_results_equal_single_path(prec, hip_fp32, hip_fp64):
    hip = hip_fp32 if prec == "fp32" else hip_fp64
    srv = serve.MorphoServer(prec, hbm_budget=64 << 20, max_batch=4)
    try:
        vols = [synth.poisson_columns(s, s + 3, lam=6, zmax=300, seed=s) for s in (20, 33, 47)]
        handles = [srv.put(v) for v in vols]
        plan = [(0, "dilation", 3.0), (1, "closing", 2.5), (2, "erosion", 1.5), (0, "opening", 2.0),
                (1, "dilation", 7.0), (2, "closing", 4.0)]
        reqs = [srv.submit(handles[i], op, r) for i, op, r in plan]
        while srv.queue:
            srv.step()
        assert len(srv.batches) >= 2  # the budget splits the six requests
        for (i, op, r), req in zip(plan, reqs):
            got = srv.get(req.result, vols[i])
            ref, _, _ = hip.apply(op, vols[i], r)
            assert np.array_equal(got.off, ref.off), (op, r)
            assert np.array_equal(got.z, ref.z), (op, r)
        assert all(req.done >= req.start >= req.arrival for req in reqs)
    finally:
        srv.close()


def test_pool_reuses_buffers(monkeypatch):
    made = []

    class FakeBuffer:
        def __init__(self, nbytes):
            self.nbytes = nbytes
            made.append(self)

        def free(self):
            pass

    monkeypatch.setattr(serve._native, "DeviceBuffer", FakeBuffer)
    pool = serve.DevicePool()
    assert pool.size_class(1) == 256 and pool.size_class(257) == 512 and pool.size_class(4096) == 4096
    a, b = pool.take(1000), pool.take(3000)
    assert (a.nbytes, b.nbytes) == (1024, 4096) and len(made) == 2
    pool.give(a)
    pool.give(b)
    c, d = pool.take(600), pool.take(2100)   # same classes: no new allocation
    assert c is a and d is b and len(made) == 2
    e = pool.take(600)                      # none free in that class
    assert e is not a and len(made) == 3 and pool.allocated == 1024 + 4096 + 1024
