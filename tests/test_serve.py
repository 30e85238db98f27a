"""Serving engine: admission, batching, plan cache and HBM budget with a
host-only backend (CPU), and on the GPU that every request's result is the
same served in the mix as alone."""
import threading
import time

import numpy as np
import pytest

from nufft_3d_amd import serve


class FakePlan:
    def __init__(self, key, pbytes):
        self.key = key
        self.pbytes = pbytes
        self.wbytes = 0
        self.closed = False
        self.users = 0

    def close(self):
        assert self.users == 0, "plan closed while a batch runs on it"
        self.closed = True

    def release_workspaces(self):
        self.wbytes = 0


class FakeBackend:
    """Sleeps in proportion to the work and returns per-coil tags
    (seed*1000 + coil) so slicing of batched results can be checked."""

    def __init__(self, sec_per_gwork=0.02):
        self.k = sec_per_gwork
        self.lock = threading.Lock()
        self.built = []
        self.max_concurrent = 0
        self.running = 0
        self.batches = []

    def make_plan(self, req):
        p = FakePlan(req.plan_key, serve.est_plan_bytes(req.N, req.M))
        with self.lock:
            self.built.append(p)
        return p

    def plan_bytes(self, plan):
        return plan.pbytes, plan.wbytes

    def run(self, plan, reqs, slot):
        assert not plan.closed
        with self.lock:
            plan.users += 1
            assert plan.users == 1, "two batches on one plan at once"
            self.running += 1
            self.max_concurrent = max(self.max_concurrent, self.running)
            self.batches.append([r.rid for r in reqs])
        nc = sum(r.nc for r in reqs)
        time.sleep(1e-3 + self.k * sum(r.work() for r in reqs) / 1e9)
        plan.wbytes = max(plan.wbytes, serve.est_workspace_bytes(reqs[0].N, reqs[0].M, nc,
                                                                 reqs[0].kind))
        out = np.concatenate([r.seed * 1000 + np.arange(r.nc) for r in reqs])
        with self.lock:
            plan.users -= 1
            self.running -= 1
        outs, o = [], 0
        for r in reqs:
            outs.append(out[o:o + r.nc])
            o += r.nc
        return outs


def small_stream(n=40, rate=2000.0, seed=3):
    return serve.make_stream(n, rate, seed=seed, sizes=(16, 24), coils=(1, 2, 4),
                             samples=(2000, 20000), traj_per_size=2, maxit=3)


def test_stream_is_seeded_and_drawn_from_the_stated_sets():
    a, b = small_stream(seed=9), small_stream(seed=9)
    assert [(r.N, r.nc, r.kind, r.M, r.arrival) for r in a] == \
           [(r.N, r.nc, r.kind, r.M, r.arrival) for r in b]
    big = serve.make_stream(400, 10.0, seed=1, traj_per_size=3)
    assert {r.N for r in big} == {64, 128, 192}
    assert {r.nc for r in big} == {1, 8, 16, 32}
    assert all(100_000 * 0.9 <= r.M <= 5_000_000 * 1.1 for r in big)
    kinds = [r.kind for r in big]
    assert 0.4 < kinds.count("pair") / len(kinds) < 0.6
    gaps = np.diff([r.arrival for r in big])
    assert np.all(gaps >= 0) and abs(np.mean(gaps) - 0.1) < 0.03
    assert len({r.plan_key for r in big}) == 9


def test_every_request_gets_its_own_result_and_batches_form():
    reqs = small_stream()
    be = FakeBackend()
    eng = serve.Engine(be, hbm_budget=200e9, slots=3, max_batch_coils=8, keep_results=True)
    st = eng.serve(reqs)
    assert st["requests"] == len(reqs)
    for r in reqs:
        assert np.array_equal(r.result, r.seed * 1000 + np.arange(r.nc))
        assert r.t_done >= r.t_admit >= r.arrival
    assert any(len(b) > 1 for b in be.batches)          # requests were batched
    byid = {r.rid: r for r in reqs}
    for b in be.batches:
        g = [byid[i] for i in b]
        assert len({(r.plan_key, r.kind, r.maxit) for r in g}) == 1
        assert sum(r.nc for r in g) <= 8
    assert be.max_concurrent > 1                          # batches overlapped
    assert st["plans_built"] == len({r.plan_key for r in reqs}) == len(be.built)
    assert st["latency_s"]["p50"] <= st["latency_s"]["p95"] <= st["latency_s"]["p99"]
    assert set(st["latency_by_class_s"]) == {"N16", "N24"}
    eng.close()


def test_no_batching_runs_requests_singly():
    reqs = small_stream(20)
    be = FakeBackend()
    eng = serve.Engine(be, slots=2, batch=False, keep_results=True)
    eng.serve(reqs)
    assert all(len(b) == 1 for b in be.batches)
    assert len(be.batches) == len(reqs)


def test_budget_is_respected_with_eviction():
    reqs = small_stream(60)
    # room for about one plan with its batch at a time
    need = [serve.est_plan_bytes(r.N, r.M) + serve.est_workspace_bytes(r.N, r.M, r.nc, r.kind)
            + serve.io_bytes(r.N, r.M, r.nc, r.kind) for r in reqs]
    budget = 1.3 * max(need)
    be = FakeBackend()
    eng = serve.Engine(be, hbm_budget=budget, slots=4, max_batch_coils=4)
    st = eng.serve(reqs)
    assert st["requests"] == len(reqs)
    assert eng.peak_bytes <= budget
    assert st["plan_evictions"] > 0
    assert st["plans_built"] > len({r.plan_key for r in reqs})   # rebuilt after eviction
    assert all(p.closed for p in be.built if p not in [e[0] for e in eng.cache.entries.values()])


def test_request_larger_than_budget_is_an_error():
    reqs = small_stream(3)
    eng = serve.Engine(FakeBackend(), hbm_budget=1e5)
    with pytest.raises(MemoryError):
        eng.serve(reqs)


def test_one_at_a_time_builds_a_plan_per_request():
    reqs = small_stream(10)
    be = FakeBackend()
    st = serve.serve_one_at_a_time(be, reqs, keep_results=True)
    assert st["plans_built"] == len(be.built) == 10
    assert all(p.closed for p in be.built)
    assert be.max_concurrent == 1
    for r in reqs:
        assert np.array_equal(r.result, r.seed * 1000 + np.arange(r.nc))


def test_percentile_is_nearest_rank():
    v = list(range(1, 101))
    assert serve._pct(v, 50) == 50 and serve._pct(v, 99) == 99 and serve._pct(v, 100) == 100
    assert serve._pct([3.0], 99) == 3.0


@pytest.mark.gpu
def test_gpu_results_same_alone_and_in_the_mix():
    import torch
    reqs = serve.make_stream(14, 1e4, seed=4, sizes=(16, 24), coils=(1, 2, 3),
                             samples=(3000, 12000), traj_per_size=2, maxit=6)
    be = serve.GpuBackend((16, 24), max_coils=8)
    eng = serve.Engine(be, slots=3, max_batch_coils=16, keep_results=True)
    st = eng.serve(reqs)
    mixed = {r.rid: r.result.clone() for r in reqs}
    batched = sum(1 for r in reqs if r.batch_coils > r.nc)
    print("served %d requests, %d in shared launches, mean batch coils %.1f"
          % (st["requests"], batched, st["mean_batch_coils"]), flush=True)
    assert batched > 0
    eng.close()
    serve.reset(reqs)
    serve.serve_one_at_a_time(be, reqs, keep_results=True)
    torch.cuda.synchronize()
    worst = 0.0
    for r in reqs:
        a, b = mixed[r.rid], r.result
        assert a.shape == b.shape
        e = float(torch.linalg.vector_norm(a - b) / max(float(torch.linalg.vector_norm(b)), 1e-30))
        worst = max(worst, e)
    print("worst relative L2 difference alone vs mixed: %.3e" % worst, flush=True)
    # the kernels treat coils independently in a fixed order; the only
    # difference is the grouping of dense-point sums (shorter segments for
    # launches of at most 4 coils), i.e. rounding
    assert worst <= 1e-6
