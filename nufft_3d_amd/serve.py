"""Continuous-batching reconstruction server on one GPU.

Requests - an fNUFT+aNUFT pair or a MINRES iNUFT of ``maxit`` iterations -
of mixed matrix size, coil count and trajectory arrive over time. The engine

* keeps device-resident plans in an LRU cache keyed by (trajectory, N);
  building a plan (sample sort, hit lists, density, deapodization) is paid
  once per trajectory instead of once per request;
* admits work against an HBM budget: the resident bytes of every cached plan
  (the library reports plan arrays and grown workspaces through
  ``nufft3d_device_bytes``) plus the input/output buffers and workspace
  growth of every batch in flight must stay within it. Idle plans are
  evicted least-recently-used first; a batch that does not fit waits for
  memory (strict FIFO on memory, so large requests are not starved);
* batches: requests waiting on the same plan and kind become extra coils
  (right-hand sides) of ONE library call. Every kernel treats coils
  independently (interpolation, the deterministic spread, FFT lines, and a
  MINRES recurrence per right-hand side with its own stopping test), so a
  request's result is the same alone and in a batch up to rounding (the
  spread groups the sums of dense k-space points in shorter segments for
  fp32 launches of at most 4 coils);
* overlaps: up to ``slots`` batches on different plans run at once, each on
  its own HIP stream driven by its own host thread (library calls release
  the GIL); a plan serves one batch at a time.

Per-request latency is arrival -> result ready (queueing included). Work is
counted in samples x coils x fwd+adj pairs (a ``maxit``-iteration iNUFT is
maxit+1 pairs: the right-hand side A'W raw plus one normal-operator
application per iteration).
"""
from __future__ import annotations

import collections
import concurrent.futures as cf
import math
import threading
import time
from dataclasses import dataclass

import numpy as np


@dataclass
class Request:
    rid: int
    traj_key: object          # identifies the trajectory (plan cache key with N)
    om: np.ndarray            # [3, M] trajectory in cycles/FOV
    N: int
    nc: int
    kind: str                 # "pair" or "inufft"
    arrival: float            # seconds after the stream starts
    maxit: int = 20
    seed: int = 0             # input data seed
    size_class: str = ""
    result: object = None
    t_admit: float = float("nan")
    t_done: float = float("nan")
    batch_coils: int = 0      # coils of the launch that served it

    @property
    def M(self):
        return int(self.om.shape[1])

    @property
    def plan_key(self):
        return (self.traj_key, self.N)

    def pairs(self):
        return 1 if self.kind == "pair" else self.maxit + 1

    def work(self):
        """samples x coils x fwd+adj pairs"""
        return self.M * self.nc * self.pairs()


# ---------------------------------------------------------------- byte models
def grid_size(N, u=2.0):
    return 2 * int(math.ceil(N * u / 2))


def est_plan_bytes(N, M, u=2.0, hits_per_sample=33.5):
    """Resident plan arrays before the plan exists (FP32 plan, J=4 radial):
    per sample 100 B of coordinates/ids/density plus 18 B per hit (2 B
    forward code, 8 B adjoint entry, 8 B adjoint weight entry), 8 B per grid
    point of row pointers, U."""
    K = grid_size(N, u)
    return int(M * (100 + 18 * hits_per_sample) + 8 * K ** 3 + 4 * N ** 3)


def est_workspace_bytes(N, M, nc, kind, u=2.0):
    """Workspaces one call of nc coils grows (complex64): grid, two FFT
    transposes, samples, and six MINRES vectors for an iNUFT."""
    K = grid_size(N, u)
    b = 8 * nc * (K ** 3 + K * N * N + K * K * N + M)
    if kind == "inufft":
        b += 8 * nc * (6 * N ** 3)
    return int(b)


def io_bytes(N, M, nc, kind):
    """Input and output buffers of a batch (plus the iNUFT's internal
    right-hand side and scratch copies)."""
    if kind == "pair":
        return int(8 * nc * (2 * N ** 3 + M))
    return int(8 * nc * (M + 4 * N ** 3))


# ---------------------------------------------------------------- backends
class GpuBackend:
    """Runs batches through the HIP library on torch-managed device memory."""

    def __init__(self, sizes, device=0, max_coils=32, seed=11, graphs=True):
        import torch
        from . import synth
        self.torch = torch
        self.graphs = graphs
        self.dev = torch.device("cuda", device)
        self.images, self.maps = {}, {}
        for N in sizes:
            img = synth.phantom_complex(N).astype(np.complex64)
            self.images[N] = torch.from_numpy(np.ascontiguousarray(img.T)).to(self.dev)
            maps = synth.coil_maps(N, max_coils, seed=seed).astype(np.complex64)
            self.maps[N] = torch.from_numpy(np.ascontiguousarray(np.transpose(maps, (3, 2, 1, 0)))
                                            ).to(self.dev)
        torch.cuda.synchronize(self.dev)
        self.max_coils = max_coils
        self._streams = {}
        self._lock = threading.Lock()

    def stream(self, slot):
        with self._lock:
            if slot not in self._streams:
                self._streams[slot] = self.torch.cuda.Stream(device=self.dev)
            return self._streams[slot]

    def make_plan(self, req):
        from .nufft_3d import nufft_3d
        plan = nufft_3d(req.om, req.N, gpu=1, device=self.dev.index or 0)
        # MINRES blocks run as HIP graphs: the library holds its device
        # allocations and releases (plan evictions, workspace trims and
        # growth, plan builds on other slots) while any thread captures, so
        # a capture is not invalidated by another thread's free; a capture
        # that still fails falls back to eager launches (bitwise the same)
        if not self.graphs:
            plan.set_graphs(False)
        return plan

    def plan_bytes(self, plan):
        return plan.device_bytes()

    def inputs(self, req):
        """Seeded input of one request on the current stream: the phantom
        times nc of the coil maps (pair), or complex Gaussian k-space (iNUFT)."""
        torch = self.torch
        if req.kind == "pair":
            c0 = req.seed % max(1, self.max_coils - req.nc + 1)
            return self.images[req.N][None] * self.maps[req.N][c0:c0 + req.nc]
        g = torch.Generator(device=self.dev)
        g.manual_seed(int(req.seed))
        return torch.randn((req.nc, req.M), dtype=torch.complex64, device=self.dev, generator=g)

    def run(self, plan, reqs, slot):
        torch = self.torch
        s = self.stream(slot)
        with torch.cuda.stream(s):
            data = [self.inputs(r) for r in reqs]
            x = data[0] if len(data) == 1 else torch.cat(data, dim=0)
            if reqs[0].kind == "pair":
                out = plan.aNUFT(plan.fNUFT(x))
            else:
                out = plan.iNUFT(x, reqs[0].maxit, 0.0)
            s.synchronize()
        outs, o = [], 0
        for r in reqs:
            outs.append(out[o:o + r.nc])
            o += r.nc
        return outs


class PlanCache:
    """LRU cache of device-resident plans with their measured bytes."""

    def __init__(self):
        self.entries = collections.OrderedDict()   # key -> [plan, plan_bytes, ws_bytes]
        self.builds = 0
        self.evictions = 0

    def resident(self):
        return sum(e[1] + e[2] for e in self.entries.values())

    def evict_idle(self, need, busy):
        """Free `need` bytes from idle plans, least recently used first:
        first their workspaces (cheap to regrow), then whole plans. Returns
        the bytes freed."""
        freed = 0
        for key, e in self.entries.items():
            if freed >= need:
                return freed
            if key in busy or e[2] == 0:
                continue
            e[0].release_workspaces()
            freed += e[2]
            e[2] = 0
        for key in list(self.entries):
            if freed >= need:
                break
            if key in busy:
                continue
            plan, pb, wb = self.entries.pop(key)
            plan.close()
            freed += pb + wb
            self.evictions += 1
        return freed


class Engine:
    """Continuous batching over a request stream under an HBM budget."""

    def __init__(self, backend, hbm_budget=200e9, slots=4, max_batch_coils=64, batch=True,
                 keep_results=False, progress=None):
        self.backend = backend
        self.budget = float(hbm_budget)
        self.slots = slots
        self.max_batch_coils = max_batch_coils
        self.batch = batch
        self.keep_results = keep_results
        self.progress = progress
        self.cache = PlanCache()
        self.peak_bytes = 0

    def _cost(self, group):
        """Bytes a batch adds to what is resident: a new plan's arrays, the
        workspace growth beyond the plan's current workspaces, and the
        batch's input/output buffers."""
        h = group[0]
        nc = sum(r.nc for r in group)
        ws = est_workspace_bytes(h.N, h.M, nc, h.kind)
        e = self.cache.entries.get(h.plan_key)
        if e is None:
            return est_plan_bytes(h.N, h.M) + ws + io_bytes(h.N, h.M, nc, h.kind)
        return max(0, ws - e[2]) + io_bytes(h.N, h.M, nc, h.kind)

    def _group(self, head, pending):
        group, coils = [head], head.nc
        if self.batch:
            for r in pending:
                if r is head or r.plan_key != head.plan_key or r.kind != head.kind:
                    continue
                if r.kind == "inufft" and r.maxit != head.maxit:
                    continue
                if coils + r.nc > self.max_batch_coils:
                    continue
                group.append(r)
                coils += r.nc
        return group

    def _work(self, group, slot, plan):
        if plan is None:
            plan = self.backend.make_plan(group[0])
        outs = self.backend.run(plan, group, slot)
        pb, wb = self.backend.plan_bytes(plan)
        return plan, pb, wb, outs

    def serve(self, requests):
        """Serve `requests` (arrival times in seconds from now). Returns stats."""
        reqs = sorted(requests, key=lambda r: r.arrival)
        pool = cf.ThreadPoolExecutor(max_workers=self.slots)
        pending = []
        inflight = {}              # future -> (key, group, reserved bytes, slot)
        busy = set()
        free_slots = list(range(self.slots))[::-1]
        reserved = 0
        nxt = 0
        done = []
        t0 = time.perf_counter()

        def now():
            return time.perf_counter() - t0

        last_report = 0.0
        try:
            while nxt < len(reqs) or pending or inflight:
                t = now()
                while nxt < len(reqs) and reqs[nxt].arrival <= t:
                    pending.append(reqs[nxt])
                    nxt += 1
                # admission: oldest request whose plan is idle, grouped with
                # every waiting request on the same plan and kind
                while pending and free_slots:
                    head = next((r for r in pending if r.plan_key not in busy), None)
                    if head is None:
                        break
                    group = self._group(head, pending)
                    cost = self._cost(group)
                    if cost > self.budget:
                        raise MemoryError("request %d needs %.1f GB, more than the %.1f GB budget"
                                          % (head.rid, cost / 1e9, self.budget / 1e9))
                    over = self.cache.resident() + reserved + cost - self.budget
                    if over > 0:
                        self.cache.evict_idle(over, busy | {head.plan_key})
                        over = self.cache.resident() + reserved + cost - self.budget
                    if over > 0 and head.plan_key in self.cache.entries and not inflight:
                        # only this batch's own cached plan is left: drop its
                        # workspaces, which the batch regrows within `cost`
                        e = self.cache.entries[head.plan_key]
                        e[0].release_workspaces()
                        e[2] = 0
                        cost = self._cost(group)
                        over = self.cache.resident() + reserved + cost - self.budget
                    if over > 0:
                        break              # wait for memory to free up
                    for r in group:
                        pending.remove(r)
                        r.t_admit = now()
                        r.batch_coils = sum(g.nc for g in group)
                    slot = free_slots.pop()
                    busy.add(head.plan_key)
                    reserved += cost
                    self.peak_bytes = max(self.peak_bytes, self.cache.resident() + reserved)
                    e = self.cache.entries.get(head.plan_key)
                    if e is not None:
                        self.cache.entries.move_to_end(head.plan_key)
                    fut = pool.submit(self._work, group, slot, e[0] if e is not None else None)
                    inflight[fut] = (head.plan_key, group, cost, slot)
                if inflight:
                    wait = 0.0005
                    if nxt < len(reqs):
                        wait = max(0.0, min(wait, reqs[nxt].arrival - now()))
                    fin, _ = cf.wait(list(inflight), timeout=wait, return_when=cf.FIRST_COMPLETED)
                    for f in fin:
                        key, group, cost, slot = inflight.pop(f)
                        plan, pb, wb, outs = f.result()
                        tdone = now()
                        if key in self.cache.entries:
                            self.cache.entries[key][1:] = [pb, wb]
                        else:
                            self.cache.entries[key] = [plan, pb, wb]
                            self.cache.builds += 1
                        for r, o in zip(group, outs):
                            r.t_done = tdone
                            r.result = o if self.keep_results else None
                            done.append(r)
                        busy.discard(key)
                        free_slots.append(slot)
                        reserved -= cost
                    if self.progress and now() - last_report > 10.0:
                        last_report = now()
                        self.progress("served %d/%d (t=%.1f s, resident %.1f GB)"
                                      % (len(done), len(reqs), last_report,
                                         self.cache.resident() / 1e9))
                elif nxt < len(reqs):
                    time.sleep(max(0.0, min(0.002, reqs[nxt].arrival - now())))
        finally:
            pool.shutdown(wait=True)
        makespan = now()
        st = summarize(done, makespan)
        st.update({"plans_built": self.cache.builds, "plan_evictions": self.cache.evictions,
                   "peak_accounted_gb": self.peak_bytes / 1e9, "slots": self.slots,
                   "batching": self.batch, "hbm_budget_gb": self.budget / 1e9,
                   "mean_batch_coils": float(np.mean([r.batch_coils for r in done])) if done
                   else 0.0})
        return st

    def close(self):
        for plan, _, _ in self.cache.entries.values():
            plan.close()
        self.cache.entries.clear()


def serve_one_at_a_time(backend, requests, keep_results=False):
    """The plain single-request form: in arrival order, each request alone on
    one stream, with its own plan built for it (``obj = nufft_3d(om,N)`` then
    fNUFT/aNUFT or iNUFT, then the plan is dropped), no batching, no overlap."""
    reqs = sorted(requests, key=lambda r: r.arrival)
    t0 = time.perf_counter()

    def now():
        return time.perf_counter() - t0

    busy_s = 0.0
    for r in reqs:
        while now() < r.arrival:
            time.sleep(min(0.002, max(0.0, r.arrival - now())))
        ts = now()
        r.t_admit = ts
        plan = backend.make_plan(r)
        out = backend.run(plan, [r], 0)[0]
        plan.close()
        r.t_done = now()
        r.batch_coils = r.nc
        r.result = out if keep_results else None
        busy_s += r.t_done - ts
    st = summarize(reqs, now())
    st["busy_s"] = busy_s
    st["capacity_msamples_coil_s"] = sum(r.work() for r in reqs) / busy_s / 1e6 if busy_s else 0.0
    st["plans_built"] = len(reqs)
    return st


def _pct(v, q):
    """Nearest-rank percentile."""
    if not v:
        return float("nan")
    v = sorted(v)
    return v[min(len(v) - 1, max(0, int(math.ceil(q / 100.0 * len(v))) - 1))]


def summarize(done, makespan):
    lat = [r.t_done - r.arrival for r in done]
    by = collections.defaultdict(list)
    for r in done:
        by[r.size_class or ("N%d" % r.N)].append(r.t_done - r.arrival)
    work = sum(r.work() for r in done)
    first = min((r.arrival for r in done), default=0.0)
    span = makespan - first
    return {
        "requests": len(done),
        "makespan_s": span,
        "work_msamples_coil": work / 1e6,
        "throughput_msamples_coil_s": work / span / 1e6 if span > 0 else float("nan"),
        "latency_s": {"p50": _pct(lat, 50), "p95": _pct(lat, 95), "p99": _pct(lat, 99)},
        "latency_by_class_s": {k: {"n": len(v), "p50": _pct(v, 50), "p95": _pct(v, 95),
                                   "p99": _pct(v, 99)} for k, v in sorted(by.items())},
    }


# ---------------------------------------------------------------- request stream
def make_stream(n_requests, rate_hz, seed=5, sizes=(64, 128, 192), coils=(1, 8, 16, 32),
                samples=(100_000, 5_000_000), traj_per_size=3, pair_frac=0.5, maxit=20):
    """Seeded request stream. Poisson arrivals at `rate_hz`; matrix size and
    coil count uniform over `sizes` and `coils`; fwd+adj pair with probability
    `pair_frac`, else a `maxit`-iteration iNUFT. Trajectories come from a
    protocol library of `traj_per_size` koosh-ball trajectories per size whose
    sample counts are log-uniform over `samples` (2N readout points per
    Fibonacci-sphere spoke), so requests share plans as scans of one
    protocol do."""
    from . import synth
    rng = np.random.default_rng(seed)
    lib = {}
    for N in sizes:
        lst = []
        for j in range(traj_per_size):
            M = math.exp(rng.uniform(math.log(samples[0]), math.log(samples[1])))
            nr = 2 * N
            ns = max(1, int(round(M / nr)))
            lst.append((("koosh", N, ns, nr), synth.koosh_ball(N, ns, nr)))
        lib[N] = lst
    t = 0.0
    reqs = []
    for i in range(n_requests):
        t += rng.exponential(1.0 / rate_hz)
        N = int(rng.choice(sizes))
        key, om = lib[N][int(rng.integers(len(lib[N])))]
        nc = int(rng.choice(coils))
        kind = "pair" if rng.uniform() < pair_frac else "inufft"
        reqs.append(Request(i, key, om, N, nc, kind, t, maxit=maxit, seed=1000 + i,
                            size_class="N%d" % N))
    return reqs


def reset(reqs):
    for r in reqs:
        r.result = None
        r.t_admit = r.t_done = float("nan")
        r.batch_coils = 0
