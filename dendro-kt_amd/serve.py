"""Request serving for the device CG (no counterpart in the reference, whose
solves are single-shot calls of HeatMat::cgSolve, FEM/examples/src/heatMat.cpp:165).

A request names a mesh (a seeded point cloud that the host refines, balances
and enumerates, as ot::DA(tree, ...) does) and a right-hand side; the server
answers with x = A^-1 b to a relative L_inf residual.  Design:

* plan cache: mesh -> (DA, device plan), keyed by the mesh spec, with LRU
  eviction when the estimated device footprint exceeds the HBM budget
  (admission control: a plan larger than the whole budget is refused);
* workers: one thread per HIP stream; solves on different streams overlap on
  the device (the C calls release the GIL), solves on the same plan are
  serialised because a plan owns one CG workspace;
* every reduction on the device runs in a fixed order, so a request's answer
  does not depend on what else is being served ("alone == in-mix").

Latencies are measured from submission to completion of each request
(end to end) and from plan availability to completion (solve).
"""
import collections
import queue
import threading
import time
from dataclasses import dataclass, field

import numpy as np


@dataclass(frozen=True)
class MeshSpec:
    dim: int
    maxdepth: int
    npoints: int
    seed: int
    order: int = 1
    ncomp: int = 4
    sigma: float = 0.05


@dataclass
class Request:
    mesh: MeshSpec
    rhs_seed: int
    tol: float = 1e-8
    mode: int = 0
    max_iter: int = 100000
    # filled in by the server
    submitted: float = 0.0
    started: float = 0.0
    solve_started: float = 0.0
    finished: float = 0.0
    plan_built: bool = False
    stats: dict = field(default_factory=dict)
    x: object = None
    error: str = ""


class PlanCache:
    def __init__(self, dkt, budget_bytes):
        self.dkt = dkt
        self.budget = budget_bytes
        self.lock = threading.Lock()
        self.entries = collections.OrderedDict()  # spec -> (da, plan, bytes, lock)
        self.building = {}

    @staticmethod
    def footprint(plan, n_nodes):
        # device maps + CG workspace (FP64 and FP32 vectors) + the request's b and x
        return plan.info()["map_bytes"] + n_nodes * (8 * 6 + 4 * 5)

    def get(self, spec):
        """(da, plan, plan_lock, built_now).  Builds on a miss; concurrent misses on
        one spec wait for a single build."""
        while True:
            with self.lock:
                if spec in self.entries:
                    self.entries.move_to_end(spec)
                    da, plan, _, lk = self.entries[spec]
                    return da, plan, lk, False
                ev = self.building.get(spec)
                if ev is None:
                    ev = threading.Event()
                    self.building[spec] = ev
                    owner = True
                else:
                    owner = False
            if not owner:
                ev.wait()
                continue
            try:
                dkt = self.dkt
                pts = dkt.gauss_points(spec.dim, spec.maxdepth, spec.npoints, spec.ncomp, spec.sigma, spec.seed)
                da = dkt.DA(dkt.balance_tree(pts, 1), spec.order)
                plan = dkt.Plan(da, dkt.OP_STIFFNESS)
                nbytes = self.footprint(plan, da.n_nodes)
                if nbytes > self.budget:
                    raise MemoryError(f"plan needs {nbytes / 1e9:.1f} GB, budget {self.budget / 1e9:.1f} GB")
                with self.lock:
                    used = sum(e[2] for e in self.entries.values())
                    while self.entries and used + nbytes > self.budget:
                        _, old = self.entries.popitem(last=False)
                        used -= old[2]
                    self.entries[spec] = (da, plan, nbytes, threading.Lock())
                    return da, plan, self.entries[spec][3], True
            finally:
                with self.lock:
                    self.building.pop(spec, None)
                ev.set()


class Server:
    def __init__(self, dkt, n_streams=4, hbm_budget_bytes=200e9):
        import torch
        self.torch = torch
        self.dkt = dkt
        self.cache = PlanCache(dkt, hbm_budget_bytes)
        self.q = queue.Queue()
        self.done = []
        self.done_lock = threading.Lock()
        self.workers = []
        for i in range(n_streams):
            t = threading.Thread(target=self._work, args=(torch.cuda.Stream(),), daemon=True)
            t.start()
            self.workers.append(t)

    def submit(self, req):
        req.submitted = time.perf_counter()
        self.q.put(req)

    def _work(self, stream):
        torch = self.torch
        while True:
            req = self.q.get()
            if req is None:
                return
            try:
                req.started = time.perf_counter()
                da, plan, lk, built = self.cache.get(req.mesh)
                req.plan_built = built
                req.solve_started = time.perf_counter()
                b = self.dkt.keyed_uniform(req.rhs_seed, da.nodes())
                b[da.boundary()] = 0.0
                with torch.cuda.stream(stream):
                    bd = torch.tensor(b, device="cuda")
                    xd = torch.zeros_like(bd)
                    with lk:
                        req.stats = plan.cg_solve_torch(bd, xd, req.mode, req.max_iter, req.tol,
                                                        stream=stream.cuda_stream)
                    req.x = xd.cpu().numpy()
                stream.synchronize()
            except Exception as e:  # reported per request
                req.error = f"{type(e).__name__}: {e}"
            req.finished = time.perf_counter()
            with self.done_lock:
                self.done.append(req)

    def drain(self, n, timeout=None):
        t0 = time.perf_counter()
        while True:
            with self.done_lock:
                if len(self.done) >= n:
                    return list(self.done)
            if timeout is not None and time.perf_counter() - t0 > timeout:
                raise TimeoutError("requests did not finish")
            time.sleep(0.002)

    def close(self):
        for _ in self.workers:
            self.q.put(None)
        for t in self.workers:
            t.join()


def percentiles(values, ps=(50, 99)):
    v = np.asarray(values, dtype=np.float64)
    return {f"p{p}": float(np.percentile(v, p)) for p in ps}
