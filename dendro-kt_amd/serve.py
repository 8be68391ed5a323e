"""Request serving for the device CG (no counterpart in the reference, whose
solves are single-shot calls of HeatMat::cgSolve, FEM/examples/src/heatMat.cpp:165).

A request names a mesh (a seeded Gaussian point cloud, or a tree refined around
a sphere, that the host refines, balances and enumerates, as ot::DA(tree, ...)
does), a kind and a right-hand side; the server answers with x = A^-1 b to a
relative L_inf residual.  Kinds:
* "poisson": b is given per node (zero on the boundary), A the
  Dirichlet-masked stiffness (HeatMat::cgSolve);
* "heat": the reference's heat-equation driver (FEM/examples/src/heatEq.cpp):
  b = M f by the mass operator with HeatVec's hooks (feVector::computeVec,
  FEM/include/feVector.h:122-168), then the same CG on the stiffness.
Design:

* plan cache: mesh -> (DA, device plan), keyed by the mesh spec, with LRU
  eviction when the estimated device footprint exceeds the HBM budget
  (admission control: a plan larger than the whole budget is refused);
* workers: one thread per HIP stream; solves on different streams overlap on
  the device (the C calls release the GIL), solves on the same plan are
  serialised because a plan owns one CG workspace;
* every reduction on the device runs in a fixed order, so a request's answer
  does not depend on what else is being served ("alone == in-mix");
* continuous batching (Server(batch=K)): each worker runs a CG engine
  (dkt_cg_engine) of K slots; a queued request joins as soon as a slot is free
  (smallest mesh first among the requests the worker holds, which bounds the
  median latency under load) and leaves when its solve stops, between chunks of 8 iterations, so one
  launch per step serves every running solve and no solve waits for a slower
  one.  Solves are FP64; each answer is bitwise the request's own FP64 solve.
  Every slot has its own CG workspace, so requests on one mesh run together.

Latencies are measured from submission to completion of each request
(end to end) and from plan availability to completion (solve).
"""
import collections
import queue
import threading
import time
from dataclasses import dataclass, field

import numpy as np

_EMPTY = object()  # no request available


@dataclass(frozen=True)
class MeshSpec:
    """shape "gauss": npoints Gaussian-mixture points (seed), one per leaf, 2:1 balanced;
    shape "sphere": cells above level 2 split everywhere and up to max_level where they
    meet the sphere (center, radius), 2:1 balanced."""
    dim: int
    maxdepth: int
    npoints: int = 0
    seed: int = 0
    order: int = 1
    ncomp: int = 4
    sigma: float = 0.05
    shape: str = "gauss"
    center: tuple = ()
    radius: float = 0.0
    max_level: int = 0

    def tree(self, dkt):
        if self.shape == "sphere":
            return dkt.balance_tree(dkt.sphere_tree(self.dim, self.maxdepth, 2, self.max_level, self.center,
                                                    self.radius), 1)
        pts = dkt.gauss_points(self.dim, self.maxdepth, self.npoints, self.ncomp, self.sigma, self.seed)
        return dkt.balance_tree(pts, 1)


@dataclass
class Request:
    mesh: MeshSpec
    rhs_seed: int
    tol: float = 1e-8
    mode: int = 0
    max_iter: int = 100000
    kind: str = "poisson"
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
        self.entries = collections.OrderedDict()  # spec -> (da, plans, bytes, lock)
        self.building = {}

    @staticmethod
    def footprint(plans, n_nodes):
        # device maps (stiffness and mass plans) + CG workspace (FP64 and FP32
        # vectors) + the request's f, b and x
        return sum(p.info()["map_bytes"] for p in plans) + n_nodes * (8 * 7 + 4 * 5)

    def get(self, spec):
        """(da, (stiffness plan, mass plan), plan_lock, built_now).  Builds on a miss;
        concurrent misses on one spec wait for a single build."""
        while True:
            with self.lock:
                if spec in self.entries:
                    self.entries.move_to_end(spec)
                    da, plans, _, lk = self.entries[spec]
                    return da, plans, lk, False
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
                da = dkt.DA(spec.tree(dkt), spec.order)
                plans = (dkt.Plan(da, dkt.OP_STIFFNESS), dkt.Plan(da, dkt.OP_MASS))
                nbytes = self.footprint(plans, da.n_nodes)
                if nbytes > self.budget:
                    raise MemoryError(f"plan needs {nbytes / 1e9:.1f} GB, budget {self.budget / 1e9:.1f} GB")
                with self.lock:
                    used = sum(e[2] for e in self.entries.values())
                    while self.entries and used + nbytes > self.budget:
                        _, old = self.entries.popitem(last=False)
                        used -= old[2]
                    self.entries[spec] = (da, plans, nbytes, threading.Lock())
                    return da, plans, self.entries[spec][3], True
            finally:
                with self.lock:
                    self.building.pop(spec, None)
                ev.set()


class Server:
    def __init__(self, dkt, n_streams=4, hbm_budget_bytes=200e9, batch=0, cache=None, max_wait=0.5):
        import torch
        self.torch = torch
        self.dkt = dkt
        self.cache = cache if cache is not None else PlanCache(dkt, hbm_budget_bytes)
        self.q = queue.Queue()
        self.batch = batch
        self.max_wait = max_wait  # batched mode: seconds after which a waiting request goes first
        self.done = []
        self.done_lock = threading.Lock()
        self.workers = []
        work = self._work_batched if batch > 0 else self._work
        for i in range(n_streams):
            t = threading.Thread(target=work, args=(torch.cuda.Stream(),), daemon=True)
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
                da, (plan, mass), lk, built = self.cache.get(req.mesh)
                req.plan_built = built
                req.solve_started = time.perf_counter()
                f = self.dkt.keyed_uniform(req.rhs_seed, da.nodes())
                if req.kind == "poisson":
                    f[da.boundary()] = 0.0
                elif req.kind != "heat":
                    raise ValueError(f"unknown request kind {req.kind}")
                with torch.cuda.stream(stream):
                    fd = torch.tensor(f, device="cuda")
                    xd = torch.zeros_like(fd)
                    with lk:
                        if req.kind == "heat":
                            bd = torch.empty_like(fd)
                            mass.matvec_torch(fd, bd, masked=2, stream=stream.cuda_stream)
                        else:
                            bd = fd
                        req.stats = plan.cg_solve_torch(bd, xd, req.mode, req.max_iter, req.tol,
                                                        stream=stream.cuda_stream)
                    req.x = xd.cpu().numpy()
                stream.synchronize()
            except Exception as e:  # reported per request
                req.error = f"{type(e).__name__}: {e}"
            req.finished = time.perf_counter()
            with self.done_lock:
                self.done.append(req)

    def _rhs(self, req, da):
        f = self.dkt.keyed_uniform(req.rhs_seed, da.nodes())
        if req.kind == "poisson":
            f[da.boundary()] = 0.0
        elif req.kind != "heat":
            raise ValueError(f"unknown request kind {req.kind}")
        return f

    def _finish(self, req):
        req.finished = time.perf_counter()
        with self.done_lock:
            self.done.append(req)

    def _next(self, block):
        """A queued request (None: shut down; _EMPTY: none)."""
        try:
            return self.q.get(block=block, timeout=0.05 if block else None)
        except queue.Empty:
            return _EMPTY

    def _work_batched(self, stream):
        """Continuous batching: requests join the worker's CG engine as slots free
        up, smallest mesh first among the waiting ones, and leave when their
        solve stops, between chunks of 8 iterations."""
        torch = self.torch
        engine = self.dkt.CgEngine(self.batch, masked=True)
        running = {}  # id(req) -> (req, b, x)
        pending = []  # (nodes, arrival, req, da, plan, mass, lock): taken from the queue, not started
        stop = False
        while True:
            # Take queued requests (at most 2 * batch waiting here, so that the
            # other workers get theirs), then start the smallest meshes first.
            while not stop and len(pending) < 2 * self.batch:
                req = self._next(block=not running and not pending)
                if req is _EMPTY:
                    break
                if req is None:
                    stop = True
                    break
                if not req.started:
                    req.started = time.perf_counter()
                try:
                    da, (plan, mass), lk, built = self.cache.get(req.mesh)
                except Exception as e:  # reported per request
                    req.error = f"{type(e).__name__}: {e}"
                    self._finish(req)
                    continue
                req.plan_built = req.plan_built or built
                pending.append((da.n_nodes, req.submitted, req, da, plan, mass, lk))
            # requests waiting longer than max_wait go first, in arrival order (no starvation)
            now = time.perf_counter()
            pending.sort(key=lambda t: (0, t[1], 0) if now - t[1] > self.max_wait else (1, t[0], t[1]))
            while pending and len(running) < self.batch:
                _, _, req, da, plan, mass, lk = pending.pop(0)
                try:
                    with torch.cuda.stream(stream):
                        req.solve_started = time.perf_counter()
                        fd = torch.tensor(self._rhs(req, da), device="cuda")
                        if req.kind == "heat":
                            # the mass plan's matvec buffers are per plan: one user at a time
                            bd = torch.empty_like(fd)
                            with lk:
                                mass.matvec_torch(fd, bd, masked=2, stream=stream.cuda_stream)
                                stream.synchronize()
                        else:
                            bd = fd
                        xd = torch.empty_like(fd)
                        engine.add(plan, bd, xd, key=id(req), max_iter=req.max_iter, tol=req.tol)
                    running[id(req)] = (req, bd, xd)
                except Exception as e:
                    req.error = f"{type(e).__name__}: {e}"
                    self._finish(req)
            if not running:
                if stop and not pending:
                    return
                continue
            try:
                with torch.cuda.stream(stream):
                    finished = engine.step(1, stream=stream.cuda_stream)
                    out = []
                    for key, st in finished:
                        req, bd, xd = running.pop(key)
                        req.stats = dict(iters=st["iters"], resid=st["resid"], refinements=0, matvecs=st["iters"])
                        req.x = xd.cpu().numpy()
                        out.append(req)
                for req in out:
                    self._finish(req)
            except Exception as e:  # every running solve fails
                for req, *_ in running.values():
                    req.error = f"{type(e).__name__}: {e}"
                    self._finish(req)
                running.clear()
                engine = self.dkt.CgEngine(self.batch, masked=True)

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


def run_open_loop(server, reqs, rate, seed=0, timeout=900):
    """Submits reqs at Poisson arrival times of the given rate (requests/s) from a
    separate thread while the server works; returns (done requests, wall seconds
    from the first arrival to the last completion)."""
    rng = np.random.default_rng(seed)
    gaps = rng.exponential(1.0 / rate, len(reqs))
    n0 = len(server.done)
    t0 = time.perf_counter()

    def feed():
        t = t0
        for r, g in zip(reqs, gaps):
            t += g
            d = t - time.perf_counter()
            if d > 0:
                time.sleep(d)
            server.submit(r)
    th = threading.Thread(target=feed, daemon=True)
    th.start()
    done = server.drain(n0 + len(reqs), timeout=timeout)[n0:]
    th.join()
    return done, max(r.finished for r in done) - t0


def percentiles(values, ps=(50, 99)):
    v = np.asarray(values, dtype=np.float64)
    return {f"p{p}": float(np.percentile(v, p)) for p in ps}
