"""Request serving for the device CG (no counterpart in the reference, whose
solves are single-shot calls of HeatMat::cgSolve, FEM/examples/src/heatMat.cpp:165).

A request names a mesh (a seeded Gaussian point cloud, or a tree refined around
a sphere, that the host refines, balances and enumerates, as ot::DA(tree, ...)
does), a kind, a precision mode and a right-hand side; the server answers with
x = A^-1 b to a relative L_inf residual.  Kinds:
* "poisson": b is given per node (zero on the boundary), A the
  Dirichlet-masked stiffness (HeatMat::cgSolve);
* "heat": the reference's heat-equation driver (FEM/examples/src/heatEq.cpp):
  b = M f by the mass operator with HeatVec's hooks (feVector::computeVec,
  FEM/include/feVector.h:122-168), then the same CG on the stiffness;
* "spacetime" (4D meshes: x, y, z and time): the stabilised space-time heat
  form of the SC19 paper (spacetime.py): b = l(f) for f keyed uniform, the
  essential boundary on the spatial boundary and at t = 0, solved by BiCGStab
  in FP64 (the form is not symmetric); these requests run on direct workers
  beside the CG engines.

Operator: by default the conforming operator sum_e C_e^T K_e C_e (symmetric,
so CG converges at every size); Server(operator="reference") serves the
reference traversal's operator (fem::matvec semantics, not symmetric at
hanging nodes: FP64 CG can stall on it at scale).

Pipeline:
* intake: builder threads take submitted requests, lease their mesh's plans
  from the plan cache (building them on a miss), form the right-hand side on
  the host and put the request on one ready list shared by all workers;
* admission (PlanCache): a mesh's device footprint is known before anything is
  allocated on the device (DKT_PLAN_DRYRUN plans give the map bytes), and
  every lease counts against the HBM budget until its request finishes: the
  maps of the leased plans plus the request's vectors.  Unleased plans are
  evicted least recently used first; a request that does not fit waits for
  leases to end (and is refused when it could never fit);
* workers: one thread per HIP stream.  Batched (Server(batch=K)): each worker
  runs a CG engine (dkt_cg_engine) of K slots, FP64 and mixed-precision
  solves side by side; a ready request joins as soon as a slot is free and
  leaves when its solve stops, between chunks of 8 iterations (continuous
  batching).  Unbatched: one solve at a time per stream;
* order: select_next picks the request a free slot takes: requests that have
  waited longer than max_wait first, oldest first; otherwise the smallest
  mesh, then the oldest;
* every reduction on the device runs in a fixed order, so a request's answer
  does not depend on what else is being served ("alone == in-mix").

Latencies are measured from submission to completion of each request
(end to end) and from the start of its solve to completion.
"""
import collections
import queue
import threading
import time
from dataclasses import dataclass, field

import numpy as np

CG_F64, CG_MIXED, CG_MIXED_RESTART, CG_MIXED_RELIABLE = 0, 1, 2, 3
KINDS = ("poisson", "heat", "spacetime")


def _spacetime():
    """The package's spacetime module (this file is loaded by path, as the package is)."""
    import importlib.util
    import os
    import sys
    if "dkt_spacetime" not in sys.modules:
        spec = importlib.util.spec_from_file_location(
            "dkt_spacetime", os.path.join(os.path.dirname(os.path.abspath(__file__)), "spacetime.py"))
        m = importlib.util.module_from_spec(spec)
        sys.modules["dkt_spacetime"] = m
        spec.loader.exec_module(m)
    return sys.modules["dkt_spacetime"]


# Device bytes per node of one request: f, b, x (FP64), the CG workspace
# (r, p, Ap in FP64; r, p, Ap, x in FP32 for the mixed modes) and the mixed
# modes' FP64 residual.
VEC_BYTES_PER_NODE = 8 * 7 + 4 * 5


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
    mode: int = CG_F64
    max_iter: int = 100000
    kind: str = "poisson"
    # filled in by the server
    submitted: float = 0.0
    started: float = 0.0
    solve_started: float = 0.0
    finished: float = 0.0
    plan_built: bool = False
    nodes: int = 0
    stats: dict = field(default_factory=dict)
    x: object = None
    error: str = ""


def select_next(ready, now, max_wait):
    """Index in `ready` (items with .req.submitted and .req.nodes) of the
    request a free slot takes next: among requests that have waited longer
    than max_wait, the oldest; otherwise the smallest mesh, then the oldest.
    None when `ready` is empty."""
    if not ready:
        return None
    aged = [i for i, it in enumerate(ready) if now - it.req.submitted > max_wait]
    if aged:
        return min(aged, key=lambda i: (ready[i].req.submitted, i))
    return min(range(len(ready)), key=lambda i: (ready[i].req.nodes, ready[i].req.submitted, i))


class AdmissionError(MemoryError):
    pass


class Lease:
    """A request's hold on a cached mesh: its DA and plans, and the bytes it
    counts against the budget until released."""

    def __init__(self, cache, key, entry, vec_bytes):
        self.cache, self.key, self.entry, self.vec_bytes = cache, key, entry, vec_bytes
        self.da, self.plans, self.lock = entry["da"], entry["plans"], entry["lock"]
        self.released = False

    def release(self):
        if not self.released:
            self.released = True
            self.cache._release(self)


class PlanCache:
    """Mesh -> (DA, stiffness plan, mass plan) under an HBM budget.

    acquire(spec) returns a Lease.  A miss builds the DA on the host and the
    plans first as dry runs (host tiling only, no device memory) to learn the
    map bytes; only when those plus the request's vectors fit the budget,
    after evicting unleased entries least recently used first, are the device
    plans built.  A request that fits the budget but not the memory left by
    current leases waits for them to end (at most wait_timeout seconds); one
    larger than the whole budget is refused with AdmissionError.

    `build(key)` may be given to replace the construction (the key is then the
    spec itself): it returns (da, map_bytes, make_plans) where make_plans()
    builds the device plans."""

    def __init__(self, dkt, budget_bytes, operator="conforming", build=None, wait_timeout=120.0):
        if operator not in ("conforming", "reference"):
            raise ValueError(f"unknown operator {operator}")
        self.dkt = dkt
        self.budget = float(budget_bytes)
        self.operator = operator
        self.build = build or self._build
        self.wait_timeout = wait_timeout
        self.cv = threading.Condition()
        self.entries = collections.OrderedDict()  # spec -> dict(da, plans, map_bytes, refs, lock)
        self.building = set()
        self.leased_vec_bytes = 0.0

    def _build(self, key):
        spec, kind = key
        dkt = self.dkt
        conf = self.operator == "conforming"
        da = dkt.DA(spec.tree(dkt), spec.order)
        if kind == "spacetime":
            st = _spacetime()
            if spec.dim != 4 or spec.order != 1:
                raise ValueError("space-time requests need a 4D p = 1 mesh")
            mats = (st.level_matrices(spec.maxdepth), st.rhs_matrices(spec.maxdepth))
            # Both plans tile the same mesh the same way (only the element
            # operator differs), so one dry run gives the maps of each.
            map_bytes = len(mats) * dkt.Plan(da, dkt.OP_CAPTURED, op_data=mats[0], conforming=conf,
                                             dry_run=True).info()["map_bytes"]

            def make():
                plans = tuple(dkt.Plan(da, dkt.OP_CAPTURED, op_data=m, conforming=conf) for m in mats)
                ids = st.dirichlet_nodes(da.nodes(), spec.maxdepth)
                for p in plans:
                    p.set_dirichlet(ids)
                return plans
            return da, map_bytes, make
        # The stiffness and mass plans share the traversal and the tiling: equal maps.
        map_bytes = 2 * dkt.Plan(da, dkt.OP_STIFFNESS, conforming=conf, dry_run=True).info()["map_bytes"]
        return da, map_bytes, lambda: (dkt.Plan(da, dkt.OP_STIFFNESS, conforming=conf),
                                       dkt.Plan(da, dkt.OP_MASS, conforming=conf))

    def used(self):
        """Bytes counted against the budget: every cached entry's maps plus the
        vectors of every live lease."""
        with self.cv:
            return sum(e["map_bytes"] for e in self.entries.values()) + self.leased_vec_bytes

    def _make_room(self, need, deadline):
        """With self.cv held: evicts unleased entries (LRU first) and waits for
        leases to end until `need` more bytes fit."""
        while True:
            used = sum(e["map_bytes"] for e in self.entries.values()) + self.leased_vec_bytes
            if used + need <= self.budget:
                return
            victim = next((k for k, e in self.entries.items() if e["refs"] == 0), None)
            if victim is not None:
                del self.entries[victim]
                continue
            left = deadline - time.monotonic()
            if left <= 0:
                raise AdmissionError(f"admission timed out: {need / 1e9:.2f} GB needed, "
                                     f"{used / 1e9:.2f} of {self.budget / 1e9:.2f} GB leased")
            self.cv.wait(min(left, 0.5))

    def acquire(self, spec, kind="poisson"):
        """(Lease, built_now) for a mesh and a request kind (Poisson and heat
        share the stiffness and mass plans; space-time has its own).  Builds on a
        miss; concurrent misses on one key wait for a single build."""
        deadline = time.monotonic() + self.wait_timeout
        spec = (spec, "spacetime" if kind == "spacetime" else "cg") if self.build == self._build else spec
        with self.cv:
            while spec in self.building:
                self.cv.wait(0.5)
            if spec in self.entries:
                e = self.entries[spec]
                vec = VEC_BYTES_PER_NODE * e["da"].n_nodes
                if e["map_bytes"] + vec > self.budget:
                    raise AdmissionError(f"mesh needs {(e['map_bytes'] + vec) / 1e9:.2f} GB, "
                                         f"budget {self.budget / 1e9:.2f} GB")
                e["refs"] += 1  # held while waiting: not evicted under us
                try:
                    self._make_room(vec, deadline)
                except Exception:
                    e["refs"] -= 1
                    raise
                self.entries.move_to_end(spec)
                self.leased_vec_bytes += vec
                return Lease(self, spec, e, vec), False
            self.building.add(spec)
        try:
            da, map_bytes, make_plans = self.build(spec)  # host only: no device memory yet
            vec = VEC_BYTES_PER_NODE * da.n_nodes
            if map_bytes + vec > self.budget:
                raise AdmissionError(f"mesh needs {(map_bytes + vec) / 1e9:.2f} GB, "
                                     f"budget {self.budget / 1e9:.2f} GB")
            with self.cv:
                self._make_room(map_bytes + vec, deadline)
                # reserve the bytes before the device allocation
                e = dict(da=da, plans=None, map_bytes=map_bytes, refs=1, lock=threading.Lock())
                self.entries[spec] = e
                self.leased_vec_bytes += vec
            try:
                e["plans"] = make_plans()
            except Exception:
                with self.cv:
                    self.entries.pop(spec, None)
                    self.leased_vec_bytes -= vec
                    self.cv.notify_all()
                raise
            return Lease(self, spec, e, vec), True
        finally:
            with self.cv:
                self.building.discard(spec)
                self.cv.notify_all()

    def _release(self, lease):
        with self.cv:
            lease.entry["refs"] -= 1
            self.leased_vec_bytes -= lease.vec_bytes
            self.cv.notify_all()


@dataclass
class Ready:
    req: Request
    lease: Lease
    f: np.ndarray


class Server:
    def __init__(self, dkt, n_streams=4, hbm_budget_bytes=200e9, batch=0, cache=None, max_wait=0.5,
                 operator="conforming", n_builders=2, n_direct=1):
        import torch
        self.torch = torch
        self.dkt = dkt
        self.cache = cache if cache is not None else PlanCache(dkt, hbm_budget_bytes, operator)
        self.intake = queue.Queue()
        self.ready = []  # CG requests (the engines' in batched mode)
        self.ready_direct = []  # batched mode: space-time requests, for the direct workers
        self.ready_cv = threading.Condition()
        self.batch = batch
        self.max_wait = max_wait
        self.closing = False
        self.done = []
        self.done_lock = threading.Lock()
        self.builders = [threading.Thread(target=self._build_loop, daemon=True) for _ in range(n_builders)]
        for t in self.builders:
            t.start()
        work = self._work_batched if batch > 0 else self._work
        self.workers = [threading.Thread(target=work, args=(torch.cuda.Stream(),), daemon=True)
                        for _ in range(n_streams)]
        if batch > 0:  # direct workers for the requests the engines do not run
            self.workers += [threading.Thread(target=self._work, args=(torch.cuda.Stream(), True), daemon=True)
                             for _ in range(n_direct)]
        for t in self.workers:
            t.start()

    def submit(self, req):
        req.submitted = time.perf_counter()
        self.intake.put(req)

    # -- intake ---------------------------------------------------------------
    def _rhs(self, req, da):
        f = self.dkt.keyed_uniform(req.rhs_seed, da.nodes())
        if req.kind == "poisson":
            f[da.boundary()] = 0.0
        elif req.kind not in KINDS:
            raise ValueError(f"unknown request kind {req.kind}")
        return f

    def _build_loop(self):
        while True:
            req = self.intake.get()
            if req is None:
                return
            req.started = time.perf_counter()
            lease = None
            try:
                if req.kind not in KINDS:
                    raise ValueError(f"unknown request kind {req.kind}")
                if req.kind == "spacetime" and req.mode != CG_F64:
                    raise ValueError("space-time requests are solved in FP64 (BiCGStab): mode must be CG_F64")
                if self.batch > 0 and req.mode not in (CG_F64, CG_MIXED, CG_MIXED_RELIABLE):
                    raise ValueError(f"batched serving runs modes CG_F64 and CG_MIXED(_RELIABLE), not {req.mode}")
                lease, built = self.cache.acquire(req.mesh, req.kind)
                req.plan_built = built
                req.nodes = lease.da.n_nodes
                item = Ready(req, lease, self._rhs(req, lease.da))
            except Exception as e:  # reported per request
                if lease is not None:
                    lease.release()
                req.error = f"{type(e).__name__}: {e}"
                self._finish(req)
                continue
            with self.ready_cv:
                direct = self.batch > 0 and req.kind == "spacetime"
                (self.ready_direct if direct else self.ready).append(item)
                self.ready_cv.notify_all()

    def _take(self, block, direct=False):
        """The next ready request by select_next (None: none, or shutting down)."""
        lst = self.ready_direct if direct else self.ready
        with self.ready_cv:
            while True:
                i = select_next(lst, time.perf_counter(), self.max_wait)
                if i is not None:
                    return lst.pop(i)
                if not block or self.closing:
                    return None
                self.ready_cv.wait(0.05)

    def _finish(self, req, lease=None):
        if lease is not None:
            lease.release()
        req.finished = time.perf_counter()
        with self.done_lock:
            self.done.append(req)

    def _mass_rhs(self, item, fd, stream):
        """b = M f with HeatVec's hooks for heat requests (b = l(f) with the
        space-time form's essential boundary for space-time ones); b = f
        otherwise."""
        if item.req.kind not in ("heat", "spacetime"):
            return fd
        bd = self.torch.empty_like(fd)
        _, mass = item.lease.plans
        with item.lease.lock:  # a plan's matvec scratch serves one caller at a time
            mass.matvec_torch(fd, bd, masked=2, stream=stream.cuda_stream)
            stream.synchronize()
        return bd

    # -- workers --------------------------------------------------------------
    def _work(self, stream, direct=False):
        torch = self.torch
        while True:
            item = self._take(block=True, direct=direct)
            if item is None:
                if self.closing:
                    return
                continue
            req = item.req
            try:
                req.solve_started = time.perf_counter()
                with torch.cuda.stream(stream):
                    fd = torch.tensor(item.f, device="cuda")
                    bd = self._mass_rhs(item, fd, stream)
                    xd = torch.zeros_like(fd)
                    plan, _ = item.lease.plans
                    with item.lease.lock:  # a plan owns one solver workspace
                        if req.kind == "spacetime":
                            st = plan.bicgstab_solve_torch(bd, xd, req.max_iter, req.tol, stream=stream.cuda_stream)
                        else:
                            st = plan.cg_solve_torch(bd, xd, req.mode, req.max_iter, req.tol,
                                                     stream=stream.cuda_stream)
                    req.stats = dict(st, converged=st["stop"] == "converged")
                    req.x = xd.cpu().numpy()
                stream.synchronize()
            except Exception as e:  # reported per request
                req.error = f"{type(e).__name__}: {e}"
            self._finish(req, item.lease)

    def _work_batched(self, stream):
        """Continuous batching: ready requests join the worker's CG engine as
        slots free up and leave when their solve stops."""
        torch = self.torch
        engine = self.dkt.CgEngine(self.batch, masked=True)
        running = {}  # id(req) -> (item, b, x)
        while True:
            while len(running) < self.batch:
                item = self._take(block=not running)
                if item is None:
                    break
                req = item.req
                try:
                    with torch.cuda.stream(stream):
                        req.solve_started = time.perf_counter()
                        fd = torch.tensor(item.f, device="cuda")
                        bd = self._mass_rhs(item, fd, stream)
                        xd = torch.empty_like(fd)
                        engine.add(item.lease.plans[0], bd, xd, key=id(req), max_iter=req.max_iter, tol=req.tol,
                                   mode=req.mode)
                    running[id(req)] = (item, bd, xd)
                except Exception as e:
                    req.error = f"{type(e).__name__}: {e}"
                    self._finish(req, item.lease)
            if not running:
                if self.closing:
                    return
                continue
            try:
                with torch.cuda.stream(stream):
                    finished = engine.step(1, stream=stream.cuda_stream)
                    out = []
                    for key, st in finished:
                        item, bd, xd = running.pop(key)
                        req = item.req
                        req.stats = dict(iters=st["iters"], resid=st["resid"], stop=st["stop"],
                                         converged=st["stop"] == "converged", refinements=0, matvecs=st["iters"])
                        if st["stop"] == "breakdown":
                            req.error = "CgBreakdown: p.Ap <= 0 or a non-finite step"
                        req.x = xd.cpu().numpy()
                        out.append(item)
                for item in out:
                    self._finish(item.req, item.lease)
            except Exception as e:  # every running solve fails
                for item, *_ in running.values():
                    item.req.error = f"{type(e).__name__}: {e}"
                    self._finish(item.req, item.lease)
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
        for _ in self.builders:
            self.intake.put(None)
        for t in self.builders:
            t.join()
        with self.ready_cv:
            self.closing = True
            self.ready_cv.notify_all()
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
