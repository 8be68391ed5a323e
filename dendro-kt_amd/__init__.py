"""Python host side of the MI355X adaptive k-tree FEM matvec.

A thin ctypes layer over the C ABI in include/dkt.h (libdkt.so, built in-tree by
build.py).  It mirrors the reference's host interfaces for this path:

* ``balance_tree`` / ``construct_tree`` / ``sort_tree``  -- SFC_Tree::distTree*
  (src/tsort.cpp:201-878)
* ``DA``        -- ot::DA<dim> (include/oda.h:40-525): CG nodes in the
  reference's order, boundary ids, and the flattened traversal
* ``Plan``      -- the device matvec, fem::matvec / feMatrix::matVec
  (FEM/include/matvec.h:231-548, FEM/include/feMatrix.h:189-259)

Device buffers are passed as raw pointers (e.g. torch tensors' ``data_ptr()``)
with a HIP stream handle; torch is plumbing here, never the compute path.  If
the library is missing, every call raises: there is no fallback.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIBNAME = "libdkt.so"

CURVE_MORTON = 0
CURVE_HILBERT = 1
OP_STIFFNESS = 0
OP_MASS = 1
OP_IDENTITY = 2
OP_CAPTURED = 3
PLAN_SIMPLE = 1
PLAN_DRYRUN = 2
PLAN_CONFORMING = 4
PLAN_TILE_SHIFT = 8
NOALLOC_SCRATCH_BYTES = 32 << 20  # DKT_NOALLOC_SCRATCH_BYTES
CG_F64 = 0
CG_MIXED = 1
CG_MIXED_RESTART = 2
CG_MIXED_RELIABLE = 3
BATCH_GRAPH = 1
BATCH_PERSISTENT = 2
BATCH_COLUMNS_SHIFT = 8  # DKT_BATCH_COLUMNS_SHIFT
BATCH_COLUMN_MAJOR = 4
ERR_BREAKDOWN = 2
STOP_NAMES = {1: "converged", 2: "max_iter", 3: "breakdown", 4: "stalled"}

_lib = None


class DktError(RuntimeError):
    pass


class CgBreakdown(DktError):
    """CG stopped because p.Ap <= 0 or a step was not finite (DKT_ERR_BREAKDOWN);
    .stats holds the solve statistics, x the last iterate."""

    def __init__(self, msg, stats):
        super().__init__(msg)
        self.stats = stats


def lib_path():
    return os.environ.get("DKT_LIB", os.path.join(HERE, LIBNAME))


def load(path=None):
    """Loads libdkt.so (once).  Raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    p = path or lib_path()
    if not os.path.exists(p):
        raise DktError(f"{p} not found: run __graft_entry__.build() (hipcc) first")
    # One HIP runtime per process: when torch is present its bundled
    # libamdhip64 (soname libamdhip64.so.7) must be loaded first so that
    # libdkt binds to the same instance; two runtimes cannot share the device,
    # and torch's stream/synchronize calls must see libdkt's kernels.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    L = ctypes.CDLL(p)
    c = ctypes
    P = c.c_void_p
    u32p = c.POINTER(c.c_uint32)
    L.dkt_last_error.restype = c.c_char_p
    L.dkt_octs_create.argtypes = [c.c_int, c.c_int, c.c_int, u32p, c.c_size_t, c.POINTER(P)]
    L.dkt_octs_destroy.argtypes = [P]
    L.dkt_octs_size.argtypes = [P]
    L.dkt_octs_size.restype = c.c_size_t
    L.dkt_octs_copy.argtypes = [P, u32p]
    L.dkt_tree_sort.argtypes = [P]
    L.dkt_tree_construct.argtypes = [P, c.c_size_t, c.POINTER(P)]
    L.dkt_tree_balance.argtypes = [P, c.c_size_t, c.POINTER(P)]
    L.dkt_tree_sphere.argtypes = [c.c_int, c.c_int, c.c_int, c.c_int, c.c_int, c.POINTER(c.c_double),
                                  c.c_double, c.POINTER(P)]
    L.dkt_points_gauss.argtypes = [c.c_int, c.c_int, c.c_int, c.c_size_t, c.c_int, c.c_double, c.c_uint64,
                                   c.POINTER(P)]
    L.dkt_keyed_uniform.argtypes = [c.c_uint64, c.c_int, u32p, c.c_size_t, c.POINTER(c.c_double)]
    L.dkt_da_create.argtypes = [P, c.c_int, c.POINTER(P)]
    L.dkt_da_destroy.argtypes = [P]
    L.dkt_da_count.argtypes = [P, c.c_int]
    L.dkt_da_count.restype = c.c_size_t
    L.dkt_da_nodes.argtypes = [P, u32p]
    L.dkt_da_elements.argtypes = [P, u32p]
    L.dkt_da_boundary.argtypes = [P, u32p]
    L.dkt_da_traversal.argtypes = [P, u32p, c.POINTER(c.c_uint8), c.POINTER(c.c_int32),
                                   c.POINTER(c.c_int32), c.POINTER(c.c_int32)]
    L.dkt_da_stale_terms.argtypes = [P, c.POINTER(c.c_int64), c.POINTER(c.c_uint8), c.POINTER(c.c_int32),
                                     c.POINTER(c.c_double)]
    L.dkt_da_write_vtu.argtypes = [P, c.c_char_p, P, c.c_char_p]
    L.dkt_da_conforming_terms.argtypes = [P, c.POINTER(c.c_int64), c.POINTER(c.c_uint8), c.POINTER(c.c_int32),
                                          c.POINTER(c.c_double)]
    _bind_device(L)
    _lib = L
    return L


def _bind_device(L):
    c = ctypes
    P = c.c_void_p
    if not hasattr(L, "dkt_plan_create"):
        return
    L.dkt_plan_create.argtypes = [P, c.c_int, P, c.c_int, c.POINTER(P)]
    L.dkt_plan_destroy.argtypes = [P]
    L.dkt_plan_matvec.argtypes = [P, c.c_int, P, P, c.c_int, c.c_int, P]
    L.dkt_plan_matvec_noalloc.argtypes = [P, c.c_int, P, P, c.c_int, c.c_int, P]
    L.dkt_plan_info.argtypes = [P, c.POINTER(c.c_double), c.c_int]
    L.dkt_plan_profile.argtypes = [P, c.c_int]
    L.dkt_plan_profile_read.argtypes = [P, c.POINTER(c.c_double), c.POINTER(c.c_double), c.POINTER(c.c_int)]
    L.dkt_cg_solve.argtypes = [P, c.c_int, P, P, c.c_int, c.c_double, c.c_int, c.POINTER(c.c_double), P]
    L.dkt_bicgstab_solve.argtypes = [P, P, P, c.c_int, c.c_double, c.c_int, c.POINTER(c.c_double), P]
    L.dkt_plan_set_dirichlet.argtypes = [P, c.POINTER(c.c_uint32), c.c_size_t]
    L.dkt_cg_solve_batch.argtypes = [c.POINTER(P), c.c_int, c.POINTER(P), c.POINTER(P), c.c_int, c.c_double, c.c_int,
                                     c.POINTER(c.c_double), P]
    L.dkt_cg_engine_create.argtypes = [c.c_int, c.c_int, c.POINTER(P)]
    L.dkt_cg_engine_add.argtypes = [P, P, P, P, c.c_int, c.c_double, c.POINTER(c.c_int)]
    L.dkt_cg_engine_add_mode.argtypes = [P, P, c.c_int, P, P, c.c_int, c.c_double, c.POINTER(c.c_int)]
    L.dkt_cg_engine_step.argtypes = [P, c.c_int, P, c.POINTER(c.c_int), c.POINTER(c.c_double), c.POINTER(c.c_int)]
    L.dkt_cg_engine_destroy.argtypes = [P]
    L.dkt_batch_create.argtypes = [c.POINTER(P), c.c_int, c.c_int, c.c_int, c.POINTER(P), c.POINTER(P), c.c_int,
                                   c.c_int, c.POINTER(P)]
    L.dkt_batch_run.argtypes = [P, P]
    L.dkt_batch_destroy.argtypes = [P]
    L.dkt_batch_status.argtypes = [P, c.POINTER(c.c_int)]


def _chk(rc):
    if rc != 0:
        raise DktError(load().dkt_last_error().decode())


def _u32(a):
    a = np.ascontiguousarray(a, dtype=np.uint32)
    return a, a.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32))


class Octs:
    """A host list of cells or points (uint32 [n, dim+1])."""

    def __init__(self, handle, dim):
        self.h = handle
        self.dim = dim

    @classmethod
    def from_array(cls, cells, dim, maxdepth, curve=CURVE_HILBERT):
        L = load()
        a, ptr = _u32(cells)
        h = ctypes.c_void_p()
        _chk(L.dkt_octs_create(dim, maxdepth, curve, ptr, a.shape[0], ctypes.byref(h)))
        return cls(h, dim)

    def __len__(self):
        return load().dkt_octs_size(self.h)

    def array(self):
        out = np.empty((len(self), self.dim + 1), dtype=np.uint32)
        _chk(load().dkt_octs_copy(self.h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32))))
        return out

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.dkt_octs_destroy(self.h)
            self.h = None


def sort_tree(octs):
    _chk(load().dkt_tree_sort(octs.h))
    return octs


def construct_tree(pts, max_pts=1):
    h = ctypes.c_void_p()
    _chk(load().dkt_tree_construct(pts.h, max_pts, ctypes.byref(h)))
    return Octs(h, pts.dim)


def balance_tree(pts, max_pts=1):
    h = ctypes.c_void_p()
    _chk(load().dkt_tree_balance(pts.h, max_pts, ctypes.byref(h)))
    return Octs(h, pts.dim)


def sphere_tree(dim, maxdepth, base_level, max_level, center, radius, curve=CURVE_HILBERT):
    c = (ctypes.c_double * dim)(*center)
    h = ctypes.c_void_p()
    _chk(load().dkt_tree_sphere(dim, maxdepth, curve, base_level, max_level, c, radius, ctypes.byref(h)))
    return Octs(h, dim)


def gauss_points(dim, maxdepth, n, ncomp, sigma, seed, curve=CURVE_HILBERT):
    h = ctypes.c_void_p()
    _chk(load().dkt_points_gauss(dim, maxdepth, curve, n, ncomp, sigma, seed, ctypes.byref(h)))
    return Octs(h, dim)


def keyed_uniform(seed, nodes):
    a, ptr = _u32(nodes)
    x = np.empty(a.shape[0], dtype=np.float64)
    _chk(load().dkt_keyed_uniform(seed, a.shape[1] - 1, ptr, a.shape[0],
                                  x.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
    return x


class DA:
    """Mesh over a balanced tree: ot::DA<dim>(tree, nEle, comm, order) on one rank."""

    def __init__(self, tree, order=1):
        self.dim = tree.dim
        self.order = order
        h = ctypes.c_void_p()
        _chk(load().dkt_da_create(tree.h, order, ctypes.byref(h)))
        self.h = h

    def count(self, what):
        return load().dkt_da_count(self.h, what)

    @property
    def n_elements(self):
        return self.count(0)

    @property
    def n_nodes(self):
        return self.count(1)

    @property
    def npe(self):
        return self.count(5)

    def nodes(self):
        out = np.empty((self.n_nodes, self.dim + 1), dtype=np.uint32)
        _chk(load().dkt_da_nodes(self.h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32))))
        return out

    def elements(self):
        out = np.empty((self.n_elements, self.dim + 1), dtype=np.uint32)
        _chk(load().dkt_da_elements(self.h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32))))
        return out

    def boundary(self):
        out = np.empty(self.count(2), dtype=np.uint32)
        _chk(load().dkt_da_boundary(self.h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32))))
        return out

    def traversal(self):
        L, H, n = self.count(3), self.count(4), self.npe
        cells = np.empty((L, self.dim + 1), dtype=np.uint32)
        child = np.empty(L, dtype=np.uint8)
        slot = np.empty((L, n), dtype=np.int32)
        hang = np.empty(L, dtype=np.int32)
        pslot = np.empty((H, n), dtype=np.int32)
        c = ctypes
        _chk(load().dkt_da_traversal(self.h, cells.ctypes.data_as(c.POINTER(c.c_uint32)),
                                     child.ctypes.data_as(c.POINTER(c.c_uint8)),
                                     slot.ctypes.data_as(c.POINTER(c.c_int32)),
                                     hang.ctypes.data_as(c.POINTER(c.c_int32)),
                                     pslot.ctypes.data_as(c.POINTER(c.c_int32))))
        nh = self.count(7)
        off = np.empty(H + 1, dtype=np.int64)
        hrank = np.empty(nh, dtype=np.uint8)
        hnode = np.empty(nh, dtype=np.int32)
        hw = np.empty(nh, dtype=np.float64)
        _chk(load().dkt_da_stale_terms(self.h, off.ctypes.data_as(c.POINTER(c.c_int64)),
                                       hrank.ctypes.data_as(c.POINTER(c.c_uint8)),
                                       hnode.ctypes.data_as(c.POINTER(c.c_int32)),
                                       hw.ctypes.data_as(c.POINTER(c.c_double))))
        return dict(cells=cells, child=child, slot=slot, hang_of=hang, pslot=pslot,
                    hc_off=off, hc_rank=hrank, hc_node=hnode, hc_w=hw)

    def write_vtu(self, path, values=None, name="u"):
        """ASCII .vtu of the mesh (dim 2 and 3): one cell per element, cell data
        "level", and the nodal field `values` (host, length n_nodes) at the
        corners (conforming interpolant at hanging corners)."""
        ptr = None
        if values is not None:
            values = np.ascontiguousarray(values, dtype=np.float64)
            if values.shape != (self.n_nodes,):
                raise DktError("values must have one entry per node")
            ptr = values.ctypes.data_as(ctypes.c_void_p)
        _chk(load().dkt_da_write_vtu(self.h, str(path).encode(), ptr, name.encode()))

    def conforming_terms(self):
        """Conforming constraints: dict(off, rank, node, w), one row of off per mesh element
        (in element order) that has ranks at points which are not nodes."""
        c = ctypes
        H, nt = self.count(9), self.count(8)
        off = np.empty(H + 1, dtype=np.int64)
        rank = np.empty(nt, dtype=np.uint8)
        node = np.empty(nt, dtype=np.int32)
        w = np.empty(nt, dtype=np.float64)
        _chk(load().dkt_da_conforming_terms(self.h, off.ctypes.data_as(c.POINTER(c.c_int64)),
                                            rank.ctypes.data_as(c.POINTER(c.c_uint8)),
                                            node.ctypes.data_as(c.POINTER(c.c_int32)),
                                            w.ctypes.data_as(c.POINTER(c.c_double))))
        return dict(off=off, rank=rank, node=node, w=w)

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.dkt_da_destroy(self.h)
            self.h = None


class Plan:
    """Device-resident matvec plan over a DA (fem::matvec / feMatrix::matVec).

    ``matvec(x_ptr, y_ptr, nrhs, dtype, masked, stream)`` takes raw device
    pointers; ``matvec_torch`` is a convenience for torch tensors (plumbing only).
    masked: 0 raw, 1 HeatMat's Dirichlet hooks (x and y), 2 HeatVec's (y only).
    """

    def __init__(self, da, op=OP_STIFFNESS, op_data=None, simple=False, tile_leaves=0, conforming=False,
                 dry_run=False):
        L = load()
        if not hasattr(L, "dkt_plan_create"):
            raise DktError("libdkt.so was built without the device plan")
        self.da = da
        self.n_nodes = da.n_nodes
        h = ctypes.c_void_p()
        ptr = None
        if op_data is not None:
            self._op_data = np.ascontiguousarray(op_data, dtype=np.float64)
            ptr = self._op_data.ctypes.data_as(ctypes.c_void_p)
        flags = (PLAN_SIMPLE if simple else 0) | (PLAN_CONFORMING if conforming else 0) | \
            (PLAN_DRYRUN if dry_run else 0)
        if tile_leaves:
            flags |= (int(tile_leaves).bit_length() - 1) << PLAN_TILE_SHIFT
        _chk(L.dkt_plan_create(da.h, op, ptr, flags, ctypes.byref(h)))
        self.h = h

    def info(self):
        keys = ["nodes", "leaves", "hanging", "npe", "map_bytes", "tiles", "shared_nodes", "partials",
                "max_tile_nodes", "readers_mean", "readers_max", "carry_terms_mean", "carry_terms_max",
                "groups_mean", "tile_nodes_mean", "hanging_entries", "cone_mean", "span_mean", "span_max",
                "tile_leaves", "max_tile_terms", "max_tile_hanging", "lds_bytes"]
        out = (ctypes.c_double * len(keys))()
        _chk(load().dkt_plan_info(self.h, out, len(keys)))
        return dict(zip(keys, list(out)))

    def profile(self, enable=True):
        """Records HIP events around the tile kernel of each subsequent matvec."""
        _chk(load().dkt_plan_profile(self.h, 1 if enable else 0))

    def profile_read(self):
        """(tile kernel ms summed, whole matvec ms summed, calls); clears the record."""
        t, tot, n = ctypes.c_double(), ctypes.c_double(), ctypes.c_int()
        _chk(load().dkt_plan_profile_read(self.h, ctypes.byref(t), ctypes.byref(tot), ctypes.byref(n)))
        return t.value, tot.value, n.value

    def matvec(self, x_ptr, y_ptr, nrhs=1, dtype=0, masked=False, stream=None):
        _chk(load().dkt_plan_matvec(self.h, dtype, ctypes.c_void_p(x_ptr), ctypes.c_void_p(y_ptr), nrhs,
                                    int(masked), ctypes.c_void_p(stream or 0)))

    def matvec_noalloc(self, x_ptr, y_ptr, nrhs=1, dtype=0, masked=False, stream=None):
        """As matvec, with the scratch in the library's fixed device array
        (dkt_plan_matvec_noalloc): no device memory is allocated."""
        _chk(load().dkt_plan_matvec_noalloc(self.h, dtype, ctypes.c_void_p(x_ptr), ctypes.c_void_p(y_ptr), nrhs,
                                            int(masked), ctypes.c_void_p(stream or 0)))

    def matvec_torch(self, x, y, masked=False, stream=None, noalloc=False):
        import torch
        if x.dtype != y.dtype or x.dtype not in (torch.float64, torch.float32):
            raise DktError("x and y must both be float64 or float32")
        if not (x.is_cuda and y.is_cuda and x.is_contiguous() and y.is_contiguous()):
            raise DktError("x and y must be contiguous device tensors")
        nrhs = 1 if x.dim() == 1 else x.shape[1]
        if x.shape[0] != self.n_nodes or y.shape != x.shape:
            raise DktError("x/y shape does not match the plan")
        s = stream if stream is not None else torch.cuda.current_stream().cuda_stream
        fn = self.matvec_noalloc if noalloc else self.matvec
        fn(x.data_ptr(), y.data_ptr(), nrhs, 0 if x.dtype == torch.float64 else 1, masked, s)

    def cg_solve(self, b_ptr, x_ptr, mode=CG_F64, max_iter=10000, tol=1e-8, masked=True, stream=None,
                 raise_on_breakdown=True):
        """HeatMat::cgSolve on the device (x: initial guess in, solution out).  Returns
        dict(iters, resid, refinements, matvecs, stop) with stop one of
        "converged", "max_iter", "breakdown", "stalled".  A breakdown raises
        CgBreakdown (carrying the statistics) unless raise_on_breakdown is False."""
        st = (ctypes.c_double * 5)()
        rc = load().dkt_cg_solve(self.h, mode, ctypes.c_void_p(b_ptr), ctypes.c_void_p(x_ptr), max_iter, tol,
                                 1 if masked else 0, st, ctypes.c_void_p(stream or 0))
        if rc not in (0, ERR_BREAKDOWN):
            _chk(rc)
        out = dict(iters=int(st[0]), resid=st[1], refinements=int(st[2]), matvecs=int(st[3]),
                   stop=STOP_NAMES.get(int(st[4]), "unknown"))
        if rc == ERR_BREAKDOWN and raise_on_breakdown:
            raise CgBreakdown(load().dkt_last_error().decode(), out)
        return out

    def cg_solve_torch(self, b, x, mode=CG_F64, max_iter=10000, tol=1e-8, masked=True, stream=None,
                       raise_on_breakdown=True):
        import torch
        if b.dtype != torch.float64 or x.dtype != torch.float64 or b.shape != (self.n_nodes,) or x.shape != b.shape:
            raise DktError("b and x must be float64 device vectors of the plan's node count")
        if not (b.is_cuda and x.is_cuda and b.is_contiguous() and x.is_contiguous()):
            raise DktError("b and x must be contiguous device tensors")
        s = stream if stream is not None else torch.cuda.current_stream().cuda_stream
        return self.cg_solve(b.data_ptr(), x.data_ptr(), mode, max_iter, tol, masked, s, raise_on_breakdown)

    def bicgstab_solve_torch(self, b, x, max_iter=10000, tol=1e-8, masked=True, stream=None,
                             raise_on_breakdown=True):
        """BiCGStab on the plan's operator (nonsymmetric operators, e.g. the
        space-time heat form); x: initial guess in, solution out.  Returns the
        same statistics as cg_solve."""
        import torch
        if b.dtype != torch.float64 or x.dtype != torch.float64 or b.shape != (self.n_nodes,) or x.shape != b.shape:
            raise DktError("b and x must be float64 device vectors of the plan's node count")
        if not (b.is_cuda and x.is_cuda and b.is_contiguous() and x.is_contiguous()):
            raise DktError("b and x must be contiguous device tensors")
        s = stream if stream is not None else torch.cuda.current_stream().cuda_stream
        st = (ctypes.c_double * 5)()
        rc = load().dkt_bicgstab_solve(self.h, ctypes.c_void_p(b.data_ptr()), ctypes.c_void_p(x.data_ptr()),
                                       max_iter, tol, 1 if masked else 0, st, ctypes.c_void_p(s or 0))
        if rc not in (0, ERR_BREAKDOWN):
            _chk(rc)
        out = dict(iters=int(st[0]), resid=st[1], refinements=0, matvecs=int(st[3]),
                   stop=STOP_NAMES.get(int(st[4]), "unknown"))
        if rc == ERR_BREAKDOWN and raise_on_breakdown:
            raise CgBreakdown(load().dkt_last_error().decode(), out)
        return out

    def set_dirichlet(self, ids):
        """Replaces the Dirichlet node set of the masked operators."""
        ids = np.ascontiguousarray(ids, dtype=np.uint32)
        _chk(load().dkt_plan_set_dirichlet(self.h, ids.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)), len(ids)))

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.dkt_plan_destroy(self.h)
            self.h = None


class CgEngine:
    """Continuous batching of CG solves (dkt_cg_engine): add() starts a solve
    (mode CG_F64 or CG_MIXED / CG_MIXED_RELIABLE) at the next step, step()
    runs chunks of 8 iterations of every running solve and returns the ones
    that stopped as [(key, dict(iters, resid, stop))].  Each solve is bitwise
    Plan.cg_solve in its mode from x = 0; x is overwritten."""

    def __init__(self, max_slots, masked=True):
        L = load()
        self.max_slots = max_slots
        h = ctypes.c_void_p()
        _chk(L.dkt_cg_engine_create(max_slots, 1 if masked else 0, ctypes.byref(h)))
        self.h = h
        self.running = {}  # slot -> (key, plan, b, x): kept alive while the solve runs
        self._fin = (ctypes.c_int * max_slots)()
        self._st = (ctypes.c_double * (3 * max_slots))()

    def add(self, plan, b, x, key=None, max_iter=10000, tol=1e-8, mode=CG_F64):
        import torch
        if b.dtype != torch.float64 or x.dtype != torch.float64 or b.shape != (plan.n_nodes,) or x.shape != b.shape:
            raise DktError("b and x must be float64 device vectors of the plan's node count")
        if not (b.is_cuda and x.is_cuda and b.is_contiguous() and x.is_contiguous()):
            raise DktError("b and x must be contiguous device tensors")
        slot = ctypes.c_int(-1)
        _chk(load().dkt_cg_engine_add_mode(self.h, plan.h.value, mode, ctypes.c_void_p(b.data_ptr()),
                                           ctypes.c_void_p(x.data_ptr()), max_iter, tol, ctypes.byref(slot)))
        self.running[slot.value] = (key, plan, b, x)
        return slot.value

    def step(self, chunks=1, stream=None):
        import torch
        s = stream if stream is not None else torch.cuda.current_stream().cuda_stream
        nf = ctypes.c_int(0)
        _chk(load().dkt_cg_engine_step(self.h, chunks, ctypes.c_void_p(s or 0), self._fin, self._st, ctypes.byref(nf)))
        out = []
        for i in range(nf.value):
            key = self.running.pop(self._fin[i])[0]
            out.append((key, dict(iters=int(self._st[3 * i]), resid=self._st[3 * i + 1],
                                  stop=STOP_NAMES.get(int(self._st[3 * i + 2]), "unknown"))))
        return out

    def __len__(self):
        return len(self.running)

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.dkt_cg_engine_destroy(self.h)
            self.h = None


def cg_solve_batch_torch(plans, bs, xs, max_iter=10000, tol=1e-8, masked=True, stream=None):
    """FP64 CG on every (plans[k], bs[k]) in lockstep (dkt_cg_solve_batch): xs[k]
    is overwritten with the solution from x = 0, bitwise what Plan.cg_solve
    returns from a zero guess.  Returns [dict(iters, resid)] per solve."""
    import torch
    n = len(plans)
    if n < 1 or len(bs) != n or len(xs) != n:
        raise DktError("plans, bs and xs must be non-empty lists of one length")
    for pl, b, x in zip(plans, bs, xs):
        if b.dtype != torch.float64 or x.dtype != torch.float64 or b.shape != (pl.n_nodes,) or x.shape != b.shape:
            raise DktError("b and x must be float64 device vectors of the plan's node count")
        if not (b.is_cuda and x.is_cuda and b.is_contiguous() and x.is_contiguous()):
            raise DktError("b and x must be contiguous device tensors")
    V = ctypes.c_void_p
    hp = (V * n)(*[pl.h.value for pl in plans])
    hb = (V * n)(*[b.data_ptr() for b in bs])
    hx = (V * n)(*[x.data_ptr() for x in xs])
    st = (ctypes.c_double * (3 * n))()
    s = stream if stream is not None else torch.cuda.current_stream().cuda_stream
    _chk(load().dkt_cg_solve_batch(hp, n, hb, hx, max_iter, tol, 1 if masked else 0, st, V(s or 0)))
    return [dict(iters=int(st[3 * k]), resid=st[3 * k + 1], stop=STOP_NAMES.get(int(st[3 * k + 2]), "unknown"))
            for k in range(n)]


class Batch:
    """Batched matvec over many plans: one launch per (dim, p) (dkt_batch_*).

    ``xs``/``ys`` are torch device tensors [N_k] or [N_k, nrhs] (one dtype);
    with ``graph=True`` the launches are captured into a hipGraph on the first
    run (needs a non-default stream) and replayed afterwards.  ``columns``
    caps the right-hand sides per tile record (a power of two; 1 runs the
    single-column kernels, a record per column; 0 = as many as fit in LDS);
    every record of a (dim, p) group runs in the same launch.  With
    ``column_major=True`` the vectors are [nrhs, N_k], one contiguous column
    per record.
    """

    def __init__(self, plans, xs, ys, masked=False, graph=False, persistent=False, columns=0, column_major=False):
        import torch
        L = load()
        n = len(plans)
        if not (n == len(xs) == len(ys)) or n == 0:
            raise DktError("plans, xs and ys must have the same nonzero length")
        dt = xs[0].dtype
        if dt not in (torch.float64, torch.float32):
            raise DktError("x must be float64 or float32")
        if column_major:  # [nrhs, N_k]
            if xs[0].dim() != 2:
                raise DktError("column-major vectors are [nrhs, N] tensors")
            nrhs = xs[0].shape[0]
        else:
            nrhs = 1 if xs[0].dim() == 1 else xs[0].shape[1]
        for p, x, y in zip(plans, xs, ys):
            n_k = x.shape[1] if column_major else x.shape[0]
            if x.dtype != dt or y.dtype != dt or x.shape != y.shape or n_k != p.n_nodes:
                raise DktError("vector shape or dtype does not match")
            cols = x.shape[0] if column_major else (1 if x.dim() == 1 else x.shape[1])
            if cols != nrhs or not (x.is_contiguous() and y.is_contiguous()):
                raise DktError("all vectors need the same nrhs and contiguous storage")
        P = ctypes.c_void_p
        self._keep = (list(plans), list(xs), list(ys))
        ph = (P * n)(*[p.h.value for p in plans])
        xp = (P * n)(*[x.data_ptr() for x in xs])
        yp = (P * n)(*[y.data_ptr() for y in ys])
        h = P()
        flags = (BATCH_GRAPH if graph else 0) | (BATCH_PERSISTENT if persistent else 0)
        if column_major:
            flags |= BATCH_COLUMN_MAJOR
        if columns:
            if columns & (columns - 1) or not 1 <= columns <= 64:
                raise DktError("columns per launch must be a power of two in [1, 64]")
            flags |= (columns.bit_length()) << BATCH_COLUMNS_SHIFT
        _chk(L.dkt_batch_create(ph, n, 0 if dt == torch.float64 else 1, nrhs, xp, yp, int(masked), flags,
                                ctypes.byref(h)))
        self.h = h

    def run(self, stream=None):
        import torch
        s = stream if stream is not None else torch.cuda.current_stream().cuda_stream
        _chk(load().dkt_batch_run(self.h, ctypes.c_void_p(s)))

    def status(self):
        """1 when a persistent launch's fix-up gave up waiting (results invalid), else 0."""
        err = ctypes.c_int(0)
        _chk(load().dkt_batch_status(self.h, ctypes.byref(err)))
        return err.value

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.dkt_batch_destroy(self.h)
            self.h = None
