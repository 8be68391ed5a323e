"""The oracle (CPU) against the reference's golden outputs and the mathematics.

The oracle is test infrastructure: oracle/matvec_oracle.c restates the
reference traversal (FEM/include/matvec.h:231-548) as a recursive C program,
and tests/plan_numpy.py interprets the product's flattened traversal in numpy.
Both must reproduce the reference's y = A x (oracle/ref_driver.cpp over the
reference sources, committed as tests/golden/*.npz) to 1e-13 relative L2, and
satisfy what the operator's mathematics fixes independently of the reference:
  * partition of unity of the hanging interpolation: A 1 = 0 at interior nodes
    for the stiffness operator (K_ref annihilates constants);
  * the identity elemental operator with x = 1 counts the elements sharing each
    node: 2^(number of axes along which the node is interior) on uniform grids
    (test/testMatvec.cpp:253-264);
  * mass operator on uniform grids: 1^T M 1 = measure of the unit cube = 1;
  * the element matrices: Q1 stiffness rows sum to zero, symmetric, scaled by
    h^(dim-2); mass entries sum to h^dim.
"""
import ctypes
import os
import subprocess

import numpy as np
import pytest

import fixtures as fx
import gen_inputs as gi
import plan_numpy as pn

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(REPO, "oracle", "_build", "liboracle.so")


@pytest.fixture(scope="session")
def oracle():
    if not os.path.exists(ORACLE_SO):
        subprocess.run(["make", "-C", os.path.join(REPO, "oracle"), "restatement"], check=True,
                       capture_output=True)
    O = ctypes.CDLL(ORACLE_SO)
    O.oracle_matvec.argtypes = [ctypes.c_int] * 4 + [ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int,
                                                     ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int]
    return O


def _oracle_y(O, meta, nodes, op, x, masked):
    O.oracle_reset()
    y = np.zeros_like(x)
    cv = 1 if meta["curve"] == "hilbert" else 0
    opc = {"poisson": 0, "heat": 0, "mass": 1, "ident": 2}[op]
    nodes = np.ascontiguousarray(nodes)
    rc = O.oracle_matvec(meta["dim"], meta["maxdepth"], meta["order"], cv, nodes.ctypes.data, len(nodes), opc,
                         x.ctypes.data, y.ctypes.data, int(masked))
    assert rc == 0
    return y


def _cases():
    out = []
    for name in fx.names():
        z, _ = fx.load(name)
        for key, op, xs, masked in fx.y_cases(z):
            out.append((name, key))
    return out


@pytest.mark.parametrize("name,key", _cases())
def test_oracle_restatement_matches_reference(dkt, oracle, name, key):
    da, z, meta = fx.build_da(dkt, name)
    nodes = da.nodes()
    _, op, xs, masked = key.split("_")
    x = fx.make_x(xs, nodes, meta["dim"])
    y = _oracle_y(oracle, meta, nodes, op, x, masked == "masked")
    ref = z[key]
    assert np.linalg.norm(y - ref) <= 1e-13 * max(np.linalg.norm(ref), 1e-300)


@pytest.mark.parametrize("name,key", [c for c in _cases() if c[0] != "sphere3d_l7"])
def test_flattened_traversal_matches_reference(dkt, name, key):
    """The product's host traversal (slots, parent slots, stale-buffer terms) interpreted in numpy."""
    da, z, meta = fx.build_da(dkt, name)
    nodes = da.nodes()
    _, op, xs, masked = key.split("_")
    x = fx.make_x(xs, nodes, meta["dim"])
    y = pn.apply(da.traversal(), x, meta["dim"], meta["order"], meta["maxdepth"],
                 "poisson" if op == "heat" else op, da.boundary() if masked == "masked" else None)
    ref = z[key]
    assert np.linalg.norm(y - ref) <= 1e-13 * max(np.linalg.norm(ref), 1e-300)


@pytest.mark.parametrize("name,key", _cases())
def test_plan_cpu_restatement_matches_reference(dkt, oracle, name, key):
    """The CPU baseline's restatement over the flattened traversal
    (oracle/plan_cpu.c), on 1 and 3 threads, against the reference's outputs."""
    import importlib.util
    spec = importlib.util.spec_from_file_location("plan_cpu", os.path.join(REPO, "oracle", "plan_cpu.py"))
    pc = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(pc)
    O = pc.load()
    da, z, meta = fx.build_da(dkt, name)
    nodes = da.nodes()
    _, op, xs, masked = key.split("_")
    x = fx.make_x(xs, nodes, meta["dim"])
    tr = da.traversal()
    ref = z[key]
    for nt in (1, 3):
        y = pc.matvec(O, tr, meta["dim"], meta["order"], da.n_nodes, x, op, nt,
                      da.boundary() if masked == "masked" else None)
        assert np.linalg.norm(y - ref) <= 1e-13 * max(np.linalg.norm(ref), 1e-300), (nt,)


@pytest.mark.parametrize("dim,p", [(2, 1), (3, 1), (4, 1), (2, 2), (3, 2)])
def test_stiffness_annihilates_constants(dkt, oracle, dim, p):
    maxd = 8 if p == 1 else 10
    if p == 1:
        pts = gi.gauss_points(dim, 150, 3, 0.07, 2, maxd)
    else:  # leaves two levels above the finest keep order-2 node coordinates integral
        pts = gi.cells_at_level(gi.gauss_points(dim, 150, 3, 0.07, 2, maxd - 2), maxd - 2, maxd)
    o = dkt.Octs.from_array(pts, dim, maxd, dkt.CURVE_HILBERT)
    da = dkt.DA(dkt.balance_tree(o, 1), p)
    nodes = da.nodes()
    meta = {"dim": dim, "maxdepth": maxd, "order": p, "curve": "hilbert"}
    y = _oracle_y(oracle, meta, nodes, "poisson", np.ones(len(nodes)), False)
    interior = np.ones(len(nodes), bool)
    interior[da.boundary()] = False
    assert np.abs(y[interior]).max() < 1e-12


@pytest.mark.parametrize("dim,lev", [(2, 3), (3, 2), (4, 2)])
def test_identity_operator_counts_elements_per_node(dkt, oracle, dim, lev):
    maxd = 6
    cells = np.array(gi.example_tree(2, dim, lev, maxd), dtype=np.uint32)
    da = dkt.DA(dkt.sort_tree(dkt.Octs.from_array(cells, dim, maxd, dkt.CURVE_HILBERT)), 1)
    nodes = da.nodes()
    meta = {"dim": dim, "maxdepth": maxd, "order": 1, "curve": "hilbert"}
    y = _oracle_y(oracle, meta, nodes, "ident", np.ones(len(nodes)), False)
    c = nodes[:, :dim].astype(np.int64)
    interior_axes = ((c > 0) & (c < (1 << maxd))).sum(axis=1)
    assert np.array_equal(y, 2.0 ** interior_axes)


@pytest.mark.parametrize("dim,lev", [(2, 4), (3, 3), (4, 2)])
def test_mass_operator_integrates_one(dkt, oracle, dim, lev):
    """Uniform grids (no hanging leaves): 1^T M 1 = 1.  With hanging leaves the reference
    drops transposed contributions whose parent rank has no parent-level node, so the
    identity does not hold there and is not asserted."""
    maxd = 6
    cells = np.array(gi.example_tree(2, dim, lev, maxd), dtype=np.uint32)
    da = dkt.DA(dkt.sort_tree(dkt.Octs.from_array(cells, dim, maxd, dkt.CURVE_HILBERT)), 1)
    nodes = da.nodes()
    meta = {"dim": dim, "maxdepth": maxd, "order": 1, "curve": "hilbert"}
    y = _oracle_y(oracle, meta, nodes, "mass", np.ones(len(nodes)), False)
    assert abs(y.sum() - 1.0) < 1e-13


@pytest.mark.parametrize("dim,p", [(2, 1), (3, 1), (4, 1), (2, 2), (3, 2)])
def test_reference_element_matrices(dim, p):
    K, _ = pn.ref_matrix(dim, p, "poisson")
    M, _ = pn.ref_matrix(dim, p, "mass")
    assert np.allclose(K, K.T, atol=0) and np.abs(K.sum(axis=1)).max() < 1e-13
    assert abs(M.sum() - 1.0) < 1e-13
    # Q1 in 2D: the textbook stiffness [[2/3,-1/6,-1/6,-1/3], ...] (axis 0 fastest)
    if dim == 2 and p == 1:
        ref = np.array([[4, -1, -1, -2], [-1, 4, -2, -1], [-1, -2, 4, -1], [-2, -1, -1, 4]]) / 6.0
        assert np.allclose(K, ref, rtol=0, atol=1e-15)
    # eigenvalues: one zero (constants), the rest positive
    ev = np.linalg.eigvalsh(K)
    assert abs(ev[0]) < 1e-12 and ev[1] > 1e-3


def test_goldens_exercise_stale_parent_buffer_reads(dkt):
    """A planted mistake: dropping the stale parent-buffer terms from the interpreter must
    break parity on fixtures whose traversal reads stale entries, so those goldens pin them."""
    checked = 0
    for name in ("ex3_3d_p1", "sphere3d_small", "gauss4d_small"):
        da, z, meta = fx.build_da(dkt, name)
        if da.count(6) == 0:
            continue
        tr = da.traversal()
        tr["hc_off"] = np.zeros_like(tr["hc_off"])
        nodes = da.nodes()
        key = "y_poisson_hash0_raw"
        x = fx.make_x("hash0", nodes, meta["dim"])
        y = pn.apply(tr, x, meta["dim"], meta["order"], meta["maxdepth"], "poisson")
        ref = z[key]
        assert np.linalg.norm(y - ref) > 1e-6 * np.linalg.norm(ref)
        checked += 1
    assert checked > 0
