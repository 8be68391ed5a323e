"""Host mesh layer (CPU): tree sort/construct/balance, CG node enumeration in
the reference's order, traversal flattening and the synthetic generators.

Pins: the reference's own outputs committed under tests/golden (elements,
nodes, boundary ids, traversal leaves and integer-decoded gathers, produced by
oracle/ref_driver.cpp built from the reference sources), the closed-form node
counts of test/testAdaptiveExamples.h:42-45,88-91,128-141, the tree
invariants of test/testTreeBalancing.cpp:607-676 and test/testTreeSort.cpp,
and the Hilbert face adjacency of test/testHCurveTable.cpp:19-40.
"""
import hashlib

import numpy as np
import pytest

import fixtures as fx
import gen_inputs as gi


def _sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


@pytest.mark.parametrize("name", fx.names())
def test_mesh_matches_reference(dkt, name):
    """Elements (SFC order), nodes (countCGNodes order + level), boundary ids, bit for bit."""
    da, z, meta = fx.build_da(dkt, name)
    for key, arr in (("elems", da.elements()), ("nodes", da.nodes()), ("bdy", da.boundary())):
        if key in z.files:
            assert arr.shape == z[key].shape and np.array_equal(arr, z[key]), key
        else:
            assert _sha(arr) == meta["sha_" + key], key


@pytest.mark.parametrize("name", [n for n in fx.names() if n != "sphere3d_l7"])
def test_traversal_matches_reference_gathers(dkt, name):
    """Traversal leaves in the reference's visiting order; every present slot equals the
    node index the reference gathers there with vecIn[i] = i (FEM/include/matvec.h:380-400)."""
    da, z, meta = fx.build_da(dkt, name)
    if "leaf_cells" not in z.files:
        pytest.skip("fixture without traversal records")
    tr = da.traversal()
    assert np.array_equal(tr["cells"], z["leaf_cells"])
    li = z["leaf_in"]
    assert tr["slot"].shape == li.shape
    m = tr["slot"] >= 0
    assert np.array_equal(tr["slot"][m].astype(np.float64), li[m])


def _num_points(ex, dim, p, end_l):
    if ex == 1:
        return (end_l - 2) * ((4 * p - 1) ** dim - (2 * p - 1) ** dim) + (4 * p + 1) ** dim
    if ex == 2:
        return ((1 << end_l) * p + 1) ** dim
    total = ((1 << end_l) * p + 1) ** dim
    for lev in range(2, end_l):
        total += (((1 << lev) - 2) * p + 1) ** dim
        total -= (((1 << (lev + 1)) - 4) * p + 1) ** dim
    return total


@pytest.mark.parametrize("ex,dim,p,end_l", [
    (1, 2, 1, 4), (1, 2, 2, 4), (1, 3, 1, 4), (1, 3, 2, 3), (1, 4, 1, 3),
    (2, 2, 1, 3), (2, 2, 2, 3), (2, 3, 1, 3), (2, 3, 2, 2), (2, 4, 1, 2),
    (3, 2, 1, 4), (3, 2, 2, 4), (3, 3, 1, 3), (3, 3, 2, 3), (3, 4, 1, 3),
])
def test_closed_form_node_counts(dkt, ex, dim, p, end_l):
    """DA::getLocalNodalSz of the example trees against the closed forms."""
    maxd = 8
    cells = np.array(gi.example_tree(ex, dim, end_l, maxd), dtype=np.uint32)
    t = dkt.sort_tree(dkt.Octs.from_array(cells, dim, maxd, dkt.CURVE_HILBERT))
    da = dkt.DA(t, p)
    assert da.n_nodes == _num_points(ex, dim, p, end_l)


def _leaves(dim, maxd, n, seed, curve):
    pts = gi.gauss_points(dim, n, 3, 0.08, seed, maxd)
    o = dkt_mod().Octs.from_array(pts, dim, maxd, curve)
    return pts, dkt_mod().balance_tree(o, 1).array(), dkt_mod().construct_tree(o, 1).array()


def dkt_mod():
    import conftest
    m = conftest.load_dkt()
    m.load()
    return m


def _boxes(cells, dim, maxd):
    lo = cells[:, :dim].astype(np.int64)
    h = (1 << (maxd - cells[:, dim].astype(np.int64)))[:, None]
    return lo, lo + h


@pytest.mark.parametrize("dim,n", [(2, 300), (3, 200), (4, 60)])
def test_balanced_tree_invariants(dkt, dim, n):
    """Complete (volumes tile the unit cube), sorted (sorting again is the identity),
    at most one point per leaf of the construction, and 2:1 balanced across faces,
    edges and corners (checkBalancingConstraint)."""
    maxd = 7
    pts, bal, con = _leaves(dim, maxd, n, 7, dkt.CURVE_HILBERT)
    for cells in (bal, con):
        vol = sum(1 << (dim * (maxd - int(l))) for l in cells[:, dim])
        assert vol == 1 << (dim * maxd)
        again = dkt.sort_tree(dkt.Octs.from_array(cells, dim, maxd, dkt.CURVE_HILBERT)).array()
        assert np.array_equal(again, cells)
    # at most one point per leaf of the construction, except at the finest level
    # (src/tsort.cpp:667-671 stops splitting there)
    lo, hi = _boxes(con, dim, maxd)
    p = pts[:, :dim].astype(np.int64)
    for i in range(len(con)):
        inside = np.all((p >= lo[i]) & (p < hi[i]), axis=1)
        assert inside.sum() <= 1 or int(con[i, dim]) == maxd
    # 2:1 across any shared boundary (closed boxes intersect)
    lo, hi = _boxes(bal, dim, maxd)
    lev = bal[:, dim].astype(np.int64)
    for i in range(len(bal)):
        touch = np.all((lo <= hi[i]) & (hi >= lo[i]), axis=1)
        assert np.all(np.abs(lev[touch] - lev[i]) <= 1)


@pytest.mark.parametrize("dim,end_l", [(2, 4), (3, 3), (4, 2)])
def test_hilbert_consecutive_leaves_share_a_face(dkt, dim, end_l):
    """Uniform grid in Hilbert order: consecutive cells differ by one step along one axis."""
    maxd = 6
    cells = np.array(gi.example_tree(2, dim, end_l, maxd), dtype=np.uint32)
    t = dkt.sort_tree(dkt.Octs.from_array(cells, dim, maxd, dkt.CURVE_HILBERT)).array()
    step = 1 << (maxd - end_l)
    d = np.abs(np.diff(t[:, :dim].astype(np.int64), axis=0))
    assert np.all((d == step).sum(axis=1) == 1) and np.all((d == 0).sum(axis=1) == dim - 1)


def test_morton_order_is_lexicographic_in_interleaved_bits(dkt):
    """Morton (CURVE_MORTON): the sort order equals sorting by interleaved coordinate bits."""
    dim, maxd, lev = 3, 5, 3
    cells = np.array(gi.example_tree(2, dim, lev, maxd), dtype=np.uint32)
    rng = np.random.default_rng(0)
    shuffled = cells[rng.permutation(len(cells))]
    t = dkt.sort_tree(dkt.Octs.from_array(shuffled, dim, maxd, dkt.CURVE_MORTON)).array()

    def key(c):
        k = 0
        for b in range(maxd - 1, -1, -1):
            for d in range(dim - 1, -1, -1):
                k = (k << 1) | ((int(c[d]) >> b) & 1)
        return k
    assert [key(c) for c in t] == sorted(key(c) for c in cells)


@pytest.mark.parametrize("dim", [2, 3, 4])
def test_gauss_points_match_numpy_generator(dkt, dim):
    a = dkt.gauss_points(dim, 9, 500, 4, 0.05, 11).array()
    b = gi.gauss_points(dim, 500, 4, 0.05, 11, 9)
    order = np.lexsort(b.T[::-1])
    a2 = a[np.lexsort(a.T[::-1])]
    assert np.array_equal(a2, b[order])


def test_sphere_tree_matches_numpy_generator(dkt):
    a = dkt.sphere_tree(3, 30, 3, 5, (0.5, 0.5, 0.5), 0.3).array()
    b = gi.sphere_tree(3, 30, 3, 5, (0.5, 0.5, 0.5), 0.3)
    a2 = a[np.lexsort(a.T[::-1])]
    b2 = np.asarray(b, dtype=np.uint32)
    b2 = b2[np.lexsort(b2.T[::-1])]
    assert np.array_equal(a2, b2)


def test_keyed_uniform_matches_numpy(dkt):
    rng = np.random.default_rng(3)
    for dim in (2, 3, 4):
        nodes = rng.integers(0, 1 << 30, size=(1000, dim + 1), dtype=np.uint64).astype(np.uint32)
        assert np.array_equal(dkt.keyed_uniform(5, nodes), gi.keyed_uniform(5, nodes[:, :dim]))


def test_traversal_slots_cover_each_leaf_node_once(dkt):
    """Every node of a leaf appears at one rank, and only hanging leaves miss ranks."""
    pts = dkt.gauss_points(3, 8, 200, 3, 0.06, 5)
    da = dkt.DA(dkt.balance_tree(pts, 1), 1)
    tr = da.traversal()
    for e in range(tr["slot"].shape[0]):
        s = tr["slot"][e]
        present = s[s >= 0]
        assert len(np.unique(present)) == len(present)
        if tr["hang_of"][e] < 0:
            assert np.all(s >= 0)


def test_node_enumeration_independent_of_maxdepth_4d(dkt):
    """The same 4D tree expressed at maxdepth 9 and 30 (coordinates scaled by
    2^21) has the same CG nodes, in the same order, with the same levels.  At
    maxdepth 30 the 4D location keys (4 x 31 coordinate bits + level) exceed 128
    bits, so this runs the wide-key path of the enumeration."""
    pts9 = dkt.gauss_points(4, 9, 3000, 4, 0.05, 3)
    da9 = dkt.DA(dkt.balance_tree(pts9, 1), 1)
    arr = pts9.array().copy()
    arr[:, :4] <<= 21
    arr[:, 4] += 21
    da30 = dkt.DA(dkt.balance_tree(dkt.Octs.from_array(arr, 4, 30, dkt.CURVE_HILBERT), 1), 1)
    n9, n30 = da9.nodes(), da30.nodes()
    assert len(n9) == len(n30)
    assert np.array_equal(n9[:, :4].astype(np.uint64) << 21, n30[:, :4].astype(np.uint64))
    assert np.array_equal(n9[:, 4], n30[:, 4])
    assert np.array_equal(da9.boundary(), da30.boundary())
