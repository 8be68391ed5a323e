"""The space-time heat form (dendro-kt_amd/spacetime.py; the formulation of
papers/sc19/background.tex:26-53), on the CPU.  The reference has no code for
it, so the element matrices are pinned by the mathematics:

* each term equals a direct 2-point Gauss quadrature (exact here) of its
  bilinear form with the 4D Q1 basis on a cube of side h, at several levels;
* the element matrices annihilate constants;
* consistency on the conforming space of an adaptive 4D tree: the Q1
  interpolant of an exact solution that is linear in t (u = t, u = 3 + 2t;
  f = d_t u) leaves a zero residual A u - b at every node off the essential
  boundary; on a uniform tree (delta constant) so do the bilinear solutions
  u = x t + y - z (f = x) and u = (x + y) t (f = x + y), which exercise the
  spatial and stabilisation terms;
* the essential boundary is the spatial boundary plus t = 0, not t = 1.
"""
import importlib.util
import itertools
import os
import sys

import numpy as np
import pytest

import conforming_numpy as cn

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _st():
    if "dkt_spacetime" in sys.modules:
        return sys.modules["dkt_spacetime"]
    spec = importlib.util.spec_from_file_location("dkt_spacetime", os.path.join(REPO, "dendro-kt_amd", "spacetime.py"))
    m = importlib.util.module_from_spec(spec)
    sys.modules["dkt_spacetime"] = m
    spec.loader.exec_module(m)
    return m


def _quadrature(h, kappa, delta0):
    """The element matrices of a(u, v) and l(v) on [0, h]^4 by Gauss quadrature."""
    g = np.array([0.5 - 0.5 / np.sqrt(3.0), 0.5 + 0.5 / np.sqrt(3.0)]) * h
    wq = np.array([0.5, 0.5]) * h
    phi = lambda k, x: (1 - x / h) if k == 0 else x / h
    dphi = lambda k, x: (-1 / h) if k == 0 else 1 / h
    delta = delta0 * h
    A = np.zeros((16, 16))
    F = np.zeros((16, 16))
    for q in itertools.product(range(2), repeat=4):
        x = [g[qi] for qi in q]
        w = np.prod([wq[qi] for qi in q])
        val = np.zeros(16)
        grad = np.zeros((16, 4))
        dtg = np.zeros((16, 3))  # d_t of the spatial gradient
        for r in range(16):
            kk = [(r >> d) & 1 for d in range(4)]
            vals = [phi(kk[d], x[d]) for d in range(4)]
            ders = [dphi(kk[d], x[d]) for d in range(4)]
            val[r] = np.prod(vals)
            for d in range(4):
                grad[r, d] = ders[d] * np.prod([vals[e] for e in range(4) if e != d])
            for d in range(3):
                dtg[r, d] = ders[d] * ders[3] * np.prod([vals[e] for e in range(3) if e != d])
        for i in range(16):  # test
            for j in range(16):  # trial
                A[i, j] += w * (grad[j, 3] * val[i] + delta * grad[j, 3] * grad[i, 3]
                                + kappa * grad[j, :3] @ grad[i, :3] + delta * kappa * grad[j, :3] @ dtg[i])
                F[i, j] += w * (val[j] * val[i] + delta * val[j] * grad[i, 3])
    return A, F


@pytest.mark.parametrize("lev", [0, 1, 3, 7])
def test_element_matrices_match_quadrature(lev):
    st = _st()
    kappa, delta0 = 0.7, 1.3
    A = st.level_matrices(lev, kappa, delta0)[lev]
    F = st.rhs_matrices(lev, delta0)[lev]
    Aq, Fq = _quadrature(2.0 ** (-lev), kappa, delta0)
    assert np.abs(A - Aq).max() <= 1e-13 * np.abs(Aq).max()
    assert np.abs(F - Fq).max() <= 1e-13 * np.abs(Fq).max()
    assert np.abs(A @ np.ones(16)).max() <= 1e-15 * np.abs(A).max()
    assert np.abs(A - A.T).max() > 1e-3 * np.abs(A).max()  # not symmetric


def _mesh(dkt, uniform):
    if uniform:
        cells = np.array([[i << 6, j << 6, k << 6, l << 6, 2] for l in range(4) for k in range(4) for j in range(4)
                          for i in range(4)], dtype=np.uint32)
        tree = dkt.sort_tree(dkt.Octs.from_array(cells, 4, 8, dkt.CURVE_HILBERT))
    else:
        tree = dkt.balance_tree(dkt.gauss_points(4, 8, 40, 2, 0.08, 3), 1)
    da = dkt.DA(tree, 1)
    return da


@pytest.mark.parametrize("uniform", [False, True])
def test_consistency_exact_solutions(dkt, uniform):
    st = _st()
    da = _mesh(dkt, uniform)
    maxd = 8
    nodes = da.nodes()
    cells = da.elements()
    kappa, delta0 = 0.6, 1.0
    A = cn.assemble(cells, nodes, 4, 1, maxd, level_mats=st.level_matrices(maxd, kappa, delta0))
    F = cn.assemble(cells, nodes, 4, 1, maxd, level_mats=st.rhs_matrices(maxd, delta0))
    dirichlet = st.dirichlet_nodes(nodes, maxd)
    free = np.setdiff1d(np.arange(len(nodes)), dirichlet)
    xs = nodes[:, :4].astype(np.float64) / (1 << maxd)
    x, y, z, t = xs.T
    cases = [(t, np.ones_like(t)), (3 + 2 * t, 2 * np.ones_like(t))]
    if uniform:
        cases += [(x * t + y - z, x), ((x + y) * t, x + y)]
    for u, f in cases:
        r = A @ u - F @ f
        assert np.abs(r[free]).max() <= 1e-13 * np.abs(F @ f).max(), (uniform, np.abs(r[free]).max())
    # and a solution that is not in the space leaves a residual (the check has teeth)
    u = t * t
    assert np.abs((A @ u - F @ (2 * t))[free]).max() > 1e-6


def test_dirichlet_nodes_are_space_boundary_and_initial_time(dkt):
    st = _st()
    da = _mesh(dkt, True)
    nodes = da.nodes()
    ids = st.dirichlet_nodes(nodes, 8)
    top = 1 << 8
    sel = np.zeros(len(nodes), bool)
    sel[ids] = True
    for i, n in enumerate(nodes):
        on = any(c in (0, top) for c in n[:3]) or n[3] == 0
        assert sel[i] == on
    assert set(ids.tolist()) < set(da.boundary().tolist())  # the t = 1 face is free
