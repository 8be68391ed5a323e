"""The space-time heat form on the device: per-level captured element
matrices (DKT_OP_CAPTURED) on the conforming operator of an adaptive 4D tree,
the essential boundary set with dkt_plan_set_dirichlet, and BiCGStab.

* the device matvec equals the assembled numpy operator (1e-12);
* BiCGStab reaches 1e-8 (relative L_inf, re-checked with the device matvec)
  and agrees with a dense solve of the free block (1e-7);
* end to end, the exact solution u = t (f = 1) is recovered by lifting its
  essential values and solving for the rest (1e-7);
* results are bitwise reproducible.
"""
import importlib.util
import os
import sys

import numpy as np
import pytest

import conforming_numpy as cn

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
pytestmark = pytest.mark.gpu


def _torch():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    return torch


def _st():
    if "dkt_spacetime" in sys.modules:
        return sys.modules["dkt_spacetime"]
    spec = importlib.util.spec_from_file_location("dkt_spacetime", os.path.join(REPO, "dendro-kt_amd", "spacetime.py"))
    m = importlib.util.module_from_spec(spec)
    sys.modules["dkt_spacetime"] = m
    spec.loader.exec_module(m)
    return m


def _setup(dkt, npts=60, seed=4):
    st = _st()
    maxd = 8
    da = dkt.DA(dkt.balance_tree(dkt.gauss_points(4, maxd, npts, 2, 0.08, seed), 1), 1)
    A = dkt.Plan(da, dkt.OP_CAPTURED, op_data=st.level_matrices(maxd, 1.0, 1.0), conforming=True)
    F = dkt.Plan(da, dkt.OP_CAPTURED, op_data=st.rhs_matrices(maxd, 1.0), conforming=True)
    dirichlet = st.dirichlet_nodes(da.nodes(), maxd)
    A.set_dirichlet(dirichlet)
    F.set_dirichlet(dirichlet)
    return st, da, A, F, dirichlet, maxd


def test_spacetime_matvec_and_bicgstab(dkt):
    torch = _torch()
    st, da, A, F, dirichlet, maxd = _setup(dkt)
    nodes = da.nodes()
    An = cn.assemble(da.elements(), nodes, 4, 1, maxd, level_mats=st.level_matrices(maxd, 1.0, 1.0))
    x = torch.tensor(dkt.keyed_uniform(3, nodes), device="cuda")
    y = torch.empty_like(x)
    A.matvec_torch(x, y)
    torch.cuda.synchronize()
    ref = An @ x.cpu().numpy()
    assert np.linalg.norm(y.cpu().numpy() - ref) <= 1e-12 * np.linalg.norm(ref)
    free = np.setdiff1d(np.arange(len(nodes)), dirichlet)
    b = dkt.keyed_uniform(7, nodes)
    b[dirichlet] = 0.0
    bd = torch.tensor(b, device="cuda")
    xs = []
    for _ in range(2):
        xd = torch.zeros_like(bd)
        stt = A.bicgstab_solve_torch(bd, xd, 5000, 1e-8)
        torch.cuda.synchronize()
        xs.append(xd.cpu().numpy())
    print(f"BiCGStab: {stt}, {len(nodes)} nodes", flush=True)
    assert stt["stop"] == "converged" and stt["resid"] <= 1e-8
    assert np.array_equal(xs[0], xs[1])
    ax = torch.empty_like(bd)
    A.matvec_torch(torch.tensor(xs[0], device="cuda"), ax, masked=True)
    torch.cuda.synchronize()
    assert float((bd - ax).abs().max() / bd.abs().max()) <= 1.01e-8
    xf = np.linalg.solve(An[np.ix_(free, free)], b[free])
    assert np.linalg.norm(xs[0][free] - xf) <= 1e-7 * np.linalg.norm(xf)
    assert np.all(xs[0][dirichlet] == 0.0)


def test_spacetime_recovers_exact_solution(dkt):
    """u = t solves d_t u - lap u = 1 with u = t on the essential boundary:
    lift u_D (u on the essential nodes, 0 elsewhere), solve the free block for
    the rest with b = (F f - A u_D) off the boundary, and compare with t."""
    torch = _torch()
    st, da, A, F, dirichlet, maxd = _setup(dkt, 90, 6)
    nodes = da.nodes()
    t = nodes[:, 3].astype(np.float64) / (1 << maxd)
    uD = np.zeros_like(t)
    uD[dirichlet] = t[dirichlet]
    f = torch.ones(len(nodes), dtype=torch.float64, device="cuda")
    Ff = torch.empty_like(f)
    F.matvec_torch(f, Ff)  # raw: every row
    AuD = torch.empty_like(f)
    A.matvec_torch(torch.tensor(uD, device="cuda"), AuD)
    b = Ff - AuD
    b[torch.tensor(dirichlet.astype(np.int64), device="cuda")] = 0.0
    w = torch.zeros_like(b)
    stt = A.bicgstab_solve_torch(b, w, 5000, 1e-10)
    torch.cuda.synchronize()
    assert stt["stop"] == "converged"
    u = uD + w.cpu().numpy()
    err = np.abs(u - t).max()
    print(f"exact solution u = t: max error {err:.2e} ({stt['iters']} iterations)", flush=True)
    assert err <= 1e-7
