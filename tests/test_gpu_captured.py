"""Captured elemental operators (DKT_OP_CAPTURED: dense per-level matrices, the
route of a user elementalMatVec behind feMatrix::matVec, FEM/include/feMatrix.h:56,
189-259) and the CG breakdown guard.

* Per-level matrices 2^(-level (dim-2)) K_ref through DKT_OP_CAPTURED reproduce
  the reference's own Poisson outputs (FP64 <= 1e-12, FP32 <= 2e-6 relative
  L2), masked and raw, and agree with the built-in stiffness plan.
* CG on an indefinite captured operator (-K) stops at its first iteration
  with the breakdown status: the kernels exit normally, x is left as it was,
  the host sees DKT_ERR_BREAKDOWN and the stop reason.
"""
import numpy as np
import pytest

import fixtures as fx
import plan_numpy as pn

pytestmark = pytest.mark.gpu


def _torch():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    return torch


def per_level(meta, op="poisson", sign=1.0):
    dim, p, maxd = meta["dim"], meta["order"], meta["maxdepth"]
    A, hp = pn.ref_matrix(dim, p, op)
    return np.stack([sign * A * 2.0 ** (-lev * hp) for lev in range(maxd + 1)])


@pytest.mark.parametrize("name", ["sphere3d_small", "sphere3d_l7", "ex1_4d_p1", "ex3_3d_p2", "gauss2d_p2"])
def test_captured_matches_reference_and_stiffness(dkt, name):
    torch = _torch()
    da, z, meta = fx.build_da(dkt, name)
    cap = dkt.Plan(da, dkt.OP_CAPTURED, op_data=per_level(meta))
    stiff = dkt.Plan(da, dkt.OP_STIFFNESS)
    nodes = da.nodes()
    ran = 0
    for key, op, xs, masked in fx.y_cases(z):
        if op not in ("poisson", "heat"):  # heat: HeatMat<3>, the 3D stiffness
            continue
        ran += 1
        ref = z[key]
        x = torch.tensor(fx.make_x(xs, nodes, meta["dim"]), dtype=torch.float64, device="cuda")
        y = torch.full_like(x, float("nan"))
        ys = torch.full_like(x, float("nan"))
        cap.matvec_torch(x, y, masked=masked)
        stiff.matvec_torch(x, ys, masked=masked)
        y32 = torch.full_like(x.float(), float("nan"))
        cap.matvec_torch(x.float(), y32, masked=masked)
        torch.cuda.synchronize()
        y, ys, y32 = y.cpu().numpy(), ys.cpu().numpy(), y32.cpu().double().numpy()
        nr = np.linalg.norm(ref)
        e64, e32 = np.linalg.norm(y - ref) / nr, np.linalg.norm(y32 - ref) / nr
        print(f"{name} {key}: captured fp64 {e64:.2e} fp32 {e32:.2e}", flush=True)
        assert e64 <= 1e-12 and e32 <= 2e-6
        assert np.linalg.norm(y - ys) <= 1e-12 * nr
    assert ran > 0


@pytest.mark.parametrize("mode", ["f64", "reliable", "restart"])
def test_cg_breakdown_on_indefinite_operator(dkt, mode):
    torch = _torch()
    da, z, meta = fx.build_da(dkt, "sphere3d_small")
    plan = dkt.Plan(da, dkt.OP_CAPTURED, op_data=per_level(meta, sign=-1.0))
    m = {"f64": dkt.CG_F64, "reliable": dkt.CG_MIXED_RELIABLE, "restart": dkt.CG_MIXED_RESTART}[mode]
    b = dkt.keyed_uniform(1, da.nodes())
    b[da.boundary()] = 0.0
    bd = torch.tensor(b, device="cuda")
    x0 = torch.tensor(dkt.keyed_uniform(5, da.nodes()) * 1e-3, device="cuda")
    x0[torch.tensor(da.boundary().astype(np.int64), device="cuda")] = 0.0
    x = x0.clone()
    with pytest.raises(dkt.CgBreakdown) as ei:
        plan.cg_solve_torch(bd, x, m, 500, 1e-8)
    torch.cuda.synchronize()
    st = ei.value.stats
    print(mode, st, flush=True)
    assert st["stop"] == "breakdown" and st["iters"] == 0
    assert np.isfinite(st["resid"]) and st["resid"] > 1e-8
    assert torch.equal(x, x0)  # the first step was refused before x changed
    # the same solve without raising returns the statistics
    x = x0.clone()
    st2 = plan.cg_solve_torch(bd, x, m, 500, 1e-8, raise_on_breakdown=False)
    assert st2 == st
    # and a positive definite solve on the same mesh afterwards still converges
    ok = dkt.Plan(da, dkt.OP_CAPTURED, op_data=per_level(meta))
    x = torch.zeros_like(bd)
    st3 = ok.cg_solve_torch(bd, x, m, 5000, 1e-8)
    assert st3["stop"] == "converged" and st3["resid"] <= 1e-8


@pytest.mark.parametrize("name", ["sphere3d_small", "sphere3d_l7"])
def test_heatmat_captured_through_shim_matches_reference(dkt, name):
    """The per-level matrices the feMatrix shim captured from the reference's own
    HeatMat<3>::elementalMatVec (tests/golden/heatmat3_p1_captured.npy, written by
    tests/golden/make_captured.py) run through DKT_OP_CAPTURED and reproduce the
    reference's HeatMat outputs on the same tree: feMatrix::matVec with HeatMat's
    Dirichlet hooks (y_heat_*_masked) and the raw traversal (y_poisson_*_raw,
    the same operator in 3D), FP64 within 1e-12 relative L2."""
    import os
    torch = _torch()
    K = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "heatmat3_p1_captured.npy"))
    da, z, meta = fx.build_da(dkt, name)
    assert meta["maxdepth"] + 1 == K.shape[0]
    plan = dkt.Plan(da, dkt.OP_CAPTURED, op_data=K)
    for key in ("y_heat_hash0_masked", "y_poisson_hash0_raw"):
        ref = z[key]
        x = torch.tensor(fx.make_x("hash0", da.nodes(), 3), dtype=torch.float64, device="cuda")
        y = torch.full_like(x, float("nan"))
        plan.matvec_torch(x, y, masked=key.endswith("masked"))
        torch.cuda.synchronize()
        err = np.linalg.norm(y.cpu().numpy() - ref) / np.linalg.norm(ref)
        print(f"{name} {key}: HeatMat captured through the shim, rel L2 {err:.2e}", flush=True)
        assert err <= 1e-12
