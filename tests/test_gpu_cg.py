"""Device CG (HeatMat::cgSolve, FEM/examples/src/heatMat.cpp:165-325) against a
CPU CG of the same algorithm over the oracle's matvec (FP64, sequential sums).

Checks: the L_inf relative residual that stops the solve, recomputed on the
CPU with the oracle's masked operator, is below tol; the device solution
agrees with the CPU solution; mixed precision reaches the same tolerance;
solves are bitwise reproducible; b = 0 returns at once.
"""
import ctypes
import os
import subprocess

import numpy as np
import pytest

import fixtures as fx
import gen_inputs as gi

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(REPO, "oracle", "_build", "liboracle.so")


def _torch():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    return torch


@pytest.fixture(scope="module")
def oracle():
    if not os.path.exists(ORACLE_SO):
        subprocess.run(["make", "-C", os.path.join(REPO, "oracle"), "restatement"], check=True, capture_output=True)
    O = ctypes.CDLL(ORACLE_SO)
    O.oracle_matvec.argtypes = [ctypes.c_int] * 4 + [ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int,
                                                     ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int]
    return O


def _op(O, meta, nodes):
    cv = 1 if meta["curve"] == "hilbert" else 0
    nodes = np.ascontiguousarray(nodes)

    def A(x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.zeros_like(x)
        O.oracle_reset()
        assert O.oracle_matvec(meta["dim"], meta["maxdepth"], meta["order"], cv, nodes.ctypes.data, len(nodes), 0,
                               x.ctypes.data, y.ctypes.data, 1) == 0
        return y
    return A


def _cpu_cg(A, b, tol, max_iter):
    """The reference's iteration (heatMat.cpp:205-300) in numpy."""
    x = np.zeros_like(b)
    r0 = b - A(x)
    p = r0.copy()
    normb = np.abs(b).max() or 1.0
    if np.abs(r0).max() / normb <= tol:
        return x, 0
    for i in range(1, max_iter + 1):
        Ap = A(p)
        alpha = (r0 @ r0) / (p @ Ap)
        x = x + alpha * p
        r1 = r0 - alpha * Ap
        if np.abs(r1).max() / normb <= tol:
            return x, i
        beta = (r1 @ r1) / (r0 @ r0)
        p = r1 + beta * p
        r0 = r1
    return x, max_iter


def _rhs(da, dim, seed):
    b = gi.keyed_uniform(seed, da.nodes()[:, :dim])
    b[da.boundary()] = 0.0
    return b


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["sphere3d_small", "ex3_3d_p1", "gauss4d_small", "ex1_2d_p2"])
def test_cg_f64_matches_cpu_cg(dkt, oracle, name):
    torch = _torch()
    da, z, meta = fx.build_da(dkt, name)
    A = _op(oracle, meta, da.nodes())
    b = _rhs(da, meta["dim"], 1)
    tol = 1e-8
    xc, itc = _cpu_cg(A, b, tol, 5000)
    plan = dkt.Plan(da, dkt.OP_STIFFNESS)
    bd = torch.tensor(b, device="cuda")
    xd = torch.zeros_like(bd)
    st = plan.cg_solve_torch(bd, xd, dkt.CG_F64, 5000, tol)
    torch.cuda.synchronize()
    x = xd.cpu().numpy()
    print(f"{name}: device {st}, cpu iterations {itc}")
    assert st["resid"] <= tol
    assert np.abs(b - A(x)).max() / np.abs(b).max() <= 1.01 * tol
    assert abs(st["iters"] - itc) <= max(2, itc // 50)
    assert np.linalg.norm(x - xc) <= 1e-6 * np.linalg.norm(xc)


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["sphere3d_small", "gauss4d_small"])
@pytest.mark.parametrize("mixed", ["auto", "reliable", "restart"])
def test_cg_mixed_reaches_fp64_tolerance(dkt, oracle, name, mixed):
    """Both mixed-precision schemes (reliable updates, restarted refinement)
    reach the FP64 tolerance on the oracle's operator and agree with FP64 CG."""
    torch = _torch()
    mode = {"auto": dkt.CG_MIXED, "reliable": dkt.CG_MIXED_RELIABLE, "restart": dkt.CG_MIXED_RESTART}[mixed]
    da, z, meta = fx.build_da(dkt, name)
    A = _op(oracle, meta, da.nodes())
    b = _rhs(da, meta["dim"], 2)
    tol = 1e-8
    plan = dkt.Plan(da, dkt.OP_STIFFNESS)
    bd = torch.tensor(b, device="cuda")
    x64 = torch.zeros_like(bd)
    plan.cg_solve_torch(bd, x64, dkt.CG_F64, 5000, tol)
    xm = torch.zeros_like(bd)
    st = plan.cg_solve_torch(bd, xm, mode, 5000, tol)
    torch.cuda.synchronize()
    x = xm.cpu().numpy()
    print(f"{name}: mixed ({mixed}) {st}")
    assert st["resid"] <= tol
    assert np.abs(b - A(x)).max() / np.abs(b).max() <= 1.01 * tol
    assert np.linalg.norm(x - x64.cpu().numpy()) <= 1e-6 * np.linalg.norm(x)


@pytest.mark.gpu
def test_cg_bitwise_reproducible_and_zero_rhs(dkt):
    torch = _torch()
    pts = dkt.gauss_points(4, 8, 3000, 4, 0.05, 7)
    da = dkt.DA(dkt.balance_tree(pts, 1), 1)
    plan = dkt.Plan(da, dkt.OP_STIFFNESS)
    b = torch.tensor(_rhs(da, 4, 3), device="cuda")
    xs = []
    modes = (dkt.CG_F64, dkt.CG_MIXED_RELIABLE, dkt.CG_MIXED_RESTART)
    for mode in modes + modes:
        x = torch.zeros_like(b)
        plan.cg_solve_torch(b, x, mode, 3000, 1e-8)
        xs.append(x.cpu().numpy())
    for k in range(len(modes)):
        assert np.array_equal(xs[k], xs[k + len(modes)])
    # residual with the device matvec (itself checked against the reference)
    r = torch.empty_like(b)
    plan.matvec_torch(torch.tensor(xs[0], device="cuda"), r, masked=True)
    torch.cuda.synchronize()
    rel = (b - r).abs().max().item() / b.abs().max().item()
    assert rel <= 1.01e-8
    z = torch.zeros_like(b)
    x = torch.zeros_like(b)
    st = plan.cg_solve_torch(z, x, dkt.CG_F64, 100, 1e-8)
    assert st["iters"] == 0 and float(x.abs().max()) == 0.0


@pytest.mark.gpu
def test_cg_batch_bitwise_equals_single_solves(dkt, oracle):
    """dkt_cg_solve_batch over meshes of different (dim, p), sizes and
    iteration counts: every solve is bitwise the single FP64 solve from x = 0
    (same iterations, residual and x), and its residual, recomputed with the
    oracle's operator, is below tol.  A second batch with a small max_iter
    stops the long solves early, again bitwise as the single solves do; a
    plan listed twice solves both systems."""
    torch = _torch()
    names = ["sphere3d_small", "ex3_3d_p1", "gauss4d_small", "ex1_2d_p2"]
    cases = []
    for i, name in enumerate(names):
        da, z, meta = fx.build_da(dkt, name)
        b = _rhs(da, meta["dim"], 11 + i)
        cases.append((name, da, meta, dkt.Plan(da, dkt.OP_STIFFNESS), torch.tensor(b, device="cuda")))
    tol = 1e-8
    for max_iter in (5000, 40):
        singles = []
        for name, da, meta, plan, bd in cases:
            x = torch.zeros_like(bd)
            st = plan.cg_solve_torch(bd, x, dkt.CG_F64, max_iter, tol)
            singles.append((st, x.cpu().numpy()))
        xs = [torch.full_like(bd, 7.0) for *_, bd in cases]  # overwritten: the batch starts from x = 0
        stats = dkt.cg_solve_batch_torch([c[3] for c in cases], [c[4] for c in cases], xs, max_iter, tol)
        torch.cuda.synchronize()
        for (name, da, meta, plan, bd), (st1, x1), stb, xb in zip(cases, singles, stats, xs):
            print(f"max_iter {max_iter} {name}: single {st1}, batch {stb}")
            assert stb["iters"] == st1["iters"] and stb["resid"] == st1["resid"]
            assert np.array_equal(xb.cpu().numpy(), x1)
            if max_iter == 5000:
                A = _op(oracle, meta, da.nodes())
                b = bd.cpu().numpy()
                assert np.abs(b - A(x1)).max() / np.abs(b).max() <= 1.01 * tol
        if max_iter == 40:
            assert any(s["iters"] == 40 for s in stats)
    # one plan twice in a batch (a workspace per solve), two right-hand sides
    b2 = torch.tensor(_rhs(cases[2][1], cases[2][2]["dim"], 99), device="cuda")
    x_a, x_b = torch.zeros_like(b2), torch.zeros_like(b2)
    plan = cases[2][3]
    st_a = plan.cg_solve_torch(cases[2][4], x_a, dkt.CG_F64, 5000, tol)
    st_b = plan.cg_solve_torch(b2, x_b, dkt.CG_F64, 5000, tol)
    xs2 = [torch.empty_like(b2), torch.empty_like(b2)]
    stats = dkt.cg_solve_batch_torch([plan, plan], [cases[2][4], b2], xs2, 5000, tol)
    assert stats[0]["iters"] == st_a["iters"] and stats[1]["iters"] == st_b["iters"]
    assert torch.equal(xs2[0], x_a) and torch.equal(xs2[1], x_b)


@pytest.mark.gpu
def test_cg_engine_solves_joining_mid_flight(dkt):
    """Continuous batching (dkt_cg_engine): solves that join while others are
    running, with their own tol and max_iter, stop with their own iteration
    count and are bitwise their single FP64 solves; a full engine refuses a
    solve, and a freed slot takes the next one."""
    torch = _torch()
    cases = []
    for i, name in enumerate(["gauss4d_small", "sphere3d_small", "ex3_3d_p1"]):
        da, z, meta = fx.build_da(dkt, name)
        cases.append((dkt.Plan(da, dkt.OP_STIFFNESS), torch.tensor(_rhs(da, meta["dim"], 21 + i), device="cuda")))
    params = [(1e-8, 5000), (1e-6, 5000), (1e-8, 13)]
    singles = []
    for (plan, bd), (tol, mi) in zip(cases, params):
        x = torch.zeros_like(bd)
        singles.append((plan.cg_solve_torch(bd, x, dkt.CG_F64, mi, tol), x.cpu().numpy()))
    eng = dkt.CgEngine(2)
    xs = [torch.empty_like(bd) for _, bd in cases]
    got = {}
    eng.add(cases[0][0], cases[0][1], xs[0], key=0, max_iter=params[0][1], tol=params[0][0])
    for key, st in eng.step(3):
        got[key] = st
    eng.add(cases[1][0], cases[1][1], xs[1], key=1, max_iter=params[1][1], tol=params[1][0])
    with pytest.raises(dkt.DktError):
        eng.add(cases[2][0], cases[2][1], xs[2], key=2)
    pending = [2]
    while len(got) < 3:
        for key, st in eng.step(1):
            got[key] = st
            if pending:
                k = pending.pop()
                eng.add(cases[k][0], cases[k][1], xs[k], key=k, max_iter=params[k][1], tol=params[k][0])
    torch.cuda.synchronize()
    for k, ((st1, x1), xb) in enumerate(zip(singles, xs)):
        print(k, "single", st1, "engine", got[k])
        assert got[k]["iters"] == st1["iters"] and got[k]["resid"] == st1["resid"]
        assert np.array_equal(xb.cpu().numpy(), x1)
    assert got[2]["iters"] == 13


@pytest.mark.gpu
def test_cg_engine_mixed_and_conforming(dkt):
    """The engine runs FP64 and mixed-precision (reliable-update) solves side
    by side, on reference and conforming plans: each solve is bitwise its own
    dkt_cg_solve from x = 0 in its mode (iterations, residual, stop reason and
    x), including solves that join while others are mid-flight, a solve
    stopped by max_iter, and two solves sharing one plan; every converged
    mixed solve reaches the FP64 tolerance on its true residual."""
    torch = _torch()
    specs = [("gauss4d_small", False, dkt.CG_MIXED_RELIABLE, 1e-8, 5000),
             ("sphere3d_small", True, dkt.CG_MIXED_RELIABLE, 1e-8, 5000),
             ("ex3_3d_p2", True, dkt.CG_F64, 1e-8, 5000),
             ("sphere3d_small", True, dkt.CG_MIXED_RELIABLE, 1e-8, 30),
             ("ex3_3d_p1", False, dkt.CG_F64, 1e-8, 5000),
             ("sphere3d_small", True, dkt.CG_F64, 1e-10, 5000)]
    plans = {}
    cases = []
    for i, (name, conf, mode, tol, mi) in enumerate(specs):
        if (name, conf) not in plans:
            da, z, meta = fx.build_da(dkt, name)
            plans[(name, conf)] = (da, meta, dkt.Plan(da, dkt.OP_STIFFNESS, conforming=conf))
        da, meta, plan = plans[(name, conf)]
        bd = torch.tensor(_rhs(da, meta["dim"], 41 + i), device="cuda")
        cases.append((plan, bd, mode, tol, mi))
    singles = []
    for plan, bd, mode, tol, mi in cases:
        x = torch.zeros_like(bd)
        singles.append((plan.cg_solve_torch(bd, x, mode, mi, tol), x.cpu().numpy()))
    eng = dkt.CgEngine(3)
    xs = [torch.full_like(bd, 3.0) for _, bd, *_ in cases]
    got = {}
    todo = list(range(len(cases)))

    def add(k):
        plan, bd, mode, tol, mi = cases[k]
        eng.add(plan, bd, xs[k], key=k, max_iter=mi, tol=tol, mode=mode)

    for _ in range(3):
        add(todo.pop(0))
    steps = 0
    while len(got) < len(cases):
        for key, st in eng.step(1 if steps % 3 else 2):
            got[key] = st
            if todo:
                add(todo.pop(0))
        steps += 1
        assert steps < 20000
    torch.cuda.synchronize()
    for k, ((st1, x1), xb) in enumerate(zip(singles, xs)):
        print(k, specs[k], "single", st1, "engine", got[k], flush=True)
        assert got[k]["iters"] == st1["iters"] and got[k]["resid"] == st1["resid"]
        assert got[k]["stop"] == st1["stop"]
        assert np.array_equal(xb.cpu().numpy(), x1)
        if specs[k][4] == 5000:
            assert st1["stop"] == "converged" and st1["resid"] <= specs[k][3]
    assert got[3]["stop"] == "max_iter" and got[3]["iters"] >= 30
