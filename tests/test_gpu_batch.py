"""Batched matvec (dkt_batch_*): many meshes of mixed (dim, p) in one launch per
(dim, p), FP32 and FP64, several right-hand sides.  Every plan's result must be
bitwise equal to its own dkt_plan_matvec; a captured hipGraph is replayed
several times with eager work in between, and every replay is checked."""
import numpy as np
import pytest


def _torch():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    return torch


def _meshes(dkt, count, seed):
    rng = np.random.default_rng(seed)
    out = []
    for k in range(count):
        dim = 3 if k % 4 else 2
        c = rng.uniform(0.2, 0.8, dim)
        r = rng.uniform(0.1, 0.3)
        lev = int(rng.integers(3, 6))
        p = 1 if k % 3 else 2
        tree = dkt.sphere_tree(dim, 30, 2, lev, tuple(c), r)
        da = dkt.DA(dkt.balance_tree(tree, 1), p)
        out.append(da)
    pts = dkt.gauss_points(4, 7, 300, 3, 0.08, seed)
    out.append(dkt.DA(dkt.balance_tree(pts, 1), 1))
    return out


@pytest.mark.gpu
@pytest.mark.parametrize("dtype,nrhs", [("f32", 8), ("f64", 3)])
def test_batch_matches_per_plan_matvec(dkt, dtype, nrhs):
    torch = _torch()
    dt = torch.float32 if dtype == "f32" else torch.float64
    das = _meshes(dkt, 12, 5)
    plans = [dkt.Plan(da, dkt.OP_STIFFNESS) for da in das]
    xs, ys, refs = [], [], []
    for k, (da, plan) in enumerate(zip(das, plans)):
        x = torch.tensor(np.stack([dkt.keyed_uniform(k * 16 + q, da.nodes()) for q in range(nrhs)], axis=1),
                         dtype=dt, device="cuda").contiguous()
        ref = torch.empty_like(x)
        plan.matvec_torch(x, ref, masked=True)
        xs.append(x)
        ys.append(torch.zeros_like(x))
        refs.append(ref)
    b = dkt.Batch(plans, xs, ys, masked=True)
    b.run()
    torch.cuda.synchronize()
    for y, ref in zip(ys, refs):
        assert torch.equal(y, ref)


@pytest.mark.gpu
def test_batch_graph_replays(dkt):
    torch = _torch()
    das = _meshes(dkt, 8, 9)
    plans = [dkt.Plan(da, dkt.OP_STIFFNESS) for da in das]
    xs = [torch.tensor(dkt.keyed_uniform(k, da.nodes()), dtype=torch.float32, device="cuda")
          for k, da in enumerate(das)]
    refs = []
    for p, x in zip(plans, xs):
        r = torch.empty_like(x)
        p.matvec_torch(x, r)
        refs.append(r)
    ys = [torch.zeros_like(x) for x in xs]
    s = torch.cuda.Stream()
    b = dkt.Batch(plans, xs, ys, graph=True)
    for rep in range(4):
        with torch.cuda.stream(s):
            for y in ys:
                y.fill_(-1.0)
            b.run(s.cuda_stream)
        s.synchronize()
        for y, ref in zip(ys, refs):
            assert torch.equal(y, ref), rep
        # eager work between replays
        for p, x, y in zip(plans, xs, ys):
            p.matvec_torch(x, y)
        torch.cuda.synchronize()


@pytest.mark.gpu
@pytest.mark.skip(reason="DKT_BATCH_PERSISTENT is switched off: its launch did not finish on the device (the run "
                         "was stopped for silence, twice) and the cause is not found")
@pytest.mark.parametrize("dtype,nrhs,graph", [("f64", 1, True), ("f32", 8, True), ("f64", 3, False)])
def test_persistent_batch_matches_per_plan(dkt, dtype, nrhs, graph):
    """DKT_BATCH_PERSISTENT: one persistent launch for every (dim, p) of the
    batch with the fix-up fused (2D/3D/4D, p = 1 and 2, several tile sizes):
    each run, eager or replayed from a one-node graph, is bitwise every plan's
    own matvec; the launch resets its own counters between runs (replays with
    eager work and other inputs in between) and reports no spin timeout."""
    torch = _torch()
    dt = torch.float32 if dtype == "f32" else torch.float64
    das = _meshes(dkt, 10, 17)
    plans = [dkt.Plan(da, dkt.OP_STIFFNESS, tile_leaves=(64, 128, 256)[k % 3]) for k, da in enumerate(das)]
    xs, ys = [], []
    for k, da in enumerate(das):
        x = torch.tensor(np.stack([dkt.keyed_uniform(k * 16 + q, da.nodes()) for q in range(nrhs)], axis=1),
                         dtype=dt, device="cuda").contiguous()
        xs.append(x if nrhs > 1 else x[:, 0].contiguous())
        ys.append(torch.zeros_like(xs[-1]))
    s = torch.cuda.Stream()
    b = dkt.Batch(plans, xs, ys, masked=True, graph=graph, persistent=True)
    for rep in range(4):
        for x in xs:  # other inputs every run
            x.mul_(-1.0 if rep % 2 else 1.0)
        refs = []
        for p, x in zip(plans, xs):
            r = torch.empty_like(x)
            p.matvec_torch(x, r, masked=True)
            refs.append(r)
        torch.cuda.synchronize()
        with torch.cuda.stream(s):
            for y in ys:
                y.fill_(float("nan"))
            b.run(s.cuda_stream)
        s.synchronize()
        assert b.status() == 0
        for k, (y, ref) in enumerate(zip(ys, refs)):
            assert torch.equal(y, ref), (rep, k)


@pytest.mark.gpu
def test_autotuner_picks_a_candidate_and_caches(dkt):
    """The autotuner times every candidate tile size that fits (with eight columns
    also every columns-per-record choice), returns the fastest per precision,
    and answers again from its cache; conforming plans are tuned separately."""
    import importlib.util as ilu
    import os
    torch = _torch()
    spec = ilu.spec_from_file_location(
        "dkt_autotune", os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                                     "dendro-kt_amd", "autotune.py"))
    at = ilu.module_from_spec(spec)
    spec.loader.exec_module(at)
    da = dkt.DA(dkt.balance_tree(dkt.sphere_tree(3, 30, 2, 6, (0.5, 0.5, 0.5), 0.3), 1), 1)
    for nrhs in (1, 8):
        r = at.autotune(dkt, da, reps=3, warmup=1, nrhs=nrhs)
        assert r["ms"] and {T for T, _ in r["ms"]} <= set(at.CANDIDATES)
        assert {c for _, c in r["ms"]} == set(at._column_candidates(nrhs))
        for prec in ("f64", "f32"):
            key = (r["best"][prec], r["columns"][prec])
            assert key in r["ms"]
            assert r["ms"][key][prec] == min(v[prec] for v in r["ms"].values())
        assert at.autotune(dkt, da, nrhs=nrhs) is r
        plan = dkt.Plan(da, dkt.OP_STIFFNESS, tile_leaves=r["best"]["f32"])
        assert plan.info()["tile_leaves"] == r["best"]["f32"]
    # the conforming operator is tuned and cached on its own
    rc = at.autotune(dkt, da, reps=3, warmup=1, conforming=True)
    assert rc is not at.autotune(dkt, da, nrhs=1) and rc is at.autotune(dkt, da, conforming=True)
    assert {T for T, _ in rc["ms"]} <= set(at.CANDIDATES)
    plan = dkt.Plan(da, dkt.OP_STIFFNESS, conforming=True, tile_leaves=rc["best"]["f64"])
    assert plan.info()["tile_leaves"] == rc["best"]["f64"]


@pytest.mark.gpu
@pytest.mark.parametrize("dtype", ["f64", "f32"])
def test_batch_columns_per_launch_bitwise(dkt, dtype):
    """Capping the columns per tile record (1 runs the single-column kernels, a
    record per column, over the same [N][nrhs] vectors; 2 and 4 run narrower
    multi-column records), and column-major [nrhs][N] vectors (one contiguous
    column per record), give every column bitwise the result of the default
    batch, for p = 1 and p = 2 trees and for conforming plans."""
    torch = _torch()
    dt = torch.float64 if dtype == "f64" else torch.float32
    das = [dkt.DA(dkt.balance_tree(dkt.sphere_tree(3, 30, 2, lev, c, 0.2), 1), p)
           for lev, c, p in ((6, (0.4, 0.5, 0.5), 1), (5, (0.6, 0.4, 0.5), 2), (6, (0.5, 0.6, 0.4), 1))]
    plans = [dkt.Plan(da, dkt.OP_STIFFNESS) for da in das]
    plans.append(dkt.Plan(das[0], dkt.OP_STIFFNESS, conforming=True))
    nrhs = 8
    gen = torch.Generator(device="cpu").manual_seed(7)
    xs = [(torch.rand(pl.n_nodes, nrhs, generator=gen, dtype=torch.float64) * 2 - 1).to(dt).cuda() for pl in plans]
    ref = [torch.empty_like(x) for x in xs]
    dkt.Batch(plans, xs, ref, masked=True).run()
    torch.cuda.synchronize()
    for cols in (1, 2, 4):
        ys = [torch.full_like(x, float("nan")) for x in xs]
        dkt.Batch(plans, xs, ys, masked=True, columns=cols).run()
        torch.cuda.synchronize()
        for k, (y, r) in enumerate(zip(ys, ref)):
            assert torch.equal(y, r), (cols, k)
    xc = [x.t().contiguous() for x in xs]  # [nrhs, N]
    yc = [torch.full_like(x, float("nan")) for x in xc]
    dkt.Batch(plans, xc, yc, masked=True, column_major=True).run()
    torch.cuda.synchronize()
    for k, (y, r) in enumerate(zip(yc, ref)):
        assert torch.equal(y.t(), r), ("column-major", k)
    with pytest.raises(dkt.DktError):
        dkt.Batch(plans, xs, ref, columns=3)
