"""The benchmark configurations beyond the matvec, at their own workloads:

* level-9 3D p = 2 mixed-precision CG (the sphere geometry r = 0.3 refined to
  level 9, Dirichlet zero, b keyed uniform seed 1): FP64 CG and mixed CG (FP32
  iteration, FP64 reliable updates) both reach a relative L_inf residual of
  1e-8, re-checked with the FP64 conforming matvec (itself pinned to the
  assembled numpy C^T K C), and the two solutions agree to 1e-6;
* 64 independent 3D trees, element counts log-uniform in [1e3, 2e5], random
  spheres, p alternating 1 and 2, 8 right-hand sides U[-1,1] in FP32, all run
  by one batch replayed from a hipGraph: every tree and column agrees with its
  FP64 plan (2e-6 relative L2; the FP64 plans are pinned to the reference's
  goldens), and column 0 of every p = 1 tree and of the p = 2 trees up to 2e4
  elements with the oracle's C restatement of the reference traversal; the
  replays agree bit for bit;
* served CG: requests on a pool of meshes, dim 3 / 4 in a 70/30 split, p = 1,
  1e3-1e6 elements log-uniform, Poisson and heat kinds, mixed precision, under
  Poisson arrivals into the continuously batched server: every request
  converges to 1e-8 and its answer is bitwise the same request served alone.
"""
import ctypes
import importlib.util
import os
import sys
import time

import numpy as np
import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
pytestmark = pytest.mark.gpu


def _torch():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    return torch


def _mod(name, file):
    if name in sys.modules:
        return sys.modules[name]
    spec = importlib.util.spec_from_file_location(name, os.path.join(REPO, "dendro-kt_amd", file))
    m = importlib.util.module_from_spec(spec)
    sys.modules[name] = m
    spec.loader.exec_module(m)
    return m


def _oracle():
    lib = os.path.join(REPO, "oracle", "_build", "liboracle.so")
    if not os.path.exists(lib):
        pytest.skip("oracle restatement not built")
    O = ctypes.CDLL(lib)
    O.oracle_matvec.argtypes = [ctypes.c_int] * 4 + [ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int,
                                                     ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int]
    return O


def test_level9_p2_mixed_cg(dkt):
    torch = _torch()
    t0 = time.time()
    da = dkt.DA(dkt.balance_tree(dkt.sphere_tree(3, 30, 3, 9, (0.5, 0.5, 0.5), 0.3), 1), 2)
    plan = dkt.Plan(da, dkt.OP_STIFFNESS, conforming=True)
    print(f"level-9 p=2 sphere: {da.n_elements} elements, {da.n_nodes} nodes, built in {time.time() - t0:.1f}s",
          flush=True)
    b = dkt.keyed_uniform(1, da.nodes())
    b[da.boundary()] = 0.0
    bd = torch.tensor(b, device="cuda")
    xs, sts = {}, {}
    for name, mode in (("f64", dkt.CG_F64), ("mixed", dkt.CG_MIXED)):
        x = torch.zeros_like(bd)
        t = time.perf_counter()
        st = plan.cg_solve_torch(bd, x, mode, 100000, 1e-8)
        torch.cuda.synchronize()
        print(f"CG {name}: {st} in {time.perf_counter() - t:.2f}s", flush=True)
        assert st["stop"] == "converged" and st["resid"] <= 1e-8
        ax = torch.empty_like(x)
        plan.matvec_torch(x, ax, masked=True)
        torch.cuda.synchronize()
        rel = float((bd - ax).abs().max() / bd.abs().max())
        assert rel <= 1.01e-8, (name, rel)
        xs[name], sts[name] = x, st
    diff = float(torch.linalg.norm(xs["mixed"] - xs["f64"]) / torch.linalg.norm(xs["f64"]))
    print(f"mixed vs FP64 solution: rel L2 {diff:.2e}", flush=True)
    assert diff <= 1e-6


def test_batch_64_trees_log_uniform(dkt):
    torch = _torch()
    wl = _mod("dkt_workloads", "workloads.py")
    O = _oracle()
    t0 = time.time()
    trees = wl.batch_trees(dkt, 64, 1e3, 2e5, seed=3)
    das = [dkt.DA(t, p) for t, p, _, _ in trees]
    sizes = np.array([da.n_elements for da in das])
    print(f"64 trees: {sizes.sum()} elements (min {sizes.min()}, median {int(np.median(sizes))}, "
          f"max {sizes.max()}), {sum(da.n_nodes for da in das)} nodes, built in {time.time() - t0:.1f}s", flush=True)
    assert sizes.min() >= 500 and sizes.max() <= 4e5
    assert np.sum(sizes < 1e4) >= 10 and np.sum(sizes > 3e4) >= 10  # spread over the range
    plans = [dkt.Plan(da, dkt.OP_STIFFNESS) for da in das]
    g = torch.Generator(device="cuda")
    g.manual_seed(11)
    nrhs = 8
    xs = [(torch.rand(da.n_nodes, nrhs, dtype=torch.float32, device="cuda", generator=g) * 2 - 1) for da in das]
    ys = [torch.full_like(x, float("nan")) for x in xs]
    s = torch.cuda.Stream()
    batch = dkt.Batch(plans, xs, ys, graph=True)
    outs = []
    with torch.cuda.stream(s):
        for _ in range(3):
            batch.run(s.cuda_stream)
            s.synchronize()
            outs.append([y.clone() for y in ys])
            for y in ys:
                y.fill_(float("nan"))
            s.synchronize()
    for rep in outs[1:]:
        assert all(torch.equal(a, b) for a, b in zip(outs[0], rep))
    ys = outs[0]
    checked_oracle = 0
    for k, (da, plan, x, y) in enumerate(zip(das, plans, xs, ys)):
        x64 = x.double().contiguous()
        y64 = torch.empty_like(x64)
        plan.matvec_torch(x64, y64)
        torch.cuda.synchronize()
        err = float(torch.linalg.norm(y.double() - y64) / torch.linalg.norm(y64))
        assert err <= 2e-6, (k, err)
        if da.order == 1 or da.n_elements <= 2e4:
            nodes = np.ascontiguousarray(da.nodes())
            xo = np.ascontiguousarray(x64[:, 0].cpu().numpy())
            yo = np.zeros_like(xo)
            O.oracle_reset()
            rc = O.oracle_matvec(3, 30, da.order, 1, nodes.ctypes.data, len(nodes), 0, xo.ctypes.data,
                                 yo.ctypes.data, 0)
            assert rc == 0
            e0 = np.linalg.norm(y[:, 0].double().cpu().numpy() - yo) / np.linalg.norm(yo)
            e64 = np.linalg.norm(y64[:, 0].cpu().numpy() - yo) / np.linalg.norm(yo)
            assert e0 <= 2e-6 and e64 <= 1e-12, (k, e0, e64)
            checked_oracle += 1
    print(f"oracle-checked trees: {checked_oracle}", flush=True)
    assert checked_oracle >= 40


def test_served_pool_1e3_to_1e6_alone_equals_in_mix(dkt):
    _torch()
    sv = _mod("dkt_serve", "serve.py")
    wl = _mod("dkt_workloads", "workloads.py")
    pool = [spec for spec, _ in wl.serve_pool(sv, 10, 1e3, 1e6, seed=5)]
    rng = np.random.default_rng(2)
    reqs = [sv.Request(pool[k % len(pool)], rhs_seed=int(rng.integers(1 << 30)), mode=sv.CG_MIXED,
                       kind="heat" if k % 2 else "poisson", max_iter=50000) for k in range(14)]
    srv = sv.Server(dkt, n_streams=2, batch=4)
    try:
        done, wall = sv.run_open_loop(srv, reqs, 200.0, seed=3, timeout=600)
    finally:
        srv.close()
    assert all(not r.error for r in done), [r.error for r in done]
    elems = sorted({(r.mesh.dim, r.nodes) for r in done})
    print(f"served {len(done)} requests in {wall:.1f}s; meshes (dim, nodes): {elems}", flush=True)
    assert max(n for _, n in elems) >= 2e5 and min(n for _, n in elems) <= 5e3
    assert {d for d, _ in elems} == {3, 4}
    assert all(r.stats["converged"] and r.stats["resid"] <= r.tol for r in done)
    alone = sv.Server(dkt, n_streams=1, batch=1, cache=srv.cache)
    try:
        for i, r in enumerate(reqs):
            alone.submit(sv.Request(r.mesh, r.rhs_seed, kind=r.kind, mode=r.mode, max_iter=r.max_iter))
            out = alone.drain(i + 1, timeout=300)[-1]
            assert not out.error, out.error
            assert out.stats["iters"] == r.stats["iters"] and out.stats["resid"] == r.stats["resid"]
            assert np.array_equal(out.x, r.x)
    finally:
        alone.close()
