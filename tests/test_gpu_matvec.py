"""Device matvec parity against the reference's own outputs (golden fixtures).

FP64: relative L2 <= 1e-12 against the reference matvec on the same tree and
input.  FP32 (float accumulation): relative L2 <= 2e-6 against the same.
Both device paths (tiled, simple) are checked.
"""
import numpy as np
import pytest

import fixtures as fx

pytestmark = pytest.mark.gpu

TOL64 = 1e-12
TOL32 = 2e-6


def _torch():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    return torch


@pytest.mark.parametrize("simple", [False, True], ids=["tiled", "simple"])
@pytest.mark.parametrize("name", fx.names())
def test_matvec_matches_reference(dkt, name, simple):
    torch = _torch()
    da, z, meta = fx.build_da(dkt, name)
    nodes = da.nodes()
    plans = {}
    for key, op, xs, masked in fx.y_cases(z):
        code = fx.op_code(dkt, op)
        if code not in plans:
            plans[code] = dkt.Plan(da, code, simple=simple)
        plan = plans[code]
        x = fx.make_x(xs, nodes, meta["dim"])
        ref = z[key]
        xd = torch.tensor(x, dtype=torch.float64, device="cuda")
        yd = torch.full_like(xd, float("nan"))
        plan.matvec_torch(xd, yd, masked=masked)
        torch.cuda.synchronize()
        y = yd.cpu().numpy()
        err = np.linalg.norm(y - ref) / max(np.linalg.norm(ref), 1e-300)
        print(f"{name} {key} simple={simple} fp64 rel_l2 {err:.2e}", flush=True)
        assert err <= TOL64, (name, key, err)
        x32 = xd.float()
        y32 = torch.full_like(x32, float("nan"))
        plan.matvec_torch(x32, y32, masked=masked)
        torch.cuda.synchronize()
        err32 = np.linalg.norm(y32.cpu().double().numpy() - ref) / max(np.linalg.norm(ref), 1e-300)
        print(f"{name} {key} simple={simple} fp32 rel_l2 {err32:.2e}", flush=True)
        assert err32 <= TOL32, (name, key, err32)


@pytest.mark.parametrize("tile", [64, 256])
@pytest.mark.parametrize("name", fx.names())
def test_tile_sizes_match_reference(dkt, name, tile):
    """The 64- and 256-leaf tilings (one and four wavefronts per tile, sibling
    groups cut at wavefront boundaries) against the reference's outputs, FP64,
    every operator and mask of the fixture; one and three columns."""
    torch = _torch()
    da, z, meta = fx.build_da(dkt, name)
    nodes = da.nodes()
    plans = {}
    for key, op, xs, masked in fx.y_cases(z):
        code = fx.op_code(dkt, op)
        if code not in plans:
            plans[code] = dkt.Plan(da, code, tile_leaves=tile)
        plan = plans[code]
        x = fx.make_x(xs, nodes, meta["dim"])
        ref = z[key]
        xd = torch.tensor(x, dtype=torch.float64, device="cuda")
        yd = torch.full_like(xd, float("nan"))
        plan.matvec_torch(xd, yd, masked=masked)
        x3 = torch.stack([xd, 2.0 * xd, -xd], dim=1).contiguous()
        y3 = torch.full_like(x3, float("nan"))
        plan.matvec_torch(x3, y3, masked=masked)
        torch.cuda.synchronize()
        y = yd.cpu().numpy()
        err = np.linalg.norm(y - ref) / max(np.linalg.norm(ref), 1e-300)
        assert err <= TOL64, (name, key, tile, err)
        y3 = y3.cpu().numpy()
        assert np.array_equal(y3[:, 0], y), (name, key, tile)
        assert np.linalg.norm(y3[:, 1] - 2.0 * ref) <= TOL64 * 2.0 * max(np.linalg.norm(ref), 1e-300)


@pytest.mark.parametrize("nrhs", [3, 8, 24])
def test_multi_rhs_columns_match_single(dkt, nrhs):
    torch = _torch()
    da, z, meta = fx.build_da(dkt, "gauss4d_small")
    nodes = da.nodes()
    plan = dkt.Plan(da, dkt.OP_STIFFNESS)
    X = np.stack([fx.make_x(f"hash{s}", nodes, meta["dim"]) for s in range(nrhs)], axis=1)
    xd = torch.tensor(X, dtype=torch.float64, device="cuda").contiguous()
    yd = torch.full_like(xd, float("nan"))
    plan.matvec_torch(xd, yd)
    for q in range(nrhs):
        xq = xd[:, q].contiguous()
        yq = torch.empty_like(xq)
        plan.matvec_torch(xq, yq)
        torch.cuda.synchronize()
        assert torch.equal(yq, yd[:, q]), q


def test_repeated_matvec_is_deterministic(dkt):
    torch = _torch()
    da, z, meta = fx.build_da(dkt, "sphere3d_l7")
    plan = dkt.Plan(da, dkt.OP_STIFFNESS)
    xd = torch.tensor(fx.make_x("hash0", da.nodes(), 3), device="cuda")
    ys = []
    for _ in range(3):
        yd = torch.empty_like(xd)
        plan.matvec_torch(xd, yd, masked=True)
        ys.append(yd)
    torch.cuda.synchronize()
    assert torch.equal(ys[0], ys[1]) and torch.equal(ys[1], ys[2])


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["ex1_2d_p1", "gauss4d_small"])
def test_output_only_mask_is_heatvec_compute_vec(dkt, name):
    """masked = 2 (HeatVec::computeVec's hooks, FEM/examples/src/heatVec.cpp:78-97): the
    raw mass matvec with the boundary rows zeroed, against the reference's raw y."""
    torch = _torch()
    da, z, meta = fx.build_da(dkt, name)
    nodes = da.nodes()
    x = fx.make_x("hash1", nodes, meta["dim"])
    ref = z["y_mass_hash1_raw"].copy()
    ref[da.boundary()] = 0.0
    plan = dkt.Plan(da, dkt.OP_MASS)
    xd = torch.tensor(x, device="cuda")
    yd = torch.empty_like(xd)
    plan.matvec_torch(xd, yd, masked=2)
    torch.cuda.synchronize()
    y = yd.cpu().numpy()
    assert np.linalg.norm(y - ref) <= 1e-12 * np.linalg.norm(ref)


def test_full_size_4d_workload_properties(dkt):
    """The benchmark's 4D workload at full size (2e6-point Gaussian-mixture tree,
    max level 9: 15,282,481 elements, 10,346,445 nodes, beyond what the oracle
    runs in seconds), through properties that hold at any size:
    - the host pipeline reproduces the mesh sizes exactly;
    - three tilings (64, 128, 256 leaves: different tiles, replay blocks,
      carry terms and partial slots) agree to FP64 rounding;
    - a repeated matvec is bitwise identical;
    - the stiffness annihilates constants (interpolated and replayed values
      of a constant are that constant, K_ref 1 = 0), to rounding relative to
      max |A x|, in FP64 and FP32;
    - masked = 1 zeroes the boundary rows, and on an x that vanishes on the
      boundary it equals the raw operator on interior rows bitwise."""
    torch = _torch()
    pts = dkt.gauss_points(4, 9, 2000000, 4, 0.05, 1)
    da = dkt.DA(dkt.balance_tree(pts, 1), 1)
    assert da.n_elements == 15282481 and da.n_nodes == 10346445
    assert da.count(3) == 19419623 and da.count(4) == 14545541
    x = torch.tensor(dkt.keyed_uniform(0, da.nodes()), device="cuda")
    ys = {}
    for tile in (64, 128, 256):
        plan = dkt.Plan(da, dkt.OP_STIFFNESS, tile_leaves=tile)
        y = torch.full_like(x, float("nan"))
        plan.matvec_torch(x, y)
        ys[tile] = y
        if tile == 128:
            y2 = torch.full_like(x, float("nan"))
            plan.matvec_torch(x, y2)
            torch.cuda.synchronize()
            assert torch.equal(y, y2)
            one = torch.ones_like(x)
            z = torch.full_like(x, float("nan"))
            plan.matvec_torch(one, z)
            z32 = torch.full_like(x, float("nan"), dtype=torch.float32)
            plan.matvec_torch(one.float(), z32)
            torch.cuda.synchronize()
            # to rounding: tile-start buffer values are sums of dyadic weights over long chains
            ymax = float(y.abs().max())
            assert float(z.abs().max()) <= 1e-13 * ymax and float(z32.abs().max()) <= 1e-5 * ymax
            bnd = torch.tensor(da.boundary(), device="cuda", dtype=torch.long)
            xi = x.clone()
            xi[bnd] = 0.0
            raw = torch.full_like(x, float("nan"))
            msk = torch.full_like(x, float("nan"))
            plan.matvec_torch(xi, raw)
            plan.matvec_torch(xi, msk, masked=True)
            torch.cuda.synchronize()
            assert float(msk[bnd].abs().max()) == 0.0
            interior = torch.ones(da.n_nodes, dtype=torch.bool, device="cuda")
            interior[bnd] = False
            assert torch.equal(raw[interior], msk[interior])
        del plan
    torch.cuda.synchronize()
    ref = ys[128]
    assert torch.isfinite(ref).all()
    for tile in (64, 256):
        rel = float(torch.linalg.norm(ys[tile] - ref) / torch.linalg.norm(ref))
        print(f"tile {tile} vs 128: rel L2 {rel:.3e}")
        assert rel <= 1e-13


def test_full_size_3d_p2_workload_properties(dkt):
    """The p = 2 workload of the CG benchmark at full size (3D sphere tree refined
    to level 9: 1,332,192 elements, 8,882,437 nodes) through size-independent
    properties: two tilings of the reference-semantics operator agree to FP64
    rounding and annihilate constants to rounding; the conforming operator is
    symmetric (x.Ay = y.Ax to rounding) and annihilates constants."""
    torch = _torch()
    tree = dkt.balance_tree(dkt.sphere_tree(3, 30, 3, 9, (0.5, 0.5, 0.5), 0.3), 1)
    da = dkt.DA(tree, 2)
    assert da.n_elements == 1332192 and da.n_nodes == 8882437
    x = torch.tensor(dkt.keyed_uniform(0, da.nodes()), device="cuda")
    xb = torch.tensor(dkt.keyed_uniform(1, da.nodes()), device="cuda")
    one = torch.ones_like(x)
    ys = {}
    for tile in (128, 256):
        plan = dkt.Plan(da, dkt.OP_STIFFNESS, tile_leaves=tile)
        y = torch.full_like(x, float("nan"))
        z = torch.full_like(x, float("nan"))
        plan.matvec_torch(x, y)
        plan.matvec_torch(one, z)
        torch.cuda.synchronize()
        assert float(z.abs().max()) <= 1e-12 * float(y.abs().max())
        ys[tile] = y
        del plan
    rel = float(torch.linalg.norm(ys[256] - ys[128]) / torch.linalg.norm(ys[128]))
    print(f"p=2 tiles 256 vs 128: rel L2 {rel:.3e}")
    assert rel <= 1e-13
    conf = dkt.Plan(da, dkt.OP_STIFFNESS, conforming=True)
    ax, ab, a1 = (torch.full_like(x, float("nan")) for _ in range(3))
    conf.matvec_torch(x, ax)
    conf.matvec_torch(xb, ab)
    conf.matvec_torch(one, a1)
    torch.cuda.synchronize()
    s1, s2 = float(torch.dot(xb, ax)), float(torch.dot(x, ab))
    print(f"conforming: x.Ay {s1:.17g} y.Ax {s2:.17g}")
    assert abs(s1 - s2) <= 1e-12 * float(torch.linalg.norm(xb) * torch.linalg.norm(ax))
    assert float(a1.abs().max()) <= 1e-12 * float(ax.abs().max())
