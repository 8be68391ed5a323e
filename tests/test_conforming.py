"""Conforming operator (this build's operator for solves at scale; the
reference's static-buffer traversal is neither symmetric nor conforming).

CPU: the C++ constraint terms agree with an independent numpy construction
(tests/conforming_numpy.py), and the assembled operator satisfies what the
mathematics fixes: symmetry, A 1 = 0 (constants are in the space), and the
patch test (a linear function is annihilated at interior nodes, Delta u = 0).
"""
import numpy as np
import pytest

import conforming_numpy as cn
import gen_inputs as gi


def _da(dkt, dim, p, seed):
    maxd = 8 if p == 1 else 10
    if p == 1:
        pts = gi.gauss_points(dim, 80 if dim < 4 else 30, 3, 0.08, seed, maxd)
    else:
        pts = gi.cells_at_level(gi.gauss_points(dim, 40 if dim == 3 else 80, 3, 0.08, seed, maxd - 2), maxd - 2, maxd)
    da = dkt.DA(dkt.balance_tree(dkt.Octs.from_array(pts, dim, maxd, dkt.CURVE_HILBERT), 1), p)
    return da, maxd


CASES = [(2, 1, 1), (3, 1, 2), (4, 1, 3), (2, 2, 4), (3, 2, 5)]


@pytest.mark.parametrize("dim,p,seed", CASES)
def test_constraint_terms_match_numpy(dkt, dim, p, seed):
    da, maxd = _da(dkt, dim, p, seed)
    rows = cn.constraint_rows(da.elements(), da.nodes(), dim, p, maxd)
    ct = da.conforming_terms()
    h = 0
    for e, row in enumerate(rows):
        constrained = [r for r, comb in enumerate(row) if len(comb) != 1 or 1.0 not in comb.values()]
        if not constrained:
            continue
        got = {}
        for k in range(ct["off"][h], ct["off"][h + 1]):
            got.setdefault(int(ct["rank"][k]), {})[int(ct["node"][k])] = ct["w"][k]
        h += 1
        assert sorted(got) == constrained
        for r in constrained:
            want = {k: v for k, v in row[r].items() if v != 0.0}
            assert set(got[r]) == set(want)
            for k, v in want.items():
                assert abs(got[r][k] - v) <= 1e-15
    assert h == len(ct["off"]) - 1


@pytest.mark.parametrize("dim,p,seed", CASES)
def test_conforming_operator_identities(dkt, dim, p, seed):
    da, maxd = _da(dkt, dim, p, seed)
    nodes = da.nodes()
    A = cn.assemble(da.elements(), nodes, dim, p, maxd)
    scale = np.abs(A).max()
    assert np.abs(A - A.T).max() <= 1e-13 * scale
    assert np.abs(A @ np.ones(len(nodes))).max() <= 1e-13 * scale
    c = nodes[:, :dim].astype(np.float64) / (1 << maxd)
    lin = 0.3 + c @ np.linspace(1.0, -0.7, dim)
    interior = np.ones(len(nodes), bool)
    interior[da.boundary()] = False
    assert np.abs((A @ lin)[interior]).max() <= 1e-12 * scale
    if len(nodes) <= 1500:
        ev = np.linalg.eigvalsh(A)
        assert ev[0] > -1e-12 * scale and ev[1] > 1e-9 * scale  # one constant null vector


def _torch():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    return torch


@pytest.mark.gpu
@pytest.mark.parametrize("dim,p,seed", CASES)
def test_conforming_device_matvec_matches_assembled(dkt, dim, p, seed):
    torch = _torch()
    da, maxd = _da(dkt, dim, p, seed)
    nodes = da.nodes()
    A = cn.assemble(da.elements(), nodes, dim, p, maxd)
    x = gi.keyed_uniform(3, nodes[:, :dim])
    plan = dkt.Plan(da, dkt.OP_STIFFNESS, conforming=True)
    xd = torch.tensor(x, device="cuda")
    yd = torch.empty_like(xd)
    plan.matvec_torch(xd, yd)
    y32 = torch.empty_like(xd.float())
    plan.matvec_torch(xd.float(), y32)
    torch.cuda.synchronize()
    ref = A @ x
    assert np.linalg.norm(yd.cpu().numpy() - ref) <= 1e-12 * np.linalg.norm(ref)
    assert np.linalg.norm(y32.double().cpu().numpy() - ref) <= 2e-6 * np.linalg.norm(ref)
    # Dirichlet-masked: D A D with zero rows on the boundary
    ym = torch.empty_like(xd)
    plan.matvec_torch(xd, ym, masked=1)
    torch.cuda.synchronize()
    m = np.ones(len(nodes), bool)
    m[da.boundary()] = False
    refm = np.where(m, A @ np.where(m, x, 0.0), 0.0)
    assert np.linalg.norm(ym.cpu().numpy() - refm) <= 1e-12 * np.linalg.norm(refm)


@pytest.mark.gpu
def test_conforming_cg_converges_where_reference_operator_stalls(dkt):
    """3D p = 2 sphere tree: CG on the conforming operator reaches 1e-8 (FP64 and mixed);
    the solution satisfies the assembled system on the CPU."""
    torch = _torch()
    da = dkt.DA(dkt.balance_tree(dkt.sphere_tree(3, 30, 2, 5, (0.5, 0.5, 0.5), 0.3), 1), 2)
    plan = dkt.Plan(da, dkt.OP_STIFFNESS, conforming=True)
    b = gi.keyed_uniform(1, da.nodes()[:, :3])
    b[da.boundary()] = 0.0
    bd = torch.tensor(b, device="cuda")
    out = {}
    for mode in (dkt.CG_F64, dkt.CG_MIXED):
        x = torch.zeros_like(bd)
        st = plan.cg_solve_torch(bd, x, mode, 20000, 1e-8)
        assert st["resid"] <= 1e-8, st
        out[mode] = x.cpu().numpy()
    r = torch.empty_like(bd)
    plan.matvec_torch(torch.tensor(out[dkt.CG_F64], device="cuda"), r, masked=1)
    torch.cuda.synchronize()
    assert float((bd - r).abs().max() / bd.abs().max()) <= 1.01e-8
    assert np.linalg.norm(out[dkt.CG_MIXED] - out[dkt.CG_F64]) <= 1e-6 * np.linalg.norm(out[dkt.CG_F64])


def _dry_conf_info(dkt, da, tile_leaves=0):
    import ctypes
    L = dkt.load()
    h = ctypes.c_void_p()
    flags = 2 | dkt.PLAN_CONFORMING
    if tile_leaves:
        flags |= (tile_leaves.bit_length() - 1) << dkt.PLAN_TILE_SHIFT
    assert L.dkt_plan_create(da.h, dkt.OP_STIFFNESS, None, flags, ctypes.byref(h)) == 0, L.dkt_last_error()
    keys = 21
    out = (ctypes.c_double * keys)()
    try:
        assert L.dkt_plan_info(h, out, keys) == 0
    finally:
        L.dkt_plan_destroy(h)
    return list(out)


@pytest.mark.parametrize("dim,p,seed", CASES)
@pytest.mark.parametrize("tile", [64, 256])
def test_conforming_tiled_layout(dkt, dim, p, seed, tile):
    """Host-side conforming tiles (dry run): the tiles cover the mesh elements with at
    most `tile` elements (fewer where LDS requires) and 2048 constraint terms each, and every node shared by
    tiles gets one partial slot per tile."""
    da, _ = _da(dkt, dim, p, seed)
    info = _dry_conf_info(dkt, da, tile)
    n_elem = len(da.elements())
    assert info[0] == da.n_nodes and info[1] == n_elem
    assert 16 <= info[19] <= tile  # halved while one FP64 column does not fit in LDS
    assert info[5] * info[19] >= n_elem
    assert 0 < info[20] <= 2048
    assert info[7] >= 2 * info[6]
    assert info[8] <= info[19] * info[3] + 1024  # local points: element nodes and hanging points


@pytest.mark.parametrize("dim,p,seed", CASES)
@pytest.mark.parametrize("conforming", [True, False])
def test_plan_maps_independent_of_operator(dkt, dim, p, seed, conforming):
    """The plans a request kind pairs on one mesh (stiffness and mass; two sets of
    captured per-level matrices) have maps of the same size, which lets the serving
    layer's admission use one dry run per mesh.  Conforming plans tile the same way
    for every operator; on the reference traversal captured matrices take the
    per-element path."""
    da, _ = _da(dkt, dim, p, seed)
    npe = (p + 1) ** dim
    caps = [np.tile(k * np.eye(npe), (10, 1, 1)).ravel() for k in (1.0, 2.0)]

    def size(op, m=None):
        return dkt.Plan(da, op, op_data=m, conforming=conforming, dry_run=True).info()["map_bytes"]

    st, ms = size(dkt.OP_STIFFNESS), size(dkt.OP_MASS)
    c1, c2 = size(dkt.OP_CAPTURED, caps[0]), size(dkt.OP_CAPTURED, caps[1])
    assert st > 0 and ms == st and c1 > 0 and c2 == c1
    if conforming:
        assert c1 == st


@pytest.mark.gpu
@pytest.mark.parametrize("dim,p,seed", CASES)
def test_conforming_tiled_matches_simple(dkt, dim, p, seed):
    """The tiled conforming kernel against the per-element kernels and the assembled
    operator: FP64 within 1e-13, every masking mode, every tile size; three
    right-hand sides give columns bitwise equal to single-RHS calls; repeated calls
    are bitwise identical."""
    torch = _torch()
    da, maxd = _da(dkt, dim, p, seed)
    nodes = da.nodes()
    n = len(nodes)
    A = cn.assemble(da.elements(), nodes, dim, p, maxd)
    m = np.ones(n, bool)
    m[da.boundary()] = False
    simple = dkt.Plan(da, dkt.OP_STIFFNESS, conforming=True, simple=True)
    X = np.stack([gi.keyed_uniform(s, nodes[:, :dim]) for s in (5, 6, 7)], axis=1)
    for tile in (64, 128, 256):
        plan = dkt.Plan(da, dkt.OP_STIFFNESS, conforming=True, tile_leaves=tile)
        for masked in (0, 1, 2):
            xd = torch.tensor(X[:, 0].copy(), device="cuda")
            ya, yb, yc = (torch.empty_like(xd) for _ in range(3))
            plan.matvec_torch(xd, ya, masked=masked)
            plan.matvec_torch(xd, yb, masked=masked)
            simple.matvec_torch(xd, yc, masked=masked)
            torch.cuda.synchronize()
            ya, yb, yc = ya.cpu().numpy(), yb.cpu().numpy(), yc.cpu().numpy()
            x = X[:, 0]
            if masked == 0:
                ref = A @ x
            elif masked == 1:
                ref = np.where(m, A @ np.where(m, x, 0.0), 0.0)
            else:
                ref = np.where(m, A @ x, 0.0)
            assert np.array_equal(ya, yb)
            assert np.linalg.norm(ya - yc) <= 1e-13 * np.linalg.norm(yc)
            assert np.linalg.norm(ya - ref) <= 1e-12 * np.linalg.norm(ref)
        Xd = torch.tensor(X, device="cuda")
        Y = torch.empty_like(Xd)
        plan.matvec_torch(Xd, Y, masked=1)
        torch.cuda.synchronize()
        for c in range(3):
            yc = torch.empty(n, dtype=torch.float64, device="cuda")
            plan.matvec_torch(Xd[:, c].contiguous(), yc, masked=1)
            torch.cuda.synchronize()
            assert torch.equal(Y[:, c], yc)


@pytest.mark.parametrize("dim,p,seed", [(2, 1, 1), (3, 1, 2), (3, 2, 5)])
def test_write_vtu_linear_field(dkt, dim, p, seed, tmp_path):
    """VTU output: one cell per element (quad / hexahedron), corner points in
    [0,1]^dim, cell levels, and a linear nodal field reproduced exactly at every
    corner, hanging corners included (the conforming interpolant of a linear
    function is the function)."""
    import xml.etree.ElementTree as ET
    da, maxd = _da(dkt, dim, p, seed)
    nodes = da.nodes()
    coef = np.linspace(0.7, -0.4, dim)
    f = 0.25 + (nodes[:, :dim] / float(1 << maxd)) @ coef
    path = tmp_path / "mesh.vtu"
    da.write_vtu(path, f, "phi")
    root = ET.parse(path).getroot()
    piece = root.find("UnstructuredGrid/Piece")
    ne = len(da.elements())
    nc = 1 << dim
    assert int(piece.get("NumberOfCells")) == ne and int(piece.get("NumberOfPoints")) == ne * nc
    arrays = {a.get("Name") or "points": np.array(a.text.split(), dtype=float) for a in piece.iter("DataArray")}
    pts = arrays["points"].reshape(-1, 3)
    assert pts.min() >= 0.0 and pts.max() <= 1.0
    assert np.all(arrays["types"] == (9 if dim == 2 else 12))
    assert np.array_equal(arrays["level"], da.elements()[:, dim].astype(float))
    want = 0.25 + pts[:, :dim] @ coef
    assert np.abs(arrays["phi"] - want).max() <= 1e-13
    # VTK corner order: the second point of every cell differs from the first along x only
    d01 = pts[1::nc] - pts[0::nc]
    assert np.all(d01[:, 0] > 0) and np.all(d01[:, 1:] == 0)


def test_write_vtu_refuses_4d(dkt):
    da, _ = _da(dkt, 4, 1, 3)
    with pytest.raises(dkt.DktError):
        da.write_vtu("/dev/null")


@pytest.mark.gpu
@pytest.mark.parametrize("dim,p,seed", CASES)
def test_conforming_captured_tiled_matches_stiffness(dkt, dim, p, seed):
    """Per-level captured element matrices (DKT_OP_CAPTURED) on a conforming plan
    run the tiled conforming kernel with a dense element product.  Fed the
    stiffness of every level, 2^(-level (dim - 2)) K_ref, they reproduce the
    closed-form conforming stiffness (FP64 within 1e-12 in every masking mode,
    FP32 within 2e-5 of it); repeated calls, multi-RHS columns and a batch with
    both plans are bitwise equal to single calls."""
    import plan_numpy as pn
    torch = _torch()
    da, maxd = _da(dkt, dim, p, seed)
    nodes = da.nodes()
    n = len(nodes)
    K, _ = pn.ref_matrix(dim, p, "poisson")
    mats = np.stack([K * 2.0 ** (-lev * (dim - 2)) for lev in range(maxd + 1)])
    cap = dkt.Plan(da, dkt.OP_CAPTURED, op_data=mats, conforming=True)
    stiff = dkt.Plan(da, dkt.OP_STIFFNESS, conforming=True)
    assert cap.info()["tiles"] > 0  # the tiled kernel, not the per-element one
    X = np.stack([gi.keyed_uniform(s, nodes[:, :dim]) for s in (8, 9)], axis=1)
    xd = torch.tensor(X[:, 0].copy(), device="cuda")
    for masked in (0, 1, 2):
        ya, yb, ys = (torch.empty_like(xd) for _ in range(3))
        cap.matvec_torch(xd, ya, masked=masked)
        cap.matvec_torch(xd, yb, masked=masked)
        stiff.matvec_torch(xd, ys, masked=masked)
        torch.cuda.synchronize()
        assert torch.equal(ya, yb)
        assert float(torch.linalg.norm(ya - ys) / torch.linalg.norm(ys)) <= 1e-12
        y32 = torch.empty(n, dtype=torch.float32, device="cuda")
        cap.matvec_torch(xd.float(), y32, masked=masked)
        torch.cuda.synchronize()
        assert float(torch.linalg.norm(y32.double() - ys) / torch.linalg.norm(ys)) <= 2e-5
    Xd = torch.tensor(X, device="cuda")
    Y = torch.empty_like(Xd)
    cap.matvec_torch(Xd, Y, masked=1)
    y1 = torch.empty(n, dtype=torch.float64, device="cuda")
    cap.matvec_torch(Xd[:, 1].contiguous(), y1, masked=1)
    torch.cuda.synchronize()
    assert torch.equal(Y[:, 1], y1)
    xb = [torch.tensor(X[:, 0].copy(), device="cuda") for _ in range(2)]
    yb = [torch.empty_like(xb[0]) for _ in range(2)]
    b = dkt.Batch([cap, stiff], xb, yb, masked=True)
    b.run()
    torch.cuda.synchronize()
    for plan, y in ((cap, yb[0]), (stiff, yb[1])):
        ref = torch.empty_like(y)
        plan.matvec_torch(xb[0], ref, masked=True)
        torch.cuda.synchronize()
        assert torch.equal(y, ref)
