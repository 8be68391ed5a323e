"""dkt_plan_matvec_noalloc (the matvec whose scratch is a fixed device array)
against dkt_plan_matvec, through the C ABI: bitwise equal results on two
calls in a row, the second on a smaller mesh, for the tiled, conforming and
simple paths, FP64 and FP32, one and several columns; a matvec whose scratch
exceeds the fixed array is refused and leaves y untouched.  The first call
also checks the reference's output on the same tree (FP64, 1e-12)."""
import numpy as np
import pytest

import fixtures as fx

pytestmark = pytest.mark.gpu


def _torch():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    return torch


def _pair(dkt, torch, plan, x, masked):
    y_ref = torch.full_like(x, float("nan"))
    y_na = torch.full_like(x, float("nan"))
    plan.matvec_torch(x, y_ref, masked=masked)
    plan.matvec_torch(x, y_na, masked=masked, noalloc=True)
    torch.cuda.synchronize()
    return y_ref.cpu().numpy(), y_na.cpu().numpy()


@pytest.mark.parametrize("kind", ["tiled", "conforming", "simple"])
def test_noalloc_equals_matvec_twice(dkt, kind):
    torch = _torch()
    cases = []
    for name in ("sphere3d_l7", "gauss4d_small"):  # larger mesh first, then a smaller one
        da, z, meta = fx.build_da(dkt, name)
        cases.append((name, da, z, meta))
    assert cases[1][1].n_nodes < cases[0][1].n_nodes
    for call, (name, da, z, meta) in enumerate(cases):
        plan = dkt.Plan(da, dkt.OP_STIFFNESS, simple=(kind == "simple"), conforming=(kind == "conforming"))
        nodes = da.nodes()
        x = torch.tensor(fx.make_x("hash0", nodes, meta["dim"]), dtype=torch.float64, device="cuda")
        for masked in (0, 1):
            y_ref, y_na = _pair(dkt, torch, plan, x, masked)
            print(f"{kind} call {call} {name} masked={masked} N={da.n_nodes}", flush=True)
            assert np.array_equal(y_ref, y_na), (kind, name, masked)
            key = f"y_poisson_hash0_{'masked' if masked else 'raw'}"
            if kind != "conforming" and key in z.files:
                ref = z[key]
                assert np.linalg.norm(y_na - ref) / np.linalg.norm(ref) <= 1e-12
        x3 = torch.stack([x, 2.0 * x, -x], dim=1).float().contiguous()
        y_ref, y_na = _pair(dkt, torch, plan, x3, 0)
        assert np.array_equal(y_ref, y_na), (kind, name, "f32 x3")


def test_noalloc_refuses_oversize(dkt):
    """A matvec needing more scratch than the fixed array is refused with an
    error; nothing runs (y keeps its contents)."""
    torch = _torch()
    pts = dkt.gauss_points(4, 8, 60000, 4, 0.05, 7)
    da = dkt.DA(dkt.balance_tree(pts, 1), 1)
    plan = dkt.Plan(da, dkt.OP_STIFFNESS, simple=True)
    # the simple path's contribution buffer: at least one value per (leaf, rank)
    need = da.count(3) * da.npe * 8
    assert need > dkt.NOALLOC_SCRATCH_BYTES, need
    x = torch.ones(da.n_nodes, dtype=torch.float64, device="cuda")
    y = torch.full_like(x, 7.0)
    with pytest.raises(dkt.DktError, match="exceeds the fixed scratch"):
        plan.matvec_torch(x, y, noalloc=True)
    torch.cuda.synchronize()
    assert bool((y == 7.0).all())
