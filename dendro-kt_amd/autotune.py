"""Tile-size autotuner for the tiled matvec (no counterpart in the reference).

The tiled kernel's tile size trades LDS footprint (occupancy) against shared
nodes between tiles.  ``autotune`` builds the plan of a sample mesh with each
candidate size, times the FP64 and FP32 matvecs with HIP events on the launch
stream, and returns the fastest size per precision.  Results are cached per
(dim, order) for the process and can be stored to a JSON file, so a server
tunes once per (dim, p, columns) and passes ``tile_leaves`` to later plans.
"""
import json
import os

CANDIDATES = (64, 128, 256)
_cache = {}


def autotune(dkt, da, candidates=CANDIDATES, reps=10, warmup=3, nrhs=1):
    """Fastest tile size per precision for matvecs with `nrhs` columns on `da`:
    dict(best={"f64": T, "f32": T}, ms={T: {"f64": ms, "f32": ms}})."""
    import torch
    key = (da.dim, da.order, nrhs)
    if key in _cache:
        return _cache[key]
    x64 = torch.tensor(dkt.keyed_uniform(0, da.nodes()), device="cuda")
    if nrhs > 1:
        x64 = torch.stack([x64 * (1.0 + 0.25 * q) for q in range(nrhs)], dim=1).contiguous()
    times = {}
    for T in candidates:
        try:
            plan = dkt.Plan(da, dkt.OP_STIFFNESS, tile_leaves=T)
        except dkt.DktError:
            continue  # does not fit in LDS
        row = {}
        for name, dt in (("f64", torch.float64), ("f32", torch.float32)):
            x = x64.to(dt)
            y = torch.empty_like(x)
            for _ in range(warmup):
                plan.matvec_torch(x, y)
            torch.cuda.synchronize()
            plan.profile(True)
            for _ in range(reps):
                plan.matvec_torch(x, y)
            _, tot, n = plan.profile_read()
            plan.profile(False)
            row[name] = tot / max(n, 1)
        times[T] = row
        del plan
    if not times:
        raise RuntimeError("no tile size fits")
    best = {name: min(times, key=lambda T: times[T][name]) for name in ("f64", "f32")}
    _cache[key] = {"best": best, "ms": times}
    return _cache[key]


def save(path):
    with open(path, "w") as f:
        json.dump({f"{k[0]}d_p{k[1]}_q{k[2]}": v for k, v in _cache.items()}, f, indent=1)


def load(path):
    if not os.path.exists(path):
        return
    with open(path) as f:
        for k, v in json.load(f).items():
            dim, rest = k.split("d_p")
            p, q = rest.split("_q")
            v["best"] = {n: int(t) for n, t in v["best"].items()}
            v["ms"] = {int(t): r for t, r in v["ms"].items()}
            _cache[(int(dim), int(p), int(q))] = v
