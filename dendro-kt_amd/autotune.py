"""Tile-size and columns-per-launch autotuner for the tiled matvec (no
counterpart in the reference).

The tiled kernel's tile size trades LDS footprint (occupancy) against shared
nodes between tiles.  With several right-hand sides a second choice appears:
how many columns one launch carries.  More columns read the plan's maps once
for all of them; fewer columns run kernels with fewer registers and more
waves per SIMD (one column runs the single-column kernels).  ``autotune``
builds the plan of a sample mesh with each candidate tile size, times the
FP64 and FP32 matvecs with HIP events on the launch stream (with nrhs > 1 as
a one-plan batch for each candidate column count), and returns the fastest
choice per precision.  Results are cached per (dim, order, nrhs, operator,
mesh size class) for the process -- the size class is the element count's
power of four, so a small mesh does not reuse the choice made on a large one
-- and can be stored to a JSON file, so a server tunes once per class and
passes ``tile_leaves`` (and ``columns`` to Batch) to later plans.
"""
import json
import math
import os

CANDIDATES = (64, 128, 256)
_cache = {}


def _column_candidates(nrhs):
    if nrhs <= 1:
        return (0,)
    out, c = [], 1
    while c < nrhs:
        out.append(c)
        c *= 2
    return tuple(out) + (0,)  # 0: as many as fit in LDS


def size_class(n_elements):
    """The mesh size class of the cache key: floor(log4(elements))."""
    return int(math.log(max(1, n_elements), 4))


def autotune(dkt, da, candidates=CANDIDATES, reps=10, warmup=3, nrhs=1, conforming=False):
    """Fastest tile size (and, for nrhs > 1, columns per tile record) per precision
    for matvecs with `nrhs` columns on `da` (reference-semantics plans, or the
    conforming operator): dict(best={"f64": T, "f32": T}, columns={"f64": c,
    "f32": c}, ms={(T, c): {"f64": ms, "f32": ms}})."""
    import torch
    key = (da.dim, da.order, nrhs, int(bool(conforming)), size_class(da.n_elements))
    if key in _cache:
        return _cache[key]
    x64 = torch.tensor(dkt.keyed_uniform(0, da.nodes()), device="cuda")
    if nrhs > 1:
        x64 = torch.stack([x64 * (1.0 + 0.25 * q) for q in range(nrhs)], dim=1).contiguous()
    times = {}
    for T in candidates:
        try:
            plan = dkt.Plan(da, dkt.OP_STIFFNESS, tile_leaves=T, conforming=conforming)
        except dkt.DktError:
            continue  # does not fit in LDS
        for c in _column_candidates(nrhs):
            row = {}
            for name, dt in (("f64", torch.float64), ("f32", torch.float32)):
                x = x64.to(dt)
                y = torch.empty_like(x)
                if nrhs == 1:
                    for _ in range(warmup):
                        plan.matvec_torch(x, y)
                    torch.cuda.synchronize()
                    plan.profile(True)
                    for _ in range(reps):
                        plan.matvec_torch(x, y)
                    _, tot, n = plan.profile_read()
                    plan.profile(False)
                    row[name] = tot / max(n, 1)
                else:
                    b = dkt.Batch([plan], [x], [y], columns=c)
                    st = torch.cuda.current_stream()
                    for _ in range(warmup):
                        b.run(st.cuda_stream)
                    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                    e0.record(st)
                    for _ in range(reps):
                        b.run(st.cuda_stream)
                    e1.record(st)
                    e1.synchronize()
                    row[name] = e0.elapsed_time(e1) / reps
                    del b
            times[(T, c)] = row
        del plan
    if not times:
        raise RuntimeError("no tile size fits")
    best, cols = {}, {}
    for name in ("f64", "f32"):
        T, c = min(times, key=lambda k: times[k][name])
        best[name], cols[name] = T, c
    _cache[key] = {"best": best, "columns": cols, "ms": times}
    return _cache[key]


def save(path):
    out = {}
    for (dim, p, q, conf, sc), v in _cache.items():
        out[f"{dim}d_p{p}_q{q}_c{conf}_s{sc}"] = dict(v, ms={f"{T}_{c}": r for (T, c), r in v["ms"].items()})
    with open(path, "w") as f:
        json.dump(out, f, indent=1)


def load(path):
    if not os.path.exists(path):
        return
    with open(path) as f:
        for k, v in json.load(f).items():
            dim, rest = k.split("d_p")
            p, rest = rest.split("_q")
            q, rest = rest.split("_c")
            conf, sc = rest.split("_s")
            v["best"] = {n: int(t) for n, t in v["best"].items()}
            v["columns"] = {n: int(c) for n, c in v.get("columns", {"f64": 0, "f32": 0}).items()}
            v["ms"] = {tuple(int(a) for a in t.split("_")): r for t, r in v["ms"].items()}
            _cache[(int(dim), int(p), int(q), int(conf), int(sc))] = v
