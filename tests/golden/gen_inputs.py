"""Synthetic mesh inputs and the node-keyed input vector, in numpy.

Test infrastructure: the fixture script feeds these to the reference driver,
and the tests check that the product's own C++ generators produce the same
arrays bit for bit.

* ``example_tree`` restates the three closed-form example trees of the
  reference tests (test/testAdaptiveExamples.h:36-171).
* ``sphere_tree`` is the refinement rule of the 3D workload: split every cell
  above ``base_level``, and cells above ``max_level`` whose box meets the
  sphere surface (min corner distance <= r <= max corner distance).
* ``gauss_points`` draws points of a K-component Gaussian mixture with integer
  arithmetic only (Irwin-Hall sum of 12 uniforms per coordinate), so numpy and
  C++ produce identical integer coordinates.
* ``keyed_uniform`` is the U[-1,1) input value of a node keyed by its integer
  coordinates (SplitMix64 finaliser), identical to oracle/npy_io.h.
"""
import numpy as np

M64 = (1 << 64) - 1
GOLDEN = 0x9E3779B97F4A7C15


def _mix64_arr(z):
    z = z.astype(np.uint64)
    with np.errstate(over="ignore"):
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return z


def keyed_uniform(seed, coords):
    """coords: uint32 array [n, dim] -> float64 [n] in [-1, 1)."""
    coords = np.asarray(coords, dtype=np.uint64)
    n, dim = coords.shape
    with np.errstate(over="ignore"):
        s = _mix64_arr(np.full(n, (seed + GOLDEN) & M64, dtype=np.uint64))
        for d in range(dim):
            s = _mix64_arr(s ^ (coords[:, d] * np.uint64(GOLDEN) + np.uint64(d)))
    return 2.0 * ((s >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)) - 1.0


def _child(anchor, lev, ch, dim, maxd):
    h = 1 << (maxd - lev - 1)
    return tuple(anchor[d] + (h if (ch >> d) & 1 else 0) for d in range(dim)), lev + 1


def example_tree(ex, dim, end_l, maxd):
    """Leaves of Example1/2/3 (test/testAdaptiveExamples.h), uint32 [n, dim+1]."""
    out = []
    nch = 1 << dim
    root = (0,) * dim
    if ex == 1:
        def corner(a, lev, ch):
            if lev >= end_l:
                out.append(list(a) + [lev])
                return
            for o in range(nch):
                if o != ch:
                    c, l = _child(a, lev, o, dim, maxd)
                    out.append(list(c) + [l])
            c, l = _child(a, lev, ch, dim, maxd)
            corner(c, l, ch)
        for ch in range(nch):
            c, l = _child(root, 0, ch, dim, maxd)
            corner(c, l, nch - 1 - ch)
    elif ex == 2:
        def fill(a, lev):
            if lev >= end_l:
                out.append(list(a) + [lev])
                return
            for ch in range(nch):
                c, l = _child(a, lev, ch, dim, maxd)
                fill(c, l)
        fill(root, 0)
    elif ex == 3:
        top = 1 << maxd

        def touching(a, lev):
            h = 1 << (maxd - lev)
            return any(a[d] == 0 or a[d] + h == top for d in range(dim))

        def sub(a, lev):
            if lev >= end_l:
                out.append(list(a) + [lev])
                return
            for ch in range(nch):
                c, l = _child(a, lev, ch, dim, maxd)
                if touching(c, l):
                    sub(c, l)
                else:
                    out.append(list(c) + [l])
        for ch in range(nch):
            c, l = _child(root, 0, ch, dim, maxd)
            sub(c, l)
    else:
        raise ValueError(ex)
    return np.array(out, dtype=np.uint32)


def sphere_split(anchor, lev, maxd, base_level, max_level, center, radius):
    """Refinement predicate of the sphere workload (floating point, IEEE double)."""
    if lev < base_level:
        return True
    if lev >= max_level:
        return False
    scale = 1.0 / float(1 << maxd)
    h = float(1 << (maxd - lev)) * scale
    mind2 = 0.0
    maxd2 = 0.0
    for d in range(len(anchor)):
        a = float(anchor[d]) * scale
        lo = a - center[d]
        hi = a + h - center[d]
        gap = max(0.0, max(lo, -hi))
        mind2 += gap * gap
        maxd2 += max(lo * lo, hi * hi)
    r2 = radius * radius
    return mind2 <= r2 and r2 <= maxd2


def sphere_tree(dim, maxd, base_level, max_level, center, radius):
    """Complete (not yet balanced) tree refined around a sphere surface."""
    out = []
    nch = 1 << dim

    def rec(a, lev):
        if lev > 0 and not sphere_split(a, lev, maxd, base_level, max_level, center, radius):
            out.append(list(a) + [lev])
            return
        for ch in range(nch):
            c, l = _child(a, lev, ch, dim, maxd)
            rec(c, l)
    rec((0,) * dim, 0)
    return np.array(out, dtype=np.uint32)


def _u32_stream(seed, count):
    idx = np.arange(count, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = _mix64_arr(np.uint64((seed * GOLDEN) & M64) ^ (idx * np.uint64(GOLDEN) + np.uint64(0x632BE59BD9B4E019)))
    return (z >> np.uint64(32)).astype(np.uint64)


def gauss_points(dim, n, ncomp, sigma, seed, maxd):
    """Points of a Gaussian mixture in [0,1)^dim as level-maxd integer cells.

    Stream layout: first ncomp*dim centre draws, then per point one component
    draw followed by 12*dim uniform draws.  Returns uint32 [n, dim+1] with
    level column = maxd.
    """
    per = 1 + 12 * dim
    u = _u32_stream(seed, ncomp * dim + n * per)
    centres = (u[: ncomp * dim].astype(np.float64) * (1.0 / 4294967296.0)).reshape(ncomp, dim)
    body = u[ncomp * dim:].reshape(n, per)
    comp = (body[:, 0] % np.uint64(ncomp)).astype(np.int64)
    sums = body[:, 1:].reshape(n, dim, 12).sum(axis=2).astype(np.int64)
    z = (sums - 6 * 4294967296).astype(np.float64) * (1.0 / 4294967296.0)
    x = centres[comp] + sigma * z
    top = float(1 << maxd)
    xi = np.floor(np.clip(x, 0.0, 1.0) * top)
    xi = np.minimum(xi, top - 1.0).astype(np.uint32)
    out = np.empty((n, dim + 1), dtype=np.uint32)
    out[:, :dim] = xi
    out[:, dim] = maxd
    return out


def cells_at_level(points, level, maxd):
    """Distinct level-`level` cells containing the given level-`level` points,
    re-expressed on a 2^maxd grid (maxd > level keeps order-2 node coordinates
    integral)."""
    dim = points.shape[1] - 1
    u = np.unique(points[:, :dim], axis=0).astype(np.uint32) << np.uint32(maxd - level)
    out = np.empty((u.shape[0], dim + 1), dtype=np.uint32)
    out[:, :dim] = u
    out[:, dim] = level
    return out
