"""Independent numpy construction of the conforming operator (test oracle).

The continuous piecewise-polynomial space on a 2:1 balanced tree: the value at
a node position of a leaf that is not a mesh node is the parent cell's
Lagrange interpolant there, recursively through coarser cells, until every
point is a node.  C_e maps mesh nodes to the element's nodal values, and
A = sum_e C_e^T (2^(-level hp) K_ref) C_e over the flattened traversal leaves
(virtual subdivisions of coarse boundary elements included: interpolated
children sum to the parent's matrix exactly).
"""
import numpy as np

import plan_numpy as pn


def constraint_rows(cells, nodes, dim, p, maxd):
    """Per mesh element: list of nPe sparse rows {node: weight}."""
    at = {tuple(int(v) for v in n[:dim]): i for i, n in enumerate(nodes)}
    ip = [pn.interp_1d(p, 0), pn.interp_1d(p, 1)]  # ip[c][k, i] = l_k(child point i)
    n1 = p + 1
    memo = {}

    def resolve(x, a, lev):
        key = (x, lev)
        if key in memo:
            return memo[key]
        if x in at:
            out = {at[x]: 1.0}
        else:
            assert lev > 0, "root node missing"
            ln = 1 << (maxd - lev)
            step, plen = ln // p, 2 * ln
            pa = tuple(ai & ~(plen - 1) for ai in a)
            c = [(a[d] - pa[d]) // ln for d in range(dim)]
            t = [(x[d] - a[d]) // step for d in range(dim)]
            out = {}
            for u in range(n1 ** dim):
                w, y, uu = 1.0, [], u
                for d in range(dim):
                    ud = uu % n1
                    uu //= n1
                    w *= ip[c[d]][ud, t[d]]
                    y.append(pa[d] + ud * (plen // p))
                if w == 0.0:
                    continue
                for k, v in resolve(tuple(y), pa, lev - 1).items():
                    out[k] = out.get(k, 0.0) + w * v
        memo[key] = out
        return out

    rows = []
    for e in range(cells.shape[0]):
        a = tuple(int(v) for v in cells[e, :dim])
        lev = int(cells[e, dim])
        step = (1 << (maxd - lev)) // p
        row = []
        for r in range(n1 ** dim):
            x, rr = [], r
            for d in range(dim):
                x.append(a[d] + (rr % n1) * step)
                rr //= n1
            row.append(resolve(tuple(x), a, lev))
        rows.append(row)
    return rows


def assemble(cells, nodes, dim, p, maxd, op="poisson", level_mats=None):
    """Dense conforming operator over the mesh elements (small meshes only);
    level_mats[level] replaces the element matrix (e.g. a nonsymmetric form)."""
    K, hp = pn.ref_matrix(dim, p, op)
    n = nodes.shape[0]
    A = np.zeros((n, n))
    rows = constraint_rows(cells, nodes, dim, p, maxd)
    for e, row in enumerate(rows):
        lev = int(cells[e, dim])
        if level_mats is not None:
            K, hp = level_mats[lev], None
        scale = 1.0 if hp is None else (2.0 ** (-float(lev))) ** hp
        C = np.zeros((len(row), n))
        for r, comb in enumerate(row):
            for k, w in comb.items():
                C[r, k] += w
        A += scale * (C.T @ K @ C)
    return A
