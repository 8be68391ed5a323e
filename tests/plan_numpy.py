"""Numpy interpreter of the flattened traversal (test helper).

Applies the per-leaf maps exported by dkt_da_traversal/dkt_da_hanging_gather
with the semantics of the reference's leaf step (FEM/include/matvec.h:380-522):
gather present nodes, interpolate missing ones from the parent-level nodes
(plus the stale parent-buffer terms), apply the elemental matrix, add
present ranks to their nodes, and add the transposed interpolation of the
missing ranks to the parent-level nodes at ranks the leaf lacks.  Used to pin
the maps against the reference's matvec outputs independently of the GPU code.
"""
import numpy as np


def one_d(p):
    if p == 1:
        K = np.array([[1.0, -1.0], [-1.0, 1.0]])
        M = np.array([[2.0, 1.0], [1.0, 2.0]]) / 6.0
    elif p == 2:
        K = np.array([[7.0, -8.0, 1.0], [-8.0, 16.0, -8.0], [1.0, -8.0, 7.0]]) / 3.0
        M = np.array([[4.0, 2.0, -1.0], [2.0, 16.0, 2.0], [-1.0, 2.0, 4.0]]) / 30.0
    else:
        raise ValueError(p)
    return K, M


def kron_axes(mats):
    """Tensor product with axis 0 fastest: mats[d] acts on axis d."""
    out = np.array([[1.0]])
    for m in mats:
        out = np.kron(m, out)
    return out


def ref_matrix(dim, p, op):
    K, M = one_d(p)
    n = (p + 1) ** dim
    if op == "poisson":
        A = np.zeros((n, n))
        for t in range(dim):
            A += kron_axes([K if d == t else M for d in range(dim)])
        return A, dim - 2
    if op == "mass":
        return kron_axes([M] * dim), dim
    if op == "ident":
        return np.eye(n), None
    raise ValueError(op)


def interp_1d(p, c):
    """ip[k, i] = l_k(child point i) on [0,1] with nodes k/p, child c."""
    n = p + 1
    t = np.arange(n) / p
    xs = 0.5 * (c + t)
    ip = np.ones((n, n))
    for k in range(n):
        for j in range(n):
            if j != k:
                ip[k, :] *= (xs - t[j]) / (t[k] - t[j])
    return ip


def interp_matrix(dim, p, c):
    """W[r, s]: weight of parent rank s in child rank r, child Morton index c."""
    return kron_axes([interp_1d(p, (c >> d) & 1).T for d in range(dim)])


def apply(tr, x, dim, p, maxd, op="poisson", masked_bdy=None):
    A, hp = ref_matrix(dim, p, op)
    x = np.asarray(x, dtype=np.float64).copy()
    if masked_bdy is not None:
        x[masked_bdy] = 0.0
    y = np.zeros_like(x)
    slot, hang, pslot = tr["slot"], tr["hang_of"], tr["pslot"]
    off, hr, hn, hw = tr["hc_off"], tr["hc_rank"], tr["hc_node"], tr["hc_w"]
    npe = slot.shape[1]
    levels = tr["cells"][:, dim].astype(np.int64)
    Ws = [interp_matrix(dim, p, c) for c in range(1 << dim)]
    for e in range(slot.shape[0]):
        s = slot[e]
        u = np.where(s >= 0, x[np.maximum(s, 0)], 0.0)
        h = hang[e]
        if h >= 0:
            ps = pslot[h]
            xp = np.where(ps >= 0, x[np.maximum(ps, 0)], 0.0)
            ui = Ws[tr["child"][e]] @ xp
            a, b = off[h], off[h + 1]
            np.add.at(ui, hr[a:b], hw[a:b] * x[hn[a:b]])
            u = np.where(s >= 0, u, ui)
        scale = 1.0 if hp is None else (2.0 ** (-float(levels[e]))) ** hp
        v = scale * (A @ u)
        pres = s >= 0
        np.add.at(y, s[pres], v[pres])
        if h >= 0:
            vm = np.where(pres, 0.0, v)
            w = Ws[tr["child"][e]].T @ vm
            ps = pslot[h]
            tgt = (ps >= 0) & (~pres)
            np.add.at(y, ps[tgt], w[tgt])
    if masked_bdy is not None:
        y[masked_bdy] = 0.0
    return y
