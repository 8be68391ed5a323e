"""The stabilised space-time heat operator of the SC19 paper
(papers/sc19/background.tex:26-53): on the 4D k-tree, three space axes
(x, y, z = tree axes 0-2) and time (axis 3), Q1 elements, the discrete form

    a(u, v) = (d_t u, v) + delta (d_t u, d_t v) + kappa (grad u, grad v)
              + delta kappa (grad u, grad d_t v),
    l(v)    = (f, v) + delta (f, d_t v),

with delta = delta0 h proportional to the element size h (the paper: "a mesh
dependent parameter proportional to the mesh size"), essential boundary on
the spatial boundary and at t = 0 (the final time t = 1 is free).  The
reference holds the formulation only (no code), so the element matrices are
built here from the 1D matrices of [0, 1] (row = test function, column =
trial function; M mass, K stiffness, C[i][j] = int phi_i phi_j'):

    A(h) = h^3 (M M M C + delta0 M M M K)
           + kappa h^2 (sum_d K_d M M M + delta0 sum_d K_d M M C^T),
    F(h) = h^4 (M M M M + delta0 M M M C^T)           (the right-hand side),

tensor products with axis 0 fastest, time last.  The operator is not
symmetric; it runs as per-level captured matrices (DKT_OP_CAPTURED) and is
solved with BiCGStab.  The paper writes the last term of a without kappa
(kappa = 1 there); with kappa it is the term that keeps the form consistent,
(d_t u - kappa lap u, v + delta d_t v) = (f, v + delta d_t v).  The tests pin
the matrices by direct quadrature of the bilinear form and by consistency:
the Q1 interpolant of an exact solution leaves a zero residual at every
node off the essential boundary.
"""
import numpy as np


def one_d():
    """1D Q1 matrices on [0, 1]: mass, stiffness, C[i][j] = int phi_i phi_j'."""
    M = np.array([[2.0, 1.0], [1.0, 2.0]]) / 6.0
    K = np.array([[1.0, -1.0], [-1.0, 1.0]])
    C = np.array([[-0.5, 0.5], [-0.5, 0.5]])
    return M, K, C


def kron_axes(mats):
    """Tensor product with axis 0 fastest: mats[d] acts on axis d."""
    out = np.array([[1.0]])
    for m in mats:
        out = np.kron(m, out)
    return out


def unit_terms():
    """The four terms of a(u, v) on the unit 4D cube (kappa = delta = 1):
    (d_t u, v), (d_t u, d_t v), (grad u, grad v), (grad u, grad d_t v)."""
    M, K, C = one_d()
    t1 = kron_axes([M, M, M, C])
    t2 = kron_axes([M, M, M, K])
    t3 = sum(kron_axes([K if d == s else M for d in range(3)] + [M]) for s in range(3))
    t4 = sum(kron_axes([K if d == s else M for d in range(3)] + [C.T]) for s in range(3))
    return t1, t2, t3, t4


def level_matrices(maxdepth, kappa=1.0, delta0=1.0):
    """[maxdepth + 1][16][16]: the element matrix of a(u, v) for an element of
    each level (h = 2^-level), row-major out = A in (DKT_OP_CAPTURED layout)."""
    t1, t2, t3, t4 = unit_terms()
    out = np.empty((maxdepth + 1, 16, 16))
    for lev in range(maxdepth + 1):
        h = 2.0 ** (-lev)
        out[lev] = h ** 3 * (t1 + delta0 * t2) + kappa * h ** 2 * (t3 + delta0 * t4)
    return out


def rhs_matrices(maxdepth, delta0=1.0):
    """[maxdepth + 1][16][16]: the element matrix of l(v) applied to the nodal
    values of f (f in the Q1 space), per level."""
    M, K, C = one_d()
    base = kron_axes([M, M, M, M])
    stab = kron_axes([M, M, M, C.T])
    out = np.empty((maxdepth + 1, 16, 16))
    for lev in range(maxdepth + 1):
        h = 2.0 ** (-lev)
        out[lev] = h ** 4 * (base + delta0 * stab)
    return out


def dirichlet_nodes(nodes, maxdepth):
    """Ids of the nodes on the spatial boundary (x, y or z at 0 or 1) or at
    t = 0: the essential boundary of the space-time form (t = 1 is free)."""
    top = 1 << maxdepth
    sp = nodes[:, :3]
    on_space = np.any((sp == 0) | (sp == top), axis=1)
    at_start = nodes[:, 3] == 0
    return np.nonzero(on_space | at_start)[0].astype(np.uint32)
