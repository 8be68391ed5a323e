"""ctypes binding of the CPU restatement over the flattened traversal
(plan_cpu.c, built into oracle/_build/liboracle.so by `make restatement`).
Test infrastructure and the CPU baseline: only tests/ and bench.py's
cpu_baseline leg use it."""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "_build", "liboracle.so")
OPS = {"poisson": 0, "heat": 0, "mass": 1, "ident": 2}


def load():
    O = ctypes.CDLL(LIB)
    P = ctypes.c_void_p
    O.oracle_plan_matvec.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_size_t, ctypes.c_size_t] + [P] * 9 + \
        [ctypes.c_int, P, P, ctypes.c_int]
    return O


def matvec(O, tr, dim, order, n_nodes, x, op="poisson", nthreads=1, masked_bdy=None):
    """y = A x with the reference's leaf semantics over the traversal dict
    returned by DA.traversal(); masked_bdy: boundary ids zeroed in x and y
    (feMatrix::matVec with HeatMat's Dirichlet hooks)."""
    x = np.ascontiguousarray(x, dtype=np.float64).copy()
    if masked_bdy is not None:
        x[masked_bdy] = 0.0
    y = np.empty_like(x)
    a = {k: np.ascontiguousarray(v) for k, v in tr.items()}
    rc = O.oracle_plan_matvec(dim, order, n_nodes, len(a["child"]), a["cells"].ctypes.data, a["child"].ctypes.data,
                              a["slot"].ctypes.data, a["hang_of"].ctypes.data, a["pslot"].ctypes.data,
                              a["hc_off"].ctypes.data, a["hc_rank"].ctypes.data, a["hc_node"].ctypes.data,
                              a["hc_w"].ctypes.data, OPS[op], x.ctypes.data, y.ctypes.data, nthreads)
    if rc != 0:
        raise RuntimeError(f"oracle_plan_matvec failed ({rc})")
    if masked_bdy is not None:
        y[masked_bdy] = 0.0
    return y
