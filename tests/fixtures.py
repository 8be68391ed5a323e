"""Loads the golden fixtures and rebuilds the same mesh with the product."""
import glob
import json
import os

import numpy as np

import gen_inputs as gi

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def names():
    return sorted(os.path.basename(f)[:-4] for f in glob.glob(os.path.join(GOLDEN, "*.npz")))


def load(name):
    z = np.load(os.path.join(GOLDEN, name + ".npz"))
    meta = json.loads(str(z["meta"]))
    return z, meta


def fixture_input(name, z):
    if "input" in z.files:
        return z["input"]
    if name == "sphere3d_l7":
        return gi.sphere_tree(3, 30, 3, 7, (0.5, 0.5, 0.5), 0.3)
    raise KeyError(name)


def build_da(dkt, name):
    z, meta = load(name)
    curve = dkt.CURVE_HILBERT if meta["curve"] == "hilbert" else dkt.CURVE_MORTON
    o = dkt.Octs.from_array(fixture_input(name, z), meta["dim"], meta["maxdepth"], curve)
    t = dkt.balance_tree(o, 1) if meta["mode"] == "balance" else dkt.sort_tree(o)
    return dkt.DA(t, meta["order"]), z, meta


def y_cases(z):
    """(key, op, x-spec, masked) for every reference output in a fixture."""
    out = []
    for k in z.files:
        if k.startswith("y_"):
            _, op, xs, mask = k.split("_")
            out.append((k, op, xs, mask == "masked"))
    return out


def make_x(xs, nodes, dim):
    if xs == "ones":
        return np.ones(nodes.shape[0])
    if xs.startswith("hash"):
        return gi.keyed_uniform(int(xs[4:]), nodes[:, :dim])
    raise ValueError(xs)


def op_code(dkt, op):
    return {"poisson": dkt.OP_STIFFNESS, "heat": dkt.OP_STIFFNESS, "mass": dkt.OP_MASS,
            "ident": dkt.OP_IDENTITY}[op]
