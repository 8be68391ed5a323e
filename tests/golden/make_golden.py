"""Regenerates the golden fixtures from the reference Dendro-KT.

Runs oracle/_ref/ref_driver_{morton,hilbert} (the reference compiled from its
own sources by oracle/Makefile; it only builds where the reference sources are
present) on the inputs of gen_inputs.py and stores, per case, a compressed
.npz with:

  elems        uint32 [E, dim+1]  balanced tree leaves in the reference's SFC order
  nodes        uint32 [N, dim+1]  DA node coordinates + level, reference node order
  bdy          uint32 [nb]        DA boundary node ids
  y_<op>_<x>_<mask>  float64 [N]  reference matvec outputs
  leaf_cells   uint32 [L, dim+1]  traversal leaves (only small cases)
  leaf_in      float64 [L, nPe]   gathered input with x_i = i (only small cases)
  A_<op>_<mask> float64 [N, N]    assembled operator by unit vectors (tiny cases)

Large arrays are replaced by SHA-256 digests (``sha_<name>``) where noted.
Usage: python tests/golden/make_golden.py [case ...]
"""
import hashlib
import json
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, HERE)
import gen_inputs as gi  # noqa: E402

REFDIR = os.path.join(REPO, "oracle", "_ref")


def sha(a):
    a = np.ascontiguousarray(a)
    return hashlib.sha256(a.tobytes()).hexdigest()


# name: (dim, maxd, order, curve, mode, input-builder, matvecs, leaves, matrices, digest_only)
def cases():
    c = {}
    c["ex1_2d_p1"] = (2, 4, 1, "morton", "tree", lambda: gi.example_tree(1, 2, 3, 4),
                      ["ident:ones:raw", "poisson:hash0:raw", "poisson:hash0:masked", "mass:hash1:raw"],
                      True, ["ident:raw", "poisson:raw"], False)
    c["ex1_2d_p1_hilbert"] = (2, 4, 1, "hilbert", "tree", lambda: gi.example_tree(1, 2, 3, 4),
                              ["poisson:hash0:raw"], True, [], False)
    c["ex3_2d_p1"] = (2, 6, 1, "hilbert", "tree", lambda: gi.example_tree(3, 2, 4, 6),
                      ["ident:ones:raw", "poisson:hash0:raw", "poisson:hash0:masked"], True,
                      ["poisson:raw"], False)
    c["ex1_3d_p1"] = (3, 6, 1, "hilbert", "tree", lambda: gi.example_tree(1, 3, 4, 6),
                      ["ident:ones:raw", "poisson:hash0:raw", "heat:hash0:masked", "heat:hash0:raw"],
                      True, [], False)
    c["ex3_3d_p1"] = (3, 5, 1, "morton", "tree", lambda: gi.example_tree(3, 3, 3, 5),
                      ["ident:ones:raw", "poisson:hash0:raw"], True, ["poisson:raw"], False)
    c["ex1_4d_p1"] = (4, 5, 1, "hilbert", "tree", lambda: gi.example_tree(1, 4, 3, 5),
                      ["ident:ones:raw", "poisson:hash0:raw", "poisson:hash0:masked", "mass:hash1:raw"],
                      True, [], False)
    c["ex2_3d_p2"] = (3, 5, 2, "morton", "tree", lambda: gi.example_tree(2, 3, 2, 5),
                      ["ident:ones:raw", "poisson:hash0:raw"], True, [], False)
    c["ex1_2d_p2"] = (2, 5, 2, "hilbert", "tree", lambda: gi.example_tree(1, 2, 3, 5),
                      ["ident:ones:raw", "poisson:hash0:raw"], True, ["ident:raw", "poisson:raw"], False)
    c["ex3_3d_p2"] = (3, 6, 2, "hilbert", "tree", lambda: gi.example_tree(3, 3, 3, 6),
                      ["poisson:hash0:raw", "mass:hash1:raw"], True, [], False)
    c["sphere3d_small"] = (3, 30, 1, "hilbert", "balance",
                           lambda: gi.sphere_tree(3, 30, 2, 5, (0.5, 0.5, 0.5), 0.3),
                           ["poisson:hash0:raw", "heat:hash0:masked", "poisson:hash0:masked"], True, [], False)
    c["sphere2d_p2"] = (2, 12, 2, "hilbert", "balance",
                        lambda: gi.sphere_tree(2, 12, 2, 6, (0.4, 0.55), 0.27),
                        ["poisson:hash0:raw", "poisson:hash0:masked"], True, [], False)
    c["gauss4d_small"] = (4, 6, 1, "hilbert", "balance", lambda: gi.gauss_points(4, 400, 4, 0.05, 7, 6),
                          ["poisson:hash0:raw", "poisson:hash0:masked", "mass:hash1:raw"], True, [], False)
    c["gauss3d_morton"] = (3, 7, 1, "morton", "balance", lambda: gi.gauss_points(3, 600, 3, 0.08, 3, 7),
                           ["poisson:hash0:raw", "ident:ones:raw"], True, [], False)
    c["gauss2d_p2"] = (2, 10, 2, "morton", "balance",
                       lambda: gi.cells_at_level(gi.gauss_points(2, 300, 2, 0.1, 5, 8), 8, 10),
                       ["poisson:hash0:raw", "ident:ones:raw"], True, [], False)
    # The 3D workload: sphere r=0.3 at the centre, base level 3, max level 7.
    c["sphere3d_l7"] = (3, 30, 1, "hilbert", "balance",
                             lambda: gi.sphere_tree(3, 30, 3, 7, (0.5, 0.5, 0.5), 0.3),
                             ["heat:hash0:masked", "poisson:hash0:raw"], False, [], True)
    return c


def run_case(name, spec, outdir):
    dim, maxd, order, curve, mode, build, matvecs, leaves, matrices, digest_only = spec
    octs = build()
    with tempfile.TemporaryDirectory() as td:
        inp = os.path.join(td, "in.npy")
        np.save(inp, octs)
        cmd = [os.path.join(REFDIR, "ref_driver_" + curve), "--dim", str(dim), "--maxdepth", str(maxd),
               "--order", str(order), "--in", inp, "--mode", mode, "--out", td]
        for m in matvecs:
            cmd += ["--matvec", m]
        for m in matrices:
            cmd += ["--matrix", m]
        if leaves:
            cmd += ["--leaves"]
        r = subprocess.run(cmd, capture_output=True, text=True, timeout=1800)
        if r.returncode != 0:
            raise RuntimeError(f"{name}: driver failed\n{r.stdout}\n{r.stderr}")
        out = {}
        meta = {"dim": dim, "maxdepth": maxd, "order": order, "curve": curve, "mode": mode,
                "driver_log": r.stdout.strip()}
        for fn in sorted(os.listdir(td)):
            if not fn.endswith(".npy") or fn == "in.npy":
                continue
            key = fn[:-4]
            arr = np.load(os.path.join(td, fn))
            if digest_only and key in ("elems", "nodes", "bdy"):
                meta["sha_" + key] = sha(arr)
                meta["n_" + key] = int(arr.shape[0])
            else:
                out[key] = arr
        if not digest_only:
            out["input"] = octs
        out["meta"] = np.array(json.dumps(meta))
        np.savez_compressed(os.path.join(outdir, name + ".npz"), **out)
        print(name, meta["driver_log"])


def main(argv):
    allc = cases()
    names = argv[1:] or list(allc)
    for n in names:
        run_case(n, allc[n], HERE)


if __name__ == "__main__":
    main(sys.argv)
