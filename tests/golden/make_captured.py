"""Writes tests/golden/heatmat3_p1_captured.npy: the per-level element matrices
of the reference's HeatEq::HeatMat<3> (p = 1, levels 0..30) captured through
the feMatrix shim (include/dkt_fematrix.hpp) by oracle/capture_heatmat.cpp,
which links the reference build of oracle/Makefile.  Needs the reference
sources; run from the repository root: python tests/golden/make_captured.py"""
import os
import shutil
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))


def main():
    subprocess.run(["make", "-C", os.path.join(REPO, "oracle"), "capture"], check=True)
    with tempfile.TemporaryDirectory() as td:
        out = os.path.join(td, "K.npy")
        subprocess.run([os.path.join(REPO, "oracle", "_ref", "capture_heatmat"), out, "30"], check=True)
        shutil.copy(out, os.path.join(HERE, "heatmat3_p1_captured.npy"))
    return 0


if __name__ == "__main__":
    sys.exit(main())
