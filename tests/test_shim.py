"""The feMatrix shim (include/dkt_fematrix.hpp) against the reference, on the
CPU: compiled with the reference's own headers and library build
(oracle/capture_heatmat.cpp), its operator capture calls the reference's
HeatEq::HeatMat<3>::elementalMatVec on unit vectors per level.  The captured
matrices equal the closed form 2^-level K_ref (the 3D Q1 stiffness) and the
committed fixture that the GPU tests run through DKT_OP_CAPTURED, and the ABI
accepts them (a host-side plan on the same tree, whose DA has the reference's
node count).  Skipped where the reference sources are absent."""
import os
import subprocess

import numpy as np
import pytest

import plan_numpy as pn

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = "/root/reference"


@pytest.fixture(scope="module")
def captured(tmp_path_factory, dkt):
    if not os.path.isdir(os.path.join(REF, "src")):
        pytest.skip("reference sources not present")
    r = subprocess.run(["make", "-C", os.path.join(REPO, "oracle"), "capture"], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-2000:]
    out = str(tmp_path_factory.mktemp("cap") / "K.npy")
    r = subprocess.run([os.path.join(REPO, "oracle", "_ref", "capture_heatmat"), out, "30"], capture_output=True,
                       text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "DA nodes reference 125, dkt 125" in r.stdout
    return np.load(out)


def test_capture_equals_closed_form(captured):
    A, hp = pn.ref_matrix(3, 1, "poisson")
    assert captured.shape == (31, 8, 8)
    for lev in range(31):
        ref = A * 2.0 ** (-lev * hp)
        assert np.abs(captured[lev] - ref).max() <= 1e-15 * np.abs(ref).max(), lev


def test_capture_equals_committed_fixture(captured):
    fixed = np.load(os.path.join(REPO, "tests", "golden", "heatmat3_p1_captured.npy"))
    assert np.array_equal(captured, fixed)


def test_shim_element_coords_match_reference_layout():
    """element_coords is fem::matvec's coordinate layout (matvec.h:459-465):
    node n lexicographic with axis 0 fastest, at 2^-level * i_d / p.  Checked
    through a tiny C++ program compiled against the header."""
    src = r'''
#include <cstdio>
#include "dkt_fematrix.hpp"
int main() {
  auto c = dkt_shim::element_coords(3, 2, 1);
  for (double v : c) printf("%.17g\n", v);
  return 0;
}
'''
    import tempfile
    with tempfile.TemporaryDirectory() as td:
        f = os.path.join(td, "t.cpp")
        open(f, "w").write(src)
        exe = os.path.join(td, "t")
        r = subprocess.run(["g++", "-std=c++14", "-D__HIP_PLATFORM_AMD__", "-I" + os.path.join(REPO, "include"),
                            "-I/opt/rocm/include", f, "-o", exe], capture_output=True, text=True)
        assert r.returncode == 0, r.stderr[-2000:]
        vals = np.array([float(v) for v in subprocess.run([exe], capture_output=True, text=True).stdout.split()])
    c = vals.reshape(27, 3)
    for n in range(27):
        idx = [(n // 3 ** d) % 3 for d in range(3)]
        assert list(c[n]) == [0.5 * i / 2 for i in idx]
