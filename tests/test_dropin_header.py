"""The C++ drop-in header (include/dftfe_chfsi.hpp) is plain C++14 over the C ABI:
it compiles with the host compiler alone (no hipcc, no deal.II), and it
declares the reference's class and method names.  No GPU calls."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "dftfe_chfsi.hpp")


def test_header_compiles_with_host_compiler(tmp_path):
    gxx = shutil.which("g++")
    if gxx is None or not os.path.isdir("/opt/rocm/include/hip"):
        pytest.skip("g++ or the HIP runtime headers are missing")
    src = tmp_path / "use.cpp"
    src.write_text(
        '#include "dftfe_chfsi.hpp"\n'
        "int main() {\n"
        "  dftfe::dftParameters p;\n"
        "  chfsi_params c = dftfe::toChfsiParams(p);\n"
        "  return c.cheby_block_size == 400 ? 0 : 1;\n"
        "}\n")
    res = subprocess.run([gxx, "-std=c++14", "-fsyntax-only", "-Wall", "-Wextra",
                          "-I/opt/rocm/include", "-D__HIP_PLATFORM_AMD__",
                          "-I", os.path.join(ROOT, "include"), str(src)],
                         capture_output=True, text=True)
    assert res.returncode == 0, res.stderr


def test_header_keeps_reference_names():
    text = open(HEADER).read()
    for name in ["class operatorDFTDeviceClass", "class kohnShamDFTOperatorDeviceClass",
                 "class chebyshevOrthogonalizedSubspaceIterationSolverDevice",
                 "class distributedDeviceVec", "double solve(", "void solveNoRR(",
                 "void reinitSpectrumBounds(", "void HX(", "void HXCheby(",
                 "void reinitkPointSpinIndex(", "const double *getSqrtMassVec()",
                 "const double *getInvSqrtMassVec()", "void computeHamiltonianMatricesAllkpt("]:
        assert name in text, name
