"""Host check of the vertex-space smoothed-aggregation AMG (product code in
ins-solver_amd/csrc/amg.cpp), compiled with g++ against a textbook P1 stiffness
matrix built in tests/cpp/amg_check.cpp: Galerkin coarse operators, R = P^T,
operator complexity, and h-robust PCG iteration counts with one V-cycle."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "ins-solver_amd", "csrc")


@pytest.fixture(scope="module")
def amg_check(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("amg") / "amg_check")
    cmd = ["g++", "-std=c++17", "-O2", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include", f"-I{CSRC}",
           os.path.join(ROOT, "tests", "cpp", "amg_check.cpp"), os.path.join(CSRC, "amg.cpp"),
           os.path.join(CSRC, "mesh.cpp"), os.path.join(CSRC, "refelem.cpp"), "-o", exe]
    subprocess.run(cmd, check=True, capture_output=True, text=True)
    return exe


def run(exe, dim, n):
    out = subprocess.run([exe, str(dim), str(n)], check=True, capture_output=True, text=True,
                         timeout=300).stdout.split()
    d, nn, rows, levels, iters, opc, gerr = out[:7]
    return {"rows": int(rows), "levels": int(levels), "iters": int(iters), "opc": float(opc),
            "gerr": float(gerr)}


@pytest.mark.parametrize("dim,sizes", [(2, (24, 96)), (3, (8, 24))])
def test_amg_galerkin_and_h_robust(amg_check, dim, sizes):
    res = [run(amg_check, dim, n) for n in sizes]
    for r in res:
        assert r["gerr"] < 1e-13, r          # A_c = P^T A P and R = P^T to rounding
        assert r["opc"] < 2.0, r              # operator complexity
        assert r["iters"] <= 20, r            # PCG with one V-cycle to 1e-10
    assert res[-1]["levels"] > 2              # the finer mesh gets a real hierarchy
    assert res[-1]["iters"] - res[0]["iters"] <= 3, res   # 16x (2D) / 27x (3D) more unknowns
