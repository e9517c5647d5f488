"""Host code and the oracle under AddressSanitizer and UndefinedBehaviorSanitizer (CPU only,
built with g++/gcc here; no device code is involved):

* tests/cpp/asan_check.cpp: mesh generators, topology, geometry, renumbering, adoption of a
  caller's face sets, the OP2 HDF5 reader on the reference's mesh files, reference elements,
  the over-integration rules and tables, and the AMG set-up;
* tests/cpp/oracle_asan.c: every oracle entry point at orders 1..4 (3D 1..3).
Any memory error or undefined behaviour aborts the program, which fails the test.
"""
import glob
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "ins-solver_amd", "csrc")
SAN = ["-O1", "-g", "-fno-omit-frame-pointer", "-fsanitize=address,undefined", "-fno-sanitize-recover=all"]
ENV = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=1", UBSAN_OPTIONS="halt_on_error=1")


def _build(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        pytest.fail("sanitizer build failed:\n" + r.stderr[-3000:])


def test_host_code_under_sanitizers(tmp_path):
    exe = str(tmp_path / "asan_check")
    srcs = [os.path.join(ROOT, "tests", "cpp", "asan_check.cpp")] + [
        os.path.join(CSRC, f) for f in ("mesh.cpp", "refelem.cpp", "oi.cpp", "hdf5_reader.cpp", "amg.cpp")]
    _build(["g++", "-std=c++17", *SAN, "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include", f"-I{CSRC}", *srcs,
            "-o", exe])
    meshes = sorted(glob.glob(os.path.join(ROOT, "tests", "golden", "meshes", "*.h5")))
    r = subprocess.run([exe, *meshes], capture_output=True, text=True, env=ENV, timeout=600)
    assert r.returncode == 0 and r.stdout.strip().endswith("ok"), r.stderr[-3000:]


def test_oracle_under_sanitizers(tmp_path):
    exe = str(tmp_path / "oracle_asan")
    _build(["gcc", "-std=gnu11", *SAN, "-fopenmp", f"-I{os.path.join(ROOT, 'oracle')}",
            os.path.join(ROOT, "tests", "cpp", "oracle_asan.c"), os.path.join(ROOT, "oracle", "ins_oracle.c"),
            "-lm", "-o", exe])
    r = subprocess.run([exe], capture_output=True, text=True, env=ENV, timeout=600)
    assert r.returncode == 0 and r.stdout.strip().endswith("ok"), r.stderr[-3000:]
