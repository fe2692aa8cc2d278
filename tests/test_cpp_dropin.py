"""The C++ drop-in API (bnmf-algs_amd/include/bnmf_algs) compiled with g++
against libbnmf_hip.so, running the reference's own bld_mult / bld_mult_cuda
cases (test/bld/bld_mult/test_bld_mult.cpp:10-90). The parameter checks run
on the CPU; the solves need the GPU."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIBDIR = os.path.join(ROOT, "bnmf-algs_amd", "lib")


def _build(tmp_path):
    if not os.path.exists(os.path.join(LIBDIR, "libbnmf_hip.so")):
        pytest.fail("libbnmf_hip.so is missing: run __graft_entry__.build() first")
    exe = str(tmp_path / "test_bld_mult_dropin")
    subprocess.run(["g++", "-std=c++17", "-O1", "-Wall", "-Wextra", "-Werror",
                    "-I", os.path.join(ROOT, "include"),
                    "-I", os.path.join(ROOT, "bnmf-algs_amd", "include"),
                    os.path.join(ROOT, "tests", "cpp", "test_bld_mult_dropin.cpp"),
                    "-L", LIBDIR, "-lbnmf_hip", f"-Wl,-rpath,{LIBDIR}",
                    "-L/opt/rocm/lib", "-Wl,-rpath,/opt/rocm/lib", "-o", exe], check=True)
    return exe


def test_dropin_parameter_checks(tmp_path):
    exe = _build(tmp_path)
    out = subprocess.run([exe, "cpu"], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    assert "0 failed" in out.stdout


@pytest.mark.gpu
def test_dropin_reference_cases(tmp_path):
    exe = _build(tmp_path)
    out = subprocess.run([exe, "gpu"], capture_output=True, text=True, timeout=300)
    print(out.stdout)
    assert out.returncode == 0, out.stdout + out.stderr
    assert "0 failed" in out.stdout
