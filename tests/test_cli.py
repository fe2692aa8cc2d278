"""The command-line example (bnmf-algs_amd/examples/main_file_factorization.cpp),
the bld_mult branch of the reference's example/main_file_factorization.cpp:
argument handling on the CPU, a factorization of a matrix file on the GPU
checked against the library's own bld_mult and bld_fact."""
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "bnmf-algs_amd", "lib", "main_file_factorization")


def _exe():
    if not os.path.exists(EXE):
        pytest.fail(f"{EXE} is missing: run __graft_entry__.build() first")
    return EXE


def test_cli_arguments(tmp_path):
    r = subprocess.run([_exe()], capture_output=True, text=True, cwd=tmp_path, timeout=60)
    assert r.returncode != 0 and "usage" in r.stdout
    r = subprocess.run([_exe(), "nmf", "x.txt", "2", "1", "5"], capture_output=True, text=True,
                       cwd=tmp_path, timeout=60)
    assert r.returncode != 0 and "bld_mult" in r.stdout
    (tmp_path / "ragged.txt").write_text("1 2 3\n4 5\n")
    r = subprocess.run([_exe(), "bld_mult", "ragged.txt", "2", "1", "5"], capture_output=True,
                       text=True, cwd=tmp_path, timeout=60)
    assert r.returncode != 0 and "cannot read" in r.stdout


@pytest.mark.gpu
def test_cli_factorization(tmp_path, hip):
    rng = np.random.default_rng(5)
    X = rng.poisson(3.0, (12, 9)).astype(np.float64)
    np.savetxt(tmp_path / "X.txt", X, fmt="%g")
    z, iters = 4, 50
    r = subprocess.run([_exe(), "bld_mult", "X.txt", str(z), "1", str(iters)], capture_output=True,
                       text=True, cwd=tmp_path, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    S = np.loadtxt(tmp_path / "S.txt").reshape(12, 9, z)
    W = np.loadtxt(tmp_path / "W.txt").reshape(12, z)
    H = np.loadtxt(tmp_path / "H.txt").reshape(z, 9)
    L = np.loadtxt(tmp_path / "L.txt").reshape(9)
    beta = [10.0] * z
    beta[-1] = 60.0
    p = hip.Params(100.0, 1.0, [0.05] * 12, beta)
    S_ref = hip.bld_mult(X, z, p, iters, True)
    assert np.array_equal(S, S_ref)          # 17 significant digits round-trip doubles
    W_ref, H_ref, L_ref = hip.bld_fact(S_ref, p)
    assert np.array_equal(W, W_ref) and np.array_equal(H, H_ref) and np.array_equal(L, L_ref)
