"""The command-line example (bnmf-algs_amd/examples/main_file_factorization.cpp),
the bld_mult, bld_add and bld_appr branches of the reference's
example/main_file_factorization.cpp: argument handling on the CPU, a
factorization of a matrix file per branch on the GPU checked against the
library's own solvers and bld_fact, and the output layout of the reference's
Eigen stream (x lines of y*z values for S, 6 significant digits, one global
column width). The layout is restated from Eigen's documented stream format;
Eigen is absent here, so its exact bytes are parity unpinned."""
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
    assert r.returncode != 0 and "bld_mult, bld_add and bld_appr" in r.stdout
    (tmp_path / "ragged.txt").write_text("1 2 3\n4 5\n")
    r = subprocess.run([_exe(), "bld_mult", "ragged.txt", "2", "1", "5"], capture_output=True,
                       text=True, cwd=tmp_path, timeout=60)
    assert r.returncode != 0 and "cannot read" in r.stdout


def _eigen_text(m):
    """Eigen's default stream output of a row-major block: %g with 6
    significant digits (the stream default), right-aligned to the widest
    value, one space between values, newline between rows, none at the end."""
    m = np.atleast_2d(m)
    txt = [[f"{v:g}" for v in row] for row in m]
    w = max(len(t) for row in txt for t in row)
    return "\n".join(" ".join(t.rjust(w) for t in row) for row in txt)


def _run(tmp_path, alg, z, iters, env=None):
    r = subprocess.run([_exe(), alg, "X.txt", str(z), "1", str(iters)], capture_output=True, text=True,
                       cwd=tmp_path, timeout=120, env=env)
    assert r.returncode == 0, r.stdout + r.stderr


def _params(hip, x, z):
    beta = [10.0] * z
    beta[-1] = 60.0
    return hip.Params(100.0, 1.0, [0.05] * x, beta)


@pytest.mark.gpu
@pytest.mark.parametrize("alg", ["bld_mult", "bld_add", "bld_appr"])
def test_cli_factorization(tmp_path, hip, alg):
    rng = np.random.default_rng(5)
    x, y, z, iters = 12, 9, 4, 50
    X = rng.poisson(3.0, (x, y)).astype(np.float64)
    np.savetxt(tmp_path / "X.txt", X, fmt="%g")
    p = _params(hip, x, z)
    if alg == "bld_mult":
        S_ref = hip.bld_mult(X, z, p, iters, True)
    elif alg == "bld_add":
        S_ref = hip.bld_add(X, z, p, iters)
    else:
        S_ref = hip.bld_appr(X, z, p, iters)[0]
    W_ref, H_ref, L_ref = hip.bld_fact(S_ref, p)
    # full precision: the values round-trip
    _run(tmp_path, alg, z, iters, env={**os.environ, "BNMF_CLI_PRECISION": "17"})
    S = np.loadtxt(tmp_path / "S.txt")
    assert S.shape == (x, y * z)
    assert np.array_equal(S.reshape(x, y, z), S_ref)
    W = np.loadtxt(tmp_path / "W.txt").reshape(x, z)
    H = np.loadtxt(tmp_path / "H.txt").reshape(z, y)
    L = np.loadtxt(tmp_path / "L.txt").reshape(y)
    assert np.array_equal(W, W_ref) and np.array_equal(H, H_ref) and np.array_equal(L, L_ref)
    # the default: the reference's stream layout and precision
    env = {k: v for k, v in os.environ.items() if k != "BNMF_CLI_PRECISION"}
    _run(tmp_path, alg, z, iters, env=env)
    assert (tmp_path / "S.txt").read_text() == _eigen_text(S_ref.reshape(x, y * z))
    assert (tmp_path / "W.txt").read_text() == _eigen_text(W_ref)
    assert (tmp_path / "H.txt").read_text() == _eigen_text(H_ref)
    assert (tmp_path / "L.txt").read_text() == _eigen_text(L_ref.reshape(1, y))
