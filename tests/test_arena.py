"""The serving HBM arena's allocator, compiled for the CPU with g++ and
checked against a brute-force occupancy map (tests/cpp/test_arena.cpp)."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_arena_allocator(tmp_path):
    exe = str(tmp_path / "test_arena")
    subprocess.run(["g++", "-std=c++17", "-O1", "-Wall", "-fsanitize=address,undefined",
                    "-I", os.path.join(ROOT, "bnmf-algs_amd", "csrc"),
                    os.path.join(ROOT, "tests", "cpp", "test_arena.cpp"), "-o", exe], check=True)
    out = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    assert "arena ok" in out.stdout
