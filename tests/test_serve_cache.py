"""CPU check of the solve server's admission bookkeeping: the HBM budget, pinned
resident factorizations and least-recently-used eviction (bola_amd/csrc/serve_cache.hpp,
the same header the library's server is built from).  In particular a cache hit whose
right-hand side needs an eviction must never evict its own factorization."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_serve_cache_bookkeeping(tmp_path):
    exe = tmp_path / "serve_cache_check"
    subprocess.run(["g++", "-O1", "-std=c++17", "-Wall", "-Werror", "-I", os.path.join(ROOT, "bola_amd", "csrc"),
                    os.path.join(ROOT, "tests", "serve_cache_check.cpp"), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True)
    assert out.returncode == 0, out.stdout + out.stderr
    assert "all checks passed" in out.stdout
