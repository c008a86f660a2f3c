"""CPU checks of the C ABI: the library loads and exports exactly what include/bola_hip.h
declares, and the Python mirror binds every declared symbol.  No GPU work is done."""
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "bola_hip.h")


def header_symbols():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(bola_hip_[a-z0-9_]+)\s*\(", text)))


@pytest.fixture(scope="module")
def built():
    import bola_amd
    if not os.path.exists(bola_amd.library_path()):
        bola_amd.build()
    return bola_amd


def test_header_declares_the_path(built):
    syms = header_symbols()
    for s in ["bola_hip_dgemm", "bola_hip_dgetrf", "bola_hip_dpotrf", "bola_hip_dtrsm", "bola_hip_dsyrk",
              "bola_hip_dlaswp", "bola_hip_idamax", "bola_hip_dgetrs", "bola_hip_dpotrs"]:
        assert s in syms and s + "_dev" in syms


def test_library_exports_every_declared_symbol(built):
    out = subprocess.run(["nm", "-D", "--defined-only", built.library_path()], capture_output=True, text=True,
                         check=True).stdout
    exported = set(re.findall(r"\bT (bola_hip_\w+)", out))
    missing = [s for s in header_symbols() if s not in exported]
    assert not missing, missing
    extra = sorted(s for s in exported if s not in header_symbols())
    assert not extra, extra


def test_python_binding_covers_header(built):
    L = built.lib()  # dlopen only; no device call
    assert sorted(built.declared_symbols()) == header_symbols()
    for s in header_symbols():
        assert hasattr(L, s)


def test_library_carries_gfx950_code_object(built):
    blob = open(built.library_path(), "rb").read()
    assert b"amdgcn-amd-amdhsa--gfx950" in blob


def test_missing_library_fails_loudly(monkeypatch, tmp_path):
    import bola_amd
    monkeypatch.setattr(bola_amd, "_SO", str(tmp_path / "nope.so"))
    monkeypatch.setattr(bola_amd, "_lib", None)
    with pytest.raises(RuntimeError):
        bola_amd.lib()
