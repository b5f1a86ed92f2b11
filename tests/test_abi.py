"""The C ABI library (CPU only: loads it and checks its exports; no GPU calls)."""
import ctypes
import re
import subprocess

import pytest

from conftest import REPO

HEADER = REPO / "include" / "tn93g.h"


def _declared():
    text = HEADER.read_text()
    return sorted(set(re.findall(r"\b(tn93g_[a-z_]+)\s*\(", text)))


def test_header_declares_the_entry_points():
    assert _declared() == sorted(["tn93g_open", "tn93g_close", "tn93g_error", "tn93g_set_upload", "tn93g_set_free",
                                  "tn93g_run", "tn93g_result_free", "tn93g_version", "tn93g_run_batch",
                                  "tn93g_set_create", "tn93g_set_append", "tn93g_set_size", "tn93g_run_bytes",
                                  "tn93g_result_certify", "tn93g_run_noalloc"])


def test_library_exports_every_declared_symbol():
    import swifttn93_amd as T
    if not T.LIB_PATH.exists():
        pytest.skip("engine not built")
    lib = ctypes.CDLL(str(T.LIB_PATH))
    for name in _declared():
        assert hasattr(lib, name), name
    out = subprocess.run(["nm", "-D", "--defined-only", str(T.LIB_PATH)], capture_output=True, text=True).stdout
    exported = {ln.split()[-1] for ln in out.splitlines() if " T " in ln and ln.split()[-1].startswith("tn93g_")}
    assert exported == set(_declared())
    lib.tn93g_version.restype = ctypes.c_char_p
    assert b"gfx950" in lib.tn93g_version()
    assert set(T.EXPORTS) == set(_declared())


def test_header_compiles_as_c():
    subprocess.run(["gcc", "-fsyntax-only", "-x", "c", "-std=c99", str(HEADER)], check=True)


def test_no_fallback_without_library(tmp_path, monkeypatch):
    """The Python binding refuses to run when the native library is missing (there is no CPU path)."""
    import swifttn93_amd as T
    monkeypatch.setattr(T, "LIB_PATH", tmp_path / "missing.so")
    monkeypatch.setattr(T, "_lib", None)
    with pytest.raises(RuntimeError):
        T.load()
