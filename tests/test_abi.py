"""The C-ABI libraries load and export every function their headers declare (no GPU calls)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared(header: str):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(asr_\w+)\s*\(", text)))


@pytest.mark.parametrize("header,lib", [("asr_host.h", "libasr_host.so"), ("asr_hip.h", "libasr_hip.so")])
def test_library_exports_declared_symbols(header, lib):
    path = os.path.join(ROOT, "asr_amd", lib)
    assert os.path.exists(path), f"{lib} not built"
    L = ctypes.CDLL(path)
    names = declared(header)
    assert len(names) >= 8
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_engine_refuses_without_library(monkeypatch, tmp_path):
    """The product path has no CPU fallback: a missing HIP library is an error."""
    from asr_amd import engine
    monkeypatch.setattr(engine, "_PKG", str(tmp_path))
    monkeypatch.setattr(engine, "_LIB", None)
    with pytest.raises(engine.EngineError):
        engine.lib()
