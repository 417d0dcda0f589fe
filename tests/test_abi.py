"""C-ABI checks that need no GPU: the library loads and exports every symbol
declared in include/*.h, and the Python binding knows every one of them."""
import glob
import os
import re

import pytest

from muesli4_amd import hip

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DECL = re.compile(r"^\s*(?:int|void\s*\*|const\s+char\s*\*)\s*(msl_[a-z0-9_]+)\s*\(", re.M)


def declared_symbols():
    names = []
    for h in sorted(glob.glob(os.path.join(ROOT, "include", "*.h"))):
        with open(h) as f:
            names += DECL.findall(f.read())
    return names


def test_headers_declare_symbols():
    names = declared_symbols()
    assert len(names) >= 30
    assert len(names) == len(set(names))


def test_library_exports_every_declared_symbol():
    L = hip.lib()
    missing = [n for n in declared_symbols() if not hasattr(L, n)]
    assert not missing, missing


def test_binding_covers_every_declared_symbol():
    missing = [n for n in declared_symbols() if n not in hip.SIGNATURES]
    assert not missing, missing
    extra = [n for n in hip.SIGNATURES if n not in declared_symbols()]
    assert not extra, extra


def test_library_is_gfx950_code_object():
    # the offload bundle inside the shared object names the target
    with open(hip.lib_path(), "rb") as f:
        blob = f.read()
    assert b"gfx950" in blob


def test_oracle_is_not_linked_by_product():
    with open(hip.lib_path(), "rb") as f:
        blob = f.read()
    assert b"orc_" not in blob and b"liboracle" not in blob


def test_missing_library_raises(tmp_path, monkeypatch):
    monkeypatch.setattr(hip, "_LIB_PATH", str(tmp_path / "nope.so"))
    monkeypatch.setattr(hip, "_lib", None)
    with pytest.raises(ImportError):
        hip.lib()
