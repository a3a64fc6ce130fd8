"""The C-ABI library builds for gfx950 and exports every entry point that
include/sdbfx.h declares (no GPU calls).  CPU only."""
import ctypes
import os
import re

import sdhash_amd

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    src = open(os.path.join(ROOT, "include", "sdbfx.h")).read()
    return sorted(set(re.findall(r"\b(sdbfx_[a-z_]+)\s*\(", src)))


def test_header_and_python_binding_agree():
    assert declared() == sorted(sdhash_amd.SYMBOLS)


def test_library_exports_declared_symbols():
    if not os.path.exists(sdhash_amd.LIB_PATH):
        from sdhash_amd import build
        build.build()
    lib = ctypes.CDLL(sdhash_amd.LIB_PATH)
    for name in declared():
        assert hasattr(lib, name), name


def test_missing_library_fails_loudly(monkeypatch, tmp_path):
    monkeypatch.setattr(sdhash_amd, "LIB_PATH", str(tmp_path / "absent.so"))
    monkeypatch.setattr(sdhash_amd, "_lib", None)
    try:
        sdhash_amd.load()
    except sdhash_amd.SdbfxError:
        return
    raise AssertionError("load() must raise without the library")


def test_name_array_is_passed_as_built():
    """A prebuilt name array goes to the library as is; a list is converted;
    an array with fewer names than inputs is refused."""
    import types
    names = [b"frag%04d" % i for i in range(5)]
    arr = sdhash_amd.name_array(names)
    holder = types.SimpleNamespace()
    assert sdhash_amd.DigestSet._names(holder, arr, 5) is arr
    assert [arr[i] for i in range(5)] == names
    conv = sdhash_amd.DigestSet._names(holder, names, 5)
    assert [conv[i] for i in range(5)] == names
    try:
        sdhash_amd.DigestSet._names(holder, arr, 6)
    except ValueError:
        pass
    else:
        raise AssertionError("a short name array was accepted")
