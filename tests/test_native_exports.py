"""The C-ABI library loads and exports every entry point the header declares
(no compute calls: this runs without a GPU)."""
import ctypes
import os
import re

from sunaengine_amd import native

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    names = set()
    for h in os.listdir(os.path.join(ROOT, "include")):
        if h.endswith(".h"):
            txt = open(os.path.join(ROOT, "include", h)).read()
            names |= set(re.findall(r"\b(sn3?_[a-z_0-9]+)\s*\(", txt))
    return names


def test_library_exports_every_declared_symbol():
    L = native.lib()
    declared = _declared()
    assert declared, "no declarations found"
    missing = [n for n in sorted(declared) if not hasattr(L, n)]
    assert not missing, missing
    assert set(native.EXPORTS) <= declared


def test_missing_library_fails_loudly(monkeypatch):
    monkeypatch.setattr(native, "_LIB", None)
    monkeypatch.setattr(native, "LIB_PATH", "/nonexistent/libsunaengine_amd.so")
    try:
        native.lib()
    except native.NativeMissing as e:
        assert "not built" in str(e)
    else:
        raise AssertionError("expected NativeMissing")


def test_digest_state_argument_errors():
    """digestState's thrown errors (src/hash/statehash.js:97-108) come back as
    SN_E_ARG before any device work, so this runs without a GPU."""
    import pytest
    bogus = 0x1000  # never dereferenced: every case fails its argument check first
    cases = [
        (dict(state_words=64, particle_count=8, stride_words=8, hashed_words=9), "hashedWords > strideWords"),
        (dict(state_words=63, particle_count=8, stride_words=8, hashed_words=8), "state buffer too small"),
        (dict(state_words=64, particle_count=8, stride_words=8, hashed_words=8, bond_ptr=bogus, bond_cap=0),
         "digestState: p.bond given without p.bondCap"),
    ]
    for kw, msg in cases:
        with pytest.raises(native.SnError) as ei:
            native.digest_state(bogus, **kw)
        assert ei.value.code == native.SN_E_ARG
        assert str(ei.value) == msg
