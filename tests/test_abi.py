"""CPU checks of the C-ABI library: it loads, exports every declared symbol,
and its reduction geometry is the one the oracle restates.  No compute call
is made (there is no GPU here)."""
import ctypes
import os
import re

from cgxx_amd import cghip
from oracle import oracle

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = []
    for fn in sorted(os.listdir(os.path.join(REPO, "include"))):
        if fn.endswith(".h"):
            text = open(os.path.join(REPO, "include", fn)).read()
            names += re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(cghip_\w+)\s*\(", text, re.M)
    return names


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(cghip.LIB_PATH)
    names = declared_symbols()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert set(names) == set(cghip.SYMBOLS), set(names) ^ set(cghip.SYMBOLS)


def test_geometry_matches_oracle():
    assert cghip.geometry() == oracle.GEOMETRY


def test_missing_library_fails_loudly(monkeypatch, tmp_path):
    monkeypatch.setattr(cghip, "LIB_PATH", str(tmp_path / "absent.so"))
    monkeypatch.setattr(cghip, "_lib", None)
    try:
        cghip.load()
    except cghip.CGHIPError as e:
        assert "no CPU fallback" in str(e)
    else:
        raise AssertionError("load() must raise without the library")


def test_dot_noalloc_rejects_bad_arguments_before_any_launch():
    """The allocation-free dot refuses a negative size or NULL pointers on
    the host, with a message for cghip_last_error(NULL), before any HIP
    call (so this runs without a GPU)."""
    L = cghip.load()
    assert L.cghip_dot_noalloc(None, None, -5, None, None) == 1
    assert b"negative" in L.cghip_last_error(None)
    assert L.cghip_dot_noalloc(None, None, 10, None, None) == 1
    assert b"NULL" in L.cghip_last_error(None)
