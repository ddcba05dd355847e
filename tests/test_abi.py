"""CPU tests of the C-ABI library: it loads, exports every symbol that
include/cudapix.h declares, and its host-side arithmetic self checks pass.
No device calls are made here."""
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_symbols():
    with open(os.path.join(ROOT, "include", "cudapix.h")) as f:
        src = f.read()
    return sorted(set(re.findall(r"\b(cpx_[a-z0-9_]+)\s*\(", src)))


def test_header_symbols_are_exported():
    import cudapix_amd as C
    L = C.lib()
    syms = header_symbols()
    assert len(syms) >= 45
    missing = [s for s in syms if not hasattr(L, s)]
    assert not missing, missing
    assert sorted(C.EXPORTS) == syms


def test_arith_selftest_host():
    import cudapix_amd as C
    assert C.lib().cpx_selftest_arith(12345, 300000) == 0
    assert C.lib().cpx_selftest_arith(99, 300000) == 0


def test_strerror_codes():
    import cudapix_amd as C
    L = C.lib()
    for code, name in C.ERRORS.items():
        assert L.cpx_strerror(code)
    assert L.cpx_strerror(-99) == b"unknown error"


def test_no_fallback_when_library_missing(tmp_path, monkeypatch):
    import cudapix_amd as C
    monkeypatch.setattr(C, "LIB_PATH", str(tmp_path / "missing.so"))
    monkeypatch.setattr(C, "_lib", None)
    try:
        C.lib()
    except ImportError as e:
        assert "no fallback" in str(e)
    else:
        raise AssertionError("expected ImportError")


def test_product_does_not_import_oracle():
    pkg = os.path.join(ROOT, "cudapix_amd")
    for dirpath, _, files in os.walk(pkg):
        for fn in files:
            if fn.endswith((".py", ".hip", ".hpp", ".h", ".cpp")):
                with open(os.path.join(dirpath, fn)) as f:
                    src = f.read()
                assert "oracle" not in src.lower(), fn
