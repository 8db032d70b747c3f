"""The C-ABI libraries load and export every entry point their headers
declare (no device calls: this runs without a GPU)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w[\w\s\*]*?\b(tb[kh]_\w+)\s*\(", text, flags=re.M)))


@pytest.mark.parametrize("header,lib", [("tbkpm.h", "libtbkpm.so"), ("tbhost.h", "libtbhost.so")])
def test_library_exports_declared_symbols(header, lib):
    names = declared(header)
    assert len(names) >= 10
    so = ctypes.CDLL(os.path.join(ROOT, "tight_binding_amd", lib))
    missing = [n for n in names if not hasattr(so, n)]
    assert not missing, missing


def test_python_bindings_cover_headers():
    from tight_binding_amd import _native as N

    assert set(declared("tbkpm.h")) == set(N.KPM_SIGNATURES)
    assert set(declared("tbhost.h")) == set(N.HOST_SIGNATURES)


def test_version_and_errors_without_device():
    from tight_binding_amd import _native as N

    lib = N.kpm()
    assert lib.tbk_version().decode().startswith("tbkpm")
    # invalid arguments are rejected before any device call
    rc = lib.tbk_cheb_first_kind_seq_braket(None, 8, None, None, 1, 1, None)
    assert rc == N.TBK_E_INVALID
    assert "null" in lib.tbk_last_error().decode()


def test_host_errors_are_codes_not_exceptions():
    from tight_binding_amd import _native as N
    from tight_binding_amd import lattice as L

    with pytest.raises(N.NativeError):
        L.UnitCell.model("no_such_model")
    c = L.UnitCell.model("square")
    with pytest.raises(N.NativeError):
        c.add_hopping("A", "Z", (0, 1), 1.0)
