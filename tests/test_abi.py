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


def test_tuning_cache_file_round_trip(tmp_path):
    """The autotuner cache's text form: load, save, load again (host only)."""
    from tight_binding_amd import _native as N
    from tight_binding_amd import kpm

    nf = len(N.tbk_tuning._fields_)
    src = tmp_path / "tuning.txt"
    lines = ["tbkpm-tuning 1 %d" % nf,
             # num_cu rows max_row_nnz bandwidth95 Rp precision complex plain nnz, tuning, ms
             "256 2097152 4 2048 64 0 0 0 8388608 " + " ".join(["0"] * nf) + " 1.165",
             "256 262144 5 512 32 0 1 0 1308672 " + " ".join(str(i % 3) for i in range(nf)) + " 0.0381"]
    src.write_text("\n".join(lines) + "\n")
    kpm.tuning_cache_clear()
    assert kpm.tuning_cache_load(str(src)) == 2
    out = tmp_path / "saved.txt"
    kpm.tuning_cache_save(str(out))
    kpm.tuning_cache_clear()
    assert kpm.tuning_cache_load(str(out)) == 2
    again = tmp_path / "again.txt"
    kpm.tuning_cache_save(str(again))
    assert out.read_text() == again.read_text()
    kpm.tuning_cache_clear()
    bad = tmp_path / "bad.txt"
    bad.write_text("tbkpm-tuning 1 %d\n1 2 3\n" % nf)
    with pytest.raises(N.NativeError):
        kpm.tuning_cache_load(str(bad))
    bad.write_text("tbkpm-tuning 1 %d\n" % (nf + 1))
    with pytest.raises(N.NativeError):
        kpm.tuning_cache_load(str(bad))
    kpm.tuning_cache_clear()
