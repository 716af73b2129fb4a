"""The C ABI library (no GPU needed): it loads, exports every entry point include/anechoic.h declares, builds the
reference's tables bit for bit, and reports errors through status codes and anc_last_error()."""
import ctypes
import re

import numpy as np
import pytest

from conftest import ROOT
from anechoic_amd import _native as N
from anechoic_amd import tables


@pytest.fixture(scope="module")
def lib():
    if not N.LIB_PATH.exists():
        from anechoic_amd import build
        build.build()
    return N.lib()


def declared_functions():
    text = (ROOT / "include" / "anechoic.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(anc_[a-z_]+)\s*\(", text)))


def test_exports_every_declared_symbol(lib):
    names = declared_functions()
    assert len(names) >= 14
    for name in names:
        assert hasattr(lib, name), name
        assert name in N.SIGNATURES, f"{name} not bound in _native.SIGNATURES"


def test_tables_match_reference(golden_tables):
    for name, g in golden_tables.items():
        tb = tables(g["schedule"])
        for k in ("fourT", "q", "kcorr", "kconst"):
            assert tb[k].tolist() == g[k], (name, k)


def test_explicit_T_matches_generic_make_tables(golden_tables):
    g = golden_tables["c1_linear_counted_S1000"]
    S = 1000
    T = np.array([12.0 + (5.0 - 12.0) * (t / 999) for t in range(S)])
    tb = tables(dict(S=S, shape="explicit", T=T, q=8.0, lam=1.05, ramp=True))
    for k in ("fourT", "q", "kcorr", "kconst"):
        assert tb[k].tolist() == g[k]


def test_np(lib):
    assert [lib.anc_np(n) for n in (1, 2000, 2048, 2049, 3000, 16384)] == [2048, 2048, 2048, 2304, 3072, 16384]


def test_error_paths(lib):
    with pytest.raises(N.AncError) as e:
        tables(dict(S=0, q=1.0))
    assert e.value.code == -1 and "S must be" in str(e.value)
    with pytest.raises(N.AncError) as e:
        tables(dict(S=10, shape="linear", t0=1.0, t1=-1.0, q=1.0))
    assert e.value.code == -3
    with pytest.raises(N.AncError):
        tables(dict(S=10, t0=1e6, t1=1.0, q=1.0))
    c = N.Sched()
    assert lib.anc_make_tables(ctypes.byref(c), None, None, None, None) == -1
    assert b"null" in lib.anc_last_error()


def test_run_entry_points_reject_null_context(lib):
    out = N.Out()
    assert lib.anc_run_plain(None, 0, 1, 0, ctypes.byref(out), None) < 0
    assert lib.anc_run_batched(None, 0, 1, None, 4, ctypes.byref(out), None) < 0
    assert lib.anc_set_tables(None, 1, None, None, None, None) == -1
    assert lib.anc_instance_bound(None, 0) < 0


def test_meanfield_entry_points_check_j_layout(lib):
    """jbits must be 8 or 4 and the row stride must cover a row of the chosen storage; checked before any device
    work (the pointers below are never dereferenced)."""
    p = ctypes.c_void_p(256)
    f = ctypes.c_float(1.0)
    n = 2000
    cases = [(5, 2016), (3, 2016), (8, 1008), (8, 2008), (4, 992), (4, 1000), (2, 496)]   # bad jbits; short rows; ldj % 16
    for jbits, ldj in cases:
        assert lib.anc_mf_gemv(2, n, p, jbits, ldj, n * ldj, p, p, None) == -1, (jbits, ldj)
        assert lib.anc_mf_forward(2, n, p, jbits, ldj, n * ldj, f, 10, p, p, 0, p, 4, p, p, p, None) == -1
        assert lib.anc_mf_backward(2, n, p, jbits, ldj, n * ldj, f, 10, p, p, 0, 4, p, p, p, p, p, p, None) == -1
    assert b"jbits" in (lib.anc_mf_gemv(2, n, p, 5, 2000, n * 2000, p, p, None), lib.anc_last_error())[1]
    assert lib.anc_mf_gemv(2, n, p, 4, 1008, n * 1008 - 1, p, p, None) == -1            # instance stride
    assert lib.anc_mf_pack_j4(2, n, p, 2000, n * 2000, p, 992, n * 992, p, None) == -1    # packed row too short
    assert lib.anc_mf_pack_j4(2, n, p, 1008, n * 1008, p, 1008, n * 1008, p, None) == -1  # int8 row too short
    assert lib.anc_mf_pack_j4(2, n, p, 2000, n * 2000, p, 1008, n * 1008, None, None) == -1
    assert lib.anc_mf_pack_j2(2, n, p, 2000, n * 2000, p, 496, n * 496, p, None) == -1      # 2-bit row too short
