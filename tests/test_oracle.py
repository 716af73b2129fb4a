"""Pins the oracle (oracle/golden.c) before it is trusted as the checker of the HIP engine.

Checks, per oracle function:
  RNG (xoshiro128**, splitmix64)   published test vectors of the algorithms (Blackman-Vigna; Steele-Lea-Flood)
  tables                            tables written by the reference headers (golden_tables.json) + closed forms
  run_trial                         full results incl. n_lin traces from the reference headers (golden_trials.json),
                                    the reference's own held known answers (reference_cases.json), and invariants
                                    (final field = b + J s recomputed; energy identity; cut identity)
  generators                        the reference fingerprints/sumw printed in its cases.txt files
  NP > 2048 generalisation          invariants only (the reference is compiled for NP = 2048): parity unpinned beyond
                                    what field recomputation, energy identity and flip/trace consistency pin.
"""
import numpy as np
import pytest

from conftest import cdiv, hex_to_words
from helpers import case_schedule, golden_instance, pad2048, reference_matrix
from oracle import oracle as O


def test_xoshiro128ss_known_vector():
    # xoshiro128** from state {1, 2, 3, 4}: the reference implementation's first outputs.
    import ctypes
    s = np.array([1, 2, 3, 4], np.uint32)
    L = O.lib()
    got = [L.orc_rng_next(s.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32))) for _ in range(6)]
    assert got == [11520, 0, 5927040, 70819200, 2031721883, 1637235492]


def test_splitmix64_known_vector():
    assert [int(v) for v in O.splitmix_seq(0, 3)] == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]


def test_seeding_aliasing_of_trial_and_lane():
    # x = seed ^ (trial << 20) ^ (lane << 48): trial bit 28 aliases lane bit 0 (sca_ref.hpp:80).
    a = O.rng_stream(5, 1 << 28, 0, 4)
    b = O.rng_stream(5, 0, 1, 4)
    c = O.rng_stream(5, 0, 2, 4)
    assert np.array_equal(a, b) and not np.array_equal(a, c)


def test_tables_match_reference(golden_tables):
    for name, g in golden_tables.items():
        tb = O.tables(g["schedule"])
        for k in ("fourT", "q", "kcorr", "kconst"):
            assert tb[k].tolist() == g[k], (name, k)


def test_tables_closed_form():
    sch = dict(S=11, t0=12.0, t1=5.0, q=8.0, lam=1.05, ramp=True)
    tb = O.tables(sch)
    T = [12.0 * (5.0 / 12.0) ** (t / 10) for t in range(11)]
    assert tb["fourT"][0] == round(4 * 12 * 65536) and tb["fourT"][-1] == round(4 * 5 * 65536)
    assert all(v == 8 * 65536 for v in tb["q"])
    assert tb["kcorr"][0] == 0 and tb["kconst"][0] == 0
    # before the ramp (t < 0.7 S) kcorr = lambda / (2 T_{t-1}) in Q8.24
    for t in range(1, 7):
        assert abs(tb["kcorr"][t] - 1.05 / (2 * T[t - 1]) * 2**24) <= 0.5 + 1e-6
    assert tb["kcorr"][-1] == 0   # ramp reaches zero at the last step
    kt = O.tables(dict(S=5, t0=10.0, t1=10.0, q=1.0, kappa=2.0))
    assert kt["kcorr"].tolist() == [0] * 5 and kt["kconst"].tolist() == [0] + [2 * 10 * 65536] * 4
    tec = O.tables(dict(S=3, t0=1.0, t1=1.0, q=1.0, jv=-8.0))
    assert tec["kconst"].tolist() == [0, 8 * 65536, 8 * 65536]
    lin = O.tables(dict(S=8, shape="linear", t0=12.0, t1=5.0, q=8.0))
    assert lin["fourT"].tolist() == [int(np.floor(4 * (12 - 7 * t / 7) * 65536 + 0.5)) for t in range(8)]


def _check_result(r, g, n):
    assert r["sum_sh"] == g["sum_sh"] and r["sum_bs"] == g["sum_bs"] and r["flips"] == g["flips"]
    assert np.array_equal(r["spins"][:32], hex_to_words(g["spins"]))
    assert r["nlin"].tolist() == g["nlin"]


def test_run_trial_matches_reference_headers(golden_trials, golden_tables):
    for key, rec in golden_trials.items():
        inst = rec["instance"]
        J, bias = golden_instance(inst)
        n = inst["n"]
        assert f"{O.fingerprint(pad2048(J, n)):016x}" == inst["fingerprint"]
        tb = {k: np.array(golden_tables[rec["schedule"]][k]) for k in ("fourT", "q", "kcorr", "kconst")}
        tb = dict(fourT=tb["fourT"].astype(np.int32), q=tb["q"].astype(np.int32), kcorr=tb["kcorr"].astype(np.int64),
                  kconst=tb["kconst"].astype(np.int32))
        for g in rec["results"]:
            r = O.run_trial(J, n, tb, rec["seed"], g["trial"], bias=bias, trace=True)
            _check_result(r, g, n)


def test_reference_held_known_answers(ref_cases):
    """Every per-trial known answer in the reference's own cases.txt suites (RTL/ASIC vectors of the golden model)."""
    checked = 0
    for suite in ref_cases:
        m = suite["matrix"]
        J, bias, scale = reference_matrix(m)
        assert f"{O.fingerprint(J):016x}" == m["fingerprint"], suite["name"]
        assert O.sumw(J, 2048) == m["sumw"]
        assert f"{scale:.12g}".startswith(m["auto_scale_printed"][:8])
        n = m["n"]
        for c in suite["cases"]:
            sch = case_schedule(c, scale)
            if c["scale_printed"] != "1":
                assert f"{scale:.12g}" == c["scale_printed"] or f"{scale:g}" == c["scale_printed"]
            tb = O.tables(sch)
            use_b = bias if (bias is not None and c["use_bias"]) else None
            for g in c["trials"]:
                r = O.run_trial(J, n, tb, c["seed"], g["trial"], bias=use_b)
                assert r["flips"] == g["flips"], (suite["name"], c["name"], g["trial"])
                if "sum_sh" in g:
                    assert r["sum_sh"] == g["sum_sh"]
                if "sum_bs" in g:
                    assert r["sum_bs"] == g["sum_bs"]
                if "energy2" in g:
                    assert -(r["sum_sh"] + r["sum_bs"]) == g["energy2"]
                if "cut" in g:
                    assert cdiv(m["sumw"] + cdiv(r["sum_sh"], 2), 2) == g["cut"]
                checked += 1
    assert checked >= 60


def _energy(J, n, bias, s):
    s = s.astype(np.int64)
    Jn = J[:n, :n].astype(np.int64)
    e = -(s @ Jn @ s) // 2
    if bias is not None:
        e -= int(np.dot(np.asarray(bias[:n], np.int64), s))
    return int(e)


@pytest.mark.parametrize("n", [300, 2000, 2600])
def test_invariants_fields_energy_cut(n):
    J, bias = O.gen_random_k(77 + n, 0.3, 3, n, 40, ld=max(2048, n), rows=max(2048, n))
    tb = O.tables(dict(S=60, t0=6.0, t1=1.0, q=2.0, lam=0.5, ramp=True))
    r = O.run_trial(J, n, tb, 99, 3, bias=bias, trace=True, want_h=True)
    s = O.spins_pm(r["spins"], n)
    h = O.fields(J, n, r["spins"], bias)
    assert np.array_equal(h, r["h"])                               # final fields = b + J s, recomputed
    assert r["sum_sh"] == int(np.dot(s.astype(np.int64), h))
    assert -(r["sum_sh"] + r["sum_bs"]) == 2 * _energy(J, n, bias, s)   # E = -(sum_sh + sum_bs) / 2
    r0 = O.run_trial(J, n, tb, 99, 3, bias=None, want_h=True)
    s0 = O.spins_pm(r0["spins"], n)
    cut = sum(-int(J[x, y]) for x in range(n) for y in range(x + 1, n) if s0[x] != s0[y]) if n <= 300 else None
    if cut is not None:
        assert cdiv(O.sumw(J, n) + cdiv(r0["sum_sh"], 2), 2) == cut   # cut = (W - H) / 2 identity
    assert 0 <= min(r["nlin"]) and max(r["nlin"]) <= n


def test_np_generalisation():
    assert [O.np_of(n) for n in (1, 2000, 2048, 2049, 3000, 16384)] == [2048, 2048, 2048, 2304, 3072, 16384]


def test_checks_are_sensitive():
    """Planted mistakes in the oracle's inputs must change the compared outputs."""
    J, bias = O.gen_random_k(5, 0.5, 1, 500, 0, ld=2048, rows=2048)
    tb = O.tables(dict(S=50, t0=4.0, t1=0.5, q=3.0, lam=0.9, ramp=True))
    a = O.run_trial(J, 500, tb, 1, 1, trace=True)
    bad = {k: v.copy() for k, v in tb.items()}
    bad["kcorr"][10] += 1 << 16
    b = O.run_trial(J, 500, bad, 1, 1, trace=True)
    assert a["nlin"][:10].tolist() == b["nlin"][:10].tolist() and a["nlin"].tolist() != b["nlin"].tolist()
    c = O.run_trial(J, 500, tb, 1, 2, trace=True)
    assert c["sum_sh"] != a["sum_sh"] or c["flips"] != a["flips"]
