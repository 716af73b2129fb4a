"""Schedule search (host logic; no GPU): the pilot grid is the protocol's (gpu/onsager_v2_20261007/src/run_protocol.sh
phase P: 120 cells in its index order, whose tables equal the reference's own build_tables), and the selection rule
is TTS99 at the lower Wilson 95 % bound with ties to the smaller point estimate (analyze.py:81-85)."""
import math

import numpy as np
import pytest

from oracle import oracle as O


@pytest.fixture(scope="module")
def S():
    from anechoic_amd import search
    return search


def test_reference_grid_order_and_size(S):
    cells = S.reference_grid()
    assert len(cells) == 120
    assert cells[0]["name"] == "kappa1.5_q6_t12_S80" and cells[5]["name"] == "kappa1.5_q6_t12_S280"
    assert cells[6]["name"] == "kappa1.5_q6_t15_S80" and cells[-1]["name"] == "lam1.05_q8_t15_S280"
    assert all(c["schedule"]["ramp"] and c["schedule"]["t1"] == 5.0 and c["schedule"]["shape"] == "geometric" for c in cells)
    assert len({c["name"] for c in cells}) == 120


@pytest.mark.skipif(not O.ref_available(), reason="reference headers not built (oracle/_ref)")
def test_grid_tables_equal_reference_build_tables(S):
    from anechoic_amd import tables
    cells = S.reference_grid()
    for k in (0, 17, 45, 77, 119):
        sch = cells[k]["schedule"]
        mine = tables(sch)
        ref = O.ref_build_tables(dict(S=sch["S"], t0=sch["t0"], t1=sch["t1"], q=sch["q"], lam=sch.get("lam", 0.0),
                                      kappa=sch.get("kappa", 0.0), ramp=True))
        for key in ("fourT", "q", "kcorr", "kconst"):
            assert np.array_equal(mine[key], ref[key]), (cells[k]["name"], key)


def _cell(S, succ, B, t, name):
    from anechoic_amd import tts
    e = tts.estimate(succ, B, t)
    e.update(name=name, score_primary=S.score(e, "primary"), score_secondary=S.score(e, "secondary"))
    return e


def test_selection_rule(S):
    from anechoic_amd import tts
    a = _cell(S, [3] * 32, 240, 1.0, "a")                  # P_batch 1 over 32 batches, 1 ms
    b = _cell(S, [3] * 32, 240, 0.5, "b")                  # same outcomes, half the time
    c = _cell(S, [0] + [2] * 31, 240, 0.3, "c")            # one empty batch, 0.3 ms
    lo = tts.wilson(31, 32)[0]
    assert c["score_primary"] == pytest.approx(0.3 * math.log(0.01) / math.log(1 - lo))
    assert b["score_primary"] == pytest.approx(a["score_primary"] / 2)
    # by hand: Wilson lower bounds 0.893 (32 of 32) and 0.843 (31 of 32): b = 0.5 ln 0.01 / ln 0.107 = 1.03 ms,
    # c = 0.3 ln 0.01 / ln 0.157 = 0.75 ms
    assert c["score_primary"] < b["score_primary"] < a["score_primary"]
    assert b["score_primary"] == pytest.approx(1.03, abs=0.01) and c["score_primary"] == pytest.approx(0.746, abs=0.01)
    assert S.select([a, b, c], "primary")["name"] == "c"
    # secondary: p at its Wilson lower bound; more successes per batch win at equal time
    d = _cell(S, [6] * 32, 240, 1.0, "d")
    assert S.select([a, d], "secondary")["name"] == "d"
    # ties on the score go to the smaller point estimate
    e1 = dict(name="x", score_primary=1.0, tts_primary=0.9)
    e2 = dict(name="y", score_primary=1.0, tts_primary=0.8)
    assert S.select([e1, e2], "primary")["name"] == "y"


def test_pilot_trial_ranges_must_not_overlap(S):
    with pytest.raises(ValueError, match="stride"):
        S.pilot(None, 0, 0, [dict(name="x", schedule={})], 0, 1, B=240, batches=64, stride=8000)
