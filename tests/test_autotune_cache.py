"""Autotuner host logic (no GPU): the JSON cache round trip and the candidate sets per instance size."""
from anechoic_amd import autotune as AT


def test_cache_round_trip(tmp_path):
    p = tmp_path / "sub" / "tune.json"
    c = AT.TuneCache(p)
    assert c.get("k") is None
    c.put("k", dict(best=dict(batched_team=16, batched_nt=0, exchange=2), ms=1.5))
    c.put("k2", dict(best=[2, 3], ms=2.0))
    d = AT.TuneCache(p)
    assert d.get("k")["best"]["batched_team"] == 16 and d.get("k2")["best"] == [2, 3]
    p.write_text("not json")
    assert AT.TuneCache(p).get("k") is None


def test_candidates_per_size():
    small = AT.batched_candidates(2000)
    assert dict(batched_team=16, batched_nt=0, exchange=2) in small
    assert all(o["batched_team"] in (16, 32) for o in small)
    assert {o["exchange"] for o in AT.batched_candidates(3000)} == {0, 2}
    assert [o["large_tile"] for o in AT.batched_candidates(16384)] == [128, 64]


def test_cache_path_from_environment(monkeypatch, tmp_path):
    monkeypatch.setenv("ANECHOIC_AUTOTUNE_CACHE", str(tmp_path / "x.json"))
    assert AT.default_cache_path() == tmp_path / "x.json"
