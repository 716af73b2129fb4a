"""Launch-shape autotuner on the GPU: every candidate shape of the batched launch gives identical results (equal to the
oracle), the tuner leaves the fastest in the engine's options and records it in its cache, and a second call is served
from the cache without timing."""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def eng():
    import torch
    if not torch.cuda.is_available():
        pytest.fail("GPU tests need a ROCm GPU")
    from anechoic_amd import Engine
    e = Engine(0)
    yield e
    e.close()


@pytest.mark.parametrize("n", [2000, 2600, 3500])
def test_batched_candidates_identical_and_cached(eng, n, tmp_path):
    from anechoic_amd import autotune as AT
    from anechoic_amd import instances as I
    if n == 2000:
        J, b = I.ising_random(20261203, 2000)[0], None
    elif n == 3500:
        J, b = O.gen_ising_random(3500, n, kbits=2, pzero=0.0, bmax=3)
    else:
        J, b = I.erdos_renyi(4, n, 0.02)
    tb = O.tables(dict(S=40, t0=12.0, t1=5.0, q=8.0, lam=1.05, ramp=True))
    eng.set_tables(tb)
    inst = eng.add_instance(J, bias=b)
    R = 80
    trials = np.arange(R, dtype=np.uint32) + 11
    ref = O.run_batch(J, n, tb, 0x7E57, trials, bias=b, trace=True, nthreads=16)
    words = O.np_of(n) // 64
    before = {k: eng.get_option(k) for k in ("batched_team", "batched_nt", "exchange", "multi_nt3", "large_tile")}
    try:
        for opts in AT.batched_candidates(n):
            AT.apply(eng, opts)
            res = eng.run(inst, 0x7E57, trials, trace=True)
            assert np.array_equal(res["spins"][:, :words].view(np.uint64), ref["spins"]), opts
            assert np.array_equal(res["sum_sh"], ref["sum_sh"]) and np.array_equal(res["nlin"], ref["nlin"]), opts
        AT.apply(eng, before)
        cache = AT.TuneCache(tmp_path / "tune.json")
        t1 = AT.tune_batched(eng, inst, R, cache=cache)
        assert not t1["cached"] and len(t1["table"]) == len(AT.batched_candidates(n))
        assert {k: int(eng.get_option(k)) for k in t1["best"]} == t1["best"]
        AT.apply(eng, before)
        t2 = AT.tune_batched(eng, inst, R, cache=AT.TuneCache(tmp_path / "tune.json"))
        assert t2["cached"] and t2["best"] == t1["best"]
        assert {k: int(eng.get_option(k)) for k in t2["best"]} == t1["best"]
    finally:
        AT.apply(eng, {k: int(v) for k, v in before.items()})
