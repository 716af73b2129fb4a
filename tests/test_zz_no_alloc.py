"""anc_run_batched_noalloc (exchange workspace in the library's static device array) against anc_run_batched: two
calls in a row on the same context, the second with fewer replicas, both bit-identical to the allocating entry point
(spins, sum s h, sum b s, flips, n_lin traces, final fields), for both team shapes; larger instances are refused."""
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


def _same(a, b):
    for k in ("spins", "sum_sh", "sum_bs", "flips", "nlin", "fields"):
        assert np.array_equal(a[k], b[k]), k


@pytest.mark.parametrize("team", [16, 32])
def test_noalloc_matches_batched_two_calls(eng, team):
    from anechoic_amd import instances as I
    J, _ = I.ising_random(20261203, 2000)
    tb = O.tables(dict(S=80, t0=12.0, t1=5.0, q=8.0, lam=1.05, ramp=True))
    eng.set_tables(tb)
    inst = eng.add_instance(J)
    old = eng.get_option("batched_team")
    eng.set_option("batched_team", team)
    try:
        for R, base in ((240, 7), (37, 9000)):   # the second call is smaller than the first
            trials = np.arange(base, base + R, dtype=np.uint32)
            ref = eng.run(inst, 0x0A11C, trials, trace=True, fields=True, kernel="batched")
            got = eng.run(inst, 0x0A11C, trials, trace=True, fields=True, kernel="batched_noalloc")
            assert int(eng.get_option("last_team")) == team
            _same(got, ref)
        orc = O.run_batch(J, 2000, tb, 0x0A11C, np.arange(9000, 9037, dtype=np.uint32), trace=True, nthreads=16)
        assert np.array_equal(got["sum_sh"], orc["sum_sh"]) and np.array_equal(got["nlin"], orc["nlin"])
    finally:
        eng.set_option("batched_team", old)


def test_noalloc_bias_instance_and_large_refusal(eng):
    from anechoic_amd import _native as N
    from anechoic_amd import instances as I
    J, b = I.erdos_renyi(77, 2600, 0.02)   # NP 2816: the 12-round 32-workgroup teams
    tb = O.tables(dict(S=40, t0=6.0, t1=1.0, q=2.0, lam=0.8, ramp=True, scale=0.05))
    eng.set_tables(tb)
    inst = eng.add_instance(J, bias=b)
    for R in (50, 20):
        trials = np.arange(R, dtype=np.uint32) + 3
        ref = eng.run(inst, 5, trials, trace=True, fields=True)
        got = eng.run(inst, 5, trials, trace=True, fields=True, kernel="batched_noalloc")
        _same(got, ref)
    Jl, _ = O.gen_ising_random(5, 4000)
    il = eng.add_instance(Jl)
    with pytest.raises(N.AncError, match="not supported"):
        eng.run(il, 1, np.arange(4, dtype=np.uint32), kernel="batched_noalloc")
