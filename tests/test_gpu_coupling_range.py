"""Coupling magnitudes, biases and decision ranges beyond the +-1 workload, on the int8 team kernels, against the oracle.

The engine's range plan picks int32 decision thresholds when it can prove them exact and int64 ones otherwise; both
instantiations are exercised here (8-round teams; the 12-round and large kernels' int64 paths are in
test_gpu_parity.py and test_gpu_large.py). Every comparison is exact: spins, sum s h, sum b s, flips, n_lin traces.
"""
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


def _assert_equal(res, ref, words):
    assert np.array_equal(res["spins"][:, :words].view(np.uint64), ref["spins"])
    assert np.array_equal(res["sum_sh"], ref["sum_sh"]) and np.array_equal(res["sum_bs"], ref["sum_bs"])
    assert np.array_equal(res["flips"], ref["flips"]) and np.array_equal(res["nlin"], ref["nlin"])


def _small_int_instance(seed, n, values, density=0.5, bmax=0):
    rng = np.random.default_rng(seed)
    vals = np.array(values, np.int8)
    up = np.where(rng.random((n, n)) < density, vals[rng.integers(0, len(vals), (n, n))], 0).astype(np.int8)
    J = np.triu(up, 1)
    J = (J + J.T).astype(np.int8)
    b = rng.integers(-bmax, bmax + 1, n).astype(np.int32) if bmax else None
    return J, b


@pytest.mark.parametrize("n,R", [(2000, 64), (1500, 100), (777, 33)])
def test_multi_bit_couplings_with_bias(eng, n, R):
    """Couplings of magnitudes 1..7 with biases, sizes below 2048 (padding slots), 16- and 32-replica work items."""
    J, b = _small_int_instance(1000 + n, n, [1, -1, 2, -2, 3, -3, 4, -4, 5, -5, 6, -6, 7, -7], density=0.6, bmax=300)
    tb = O.tables(dict(S=60, t0=40.0, t1=3.0, q=10.0, lam=0.9, ramp=True, scale=0.5))
    eng.set_tables(tb)
    inst = eng.add_instance(J, bias=b)
    trials = np.arange(R, dtype=np.uint32) * 3 + 7
    res = eng.run(inst, 4321, trials, trace=True, fields=True)
    words = O.np_of(n) // 64
    _assert_equal(res, O.run_batch(J, n, tb, 4321, trials, bias=b, trace=True, nthreads=16), words)
    h = O.fields(J, n, res["spins"][R - 1, :words].view(np.uint64), b)
    assert np.array_equal(res["fields"][R - 1, :n].astype(np.int64), h)


def test_dense_large_magnitude_rows(eng):
    """Rows of couplings +100 beside a body of +-6: field sums near 2^18, every anneal still exact."""
    n = 2048
    rng = np.random.default_rng(5)
    up = np.where(rng.random((n, n)) < 0.5, 6, -6).astype(np.int8)
    J = np.triu(up, 1)
    J = (J + J.T).astype(np.int8)
    J[:64, :] = np.where(np.eye(n, dtype=bool)[:64], 0, 100)
    J[:, :64] = J[:64, :].T
    tb = O.tables(dict(S=40, t0=3000.0, t1=500.0, q=50.0, lam=0.7, ramp=True))
    eng.set_tables(tb)
    inst = eng.add_instance(J)
    trials = np.arange(48, dtype=np.uint32)
    res = eng.run(inst, 8, trials, trace=True, fields=True)
    _assert_equal(res, O.run_batch(J, n, tb, 8, trials, trace=True, nthreads=16), 32)
    h = O.fields(J, n, res["spins"][0].view(np.uint64))
    assert np.array_equal(res["fields"][0].astype(np.int64), h)


@pytest.mark.parametrize("kernel", ["batched", "plain"])
def test_int64_thresholds_8_round_teams(eng, kernel):
    """Tables whose decision constants exceed int32 (kcorr = 2^31: a reachable n_lin puts the correction beyond
    int32), so the range plan selects the int64-threshold team kernel; batched and single-anneal launches."""
    n = 900
    J, b = _small_int_instance(9, n, [1, -1, 3, -3], density=0.4, bmax=1000)
    S = 40
    T = np.linspace(6.0, 1.0, S)
    tb = dict(fourT=np.round(4 * T * 65536).astype(np.int32), q=np.full(S, 3 * 65536, np.int32),
              kcorr=np.full(S, 1 << 31, np.int64), kconst=np.zeros(S, np.int32))
    assert (n * (1 << 31)) >> 8 >= 2 ** 31 - 1
    eng.set_tables(tb)
    inst = eng.add_instance(J, bias=b)
    trials = np.arange(36 if kernel == "batched" else 3, dtype=np.uint32) + 50
    ref = O.run_batch(J, n, tb, 21, trials, bias=b, trace=True, nthreads=16)
    assert ref["nlin"].max() > 0
    res = eng.run(inst, 21, trials, trace=True, kernel=kernel)
    assert int(eng.get_option("last_wide")) == 1
    _assert_equal(res, ref, O.np_of(n) // 64)


def test_multi_instance_multi_bit_items(eng):
    """One launch over instances of different coupling ranges (int2 ER, torus, +-2/+-4 dense), each equal to the
    oracle."""
    from anechoic_amd import instances as I
    insts = [I.erdos_renyi(301, 700, 0.05), I.torus(302, *I.TORUS_SHAPES[1000]), _small_int_instance(303, 2000, [2, -2, 4])]
    tb = O.tables(dict(S=50, t0=6.0, t1=1.0, q=2.0, lam=0.9, ramp=True, scale=0.05))
    eng.set_tables(tb)
    ids = [eng.add_instance(J, bias=b) for J, b in insts]
    counts = [21, 40, 35]
    trials = np.arange(sum(counts), dtype=np.uint32) * 11 + 2
    res = eng.run_multi(list(zip(ids, counts)), 606, trials, trace=True)
    row = 0
    for (J, b), cnt in zip(insts, counts):
        ref = O.run_batch(J, J.shape[0], tb, 606, trials[row:row + cnt], bias=b, trace=True, nthreads=16)
        sl = slice(row, row + cnt)
        words = O.np_of(J.shape[0]) // 64
        assert np.array_equal(res["spins"][sl, :words].view(np.uint64), ref["spins"])
        assert np.array_equal(res["sum_sh"][sl], ref["sum_sh"]) and np.array_equal(res["nlin"][sl], ref["nlin"])
        assert np.array_equal(res["flips"][sl], ref["flips"]) and np.array_equal(res["sum_bs"][sl], ref["sum_bs"])
        row += cnt


def test_huge_field_bound_runs_on_32_member_teams(eng):
    """A field bound above 2^31 / 12 (biases of 2e8) is outside the 16-member team kernel's exact range (its odd crumb
    chain accumulates 4 J' N): the default shape rule runs it on 32-member teams, with results equal to the oracle,
    and forcing 16-member teams is refused."""
    from anechoic_amd import _native as N
    n = 1000
    J, _ = _small_int_instance(31, n, [1, -1], density=0.5)
    b = np.zeros(n, np.int32)
    b[::97] = 200_000_000
    b[5::89] = -200_000_000
    tb = O.tables(dict(S=30, t0=12.0, t1=5.0, q=8.0, lam=1.05, ramp=True))
    eng.set_tables(tb)
    inst = eng.add_instance(J, bias=b)
    assert eng.bound(inst) * 12 >= 2 ** 31
    trials = np.arange(40, dtype=np.uint32)
    res = eng.run(inst, 17, trials, trace=True)
    assert int(eng.get_option("last_team")) == 32
    _assert_equal(res, O.run_batch(J, n, tb, 17, trials, bias=b, trace=True, nthreads=16), O.np_of(n) // 64)
    old = eng.get_option("batched_team")
    eng.set_option("batched_team", 16)
    try:
        with pytest.raises(N.AncError, match="16-member teams"):
            eng.run(inst, 17, trials)
    finally:
        eng.set_option("batched_team", old)
