"""The opt-in fp4 team kernel (option "fp4" = 1; couplings in {0, +-1, +-2, +-3, +-4, +-6}, n <= 2048) against the
oracle, next to the default int8 kernel on the same inputs.

The fp4 kernel multiplies J' and the exchanged spins as e2m1 MFMA operands with fp32 sums; every value is exact, so
its results must equal the oracle's bit for bit, as the int8 kernel's do. Each test checks which kernel ran
(option "last_fp4") so that a silent switch of paths cannot pass unnoticed.
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


@pytest.fixture(scope="module")
def k2000():
    from anechoic_amd import instances as I
    J, _ = I.ising_random(20261203, 2000)
    return J


def _run_path(eng, fp4, inst, seed, trials, **kw):
    old = eng.get_option("fp4")
    eng.set_option("fp4", fp4)
    try:
        res = eng.run(inst, seed, trials, **kw)
        used = int(eng.get_option("last_fp4"))
    finally:
        eng.set_option("fp4", old)
    return res, used


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


@pytest.mark.parametrize("mode", ["counted", "kappa"])
def test_c2_batch_fp4_and_int8_vs_oracle(eng, k2000, mode):
    """R = 240 of the N = 2000 +-1 workload on both kernels: every anneal, step trace and final field equal."""
    sch = (dict(S=120, t0=12.0, t1=5.0, q=8.0, lam=1.05, ramp=True) if mode == "counted"
           else dict(S=120, t0=15.0, t1=5.0, q=8.0, kappa=2.0, ramp=True))
    tb = O.tables(sch)
    eng.set_tables(tb)
    inst = eng.add_instance(k2000)
    trials = np.arange(240, dtype=np.uint32) * 5 + 3
    ref = O.run_batch(k2000, 2000, tb, 0x5EED, trials, trace=True, nthreads=16)
    for fp4 in (1, 0):
        res, used = _run_path(eng, fp4, inst, 0x5EED, trials, trace=True, fields=True)
        assert used == fp4
        _assert_equal(res, ref, 32)
        h = O.fields(k2000, 2000, res["spins"][239].view(np.uint64))
        assert np.array_equal(res["fields"][239].astype(np.int64), h)


def test_plain_kernel_fp4_vs_oracle(eng, k2000):
    tb = O.tables(dict(S=100, t0=12.0, t1=5.0, q=8.0, lam=1.05, ramp=True))
    eng.set_tables(tb)
    inst = eng.add_instance(k2000)
    trials = np.array([0, 1, 99], dtype=np.uint32)
    res, used = _run_path(eng, 1, inst, 11, trials, trace=True, kernel="plain")
    assert used == 1
    _assert_equal(res, O.run_batch(k2000, 2000, tb, 11, trials, trace=True, nthreads=16), 32)


@pytest.mark.parametrize("n,R", [(2000, 64), (1500, 100), (777, 33)])
def test_every_fp4_magnitude_with_bias(eng, n, R):
    """Couplings drawn from every value fp4 holds exactly (0, +-1, +-2, +-3, +-4, +-6), with biases, sizes below 2048
    (padding slots) and 16- or 32-replica work items."""
    J, b = _small_int_instance(1000 + n, n, [1, -1, 2, -2, 3, -3, 4, -4, 6, -6], density=0.6, bmax=300)
    tb = O.tables(dict(S=60, t0=40.0, t1=3.0, q=10.0, lam=0.9, ramp=True, scale=0.5))
    eng.set_tables(tb)
    inst = eng.add_instance(J, bias=b)
    trials = np.arange(R, dtype=np.uint32) * 3 + 7
    res, used = _run_path(eng, 1, inst, 4321, trials, trace=True, fields=True)
    assert used == 1
    words = O.np_of(n) // 64
    _assert_equal(res, O.run_batch(J, n, tb, 4321, trials, bias=b, trace=True, nthreads=16), words)
    h = O.fields(J, n, res["spins"][R - 1, :words].view(np.uint64), b)
    assert np.array_equal(res["fields"][R - 1, :n].astype(np.int64), h)


def test_dense_max_magnitude_rows(eng):
    """Every coupling +-6 (the largest fp4 integer): the fp32 partial sums reach 6 (n - 1) 2^tau; still exact."""
    n = 2048
    rng = np.random.default_rng(5)
    up = np.where(rng.random((n, n)) < 0.5, 6, -6).astype(np.int8)
    J = np.triu(up, 1)
    J = (J + J.T).astype(np.int8)
    J[:64, :] = np.where(np.eye(n, dtype=bool)[:64], 0, 6)   # rows of all +6 (and their columns)
    J[:, :64] = J[:64, :].T
    tb = O.tables(dict(S=40, t0=3000.0, t1=500.0, q=50.0, lam=0.7, ramp=True))
    eng.set_tables(tb)
    inst = eng.add_instance(J)
    trials = np.arange(48, dtype=np.uint32)
    res, used = _run_path(eng, 1, inst, 8, trials, trace=True, fields=True)
    assert used == 1
    _assert_equal(res, O.run_batch(J, n, tb, 8, trials, trace=True, nthreads=16), 32)
    h = O.fields(J, n, res["spins"][0].view(np.uint64))
    assert np.array_equal(res["fields"][0].astype(np.int64), h)


def test_instance_outside_fp4_runs_int8(eng):
    """A single coupling of 5 (not an e2m1 value) keeps the instance on the int8 kernel, with the same results."""
    J, b = _small_int_instance(77, 1200, [1, -1, 2, -2], density=0.3, bmax=20)
    J[10, 500] = J[500, 10] = 5
    tb = O.tables(dict(S=50, t0=8.0, t1=1.0, q=3.0, kappa=1.5, ramp=True, scale=0.3))
    eng.set_tables(tb)
    inst = eng.add_instance(J, bias=b)
    trials = np.arange(40, dtype=np.uint32)
    res, used = _run_path(eng, 1, inst, 3, trials, trace=True)
    assert used == 0
    _assert_equal(res, O.run_batch(J, 1200, tb, 3, trials, bias=b, trace=True, nthreads=16), O.np_of(1200) // 64)


def test_int64_thresholds_both_kernels(eng):
    """Tables whose decision constants exceed int32 (the int64-threshold kernels), on both kernels."""
    n = 900
    J, b = _small_int_instance(9, n, [1, -1, 3, -3], density=0.4, bmax=1000)
    S = 40
    T = np.linspace(6.0, 1.0, S)
    tb = dict(fourT=np.round(4 * T * 65536).astype(np.int32), q=np.full(S, 3 * 65536, np.int32),
              kcorr=np.full(S, 1 << 31, np.int64), kconst=np.zeros(S, np.int32))
    assert (n * (1 << 31)) >> 8 >= 2 ** 31 - 1   # a reachable n_lin puts the correction beyond int32
    eng.set_tables(tb)
    inst = eng.add_instance(J, bias=b)
    trials = np.arange(36, dtype=np.uint32) + 50
    ref = O.run_batch(J, n, tb, 21, trials, bias=b, trace=True, nthreads=16)
    assert ref["nlin"].max() > 0
    for fp4 in (1, 0):
        res, used = _run_path(eng, fp4, inst, 21, trials, trace=True)
        assert used == fp4
        _assert_equal(res, ref, O.np_of(n) // 64)


def test_multi_instance_fp4_items(eng):
    """A multi-instance run of eligible n <= 2048 instances runs the fp4 kernel; one ineligible instance in the batch
    moves the whole launch to the int8 kernel. Both equal the oracle."""
    from anechoic_amd import instances as I
    insts = [I.erdos_renyi(301, 700, 0.05), I.torus(302, *I.TORUS_SHAPES[1000]), _small_int_instance(303, 2000, [2, -2, 4])]
    sch = dict(S=50, t0=6.0, t1=1.0, q=2.0, lam=0.9, ramp=True, scale=0.05)
    tb = O.tables(sch)
    eng.set_tables(tb)
    ids = [eng.add_instance(J, bias=b) for J, b in insts]
    counts = [21, 40, 35]
    trials = np.arange(sum(counts), dtype=np.uint32) * 11 + 2
    refs = [O.run_batch(J, J.shape[0], tb, 606, trials[sum(counts[:i]):sum(counts[:i + 1])], bias=b, trace=True, nthreads=16)
            for i, (J, b) in enumerate(insts)]

    def check(res):
        row = 0
        for (J, b), cnt, ref in zip(insts, counts, refs):
            sl = slice(row, row + cnt)
            words = O.np_of(J.shape[0]) // 64
            assert np.array_equal(res["spins"][sl, :words].view(np.uint64), ref["spins"])
            assert np.array_equal(res["sum_sh"][sl], ref["sum_sh"]) and np.array_equal(res["nlin"][sl], ref["nlin"])
            assert np.array_equal(res["flips"][sl], ref["flips"]) and np.array_equal(res["sum_bs"][sl], ref["sum_bs"])
            row += cnt

    eng.set_option("fp4", 1)
    try:
        res = eng.run_multi(list(zip(ids, counts)), 606, trials, trace=True)
        assert int(eng.get_option("last_fp4")) == 1
        check(res)
        Jx, bx = insts[2]
        Jx = Jx.copy()
        Jx[3, 4] = Jx[4, 3] = 5
        insts[2] = (Jx, bx)
        refs[2] = O.run_batch(Jx, 2000, tb, 606, trials[61:], bias=bx, trace=True, nthreads=16)
        ids[2] = eng.add_instance(Jx, bias=bx)
        res = eng.run_multi(list(zip(ids, counts)), 606, trials, trace=True)
        assert int(eng.get_option("last_fp4")) == 0
        check(res)
    finally:
        eng.set_option("fp4", 0)
