"""Large instances (NP(n) > 3072, up to N = 16384): the streamed int8 MFMA kernel.

Checks (all exact):
  * sampled replicas against the oracle (beyond NP = 2048 the oracle's generalisation NP = 256 ceil(n / 256) is the
    reference: parity unpinned by the reference headers): spins, sum_sh, sum_bs, flips, and returned fields;
  * the full N = 16384, R = 1024 shape through size-independent invariants for every replica: returned fields equal
    b + J s of the returned spins (float32 products on the GPU, exact for these integers), sum_sh = sum s h,
    sum_bs = sum b s, spins of padding slots clear;
  * every kernel slot count (4096, 8192, 16384) with and without phantom rounds (NP(n) < kernel slot count), with
    biases and a partial last work item, on 128- and on 64-replica work items (option large_tile);
  * the counted correction with the full n_lin trace (instance-wide band counts reduced every step) against the
    oracle.
"""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

KAPPA = dict(shape="linear", t0=15.0, t1=5.0, q=8.0, kappa=2.0, ramp=True)


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
def j16k():
    J, _ = O.gen_ising_random(20261203, 16384)   # gen_ising random, K = 2: the +-1 complete graph
    return J


def _check_vs_oracle(res, J, n, tb, seed, trials, rows, bias=None):
    words = O.np_of(n) // 64
    rows = np.asarray(rows)
    ref = O.run_batch(J, n, tb, seed, trials[rows], bias=bias, nthreads=16)
    assert np.array_equal(res["spins"][rows, :words].view(np.uint64), ref["spins"]), n
    assert np.array_equal(res["sum_sh"][rows], ref["sum_sh"]), n
    assert np.array_equal(res["sum_bs"][rows], ref["sum_bs"]), n
    assert np.array_equal(res["flips"][rows], ref["flips"]), n


def _invariants_gpu(res, J, n, bias=None):
    import torch
    words = O.np_of(n) // 64
    sp = torch.from_numpy(res["spins"][:, :words].copy()).cuda()
    bits = ((sp.unsqueeze(-1) >> torch.arange(64, device="cuda")) & 1).reshape(sp.shape[0], -1)
    s = (2 * bits - 1).to(torch.float32)[:, :n]
    # padding slots of the packed spins stay clear
    assert int(bits[:, n:].sum()) == 0
    Jt = torch.from_numpy(np.ascontiguousarray(J[:n, :n])).cuda().to(torch.float32)
    h = (s @ Jt).round().to(torch.int64)
    b = torch.zeros(n, dtype=torch.int64, device="cuda") if bias is None else torch.from_numpy(bias.astype(np.int64)).cuda()
    h = h + b
    f = torch.from_numpy(res["fields"][:, :n].astype(np.int64)).cuda()
    assert torch.equal(f, h)
    si = s.to(torch.int64)
    assert np.array_equal((si * h).sum(1).cpu().numpy(), res["sum_sh"])
    assert np.array_equal((si * b).sum(1).cpu().numpy(), res["sum_bs"])


def test_n16384_kappa_vs_oracle(eng, j16k):
    n = 16384
    sch = dict(KAPPA, S=24)
    tb = O.tables(sch)
    eng.set_tables(tb)
    inst = eng.add_instance(j16k)
    trials = (np.arange(256, dtype=np.uint32) * 5 + 3).astype(np.uint32)
    res = eng.run(inst, 0xB16B, trials, fields=True)
    _check_vs_oracle(res, j16k, n, tb, 0xB16B, trials, [0, 1, 63, 64, 127, 128, 200, 255])
    for r in (0, 255):
        h = O.fields(j16k, n, res["spins"][r].view(np.uint64))
        assert np.array_equal(res["fields"][r].astype(np.int64), h)
    assert int(res["flips"].min()) > 0


def test_n16384_r1024_full_shape_invariants(eng, j16k):
    n, R = 16384, 1024
    sch = dict(KAPPA, S=40)
    tb = O.tables(sch)
    eng.set_tables(tb)
    inst = eng.add_instance(j16k)
    trials = np.arange(R, dtype=np.uint32) + np.uint32(1000)
    res = eng.run(inst, 77, trials, fields=True)
    _invariants_gpu(res, j16k, n)
    _check_vs_oracle(res, j16k, n, tb, 77, trials, [0, 300, 511, 512, 1023])


@pytest.fixture
def tile(eng, request):
    """Sets the large kernel's replicas per work item for one test, and checks afterwards that it was used."""
    old = int(eng.get_option("large_tile"))
    eng.set_option("large_tile", request.param)
    yield request.param
    used = int(eng.get_option("last_large_tile"))
    eng.set_option("large_tile", old)
    assert used == request.param


@pytest.mark.parametrize("tile", [128, 64], indirect=True)
@pytest.mark.parametrize("n", [3500, 4096, 5000, 8192, 9000])
def test_large_slot_counts_with_bias_vs_oracle(eng, n, tile):
    """Kernel slot counts 4096 (n = 3500: 14 of 16 rounds drawn; 4096), 8192 (n = 5000: 20 of 32; 8192) and
    16384 (n = 9000: 36 of 64), with biases, against the oracle for every replica of a partial last item (130
    replicas: items of 128 + 2 or 64 + 64 + 2)."""
    J, b = O.gen_ising_random(1000 + n, n, kbits=2, pzero=0.0, bmax=3)
    sch = dict(KAPPA, S=20, t0=10.0, t1=2.0)
    tb = O.tables(sch)
    eng.set_tables(tb)
    inst = eng.add_instance(J, bias=b)
    trials = (np.arange(130, dtype=np.uint32) * 11 + 7).astype(np.uint32)
    res = eng.run(inst, 4321, trials, fields=True)
    _check_vs_oracle(res, J, n, tb, 4321, trials, np.arange(130), bias=b)
    _invariants_gpu(res, J, n, bias=b)


@pytest.mark.parametrize("tile", [128, 64], indirect=True)
@pytest.mark.parametrize("n,R", [(5000, 130), (16384, 64)])
def test_large_counted_correction_with_trace_vs_oracle(eng, n, R, tile):
    """The counted correction at n > 3072: the band count of every replica is reduced over the whole instance every
    step; spins, sums, flips and the full n_lin trace equal the oracle's."""
    J, b = O.gen_ising_random(77 + n, n, kbits=2, pzero=0.0, bmax=2)
    scale = float(np.sqrt(n / 2000.0))
    sch = dict(S=16, shape="linear", t0=12.0 * scale, t1=5.0 * scale, q=8.0 * scale, lam=1.05, ramp=True)
    tb = O.tables(sch)
    eng.set_tables(tb)
    inst = eng.add_instance(J, bias=b)
    trials = (np.arange(R, dtype=np.uint32) * 3 + 9).astype(np.uint32)
    res = eng.run(inst, 555, trials, trace=True)
    rows = np.arange(R) if n <= 5000 else np.array([0, 1, 31, 63])
    words = O.np_of(n) // 64
    ref = O.run_batch(J, n, tb, 555, trials[rows], bias=b, trace=True, nthreads=16)
    assert np.array_equal(res["spins"][rows, :words].view(np.uint64), ref["spins"]), n
    assert np.array_equal(res["sum_sh"][rows], ref["sum_sh"]) and np.array_equal(res["sum_bs"][rows], ref["sum_bs"])
    assert np.array_equal(res["flips"][rows], ref["flips"]) and np.array_equal(res["nlin"][rows], ref["nlin"])
    assert int(res["nlin"].min()) >= 0 and int(res["nlin"][:, 0].max()) > 0


def test_large_refused_by_multi_instance_runs(eng):
    from anechoic_amd import _native as N
    J, _ = O.gen_ising_random(5, 4000)
    inst = eng.add_instance(J)
    eng.set_tables(O.tables(dict(KAPPA, S=10)))
    with pytest.raises(N.AncError, match="anc_run_batched"):
        eng.run_multi([(inst, 4)], 1)


@pytest.mark.parametrize("n,kbits,pzero,bmax", [(3000, 2, 0.0, 0), (777, 4, 0.5, 300), (2048, 8, 0.9, 5)])
def test_device_generator_matches_reference_generator(n, kbits, pzero, bmax):
    """anc_gen_ising against the oracle's C restatement of gen_ising random and the numpy generator (whose N = 2000
    output carries the reference generator's fingerprint in the golden fixtures, checked in test_oracle.py); the
    padding columns are zeroed."""
    from anechoic_amd import instances as I
    J, b = I.ising_random_device(99 + n, n, kbits=kbits, pzero=pzero, bmax=bmax, ld=n + 13)
    Jr, br = O.gen_ising_random(99 + n, n, kbits=kbits, pzero=pzero, bmax=bmax)
    Jn, bn = I.ising_random(99 + n, n, kbits=kbits, pzero=pzero, bmax=bmax)
    Jh = J.cpu().numpy()
    assert np.array_equal(Jh[:, :n], Jr[:n, :n]) and np.array_equal(Jh[:, :n], Jn)
    assert not Jh[:, n:].any()
    assert np.array_equal(b.cpu().numpy(), br[:n]) and np.array_equal(b.cpu().numpy(), bn)


@pytest.mark.parametrize("mode", ["counted", "kappa"])
def test_large_int64_thresholds_vs_oracle(eng, mode):
    """n = 5000 with tables whose decision constants the range plan cannot prove within int32, so the large kernel
    runs its int64-threshold instantiation: counted (kcorr = 2^27: n_lin up to n puts the correction beyond int32,
    with the full n_lin trace) and temperature-scaled (one step with kconst near 2^31). Every replica of a partial
    second work item against the oracle."""
    n, R, S = 5000, 130, 16
    J, b = O.gen_ising_random(4242, n, kbits=2, pzero=0.0, bmax=2)
    scale = float(np.sqrt(n / 2000.0))
    T = np.linspace(12.0 * scale, 5.0 * scale, S)
    fourT = np.round(4 * T * 65536).astype(np.int32)
    q = np.full(S, int(round(8.0 * scale * 65536)), np.int32)
    if mode == "counted":
        kcorr = np.full(S, 1 << 27, np.int64)
        kcorr[0] = 0
        kconst = np.zeros(S, np.int32)
        assert (n * (1 << 27)) >> 8 >= 2 ** 31 - 1
    else:
        kcorr = np.zeros(S, np.int64)
        kconst = np.concatenate([[0], np.round(2.0 * T[:-1] * 65536)]).astype(np.int32)
        kconst[7] = 2 ** 31 - (1 << 20)
    tb = dict(fourT=fourT, q=q, kcorr=kcorr, kconst=kconst)
    eng.set_tables(tb)
    inst = eng.add_instance(J, bias=b)
    trials = (np.arange(R, dtype=np.uint32) * 7 + 1).astype(np.uint32)
    res = eng.run(inst, 0x64B, trials, trace=True, fields=True)
    assert int(eng.get_option("last_wide")) == 1
    words = O.np_of(n) // 64
    ref = O.run_batch(J, n, tb, 0x64B, trials, bias=b, trace=True, nthreads=16)
    assert np.array_equal(res["spins"][:, :words].view(np.uint64), ref["spins"])
    assert np.array_equal(res["sum_sh"], ref["sum_sh"]) and np.array_equal(res["sum_bs"], ref["sum_bs"])
    assert np.array_equal(res["flips"], ref["flips"]) and np.array_equal(res["nlin"], ref["nlin"])
    if mode == "counted":
        assert int(ref["nlin"][:, 1:].max()) > 0   # the int64 corrections were reached
    _invariants_gpu(res, J, n, bias=b)
