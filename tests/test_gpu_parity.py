"""Bit-exact parity of the HIP engine (plain and batched team kernels) with the golden model.

Every comparison is exact: final spins, sum_sh, sum_bs, flips and the full per-step n_lin trace, against
  * the reference headers' own outputs (golden_trials.json, written in the development container),
  * the reference-held known answers (reference_cases.json),
  * the oracle (C restatement, run on the host) for full R = 240 batches of the N = 2000 workloads.
"""
import numpy as np
import pytest

from conftest import cdiv, hex_to_words
from helpers import case_schedule, golden_instance, reference_matrix
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


def _tables(golden_tables, name):
    g = golden_tables[name]
    return dict(fourT=np.array(g["fourT"], np.int32), q=np.array(g["q"], np.int32), kcorr=np.array(g["kcorr"], np.int64),
                kconst=np.array(g["kconst"], np.int32))


def _assert_same(res, r, g, ctx):
    assert int(res["sum_sh"][r]) == g["sum_sh"], ctx
    assert int(res["sum_bs"][r]) == g["sum_bs"], ctx
    assert int(res["flips"][r]) == g["flips"], ctx
    assert np.array_equal(res["spins"][r].view(np.uint64), hex_to_words(g["spins"])), ctx
    assert res["nlin"][r].tolist() == g["nlin"], ctx


def test_plain_kernel_c1_full_trace(eng, k2000, golden_tables, golden_trials):
    """N = 2000, +-1 complete graph, counted Onsager, linear T 12 -> 5, S = 1000: one anneal per launch."""
    rec = golden_trials["c1_plain_kernel"]
    eng.set_tables(_tables(golden_tables, rec["schedule"]))
    inst = eng.add_instance(k2000)
    res = eng.run(inst, rec["seed"], [g["trial"] for g in rec["results"]], trace=True, kernel="plain")
    for r, g in enumerate(rec["results"]):
        _assert_same(res, r, g, ("plain", g["trial"]))


@pytest.mark.parametrize("key", ["c2_counted", "c2_kappa", "c2_plain", "c2_tec", "c2_linear_kappa"])
def test_batched_matches_reference_headers(eng, k2000, golden_tables, golden_trials, key):
    rec = golden_trials[key]
    eng.set_tables(_tables(golden_tables, rec["schedule"]))
    inst = eng.add_instance(k2000)
    trials = [g["trial"] for g in rec["results"]]
    res = eng.run(inst, rec["seed"], trials, trace=True, kernel="batched")
    for r, g in enumerate(rec["results"]):
        _assert_same(res, r, g, (key, g["trial"]))


@pytest.mark.parametrize("key", ["bias_er_800_counted", "bias_er_800_kappa", "bias_torus_1000_counted",
                                 "bias_torus_1000_kappa", "bias_er_2000_counted", "bias_er_2000_kappa"])
def test_batched_bias_instances(eng, golden_tables, golden_trials, key):
    rec = golden_trials[key]
    J, bias = golden_instance(rec["instance"])
    eng.set_tables(_tables(golden_tables, rec["schedule"]))
    inst = eng.add_instance(J, bias=bias)
    res = eng.run(inst, rec["seed"], [g["trial"] for g in rec["results"]], trace=True)
    for r, g in enumerate(rec["results"]):
        _assert_same(res, r, g, (key, g["trial"]))


def test_reference_held_known_answers(eng, ref_cases):
    """Every known answer of the reference's cases.txt suites, through the batched kernel."""
    checked = 0
    for suite in ref_cases:
        m = suite["matrix"]
        J, bias, scale = reference_matrix(m)
        n = m["n"]
        inst_b = eng.add_instance(J[:n, :n].copy(), bias=bias) if bias is not None else None
        inst_0 = eng.add_instance(J[:n, :n].copy())
        for c in suite["cases"]:
            from anechoic_amd import tables
            eng.set_schedule(case_schedule(c, scale))
            use_b = bias is not None and c["use_bias"]
            trials = [g["trial"] for g in c["trials"]]
            res = eng.run(inst_b if use_b else inst_0, c["seed"], trials)
            for r, g in enumerate(c["trials"]):
                ctx = (suite["name"], c["name"], g["trial"])
                assert int(res["flips"][r]) == g["flips"], ctx
                if "sum_sh" in g:
                    assert int(res["sum_sh"][r]) == g["sum_sh"], ctx
                if "sum_bs" in g:
                    assert int(res["sum_bs"][r]) == g["sum_bs"], ctx
                if "energy2" in g:
                    assert -(int(res["sum_sh"][r]) + int(res["sum_bs"][r])) == g["energy2"], ctx
                if "cut" in g:
                    assert cdiv(m["sumw"] + cdiv(int(res["sum_sh"][r]), 2), 2) == g["cut"], ctx
                checked += 1
        eng._inst.clear()
        from anechoic_amd import _native as N
        N.check(N.lib().anc_clear_instances(eng._ctx))
    assert checked >= 60


@pytest.fixture(params=[16, 32], ids=["team16", "team32"])
def team(eng, request):
    """The two team shapes of the batched launch for n <= 2048: 16-member teams (16-replica items, two teams per XCD,
    the default) and 32-member teams (one per XCD, 16- or 32-replica items)."""
    old = eng.get_option("batched_team")
    eng.set_option("batched_team", request.param)
    yield request.param
    eng.set_option("batched_team", old)


@pytest.mark.parametrize("mode", ["counted", "kappa"])
def test_batched_r240_all_replicas_vs_oracle(eng, k2000, mode, team):
    """The full R = 240 batch of the N = 2000 workload, every replica and every step checked against the oracle, on
    both team shapes."""
    sch = (dict(S=150, t0=12.0, t1=5.0, q=8.0, lam=1.05, ramp=True) if mode == "counted"
           else dict(S=150, t0=15.0, t1=5.0, q=8.0, kappa=2.0, ramp=True))
    tb = O.tables(sch)
    eng.set_tables(tb)
    inst = eng.add_instance(k2000)
    trials = np.arange(240, dtype=np.uint32) * 7 + 11
    res = eng.run(inst, 0xABCDEF, trials, trace=True, fields=True)
    assert int(eng.get_option("last_team")) == team
    ref = O.run_batch(k2000, 2000, tb, 0xABCDEF, trials, trace=True, nthreads=16)
    assert np.array_equal(res["spins"].view(np.uint64), ref["spins"])
    assert np.array_equal(res["sum_sh"], ref["sum_sh"]) and np.array_equal(res["flips"], ref["flips"])
    assert np.array_equal(res["sum_bs"], ref["sum_bs"])
    assert np.array_equal(res["nlin"], ref["nlin"])
    for r in (0, 117, 239):   # returned fields equal a from-scratch recomputation b + J s
        h = O.fields(k2000, 2000, res["spins"][r].view(np.uint64))
        assert np.array_equal(res["fields"][r].astype(np.int64), h)


@pytest.mark.parametrize("R", [1, 17, 100, 257])
def test_batched_uneven_replica_counts(eng, R, team):
    J, bias = O.gen_random_k(4242, 0.2, 2, 1100, 30, ld=1100, rows=1100)
    tb = O.tables(dict(S=40, t0=5.0, t1=1.0, q=2.0, lam=0.8, ramp=True, scale=0.3))
    eng.set_tables(tb)
    inst = eng.add_instance(J, bias=bias)
    trials = (np.arange(R, dtype=np.uint32) * 2654435761) & 0x0FFFFFFF
    res = eng.run(inst, 31337, trials, trace=True)
    assert int(eng.get_option("last_team")) == team
    ref = O.run_batch(J, 1100, tb, 31337, trials, bias=bias, trace=True, nthreads=16)
    assert np.array_equal(res["spins"].view(np.uint64), ref["spins"])
    assert np.array_equal(res["sum_sh"], ref["sum_sh"]) and np.array_equal(res["sum_bs"], ref["sum_bs"])
    assert np.array_equal(res["flips"], ref["flips"]) and np.array_equal(res["nlin"], ref["nlin"])


def test_instance_validation_errors(eng):
    from anechoic_amd import _native as N
    J = np.zeros((64, 64), np.int8)
    J[1, 2] = 1
    with pytest.raises(N.AncError) as e:
        eng.add_instance(J)
    assert e.value.code == -2
    J[2, 1] = 1
    J[3, 3] = 1
    with pytest.raises(N.AncError) as e:
        eng.add_instance(J)
    assert e.value.code == -2


def test_multi_instance_uneven_batch_vs_oracle(eng):
    """Uneven batch of ER and torus instances (int2 J on {-2,-1,+1,+2} and int2 bias) in one launch."""
    from anechoic_amd import instances as I
    insts = [I.erdos_renyi(101, 800, 0.05), I.torus(102, *I.TORUS_SHAPES[1000]), I.erdos_renyi(103, 2000, 0.01),
             I.torus(104, *I.TORUS_SHAPES[800])]
    sch = dict(S=80, t0=6.0, t1=1.0, q=2.0, kappa=1.5, ramp=True, scale=0.05)
    tb = O.tables(sch)
    eng.set_tables(tb)
    ids = [eng.add_instance(J, bias=b) for J, b in insts]
    counts = [40, 17, 64, 3]
    trials = np.arange(sum(counts), dtype=np.uint32) + 1000
    res = eng.run_multi(list(zip(ids, counts)), 99, trials, trace=True)
    row = 0
    for (J, b), cnt in zip(insts, counts):
        n = J.shape[0]
        ref = O.run_batch(J, n, tb, 99, trials[row:row + cnt], bias=b, trace=True, nthreads=16)
        sl = slice(row, row + cnt)
        assert np.array_equal(res["spins"][sl, :32].view(np.uint64), ref["spins"]), (n, row)
        assert np.array_equal(res["sum_sh"][sl], ref["sum_sh"]) and np.array_equal(res["sum_bs"][sl], ref["sum_bs"])
        assert np.array_equal(res["flips"][sl], ref["flips"]) and np.array_equal(res["nlin"][sl], ref["nlin"])
        row += cnt


def test_np3072_team_shape_and_mixed_classes_vs_oracle(eng):
    """n = 3000 (NP = 3072, 12 rounds; one-XCD teams) alone and mixed with NP = 2048 instances in one call.
    Beyond NP = 2048 the oracle's generalisation is the reference (parity unpinned by the reference headers)."""
    from anechoic_amd import instances as I
    insts = [I.torus(201, *I.TORUS_SHAPES[3000]), I.erdos_renyi(202, 1000, 0.03), I.erdos_renyi(203, 2600, 0.01)]
    sch = dict(S=60, t0=6.0, t1=1.0, q=2.0, lam=0.8, ramp=True, scale=0.05)
    tb = O.tables(sch)
    eng.set_tables(tb)
    ids = [eng.add_instance(J, bias=b) for J, b in insts]
    counts = [33, 20, 7]
    trials = np.arange(sum(counts), dtype=np.uint32) * 3 + 5
    res = eng.run_multi(list(zip(ids, counts)), 4242, trials, trace=True, fields=True)
    row = 0
    for (J, b), cnt in zip(insts, counts):
        n = J.shape[0]
        words = O.np_of(n) // 64
        ref = O.run_batch(J, n, tb, 4242, trials[row:row + cnt], bias=b, trace=True, nthreads=16)
        sl = slice(row, row + cnt)
        assert np.array_equal(res["spins"][sl, :words].view(np.uint64), ref["spins"]), (n, row)
        assert np.array_equal(res["sum_sh"][sl], ref["sum_sh"]) and np.array_equal(res["sum_bs"][sl], ref["sum_bs"])
        assert np.array_equal(res["flips"][sl], ref["flips"]) and np.array_equal(res["nlin"][sl], ref["nlin"])
        h = O.fields(J, n, res["spins"][row, :words].view(np.uint64), b)
        assert np.array_equal(res["fields"][row, :n].astype(np.int64), h)
        row += cnt
    single = eng.run(ids[0], 4242, trials[:33], trace=True)
    assert np.array_equal(single["sum_sh"], res["sum_sh"][:33]) and np.array_equal(single["nlin"], res["nlin"][:33])


@pytest.mark.parametrize("xflavor", [1, 2, 3])
def test_exchange_flavours_give_identical_results(eng, k2000, xflavor):
    tb = O.tables(dict(S=50, t0=12.0, t1=5.0, q=8.0, lam=1.05, ramp=True))
    eng.set_tables(tb)
    inst = eng.add_instance(k2000)
    old = eng.get_option("xflavor")
    eng.set_option("xflavor", xflavor)
    try:
        res = eng.run(inst, 77, np.arange(64, dtype=np.uint32), trace=True)
        assert int(eng.get_option("xflavor")) == xflavor   # the requested exchange produced the results
    finally:
        eng.set_option("xflavor", old)
    ref = O.run_batch(k2000, 2000, tb, 77, np.arange(64, dtype=np.uint32), trace=True, nthreads=16)
    assert np.array_equal(res["sum_sh"], ref["sum_sh"]) and np.array_equal(res["nlin"], ref["nlin"])


@pytest.mark.parametrize("xflavor", [3, 1])
def test_r240_batch_both_exchange_scopes_vs_oracle(eng, k2000, xflavor):
    """R = 240 (32-replica work items) with the L2-local and the device-scope exchange: every anneal equal to the
    oracle, with full n_lin traces and the final fields of the last replica."""
    tb = O.tables(dict(S=60, t0=12.0, t1=5.0, q=8.0, lam=1.05, ramp=True))
    eng.set_tables(tb)
    inst = eng.add_instance(k2000)
    old = eng.get_option("xflavor")
    eng.set_option("xflavor", xflavor)
    try:
        trials = (np.arange(240, dtype=np.uint32) * 13 + 1).astype(np.uint32)
        res = eng.run(inst, 0xA11, trials, trace=True, fields=True)
        assert int(eng.get_option("xflavor")) == xflavor
    finally:
        eng.set_option("xflavor", old)
    ref = O.run_batch(k2000, 2000, tb, 0xA11, trials, trace=True, nthreads=16)
    assert np.array_equal(res["spins"][:, :32].view(np.uint64), ref["spins"])
    assert np.array_equal(res["sum_sh"], ref["sum_sh"]) and np.array_equal(res["flips"], ref["flips"])
    assert np.array_equal(res["nlin"], ref["nlin"])
    h = O.fields(k2000, 2000, res["spins"][239].view(np.uint64))
    assert np.array_equal(res["fields"][239].astype(np.int64), h)


@pytest.fixture(params=[0, 1], ids=["flip_bytes", "spin_nibbles"])
def other_exchange(eng, request):
    """The exchanges other than the default spin crumbs (option "exchange" 2): the flip matrix (0: tagged {-2, 0, 2}
    bytes, accumulated field) and the spin nibbles (1: N = s + 1 + tau in four bits)."""
    old = eng.get_option("exchange")
    eng.set_option("exchange", request.param)
    yield eng
    eng.set_option("exchange", old)


@pytest.mark.parametrize("R", [240, 100])
def test_other_exchange_batch_vs_oracle(other_exchange, k2000, R):
    """The N = 2000 workload (32- and 16-replica work items) on the flip-matrix and spin-nibble exchanges."""
    eng = other_exchange
    tb = O.tables(dict(S=80, t0=12.0, t1=5.0, q=8.0, lam=1.05, ramp=True))
    eng.set_tables(tb)
    inst = eng.add_instance(k2000)
    trials = (np.arange(R, dtype=np.uint32) * 19 + 4).astype(np.uint32)
    res = eng.run(inst, 0xF11, trials, trace=True, fields=True)
    ref = O.run_batch(k2000, 2000, tb, 0xF11, trials, trace=True, nthreads=16)
    assert np.array_equal(res["spins"][:, :32].view(np.uint64), ref["spins"])
    assert np.array_equal(res["sum_sh"], ref["sum_sh"]) and np.array_equal(res["flips"], ref["flips"])
    assert np.array_equal(res["nlin"], ref["nlin"])
    h = O.fields(k2000, 2000, res["spins"][R - 1].view(np.uint64))
    assert np.array_equal(res["fields"][R - 1].astype(np.int64), h)


def test_other_exchange_mixed_shapes_vs_oracle(other_exchange):
    """Both team shapes (n = 3000: 64-workgroup teams with the flip matrix, 32-workgroup 12-round teams with spin
    nibbles; n = 1000: 32-workgroup teams) in one mixed launch."""
    eng = other_exchange
    from anechoic_amd import instances as I
    insts = [I.torus(211, *I.TORUS_SHAPES[3000]), I.erdos_renyi(212, 1000, 0.03)]
    tb = O.tables(dict(S=50, t0=6.0, t1=1.0, q=2.0, kappa=1.5, ramp=True, scale=0.05))
    eng.set_tables(tb)
    ids = [eng.add_instance(J, bias=b) for J, b in insts]
    counts = [40, 24]
    trials = np.arange(sum(counts), dtype=np.uint32) * 5 + 9
    res = eng.run_multi(list(zip(ids, counts)), 515, trials, trace=True)
    row = 0
    for (J, b), cnt in zip(insts, counts):
        n = J.shape[0]
        words = O.np_of(n) // 64
        ref = O.run_batch(J, n, tb, 515, trials[row:row + cnt], bias=b, trace=True, nthreads=16)
        sl = slice(row, row + cnt)
        assert np.array_equal(res["spins"][sl, :words].view(np.uint64), ref["spins"]), (n, row)
        assert np.array_equal(res["sum_sh"][sl], ref["sum_sh"]) and np.array_equal(res["sum_bs"][sl], ref["sum_bs"])
        assert np.array_equal(res["flips"][sl], ref["flips"]) and np.array_equal(res["nlin"][sl], ref["nlin"])
        row += cnt
    single = eng.run(ids[0], 515, trials[:40], trace=True)
    assert np.array_equal(single["sum_sh"], res["sum_sh"][:40]) and np.array_equal(single["nlin"], res["nlin"][:40])


@pytest.mark.parametrize("n,R,xflavor,nt3", [(3000, 100, 3, 1), (3000, 100, 1, 1), (2600, 40, 3, 1), (3000, 100, 3, 2),
                                             (3000, 100, 1, 2), (2600, 40, 3, 2)])
def test_np3072_one_xcd_teams_vs_oracle(eng, n, R, xflavor, nt3):
    """n = 2049..3072 with the spin exchange: 32-workgroup teams of 12 rounds (96 rows per member) on one XCD each,
    16-replica work items (all J' fragments in registers) and 32-replica ones (the last load block's fragments in LDS);
    both exchange flavours, partial last items, and phantom rounds (n = 2600: NP 2816 < 3072)."""
    from anechoic_amd import instances as I
    J, b = (I.torus(301, *I.TORUS_SHAPES[3000]) if n == 3000 else I.erdos_renyi(302, n, 0.02))
    tb = O.tables(dict(S=60, t0=6.0, t1=1.0, q=2.0, lam=0.8, ramp=True, scale=0.05))
    eng.set_tables(tb)
    inst = eng.add_instance(J, bias=b)
    old = {k: eng.get_option(k) for k in ("xflavor", "multi_nt3", "batched_nt")}
    eng.set_option("xflavor", xflavor)
    eng.set_option("multi_nt3", nt3)
    eng.set_option("batched_nt", nt3)
    try:
        trials = (np.arange(R, dtype=np.uint32) * 5 + 17).astype(np.uint32)
        res = eng.run(inst, 0x3072, trials, trace=True, fields=True)
        assert int(eng.get_option("xflavor")) == xflavor and int(eng.get_option("last_nt3")) == nt3
    finally:
        for k, v in old.items():
            eng.set_option(k, v)
    ref = O.run_batch(J, n, tb, 0x3072, trials, bias=b, trace=True, nthreads=16)
    words = O.np_of(n) // 64
    assert np.array_equal(res["spins"][:, :words].view(np.uint64), ref["spins"])
    assert np.array_equal(res["sum_sh"], ref["sum_sh"]) and np.array_equal(res["sum_bs"], ref["sum_bs"])
    assert np.array_equal(res["flips"], ref["flips"]) and np.array_equal(res["nlin"], ref["nlin"])
    h = O.fields(J, n, res["spins"][R - 1, :words].view(np.uint64), b)
    assert np.array_equal(res["fields"][R - 1, :n].astype(np.int64), h)


@pytest.mark.parametrize("nt3", [1, 2])
def test_np3072_one_xcd_teams_int64_thresholds(eng, nt3):
    """The 12-round one-XCD team with decision constants beyond int32 (the int64-threshold instantiations, 16- and
    32-replica items)."""
    from anechoic_amd import instances as I
    n = 2900
    J, b = I.erdos_renyi(303, n, 0.01)
    S = 30
    T = np.linspace(4.0, 1.0, S)
    tb = dict(fourT=np.round(4 * T * 65536).astype(np.int32), q=np.full(S, 2 * 65536, np.int32),
              kcorr=np.full(S, 1 << 31, np.int64), kconst=np.zeros(S, np.int32))
    assert (n * (1 << 31)) >> 8 >= 2 ** 31 - 1
    eng.set_tables(tb)
    inst = eng.add_instance(J, bias=b)
    trials = np.arange(33, dtype=np.uint32) + 3
    old = {k: eng.get_option(k) for k in ("multi_nt3", "batched_nt")}
    eng.set_option("multi_nt3", nt3)
    eng.set_option("batched_nt", nt3)
    try:
        res = eng.run(inst, 99, trials, trace=True)
        assert int(eng.get_option("last_nt3")) == nt3
    finally:
        for k, v in old.items():
            eng.set_option(k, v)
    assert int(eng.get_option("last_wide")) == 1
    ref = O.run_batch(J, n, tb, 99, trials, bias=b, trace=True, nthreads=16)
    assert ref["nlin"].max() > 0
    assert np.array_equal(res["spins"][:, :O.np_of(n) // 64].view(np.uint64), ref["spins"])
    assert np.array_equal(res["sum_sh"], ref["sum_sh"]) and np.array_equal(res["nlin"], ref["nlin"])


@pytest.mark.parametrize("xcd2,nt3", [(1, 1), (3, 1), (7, 1), (8, 1), (3, 2), (5, 2), (8, 2)])
def test_mixed_launch_every_xcd_split(eng, xcd2, nt3):
    """Forced XCD splits of the mixed launch (spin exchange: any 1..7 XCDs for the n <= 2048 teams, or 8: every XCD
    runs both classes in turn) with 16- and 32-replica items of the n = 2049..3072 instances, each equal to the
    oracle."""
    from anechoic_amd import instances as I
    insts = [I.erdos_renyi(311, 3000, 0.004), I.torus(312, *I.TORUS_SHAPES[800]), I.erdos_renyi(313, 2500, 0.01)]
    tb = O.tables(dict(S=40, t0=6.0, t1=1.0, q=2.0, kappa=1.5, ramp=True, scale=0.05))
    eng.set_tables(tb)
    ids = [eng.add_instance(J, bias=b) for J, b in insts]
    counts = [50, 70, 20]
    trials = np.arange(sum(counts), dtype=np.uint32) + 700
    old = {k: eng.get_option(k) for k in ("mixed_xcd2", "multi_nt3")}
    eng.set_option("mixed_xcd2", xcd2)
    eng.set_option("multi_nt3", nt3)
    try:
        res = eng.run_multi(list(zip(ids, counts)), 31, trials, trace=True)
        assert int(eng.get_option("last_xcd2")) == xcd2 and int(eng.get_option("last_nt3")) == nt3
    finally:
        for k, v in old.items():
            eng.set_option(k, v)
    row = 0
    for (J, b), cnt in zip(insts, counts):
        n = J.shape[0]
        ref = O.run_batch(J, n, tb, 31, trials[row:row + cnt], bias=b, trace=True, nthreads=16)
        sl = slice(row, row + cnt)
        assert np.array_equal(res["spins"][sl, :O.np_of(n) // 64].view(np.uint64), ref["spins"]), (n, row)
        assert np.array_equal(res["sum_sh"][sl], ref["sum_sh"]) and np.array_equal(res["nlin"][sl], ref["nlin"])
        assert np.array_equal(res["flips"][sl], ref["flips"]) and np.array_equal(res["sum_bs"][sl], ref["sum_bs"])
        row += cnt


@pytest.mark.parametrize("wide", [False, True])
def test_small_instances_on_four_round_teams_vs_oracle(eng, wide):
    """Multi-instance batches of n <= 1024 instances run on 4-round teams (J' laid out with 1024 slots; each lane still
    draws the 8 numbers per step of the NP = 2048 slot layout, rounds 4 .. 7 being padding), in turn with NP = 2048
    and NP = 3072 items on every XCD: equal to the oracle, with int32 and with int64 decision thresholds, and identical
    to the 8-round teams (option small_kr4 0). n = 1024 is the last 4-round size, 1025 the first 8-round one."""
    from anechoic_amd import instances as I
    insts = [I.erdos_renyi(401, 1000, 0.03), I.torus(402, *I.TORUS_SHAPES[800]), I.erdos_renyi(403, 1024, 0.02),
             I.erdos_renyi(404, 1025, 0.02), I.erdos_renyi(405, 300, 0.05), I.erdos_renyi(406, 2900, 0.01)]
    S = 40
    if wide:
        T = np.linspace(4.0, 1.0, S)
        tb = dict(fourT=np.round(4 * T * 65536).astype(np.int32), q=np.full(S, 2 * 65536, np.int32),
                  kcorr=np.full(S, 1 << 31, np.int64), kconst=np.zeros(S, np.int32))
    else:
        tb = O.tables(dict(S=S, t0=6.0, t1=1.0, q=2.0, lam=0.8, ramp=True, scale=0.05))
    eng.set_tables(tb)
    ids = [eng.add_instance(J, bias=b) for J, b in insts]
    counts = [45, 20, 33, 16, 7, 20]
    trials = np.arange(sum(counts), dtype=np.uint32) * 3 + 101
    groups = list(zip(ids, counts))
    old = {k: eng.get_option(k) for k in ("small_kr4", "mixed_xcd2")}
    try:
        eng.set_option("mixed_xcd2", 8)
        res = eng.run_multi(groups, 77, trials, trace=True, fields=True)
        assert int(eng.get_option("last_kr4")) == 1 and int(eng.get_option("last_wide")) == int(wide)
        eng.set_option("small_kr4", 0)
        res8 = eng.run_multi(groups, 77, trials, trace=True, fields=True)
        assert int(eng.get_option("last_kr4")) == 0
        # only small instances: every XCD runs 4-round teams
        eng.set_option("small_kr4", 1)
        only = eng.run_multi(groups[:2], 77, trials[:65], trace=True)
        assert int(eng.get_option("last_kr4")) == 1
    finally:
        for k, v in old.items():
            eng.set_option(k, v)
    for k in ("spins", "sum_sh", "sum_bs", "flips", "nlin", "fields"):
        assert np.array_equal(res[k], res8[k]), k
    for k in ("spins", "sum_sh", "sum_bs", "flips", "nlin"):
        assert np.array_equal(only[k], res[k][:65, :only[k].shape[1]] if k == "spins" else res[k][:65]), k
    row = 0
    for (J, b), cnt in zip(insts, counts):
        n = J.shape[0]
        ref = O.run_batch(J, n, tb, 77, trials[row:row + cnt], bias=b, trace=True, nthreads=16)
        sl = slice(row, row + cnt)
        assert np.array_equal(res["spins"][sl, :O.np_of(n) // 64].view(np.uint64), ref["spins"]), (n, row)
        assert np.array_equal(res["sum_sh"][sl], ref["sum_sh"]) and np.array_equal(res["sum_bs"][sl], ref["sum_bs"]), n
        assert np.array_equal(res["flips"][sl], ref["flips"]) and np.array_equal(res["nlin"][sl], ref["nlin"]), n
        h = O.fields(J, n, res["spins"][row, :O.np_of(n) // 64].view(np.uint64), b)
        assert np.array_equal(res["fields"][row, :n].astype(np.int64), h), n
        row += cnt


def test_team16_int32_without_24bit_split_vs_oracle(eng):
    """The 16-member kernel's int32 decisions with the counted correction's full product n_lin kcorr >> 8 (the range
    plan proves int32 for n = 100 and kcorr = 2^32, but |kcorr >> 8| = 2^24 rules out the 24-bit split): equal to the
    oracle, replicas in two items (one partly filled)."""
    from anechoic_amd import instances as I
    n = 100
    J, b = I.erdos_renyi(501, n, 0.3)
    S = 30
    T = np.linspace(4.0, 1.0, S)
    tb = dict(fourT=np.round(4 * T * 65536).astype(np.int32), q=np.full(S, 2 * 65536, np.int32),
              kcorr=np.full(S, 1 << 32, np.int64), kconst=np.zeros(S, np.int32))
    assert (n * (1 << 32)) >> 8 < 2 ** 31 - 4 * 65536 and (1 << 32) >> 8 >= 1 << 23
    eng.set_tables(tb)
    inst = eng.add_instance(J, bias=b)
    trials = np.arange(27, dtype=np.uint32) * 5 + 11
    old = eng.get_option("batched_team")
    eng.set_option("batched_team", 16)
    try:
        res = eng.run(inst, 123, trials, trace=True, fields=True)
        assert int(eng.get_option("last_team")) == 16 and int(eng.get_option("last_wide")) == 0
    finally:
        eng.set_option("batched_team", old)
    ref = O.run_batch(J, n, tb, 123, trials, bias=b, trace=True, nthreads=16)
    assert ref["nlin"].max() > 0
    assert np.array_equal(res["spins"][:, :O.np_of(n) // 64].view(np.uint64), ref["spins"])
    assert np.array_equal(res["sum_sh"], ref["sum_sh"]) and np.array_equal(res["sum_bs"], ref["sum_bs"])
    assert np.array_equal(res["flips"], ref["flips"]) and np.array_equal(res["nlin"], ref["nlin"])
    h = O.fields(J, n, res["spins"][0, :O.np_of(n) // 64].view(np.uint64), b)
    assert np.array_equal(res["fields"][0, :n].astype(np.int64), h)
