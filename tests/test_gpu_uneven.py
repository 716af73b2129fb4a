"""The uneven-size workload: 32 sparse instances with int2 couplings and biases (N from {800, 1000, 2000, 3000}, half
Erdos-Renyi at 0.4-6 % density, half 2-D tori), 64 replicas each = 2048 anneals in one persistent launch that hosts
both team shapes (32-workgroup teams of 8 rounds for N <= 2048, 32-workgroup teams of 12 rounds for N = 3000, one
XCD each), eagerly and replayed from a captured hipGraph.

Checks (all exact):
  * sampled replicas of every instance (the first and the last of its 64) against the oracle: spins, sum_sh,
    sum_bs, flips and the full n_lin trace (beyond NP = 2048 the oracle's generalisation is the reference: parity
    unpinned by the reference headers);
  * every one of the 2048 anneals against size-independent invariants: the returned fields equal b + J s of the
    returned spins, sum_sh = sum s h and sum_bs = sum b s;
  * graph replays against the eager launch, with eager work and synchronisations between replays and trial ids
    changed in place;
  * the team split (XCDs per shape) and the replica tile per work item do not change any result.
"""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

SCHED = dict(S=48, t0=5.0, t1=0.8, q=2.0, lam=0.8, ramp=True, scale=0.05)
SEED = 0xC3C3
REPL = 64


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
def workload(eng):
    from anechoic_amd import instances as I
    insts = I.uneven_batch()
    tb = O.tables(SCHED)
    eng.set_tables(tb)
    ids = [eng.add_instance(d["J"], bias=d["bias"]) for d in insts]
    groups = [(i, REPL) for i in ids]
    trials = (np.arange(REPL * len(insts), dtype=np.uint32) * 7 + 11).astype(np.uint32)
    return insts, tb, groups, trials


def _invariants(res, insts):
    row = 0
    for d in insts:
        J, b, n = d["J"], d["bias"].astype(np.int64), d["n"]
        words = O.np_of(n) // 64
        sp = res["spins"][row:row + REPL, :words].view(np.uint64)
        s = np.stack([O.spins_pm(sp[i], n) for i in range(REPL)]).astype(np.int64)
        # float64 products are exact here (|h| <= 2 n + 2) and use BLAS
        h = np.rint(s.astype(np.float64) @ J[:n, :n].astype(np.float64)).astype(np.int64) + b
        assert np.array_equal(res["fields"][row:row + REPL, :n].astype(np.int64), h), (d["kind"], n)
        assert np.array_equal(res["sum_sh"][row:row + REPL], (s * h).sum(1)), (d["kind"], n)
        assert np.array_equal(res["sum_bs"][row:row + REPL], (s * b).sum(1)), (d["kind"], n)
        row += REPL


def test_uneven_workload_in_one_mixed_launch_vs_oracle(eng, workload):
    insts, tb, groups, trials = workload
    sizes = {d["n"] for d in insts}
    assert 3000 in sizes and sizes & {800, 1000, 2000}, sizes
    res = eng.run_multi(groups, SEED, trials, trace=True, fields=True)
    # both team shapes in the one launch: an XCD split, or (8) every XCD running both classes in turn
    assert 1 <= eng.get_option("last_xcd2") <= 8 and eng.get_option("last_nt3") > 0
    _invariants(res, insts)
    for gi, d in enumerate(insts):
        n = d["n"]
        words = O.np_of(n) // 64
        rows = np.array([gi * REPL, gi * REPL + REPL - 1])
        ref = O.run_batch(d["J"], n, tb, SEED, trials[rows], bias=d["bias"], trace=True, nthreads=16)
        assert np.array_equal(res["spins"][rows, :words].view(np.uint64), ref["spins"]), (gi, n)
        assert np.array_equal(res["sum_sh"][rows], ref["sum_sh"]) and np.array_equal(res["sum_bs"][rows], ref["sum_bs"])
        assert np.array_equal(res["flips"][rows], ref["flips"]) and np.array_equal(res["nlin"][rows], ref["nlin"])
    assert int(res["flips"].sum()) > 0


def test_graph_replays_match_eager(eng, workload):
    import torch
    insts, tb, groups, trials = workload
    R = REPL * len(insts)
    words = O.np_of(max(d["n"] for d in insts)) // 64
    trials_b = (trials + np.uint32(100_000)).astype(np.uint32)
    eager_a = eng.run_multi(groups, SEED, trials, trace=True)
    eager_b = eng.run_multi(groups, SEED, trials_b, trace=True)
    tdev = torch.from_numpy(trials.view(np.int32)).to(eng.device)
    out = eng.alloc_out_rows(R, words, trace=True)
    run = eng.capture_multi(groups, SEED, tdev, out)
    try:
        def check(ref, label):
            for k in ("spins", "sum_sh", "sum_bs", "flips", "nlin"):
                assert np.array_equal(out[k].cpu().numpy(), ref[k]), (label, k)

        run.launch()
        eng.sync()
        check(eager_a, "replay 1")
        # eager work on the same context between replays (it reuses the exchange workspace)
        other = eng.run_multi(groups[:3], SEED + 1, None, trace=False)
        assert other["flips"].shape[0] == 3 * REPL
        tdev.copy_(torch.from_numpy(trials_b.view(np.int32)))
        run.launch()
        eng.sync()
        check(eager_b, "replay 2")
        tdev.copy_(torch.from_numpy(trials.view(np.int32)))
        for i in range(2):
            run.launch()
            eng.sync()
            check(eager_a, f"replay {3 + i}")
    finally:
        run.destroy()


def test_team_split_and_replica_tile_do_not_change_results(eng, workload):
    insts, tb, groups, trials = workload
    sub = [groups[i] for i in range(8)]
    R = REPL * len(sub)
    base = eng.run_multi(sub, SEED, trials[:R], trace=True)
    old = {k: eng.get_option(k) for k in ("mixed_ratio", "multi_nt", "multi_nt3", "small_kr4")}
    try:
        for ratio, nt, nt3, kr4 in ((0.3, 2, 1, 1), (6.0, 2, 2, 0), (1.5, 1, 1, 1), (1.0, 1, 2, 0), (2.0, 2, 2, 1)):
            eng.set_option("mixed_ratio", ratio)
            eng.set_option("multi_nt", nt)
            eng.set_option("multi_nt3", nt3)
            eng.set_option("small_kr4", kr4)
            res = eng.run_multi(sub, SEED, trials[:R], trace=True)
            for k in ("spins", "sum_sh", "sum_bs", "flips", "nlin"):
                assert np.array_equal(res[k], base[k]), (ratio, nt, nt3, kr4, k)
    finally:
        for k, v in old.items():
            eng.set_option(k, v)
