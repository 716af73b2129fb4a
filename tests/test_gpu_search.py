"""Schedule search on the GPU: every pilot cell's per-batch success counts (host-rescored cuts) equal a recount from
anneals run again through the host-returning path on the same trial ids, no host cut differs from the engine's own,
the cells use disjoint trial ranges, and the held-out run uses ids outside all of them."""
import numpy as np
import pytest


pytestmark = pytest.mark.gpu


def test_pilot_counts_and_holdout(k2000_engine):
    eng, inst, J = k2000_engine
    from anechoic_amd import search, tts
    n = 2000
    sumw = int(-np.triu(J.astype(np.int64), 1).sum())
    target = tts.energy_target_cut(sumw, n) - 120
    cells = search.reference_grid(modes=(("kappa", 2.0), ("lam", 1.05)), qs=(8.0,), t0s=(15.0,), steps=(40, 80))
    B, batches, base, stride, seed = 32, 4, 3_000_000, 8_000, 0x77
    res = search.pilot(eng, inst, J, cells, target, seed, B=B, batches=batches, base=base, stride=stride)
    assert [r["name"] for r in res] == [c["name"] for c in cells]
    for k, (c, r) in enumerate(zip(cells, res)):
        eng.set_schedule(c["schedule"])
        counts = []
        for i in range(batches):
            trials = np.arange(base + stride * k + B * i, base + stride * k + B * (i + 1), dtype=np.uint32)
            out = eng.run(inst, seed, trials)
            counts.append(sum(tts.cut_from_sum_sh(sumw, int(v)) >= target for v in out["sum_sh"]))
        assert r["successes"] == sum(counts) and r["batches"] == batches and r["t_batch"] > 0
        assert r["p"] == pytest.approx(sum(counts) / (B * batches))
        assert r["cut_mismatches"] == 0
    sel = search.select(res, "primary")
    ho = search.holdout(eng, inst, J, cells[sel["index"]], target, seed, B=B, batches=8, base=4_000_000)
    assert ho["batches"] == 8 and ho["name"] == sel["name"] and ho["cut_mismatches"] == 0


@pytest.fixture(scope="module")
def k2000_engine():
    import torch
    if not torch.cuda.is_available():
        pytest.fail("GPU tests need a ROCm GPU")
    from anechoic_amd import Engine, instances
    J, _ = instances.ising_random(20261203, 2000)
    e = Engine(0)
    inst = e.add_instance(J)
    yield e, inst, J
    e.close()
