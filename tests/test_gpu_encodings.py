"""Encoded problems through the engine: a graph-partitioning and a TSP instance (Lucas forms, with biases) annealed
on the GPU, bit-identical to the oracle, and their returned spins scored with the encoding's own relation
H = sigma f + C (energy from sum_sh and sum_bs against f from the decoded partition / tour)."""
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


def _anneal_vs_oracle(eng, J, b, sch, R, seed):
    n = J.shape[0]
    tb = O.tables(sch)
    eng.set_tables(tb)
    inst = eng.add_instance(J, bias=b)
    trials = (np.arange(R, dtype=np.uint32) * 7 + 3).astype(np.uint32)
    res = eng.run(inst, seed, trials, trace=True)
    ref = O.run_batch(J, n, tb, seed, trials, bias=b, trace=True, nthreads=16)
    words = O.np_of(n) // 64
    assert np.array_equal(res["spins"][:, :words].view(np.uint64), ref["spins"])
    assert np.array_equal(res["sum_sh"], ref["sum_sh"]) and np.array_equal(res["sum_bs"], ref["sum_bs"])
    assert np.array_equal(res["flips"], ref["flips"]) and np.array_equal(res["nlin"], ref["nlin"])
    spins = np.stack([O.spins_pm(res["spins"][r, :words].view(np.uint64), n) for r in range(R)]).astype(np.int64)
    energy = -(res["sum_sh"] + res["sum_bs"]) // 2   # E = -(sum s h + sum b s)/2 with h = b + J s
    return spins, energy


def test_graph_partition_instance(eng):
    from anechoic_amd import instances as I
    n, k, A, B = 600, 300, 2, 1
    J, b, m = I.graph_partition(9, n, 3, k, A, B)
    sch = dict(S=120, shape="linear", t0=20.0, t1=1.0, q=2.0, kappa=1.0, ramp=True)
    spins, energy = _anneal_vs_oracle(eng, J, b, sch, 48, 0x6B)
    cut, size = I.gp_cut(m["edges"], spins)
    assert np.array_equal(energy, 2 * (B * cut + A * (size - k) ** 2) + m["C"])


def test_tsp_instance(eng):
    from anechoic_amd import instances as I
    cities, A, B = 20, 12, 1
    J, b, m = I.tsp(5, cities, 4, A, B)
    sch = dict(S=150, shape="linear", t0=40.0, t1=2.0, q=4.0, lam=0.5, ramp=True)
    spins, energy = _anneal_vs_oracle(eng, J, b, sch, 40, 0x75)
    mm = cities - 1
    x = (spins > 0).reshape(-1, mm, mm)
    W = m["W"].astype(np.int64)
    length = x[:, :, 0] @ W[0, 1:] + x[:, :, mm - 1] @ W[1:, 0]
    for p in range(mm - 1):
        length += np.einsum("ru,rv,uv->r", x[:, :, p].astype(np.int64), x[:, :, p + 1].astype(np.int64), W[1:, 1:])
    f = A * ((1 - x.sum(2)) ** 2).sum(1) + A * ((1 - x.sum(1)) ** 2).sum(1) + B * length
    assert np.array_equal(energy, 4 * f + m["C"])
    valid = I.tsp_tour_length(m["W"], spins) >= 0
    assert np.array_equal(I.tsp_tour_length(m["W"], spins)[valid], length[valid])
