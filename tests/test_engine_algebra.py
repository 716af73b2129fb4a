"""CPU check of the team kernel's arithmetic design against the oracle (no GPU needed).

The HIP kernel does not evaluate the reference's int64 z directly; it uses algebraically equivalent integer forms:
  * permuted slot numbering x' = l*KR + k (x = k*256 + l);
  * D'_u = D_u + tau_u (tau = step-parity bit ORed into even bytes), D_0 = s(0) - 1, and
    H = bias + J'(sum_u D'_u) - (tau_count - 1) rowsum(J');
  * flip  <=>  v <= floor((r - c - 1) / 2^16),  band  <=>  floor((-2T - c)/2^16) + 1 <= v <= floor((2T - c - 1)/2^16),
    with v = s h, c = q -/+ corr and r = a floor(fourT/2^16) + floor(a (fourT mod 2^16) / 2^16), a = u16 - 32768.
This test replays one anneal with exactly those formulas in numpy/Python integers and compares every n_lin, the
flips, the spins and sum s h with the oracle.
"""
import numpy as np
import pytest

from oracle import oracle as O

LANES = 256


def tau_of(u):
    return ((u >> 1) & 1) ^ 1


def emulate(J, n, bias, tb, seed, trial):
    np_ = O.np_of(n)
    KR = np_ // LANES
    xs = np.arange(np_)
    perm = (xs % LANES) * KR + xs // LANES            # x -> x'
    inv = np.empty(np_, np.int64)
    inv[perm] = xs                                     # x' -> x
    Jpad = np.zeros((np_, np_), np.int64)
    Jpad[:n, :n] = J[:n, :n]
    Jp = Jpad[np.ix_(inv, inv)]
    rowsum = Jp.sum(1)
    S = len(tb["fourT"])
    draws = np.stack([O.rng_stream(seed, trial, l, KR * (S + 1)) for l in range(LANES)]).astype(np.int64)   # [lane][d]
    valid = inv < n
    b = np.zeros(np_, np.int64)
    if bias is not None:
        b[valid] = np.asarray(bias, np.int64)[inv[valid]]
    lane_of, k_of = inv % LANES, inv // LANES          # per x'
    s = np.ones(np_, np.int64)
    top = draws[lane_of, k_of] >> 31
    s[valid & (top == 0)] = -1
    sp = np.ones(np_, np.int64)
    D0 = np.where(valid, s - 1, 0)
    acc = Jp @ (D0 + tau_of(0))
    tc = tau_of(0)
    nlin_prev, flips, trace = 0, 0, []
    for t in range(S + 1):
        H = b + acc - (tc - 1) * rowsum
        if t == S:
            break
        corr = ((nlin_prev * int(tb["kcorr"][t])) >> 8) + int(tb["kconst"][t])
        f4, q = int(tb["fourT"][t]), int(tb["q"][t])
        twoT = f4 >> 1
        nlin, D = 0, np.zeros(np_, np.int64)
        for xp in np.nonzero(valid)[0]:
            u16 = int(draws[lane_of[xp], KR * (t + 1) + k_of[xp]]) >> 16
            a = u16 - 32768
            rr = a * (f4 >> 16) + ((a * (f4 & 0xFFFF)) >> 16)
            assert rr == (a * f4) >> 16
            c = q - corr if s[xp] == sp[xp] else q + corr
            v = int(s[xp] * H[xp])
            if ((-twoT - c) >> 16) + 1 <= v <= ((twoT - c - 1) >> 16):
                nlin += 1
            if v <= ((rr - c - 1) >> 16):
                D[xp] = -2 * s[xp]
        sp = s.copy()
        s = s + D
        flips += int(np.count_nonzero(D))
        u = t + 1
        acc = acc + Jp @ (D + tau_of(u))
        tc += tau_of(u)
        nlin_prev = nlin
        trace.append(nlin)
    sv = np.where(valid, s, 0)
    sum_sh = int((sv * H).sum())
    spins_x = np.ones(np_, np.int64)
    spins_x[inv] = s
    return dict(nlin=trace, flips=flips, sum_sh=sum_sh, s=spins_x[:n])


@pytest.mark.parametrize("n,bmax,sched", [
    (300, 0, dict(S=12, t0=4.0, t1=1.0, q=2.0, lam=0.9, ramp=True)),
    (2000, 0, dict(S=6, t0=12.0, t1=5.0, q=8.0, kappa=2.0, ramp=True)),
    (700, 30, dict(S=10, t0=5.0, t1=1.0, q=2.0, jv=-1.0)),
])
def test_kernel_arithmetic_equals_oracle(n, bmax, sched):
    J, bias = O.gen_random_k(1234 + n, 0.4, 2, n, bmax, ld=n, rows=n)
    bias = bias if bmax else None
    tb = O.tables(sched)
    e = emulate(J, n, bias, tb, 777, 42)
    r = O.run_trial(J, n, tb, 777, 42, bias=bias, trace=True)
    assert e["nlin"] == r["nlin"].tolist()
    assert e["flips"] == r["flips"] and e["sum_sh"] == r["sum_sh"]
    assert np.array_equal(e["s"], O.spins_pm(r["spins"], n))
