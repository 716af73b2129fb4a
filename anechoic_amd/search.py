"""Schedule search on the discrete engine: pilot grid, pre-registered selection, held-out confirmation.

Follows the reference protocol's phases P and H (gpu/onsager_v2_20261007/src/run_protocol.sh:77-115,
PROTOCOL.md sections 8-9):
* every pilot cell (one schedule at one batch size B) runs `batches` launches of B anneals on trial ids of its own
  range (base + B i + r), and is timed per launch with HIP events on the launch stream;
* the selection score is pre-registered (analyze.py:81-85): TTS99 at the lower Wilson 95 % bound of P_batch
  (primary) or of p (secondary), ties to the smaller point estimate;
* the selected cell is re-run on held-out trial ids, and only those numbers are reported for it;
* successes are counted from cuts recomputed on the host from the returned spins (anc_score), and every anneal whose
  host cut differs from the engine's own is counted (sca_gpu.cu:717-739 does the same with host::scalar_cut).
The reference grid (`reference_grid`) is the protocol's: end-ramped temperature-scaled (kappa 1.5, 1.75, 2.0) or
counted (lambda 0.9, 1.05) correction, q in {6, 8}, T_init in {12, 15}, S in {80, ..., 280}, geometric T to 5.
Schedules proposed by the mean-field gradient (`meanfield_candidates`: knots and kappa tuned per S by anechoic_amd.tune)
enter the pilot as further cells and are selected or not by the same rule.
"""
import itertools

import numpy as np
import torch

from . import instance_io as IO
from . import tts as T

REF_MODES = (("kappa", 1.5), ("kappa", 1.75), ("kappa", 2.0), ("lam", 0.9), ("lam", 1.05))


def reference_grid(modes=REF_MODES, qs=(6.0, 8.0), t0s=(12.0, 15.0), steps=(80, 120, 160, 200, 240, 280), t1=5.0):
    """The protocol's pilot grid, in its index order (mode, q, T_init, S)."""
    cells = []
    for (mode, v), q, t0, S in itertools.product(modes, qs, t0s, steps):
        sch = dict(S=S, shape="geometric", t0=t0, t1=t1, q=q, ramp=True)
        sch["kappa" if mode == "kappa" else "lam"] = v
        cells.append(dict(name=f"{mode}{v:g}_q{q:g}_t{t0:g}_S{S}", schedule=sch))
    return cells


def meanfield_candidates(prob, m0, steps, q=8.0, kappa0=2.0, t0=15.0, t1=5.0, keep=(2, 4), lr=0.3):
    """Schedules proposed by the mean-field gradient: for each S, the Adam snapshots `keep` of the 16 temperature knots
    and kappa tuned on the differentiable relaxation (anechoic_amd.tune) from the grid's geometric T t0 -> t1 (as 16
    knots), turned into explicit-T discrete schedules. The pilot then judges them beside the grid cells."""
    from . import meanfield as M
    from . import tune
    out = []
    for S in steps:
        theta0 = t0 * (t1 / t0) ** torch.linspace(0.0, 1.0, 16, device="cuda")
        _, _, _, snaps = tune.trajectory(prob, m0, S, q, theta0, kappa0, mode="kappa", iters=max(keep), lr=lr, keep=keep)
        for i, th, kap in snaps:
            out.append(dict(name=f"meanfield_S{S}_adam{i}", schedule=M.discrete_schedule(th, kap, S, mode="kappa", q=q)))
    return out


def _cuts(sum_sh, sumw):
    """cut = (sumw + sum_sh / 2) / 2 with C integer division (sca_ref.hpp:154), on the device."""
    half = torch.div(sum_sh, 2, rounding_mode="trunc")
    return torch.div(half + int(sumw), 2, rounding_mode="trunc")


def run_cell(engine, inst, J, schedule, B, batches, trial_base, target, seed):
    """`batches` launches of B anneals (trial ids trial_base + B i + r). Returns (successes per batch, mean device ms per
    launch, cut mismatches). Success = cut >= target, with the cut recomputed on the host from J and the returned spins
    (anc_score, the counterpart of the reference's host::scalar_cut, sca_gpu.cu:717-739); cut mismatches counts the
    anneals whose host cut differs from the engine's own (sumw + sum_sh / 2) / 2."""
    engine.set_schedule(schedule)
    n = J.shape[0]
    sumw = int(-np.triu(J[:n, :n].astype(np.int64), 1).sum())
    out = engine.alloc_out(inst, B)
    ids = torch.arange(B, dtype=torch.int32, device="cuda")
    spins = torch.empty((batches,) + tuple(out["spins"].shape), dtype=out["spins"].dtype, device="cuda")
    dev_cut = torch.empty((batches, B), dtype=torch.int64, device="cuda")
    st = torch.cuda.current_stream()
    evs = []
    for i in range(batches):
        trials = ids + int(trial_base + B * i)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(st)
        engine.launch(inst, seed, trials, out, kernel="batched")
        e1.record(st)
        evs.append((e0, e1))
        spins[i].copy_(out["spins"])
        dev_cut[i] = _cuts(out["sum_sh"], sumw)
    torch.cuda.synchronize()
    engine.sync()
    t = float(np.mean([a.elapsed_time(b) for a, b in evs]))
    host_cut = IO.score(J[:n, :n], spins.reshape(batches * B, -1).cpu().numpy(), nthreads=16)["cut"].reshape(batches, B)
    mismatches = int((host_cut != dev_cut.cpu().numpy()).sum())
    return [int(x) for x in (host_cut >= target).sum(axis=1)], t, mismatches


def score(est, which):
    """Pre-registered selection score: TTS99 at the lower Wilson 95 % bound (analyze.py:81-85)."""
    if which == "primary":
        return T.tts_primary(est["P_batch_wilson95"][0], est["t_batch"])
    return T.tts_secondary(est["p_wilson95"][0], est["B"], est["t_batch"])


def pilot(engine, inst, J, cells, target, seed, B=240, batches=32, base=3_000_000, stride=8_000):
    """Runs every cell on instance `inst` (host matrix J); cell k uses trial ids base + stride k onward (disjoint while
    B batches <= stride)."""
    if B * batches > stride:
        raise ValueError("pilot: B batches must not exceed the trial-id stride between cells")
    res = []
    for k, c in enumerate(cells):
        succ, t, bad = run_cell(engine, inst, J, c["schedule"], B, batches, base + stride * k, target, seed)
        est = T.estimate(succ, B, t)
        est.update(name=c["name"], index=k, score_primary=score(est, "primary"), score_secondary=score(est, "secondary"),
                   cut_mismatches=bad)
        res.append(est)
    return res


def select(results, which="primary"):
    """The cell with the smallest score (ties: smaller point TTS)."""
    key = "tts_primary" if which == "primary" else "tts_secondary"
    return min(results, key=lambda e: (e["score_" + which], e[key]))


def holdout(engine, inst, J, cell, target, seed, B=240, batches=128, base=4_000_000):
    """The selected cell on held-out trial ids; the only numbers reported for it."""
    succ, t, bad = run_cell(engine, inst, J, cell["schedule"], B, batches, base, target, seed)
    est = T.estimate(succ, B, t)
    est.update(name=cell["name"], cut_mismatches=bad)
    return est
