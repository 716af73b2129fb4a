"""Schedule tuning with the mean-field gradient, judged by the bit-exact discrete engine.

``tune`` minimises the mean final energy per spin of the mean-field relaxation (anechoic_amd.meanfield) over the 16
temperature knots and the correction strength with Adam; the gradient comes from the hand-written adjoint kernel.
``evaluate_discrete`` runs the tuned schedule (explicit T per step, same kappa / lambda rule) and a baseline schedule
through the batched discrete engine and compares cuts and success rates -- the mean-field objective is never used
as the verdict.

``transfer`` is the protocol that carries a tuning over to the discrete engine: the Adam trajectory's snapshots
(the starting schedule included) are candidate schedules; a discrete pilot on trial ids of its own selects one by
the pre-registered score of the schedule search (anechoic_amd.search: TTS99 at the lower Wilson bound); the
selected schedule and the starting one then run on the same held-out trial ids, and both success rates are reported
with their Wilson 95 % intervals. A gradient step that helps the relaxation but hurts the discrete engine is thus
not selected over the start unless the pilot says otherwise, and the held-out numbers show what it did.
"""
import numpy as np
import torch

from . import instance_io as IO
from . import meanfield as M
from .tts import cut_from_sum_sh


def trajectory(prob, m0, S, q, theta0, strength0, mode="kappa", gamma=1.0, iters=20, lr=0.05, ckpt=32, tmin=0.5,
               keep=()):
    """Adam on the mean-field loss. Returns (theta, strength, history of losses, snapshots): theta = the 16 knots after
    `iters` steps (device tensor), strength a scalar, snapshots = [(i, theta_i, strength_i)] for every i in `keep`
    (0 = the start)."""
    theta = theta0.detach().clone().float().cuda().requires_grad_()
    stren = torch.tensor(float(strength0), device="cuda", requires_grad=True)
    opt = torch.optim.Adam([theta, stren], lr=lr)
    hist, snaps = [], []
    if 0 in keep:
        snaps.append((0, theta.detach().clone(), float(stren.detach())))
    for i in range(iters):
        opt.zero_grad()
        L = M.loss(prob, theta, stren, m0, q, S, mode=mode, gamma=gamma, ckpt=ckpt)
        L.backward()
        opt.step()
        with torch.no_grad():
            theta.clamp_(min=tmin)
            stren.clamp_(min=0.0)
        hist.append(float(L.detach()))
        if i + 1 in keep:
            snaps.append((i + 1, theta.detach().clone(), float(stren.detach())))
    return theta.detach(), float(stren.detach()), hist, snaps


def tune(prob, m0, S, q, theta0, strength0, mode="kappa", gamma=1.0, iters=20, lr=0.05, ckpt=32, tmin=0.5):
    """Returns (theta, strength, history of losses). theta: 16 knots (device tensor); strength: scalar."""
    th, st, hist, _ = trajectory(prob, m0, S, q, theta0, strength0, mode, gamma, iters, lr, ckpt, tmin)
    return th, st, hist


def transfer(engine, inst, J, prob, m0, S, q, theta0, strength0, target, seed, keep=(0, 2, 4, 6, 8), lr=0.3,
             mode="kappa", B=240, pilot_batches=8, holdout_batches=32, pilot_base=6_000_000, holdout_base=7_000_000,
             ckpt=32):
    """The tuning-transfer protocol (module docstring) on discrete instance `inst` (host matrix J) at success threshold
    `target`: candidates = the trajectory's snapshots `keep` (must contain 0, the start), pilot of `pilot_batches`
    batches of B anneals per candidate (disjoint trial ids from pilot_base), selection by the secondary score (per-anneal
    p at its lower Wilson bound; every candidate has the same S), then the selected and the start on the same
    `holdout_batches` held-out batches (trial ids from holdout_base)."""
    from . import search
    from .tts import wilson
    if 0 not in keep:
        raise ValueError("transfer: the snapshots must include the start (0)")
    _, _, hist, snaps = trajectory(prob, m0, S, q, theta0, strength0, mode=mode, iters=max(keep), lr=lr, ckpt=ckpt,
                                   keep=tuple(keep))
    cells = [dict(name=f"adam{i}", schedule=M.discrete_schedule(th, st, S, mode=mode, q=q), strength=st,
                  knots=[round(float(x), 4) for x in th.cpu()]) for i, th, st in snaps]
    res = search.pilot(engine, inst, J, cells, target, seed, B=B, batches=pilot_batches, base=pilot_base)
    sel = search.select(res, "secondary")

    def held_out(cell):
        ho = search.holdout(engine, inst, J, cell, target, seed, B=B, batches=holdout_batches, base=holdout_base)
        lo, hi = wilson(ho["successes"], ho["batches"] * B)
        return dict(name=cell["name"], p=round(ho["p"], 5), p_wilson95=[round(lo, 5), round(hi, 5)],
                    successes=ho["successes"], anneals=ho["batches"] * B, cut_mismatches=ho["cut_mismatches"],
                    t_batch_ms=round(ho["t_batch"], 4))
    start = held_out(cells[0])
    chosen = start if sel["index"] == 0 else held_out(cells[sel["index"]])
    return dict(S=S, candidates=[dict(name=c["name"], strength=round(c["strength"], 4), knots=c["knots"]) for c in cells],
                loss_history=[round(x, 6) for x in hist],
                pilot=[dict(name=e["name"], p=round(e["p"], 5), p_wilson95=[round(x, 5) for x in e["p_wilson95"]],
                            cut_mismatches=e["cut_mismatches"]) for e in res],
                selected=sel["name"], holdout_start=start, holdout_selected=chosen,
                trial_ids=dict(pilot_base=pilot_base, holdout_base=holdout_base, B=B, pilot_batches=pilot_batches,
                               holdout_batches=holdout_batches))


def evaluate_discrete(engine, inst, J, sched, seed, R=240, target=None, trial0=0):
    """Runs R anneals of `sched` through the batched engine; returns mean/max cut (recomputed on the host from the
    returned spins), the success rate and the number of anneals whose engine cut differs from the host's."""
    engine.set_schedule(sched)
    n = J.shape[0]
    sumw = int(-np.triu(J[:n, :n].astype(np.int64), 1).sum())
    res = engine.run(inst, seed, np.arange(trial0, trial0 + R, dtype=np.uint32))
    cuts = IO.score(J[:n, :n], res["spins"], nthreads=16)["cut"]
    dev = np.array([cut_from_sum_sh(sumw, v) for v in res["sum_sh"]])
    out = dict(mean_cut=float(cuts.mean()), max_cut=int(cuts.max()), R=R, cut_mismatches=int((cuts != dev).sum()))
    if target is not None:
        out["p_target"] = float((cuts >= target).mean())
    return out
