"""Schedule tuning with the mean-field gradient, judged by the bit-exact discrete engine.

``tune`` minimises the mean final energy per spin of the mean-field relaxation (anechoic_amd.meanfield) over the 16
temperature knots and the correction strength with Adam; the gradient comes from the hand-written adjoint kernel.
``evaluate_discrete`` runs the tuned schedule (explicit T per step, same kappa / lambda rule) and a baseline schedule
through the batched discrete engine and compares cuts and success rates -- the mean-field objective is never used
as the verdict.
"""
import numpy as np
import torch

from . import instance_io as IO
from . import meanfield as M
from .tts import cut_from_sum_sh


def tune(prob, m0, S, q, theta0, strength0, mode="kappa", gamma=1.0, iters=20, lr=0.05, ckpt=32, tmin=0.5):
    """Returns (theta, strength, history of losses). theta: 16 knots (device tensor); strength: scalar."""
    theta = theta0.detach().clone().float().cuda().requires_grad_()
    stren = torch.tensor(float(strength0), device="cuda", requires_grad=True)
    opt = torch.optim.Adam([theta, stren], lr=lr)
    hist = []
    for _ in range(iters):
        opt.zero_grad()
        L = M.loss(prob, theta, stren, m0, q, S, mode=mode, gamma=gamma, ckpt=ckpt)
        L.backward()
        opt.step()
        with torch.no_grad():
            theta.clamp_(min=tmin)
            stren.clamp_(min=0.0)
        hist.append(float(L.detach()))
    return theta.detach(), float(stren.detach()), hist


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
