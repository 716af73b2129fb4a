"""Shared test helpers: regenerating the reference-held and build-defined instances with the oracle's generators."""
import numpy as np

from anechoic_amd import instances as I
from oracle import oracle as O


def reference_matrix(m):
    """Rebuilds a reference fixture matrix (2048 x 2048 dense, reference layout) and its bias from the spec."""
    kind, n = m["kind"], m["n"]
    if kind == "mb3":
        J, b = O.gen_random_k(m["seed"], m["density"], m["maxabs"], n, m["bmax"])
        return J, (b if m["bias_mode"] else None), O.j2_scale(J, n, n)
    if kind == "mbb":
        J, _ = O.gen_random_k(m["seed"], m["density"], m["maxabs"], n, 0)
        b = O.gen_mbb_bias(O.fingerprint(J), m["maxabs"], n)
        return J, b, O.j2_scale(J, 2000, 2000)
    J, _ = O.gen_random_k(m["seed"], m["density"], m["maxabs"], 2000, 0)   # mb::random_k (N = 2000)
    return J, None, O.j2_scale(J, 2000, 2000)


def case_schedule(c, auto_scale):
    scale = auto_scale if c["scale_printed"] != "1" else 1.0
    return dict(S=c["S"], t0=c["t0"], t1=c["t1"], q=c["q"], lam=c["lam"], jv=c["jv"], kappa=c["kappa"], ramp=bool(c["ramp"]),
                scale=scale)


def golden_instance(inst):
    """Regenerates a golden_trials.json instance: (J n x n, bias or None)."""
    g, n = inst["generator"], inst["n"]
    if g == "ising_random":
        J, _ = I.ising_random(inst["seed"], n, inst["kbits"], inst["pzero"])
        return J, None
    if g == "er":
        return I.erdos_renyi(inst["seed"], n, inst["density"])
    if g == "torus":
        return I.torus(inst["seed"], *I.TORUS_SHAPES[n])
    raise ValueError(g)


def pad2048(J, n):
    P = np.zeros((2048, 2048), np.int8)
    P[:n, :n] = J[:n, :n]
    return P
