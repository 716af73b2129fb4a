"""Writes the golden fixtures under tests/golden/ (run in the development container, where the reference exists).

1. reference_cases.json -- the known answers the reference itself holds for this path, parsed from its cases.txt files
   (RTL/ASIC stream suites generated by the golden model; generator parameters, seeds, trial ids and per-trial
   sum_sh / sum_bs / cut / flips). Only data is kept: matrix spec, schedule, seeds and expected values.
2. golden_tables.json / golden_trials.json -- outputs of the reference's own headers (oracle/_ref/libref.so, compiled
   from the reference sources by oracle/Makefile) on this build's workloads: tables for every schedule shape used,
   and full anneal results (spins, sum_sh, sum_bs, flips, complete n_lin traces) for N <= 2048 instances.

Usage: python tests/golden/make_golden.py   (needs /root/reference and `make -C oracle ref`)
"""
import json
import re
import sys
import tempfile
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))
from anechoic_amd import instances as I  # noqa: E402
from oracle import oracle as O  # noqa: E402

REF = Path("/root/reference")
OUT = Path(__file__).resolve().parent

CASE_FILES = [
    ("k2_bias", "asic/gf22_engine_v64_mb_20261008/sim/vec_cases/k2_bias.cases.txt", "mbb"),
    ("k4_bias", "asic/gf22_engine_v64_mb_20261008/sim/vec_cases/k4_bias.cases.txt", "mbb"),
    ("rand4_short_k4", "fpga/v80_sca/results/multibit_20261007/sim/simpack_20261007a/rand4_short_k4/cases.txt", "mb"),
    ("rand4_long_k4", "fpga/v80_sca/results/multibit_20261007/sim/simpack_20261007a/rand4_long_k4/cases.txt", "mb"),
    ("rand8_short_k8", "fpga/v80_sca/results/multibit_20261007/sim/simpack_20261007a/rand8_short_k8/cases.txt", "mb"),
    ("k4b_rand4_n777", "fpga/v80_sca/results/multibit_20261007/sim/simpack_20261008_r3/k4b_rand4_n777/cases.txt", "mb3"),
    ("k4b_rand4_n2048", "fpga/v80_sca/results/multibit_20261007/sim/simpack_20261008_r3/k4b_rand4_n2048/cases.txt", "mb3"),
]


def kv(tokens):
    d = {}
    for i in range(0, len(tokens) - 1, 2):
        d[tokens[i]] = tokens[i + 1]
    return d


def parse_cases(name, rel, kind):
    lines = (REF / rel).read_text().splitlines()
    head = lines[0].split()
    spec = head[head.index("matrix") + 1]
    hd = kv(head)
    parts = spec.split(":")
    m = dict(kind=kind, spec=spec, seed=int(parts[1]), density=float(parts[2]), maxabs=int(parts[3]),
             n=int(parts[4]) if len(parts) > 4 else 2000, bmax=int(parts[5]) if len(parts) > 5 else 0,
             fingerprint=hd["fingerprint"], sumw=int(hd["sumw"]), auto_scale_printed=hd["auto_scale"],
             bias_mode=int(hd.get("bias_mode", "0")))
    cases = []
    cur = None
    for ln in lines[1:]:
        t = ln.split()
        if not t:
            continue
        if t[0] == "case":
            d = kv(t[3:])
            cur = dict(index=int(t[1]), name=t[2], S=int(d["S"]), t0=float(d["t0"]), t1=float(d["t1"]), q=float(d["q"]),
                       lam=float(d["lambda"]), ramp=int(d["ramp"]), jv=float(d["jv"]), kappa=float(d["kappa"]),
                       scale_printed=d["scale"], trace=int(d["trace"]), keep=int(d["keep"]), seed=int(d["seed"]),
                       offset=int(d["offset"]), use_bias=int(d.get("bias", m["bias_mode"])), trials=[])
            cases.append(cur)
        elif t[0] == "trial":
            d = kv(t[2:])
            tr = dict(trial=int(t[1]), flips=int(d["flips"]))
            for k in ("sum_sh", "sum_bs", "cut", "energy2"):
                if k in d:
                    tr[k] = int(d[k])
            cur["trials"].append(tr)
    return dict(name=name, source=rel, matrix=m, cases=cases)


def hexwords(w):
    return "".join(f"{int(x):016x}" for x in np.asarray(w, np.uint64))


def trial_record(n, J2048, bias, tb, seed, trial):
    r = O.ref_run_trial_bias(n, J2048, tb, seed, trial, bias=bias, trace=True)
    return dict(trial=int(trial), spins=hexwords(r["spins"]), sum_sh=int(r["sum_sh"]), sum_bs=int(r["sum_bs"]),
                flips=int(r["flips"]), nlin=[int(v) for v in r["nlin"]])


def pad2048(J, n):
    P = np.zeros((2048, 2048), np.int8)
    P[:n, :n] = J[:n, :n]
    return P


SCHEDULES = {
    # N = 2000 workloads (pre-declared): counted Onsager O1 family and temperature-scaled kappa = 2.0, both with ramp
    "c1_linear_counted_S1000": dict(S=1000, shape="linear", t0=12.0, t1=5.0, q=8.0, lam=1.05, ramp=True),
    "c2_counted_S250": dict(S=250, t0=12.0, t1=5.0, q=8.0, lam=1.05, ramp=True),
    "c2_kappa_S250": dict(S=250, t0=15.0, t1=5.0, q=8.0, kappa=2.0, ramp=True),
    "c2_plain_S100": dict(S=100, t0=30.0, t1=5.0, q=4.0),
    "c2_tec_S100": dict(S=100, t0=20.0, t1=5.0, q=8.0, jv=-8.0),
    "ref_O1_S960": dict(S=960, t0=12.0, t1=5.0, q=8.0, lam=1.05, ramp=True),
    "ref_X5_S280": dict(S=280, t0=15.0, t1=5.0, q=8.0, kappa=1.75, ramp=True),
    "bias_counted_S200": dict(S=200, t0=6.0, t1=1.0, q=2.0, lam=0.9, ramp=True, scale=0.05),
    "bias_kappa_S200": dict(S=200, t0=6.0, t1=1.0, q=2.0, kappa=1.5, ramp=True, scale=0.05),
    "linear_kappa_S300": dict(S=300, shape="linear", t0=15.0, t1=5.0, q=8.0, kappa=2.0, ramp=True),
}


def ref_tables(sch):
    if sch.get("shape", "geom") == "geom":
        return O.ref_mb_tables(sch)
    # linear temperatures through sca::make_tables; the kappa override is mb::tables' rule applied to the same T
    S = int(sch["S"])
    den = max(1, S - 1)
    T = np.array([sch["t0"] + (sch["t1"] - sch["t0"]) * (t / den) for t in range(S)], np.float64)
    lam = np.full(S, float(sch.get("lam", 0.0)))
    if sch.get("ramp"):
        for t in range(S):
            if t >= 0.7 * S:
                lam[t] = sch.get("lam", 0.0) * (S - 1 - t) / max(1.0, S - 1 - 0.7 * S)
    if sch.get("scale", 1.0) != 1.0:
        lam = lam * sch["scale"]
    tb = O.ref_make_tables(T, np.full(S, float(sch["q"])), lam, float(sch.get("jv", 0.0)))
    if sch.get("kappa", 0.0) > 0:
        for t in range(1, S):
            rf = (S - 1 - t) / max(1.0, S - 1 - 0.7 * S) if (sch.get("ramp") and t >= 0.7 * S) else 1.0
            v = sch["kappa"] * sch.get("scale", 1.0) * T[t - 1] * rf * 65536.0
            tb["kconst"][t] = int(np.floor(v + 0.5)) if v >= 0 else -int(np.floor(-v + 0.5))
            tb["kcorr"][t] = 0
    return tb


def main():
    assert O.ref_available(), "build oracle/_ref first: make -C oracle ref"
    # 1. reference-held known answers
    cases = [parse_cases(*c) for c in CASE_FILES]
    (OUT / "reference_cases.json").write_text(json.dumps(dict(
        note="Known answers held by the reference (cases.txt of its RTL/ASIC vector suites, produced by "
             "sca::run_trial / run_trial_bias). Seeds 0xC0FFEE+case, trial ids offset+k.",
        suites=cases), indent=1) + "\n")
    # 2. tables
    tabs = {}
    for name, sch in SCHEDULES.items():
        tb = ref_tables(sch)
        tabs[name] = dict(schedule=sch, **{k: [int(v) for v in tb[k]] for k in tb})
    # build_tables (flag path, scale 1) equals mb::tables for geometric schedules
    for name in ("ref_O1_S960", "ref_X5_S280", "c2_tec_S100"):
        bt = O.ref_build_tables(SCHEDULES[name])
        assert all(tabs[name][k] == [int(v) for v in bt[k]] for k in bt), name
    (OUT / "golden_tables.json").write_text(json.dumps(tabs) + "\n")
    # 3. anneals of this build's workloads through the reference headers
    trials = {}
    J, _ = I.ising_random(20261203, 2000)
    with tempfile.TemporaryDirectory() as d:
        Jr, _ = O.ref_gen_ising_random(20261203, 2000, 2, 0.0, 0, d)
    assert np.array_equal(J, Jr)
    P = pad2048(J, 2000)
    inst_k2000 = dict(generator="ising_random", seed=20261203, n=2000, kbits=2, pzero=0.0,
                      fingerprint=f"{O.fingerprint(P):016x}", sumw=O.sumw(P, 2000))

    def add(key, inst, n, Pm, bias, sched_name, seed, trial_ids):
        sch = SCHEDULES[sched_name]
        tb = ref_tables(sch)
        trials[key] = dict(instance=inst, schedule=sched_name, seed=seed,
                           results=[trial_record(n, Pm, bias, tb, seed, t) for t in trial_ids])

    add("c1_plain_kernel", inst_k2000, 2000, P, None, "c1_linear_counted_S1000", 0xC1C1, [0])
    add("c2_counted", inst_k2000, 2000, P, None, "c2_counted_S250", 0xC2C2, [0, 1, 2, 3, 239])
    add("c2_kappa", inst_k2000, 2000, P, None, "c2_kappa_S250", 0xC2C2, [0, 1, 2, 3, 239])
    add("c2_plain", inst_k2000, 2000, P, None, "c2_plain_S100", 7, [5, 6])
    add("c2_tec", inst_k2000, 2000, P, None, "c2_tec_S100", 8, [9, (1 << 28) + 2])
    add("c2_linear_kappa", inst_k2000, 2000, P, None, "linear_kappa_S300", 9, [0])
    for (kind, n, seed) in (("er", 800, 31), ("torus", 1000, 32), ("er", 2000, 33)):
        if kind == "er":
            Jb, b = I.erdos_renyi(seed, n, 0.02)
        else:
            Jb, b = I.torus(seed, *I.TORUS_SHAPES[n])
        Pb = pad2048(Jb, n)
        inst = dict(generator=kind, seed=seed, n=n, density=0.02 if kind == "er" else None,
                    fingerprint=f"{O.fingerprint(Pb):016x}")
        add(f"bias_{kind}_{n}_counted", inst, n, Pb, b, "bias_counted_S200", 0xB1A5 + n, [0, 1])
        add(f"bias_{kind}_{n}_kappa", inst, n, Pb, b, "bias_kappa_S200", 0xB1A6 + n, [2])
    (OUT / "golden_trials.json").write_text(json.dumps(trials) + "\n")
    print("wrote", [p.name for p in OUT.glob("*.json")])


if __name__ == "__main__":
    main()
