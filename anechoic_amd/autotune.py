"""Launch-shape autotuner for the multi-instance team launches.

Candidates: the replica tile of a work item of the N <= 2048 instances (16 or 32 replicas: the N-tile count of the
MFMA B operand) and, for a batch that needs both team shapes, the split of the 8 XCDs between the teams of the
N <= 2048 instances and those of the 2048 < N <= 3072 ones (one XCD per team with the default spin exchange, XCD
pairs with the flip exchange). Each candidate runs the caller's own workload with the engine's current tables and is
timed with HIP events on the launch stream; the fastest is left in the engine's options ("multi_nt", "mixed_xcd2").
Every candidate computes the same results (the shapes only change which workgroups anneal which replicas).
"""
import torch

from .engine import np_of


def tune_multi(eng, groups, seed=0x7E57, repeats=2, stream=None):
    """Times every candidate shape on `groups` = [(instance id, replicas), ...]; returns
    dict(best=(multi_nt, mixed_xcd2), ms=..., table=[(multi_nt, mixed_xcd2, ms), ...])."""
    st = stream or torch.cuda.current_stream()
    nps = [np_of(eng._inst[g][0]) for g, _ in groups]
    both = any(p > 2048 for p in nps) and any(p <= 2048 for p in nps)
    spin = int(eng.get_option("exchange")) != 0
    splits = (tuple(range(1, 8)) if spin else (2, 4, 6)) if both else (0,)
    R = int(sum(c for _, c in groups))
    out = eng.alloc_out_rows(R, max(nps) // 64)
    table = []
    for nt in (1, 2):
        for x2 in splits:
            eng.set_option("multi_nt", nt)
            eng.set_option("mixed_xcd2", x2)
            eng.launch_multi(groups, seed, None, out, stream=st)   # untimed first launch
            ms = []
            for _ in range(repeats):
                e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                e0.record(st)
                eng.launch_multi(groups, seed, None, out, stream=st)
                e1.record(st)
                e1.synchronize()
                ms.append(e0.elapsed_time(e1))
            eng.sync(st)
            table.append((nt, x2, min(ms)))
    best = min(table, key=lambda r: r[2])
    eng.set_option("multi_nt", best[0])
    eng.set_option("mixed_xcd2", best[1])
    return dict(best=best[:2], ms=best[2], table=table)
