"""Launch-shape autotuner for the team launches, with a persistent cache.

The reference picks among compiled launch variants per run (gpu/onsager_v2_20261007/src/sca_gpu.cu:549-593: cluster
size CS and group count G; gpu/multibit_bias_20261008/src/sca_gpu_bias.cu:590-593: K, R, PREG, G). This engine has
the analogous choices at run time, as context options; every candidate computes the same results (the shapes only
change which workgroups anneal which replicas and how the spins travel), so the tuner times the caller's own workload
and keeps the fastest:

* ``tune_batched`` (anc_run_batched): for n <= 2048 the team shape -- 16-member teams with 16-replica items
  (``batched_team`` 16) or 32-member teams with 16- or 32-replica items (``batched_team`` 32, ``batched_nt`` 1 / 2) --
  and the exchange encoding (spin crumbs, spin nibbles or flip bytes: option ``exchange``); for 2048 < n <= 3072 the
  exchange encoding and replica tile (32-workgroup 12-round teams with spin nibbles and 16- or 32-replica items,
  ``multi_nt3``; or 64-workgroup teams with flip bytes); for
  n > 3072 the large kernel's replica tile (``large_tile`` 128 or 64: 64-replica items keep every team busy when R is
  small, at twice the J' stream per replica).
* ``tune_multi`` (anc_run_multi): the replica tiles of the n <= 2048 items (``multi_nt``) and of the 2048 < n <= 3072
  items (``multi_nt3``, spin exchanges) and, when both team shapes are present, the split of the 8 XCDs between them
  (``mixed_xcd2``; 8 = every XCD runs both classes in turn).

Every candidate first gets an untimed launch (code object load, clock ramp-up); then `repeats` rounds each time every
candidate once, in turn, with HIP events on the launch stream, and candidates are ranked by their median, so that a
drift of the clock or of the memory state during tuning does not favour one candidate. Choices are
cached in a JSON file keyed by device name, workload shape (n or the group list, R) and correction mode (counted when
any kcorr is nonzero, else temperature-scaled / plain), so a later process applies them without timing again.
"""
import json
import os
import tempfile
from pathlib import Path

import numpy as np
import torch

from .engine import np_of

OPTION_KEYS = ("batched_team", "batched_nt", "exchange", "multi_nt", "multi_nt3", "mixed_xcd2", "large_tile")


def default_cache_path():
    base = os.environ.get("ANECHOIC_AUTOTUNE_CACHE")
    if base:
        return Path(base)
    root = os.environ.get("XDG_CACHE_HOME") or str(Path.home() / ".cache")
    return Path(root) / "anechoic_amd" / "autotune.json"


class TuneCache:
    """Choices by key in a JSON file (read once, rewritten atomically on every put)."""

    def __init__(self, path=None):
        self.path = Path(path) if path is not None else default_cache_path()
        try:
            self.data = json.loads(self.path.read_text())
            if not isinstance(self.data, dict):
                self.data = {}
        except (OSError, ValueError):
            self.data = {}

    def get(self, key):
        return self.data.get(key)

    def put(self, key, value):
        self.data[key] = value
        self.path.parent.mkdir(parents=True, exist_ok=True)
        fd, tmp = tempfile.mkstemp(dir=str(self.path.parent), prefix=".autotune.", suffix=".json")
        with os.fdopen(fd, "w") as f:
            json.dump(self.data, f, indent=1, sort_keys=True)
        os.replace(tmp, self.path)


def _device_name(eng):
    return torch.cuda.get_device_name(eng.device).replace("|", "/")


def _mode(eng):
    return "counted" if np.any(np.asarray(eng.tables["kcorr"]) != 0) else "kappa"


def batched_key(eng, inst, R):
    n = eng._inst[inst][0]
    return f"batched|{_device_name(eng)}|n={n}|R={int(R)}|mode={_mode(eng)}"


def multi_key(eng, groups):
    sizes = ",".join(f"{eng._inst[g][0]}x{int(c)}" for g, c in groups)
    return f"multi3|{_device_name(eng)}|groups={sizes}|mode={_mode(eng)}"


def apply(eng, options):
    for k, v in options.items():
        if k in OPTION_KEYS:
            eng.set_option(k, v)


def _time_round_robin(candidates, setup, launch, st, repeats):
    """Median launch time of every candidate (setup(c) then launch()), timed round-robin after one untimed pass."""
    for c in candidates:
        setup(c)
        launch()
    times = [[] for _ in candidates]
    for _ in range(repeats):
        for i, c in enumerate(candidates):
            setup(c)
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record(st)
            launch()
            e1.record(st)
            e1.synchronize()
            times[i].append(e0.elapsed_time(e1))
    return [float(np.median(t)) for t in times]


def batched_candidates(n):
    """Option sets for anc_run_batched on an instance of n spins."""
    np_ = np_of(n)
    if np_ <= 2048:
        return [dict(batched_team=16, batched_nt=0, exchange=2), dict(batched_team=32, batched_nt=1, exchange=2),
                dict(batched_team=32, batched_nt=2, exchange=2), dict(batched_team=32, batched_nt=2, exchange=1),
                dict(batched_team=32, batched_nt=2, exchange=0)]
    if np_ <= 3072:
        return [dict(batched_team=0, batched_nt=1, exchange=2, multi_nt3=1), dict(batched_team=0, batched_nt=2, exchange=2, multi_nt3=2),
                dict(batched_team=0, batched_nt=1, exchange=0), dict(batched_team=0, batched_nt=2, exchange=0)]
    return [dict(large_tile=128), dict(large_tile=64)]


def tune_batched(eng, inst, R, seed=0x7E57, repeats=5, cache=None, stream=None, force=False):
    """Picks the fastest launch shape for `R` anneals of instance `inst` under the engine's current tables and leaves it
    in the engine's options. Returns dict(best=options, ms, table=[(options, ms), ...], cached=bool, key)."""
    cache = cache if cache is not None else TuneCache()
    key = batched_key(eng, inst, R)
    hit = None if force else cache.get(key)
    if hit is not None:
        apply(eng, hit["best"])
        return dict(hit, cached=True, key=key)
    st = stream or torch.cuda.current_stream()
    out = eng.alloc_out(inst, R)
    trials = torch.arange(R, dtype=torch.int32, device=eng.device)
    old = {k: int(eng.get_option(k)) for k in ("batched_team", "batched_nt", "exchange", "multi_nt3", "large_tile")}
    cands = batched_candidates(eng._inst[inst][0])
    try:
        ms = _time_round_robin(cands, lambda o: apply(eng, o),
                               lambda: eng.launch(inst, seed, trials, out, kernel="batched", stream=st), st, repeats)
        eng.sync(st)
    finally:
        apply(eng, old)
    table = [(dict(o), m) for o, m in zip(cands, ms)]
    best = min(table, key=lambda r: r[1])
    rec = dict(best=best[0], ms=best[1], table=[[o, m] for o, m in table])
    cache.put(key, rec)
    apply(eng, best[0])
    return dict(rec, cached=False, key=key)


def tune_multi(eng, groups, seed=0x7E57, repeats=3, stream=None, cache=None, force=False):
    """Times every candidate shape on `groups` = [(instance id, replicas), ...] and leaves the fastest in the engine's
    options; returns dict(best=(multi_nt, multi_nt3, mixed_xcd2), ms=..., table=[(multi_nt, multi_nt3, mixed_xcd2, ms),
    ...], cached, key)."""
    cache = cache if cache is not None else TuneCache()
    key = multi_key(eng, groups)
    hit = None if force else cache.get(key)

    def setup(c):
        eng.set_option("multi_nt", c[0])
        eng.set_option("multi_nt3", c[1])
        eng.set_option("mixed_xcd2", c[2])
    if hit is not None:
        setup(hit["best"])
        return dict(hit, best=tuple(hit["best"]), cached=True, key=key)
    st = stream or torch.cuda.current_stream()
    nps = [np_of(eng._inst[g][0]) for g, _ in groups]
    has2, has3 = any(p <= 2048 for p in nps), any(p > 2048 for p in nps)
    spin = int(eng.get_option("exchange")) != 0
    splits = (tuple(range(1, 9)) if spin else (2, 4, 6)) if has2 and has3 else (0,)   # 8: both classes in turn
    R = int(sum(c for _, c in groups))
    out = eng.alloc_out_rows(R, max(nps) // 64)
    cands = [(nt, nt3, x2) for nt in ((1, 2) if has2 else (2,)) for nt3 in ((1, 2) if has3 and spin else (1,))
             for x2 in splits]
    ms = _time_round_robin(cands, setup, lambda: eng.launch_multi(groups, seed, None, out, stream=st), st, repeats)
    eng.sync(st)
    table = [(*c, m) for c, m in zip(cands, ms)]
    best = min(table, key=lambda r: r[3])
    setup(best)
    rec = dict(best=list(best[:3]), ms=best[3], table=[list(r) for r in table])
    cache.put(key, rec)
    return dict(rec, best=tuple(best[:3]), cached=False, key=key)
