"""bench.py's host-side contract (no GPU): the result line is strict JSON, and more than one GPU is refused with a
message and exit status 2."""
import importlib.util
import json
import math
import subprocess
import sys

from conftest import ROOT


def _bench():
    spec = importlib.util.spec_from_file_location("bench_under_test", ROOT / "bench.py")
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_non_finite_values_become_null():
    b = _bench()
    out = b._finite({"tts": math.inf, "cells": [{"score": -math.inf, "p": 0.5}, math.nan], "n": 3, "s": "x"})
    assert out == {"tts": None, "cells": [{"score": None, "p": 0.5}, None], "n": 3, "s": "x"}
    json.dumps(out, allow_nan=False)


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--gpus", "2"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr and r.stdout == ""
