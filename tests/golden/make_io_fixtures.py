"""Writes the instance-file fixtures under tests/golden/ (development container only: needs /root/reference and
`make -C oracle ref`).

* io_ref_n37.jint8 / io_ref_n37.bias -- written by the reference's own instance generator (gen_ising random, built from
  gpu/multibit_bias_20261008/src/gen_ising.cpp into oracle/_ref/gen_ising; its files go through jmat::write_jint8 and
  jmat::write_bias), copied unchanged.
* io_small.gset -- a small G-set edge list (text written here; the reference has no G-set writer).
* io_fixtures.json -- what the reference's instance layer (jmat.hpp, oracle/_ref/libref_io.so) returns for them:
  the matrices and biases it reads, jmat::finish statistics, and scalar_cut / energy / field_mismatches of fixed spin
  vectors (with fields equal to b + J s except for three planted differences).

Usage: python tests/golden/make_io_fixtures.py
"""
import json
import shutil
import sys
import tempfile
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))
from oracle import oracle as O  # noqa: E402

OUT = Path(__file__).resolve().parent
GEN = dict(seed=20261203, n=37, kbits=4, pzero=0.3, bmax=40)
GSET_TEXT = "6 7\n1 2 1\n2 3 -1\n3 4 2\n4 5 1\n5 6 -3\n6 1 1\n1 4 5\n"


def main():
    assert O.ref_io_available() and O.REF_GEN_ISING.exists(), "build the reference drivers first: make -C oracle ref"
    with tempfile.TemporaryDirectory() as td:
        O.ref_gen_ising_random(GEN["seed"], GEN["n"], GEN["kbits"], GEN["pzero"], GEN["bmax"], td)
        prefix = Path(td) / f"gi_{GEN['seed']}_{GEN['n']}_{GEN['kbits']}"
        shutil.copyfile(str(prefix) + ".jint8", OUT / "io_ref_n37.jint8")
        shutil.copyfile(str(prefix) + ".bias", OUT / "io_ref_n37.bias")
    (OUT / "io_small.gset").write_text(GSET_TEXT)

    J, st = O.ref_io_load("jint8", OUT / "io_ref_n37.jint8")
    b = O.ref_io_load_bias(OUT / "io_ref_n37.bias", J.shape[0])
    Jg, stg = O.ref_io_load("gset", OUT / "io_small.gset")
    rng = np.random.default_rng(7)
    scores = []
    for k in range(4):
        words = rng.integers(0, 2**63, size=1, dtype=np.int64).astype(np.uint64)
        s = O.spins_pm(words, J.shape[0]).astype(np.int64)
        h = (b + J.astype(np.int64) @ s).astype(np.int32)
        h[[1, 5, 30]] += k + 1
        r = O.ref_io_score(J, words, bias=b, fields=h)
        scores.append(dict(spins=[int(w) for w in words], fields=h.tolist(), **r))
    rg = O.ref_io_score(Jg, np.array([0b101101], np.uint64))
    doc = dict(
        generator=dict(tool="gen_ising random", **GEN),
        jint8=dict(file="io_ref_n37.jint8", n=int(J.shape[0]), J=J.tolist(), stats=st),
        bias=dict(file="io_ref_n37.bias", b=b.tolist()),
        scores=scores,
        gset=dict(file="io_small.gset", J=Jg.tolist(), stats=stg, spins=0b101101, **rg),
    )
    (OUT / "io_fixtures.json").write_text(json.dumps(doc, indent=1) + "\n")
    print("wrote io fixtures:", [p.name for p in sorted(OUT.glob("io_*"))])


if __name__ == "__main__":
    main()
