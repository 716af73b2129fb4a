"""Writes tests/golden/encodings.json: instances made by the reference's own generator (`gen_ising gp` and
`gen_ising tsp`, built from gpu/multibit_bias_20261008/src/gen_ising.cpp into oracle/_ref/gen_ising) with its
metadata (the constant C of H = sigma f + C, the edge list or the distance matrix and the result of the tool's own
brute-force check). Development container only: needs /root/reference and `make -C oracle ref`.

Usage: python tests/golden/make_encoding_fixtures.py
"""
import json
import sys
import tempfile
from pathlib import Path

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))
from oracle import oracle as O  # noqa: E402

OUT = Path(__file__).resolve().parent / "encodings.json"
GP = [dict(seed=11, n=16, deg=3, k=8, A=2, B=1), dict(seed=12, n=24, deg=4, k=9, A=4, B=1)]
TSP = [dict(seed=21, cities=4, wmax=3, A=4, B=1), dict(seed=22, cities=6, wmax=5, A=9, B=1)]


def main():
    assert O.REF_GEN_ISING.exists(), "build the reference tools first: make -C oracle ref"
    doc = dict(gp=[], tsp=[])
    with tempfile.TemporaryDirectory() as td:
        for a in GP:
            J, b, meta, edges = O.ref_gen_ising_gp(a["seed"], a["n"], a["deg"], a["k"], a["A"], a["B"], td,
                                                   brute=a["n"] <= 24)
            doc["gp"].append(dict(args=a, J=J.tolist(), b=b.tolist(), C=meta["C"], edges=edges.tolist(),
                                  brute=meta.get("brute")))
        for a in TSP:
            J, b, meta = O.ref_gen_ising_tsp(a["seed"], a["cities"], a["wmax"], a["A"], a["B"], td,
                                             brute=(a["cities"] - 1) ** 2 <= 30)
            doc["tsp"].append(dict(args=a, J=J.tolist(), b=b.tolist(), C=meta["C"], W=meta["W"], brute=meta.get("brute")))
    OUT.write_text(json.dumps(doc) + "\n")
    print("wrote", OUT, OUT.stat().st_size, "bytes")


if __name__ == "__main__":
    main()
