"""Instance files, jmat::finish statistics and the independent scorers (host code of libanechoic_mi355x; no GPU).

Pinned by:
* files written by the reference's own writers (gen_ising random -> jmat::write_jint8 / write_bias) and what the
  reference's instance layer reads from them, committed under tests/golden/ (make_io_fixtures.py);
* the reference's instance layer itself (jmat.hpp compiled from the reference sources into oracle/_ref/libref_io.so,
  development container only): byte-identical writers, identical reads, statistics and scores, and the same files
  refused;
* identities on the golden anneal records (energy = -(sum_sh + sum_bs)/2, cut = (sumw + sum_sh/2)/2) and a
  hand-checked G-set example (cut and energy of a 6-vertex graph worked out by hand).
"""
import json
import struct

import numpy as np
import pytest

from conftest import GOLDEN, hex_to_words
from oracle import oracle as O


@pytest.fixture(scope="module")
def io():
    from anechoic_amd import _native as N
    if not N.LIB_PATH.exists():
        from anechoic_amd import build
        build.build()
    from anechoic_amd import instance_io
    return instance_io


@pytest.fixture(scope="module")
def fx():
    return json.loads((GOLDEN / "io_fixtures.json").read_text())


needs_ref = pytest.mark.skipif(not O.ref_io_available(), reason="reference instance layer not built (oracle/_ref)")


# ----------------------------------------------------------------------------------------- reference-written fixtures
def test_reads_files_written_by_the_reference(io, fx):
    J = io.load("jint8", GOLDEN / fx["jint8"]["file"])
    assert np.array_equal(J, np.array(fx["jint8"]["J"], np.int8))
    assert io.detect(GOLDEN / fx["jint8"]["file"]) == "jint8"
    st = io.stats(J)
    for k, v in fx["jint8"]["stats"].items():
        assert st[k] == pytest.approx(v, rel=0, abs=0), k
    b = io.load_bias(GOLDEN / fx["bias"]["file"], J.shape[0])
    assert b.tolist() == fx["bias"]["b"]
    # the generator the file came from (oracle restatement of gen_ising random) gives the same instance
    Jo, bo = O.gen_ising_random(fx["generator"]["seed"], fx["generator"]["n"], kbits=fx["generator"]["kbits"],
                                pzero=fx["generator"]["pzero"], bmax=fx["generator"]["bmax"])
    n = J.shape[0]
    assert np.array_equal(Jo[:n, :n], J) and np.array_equal(bo[:n], b)


def test_writers_reproduce_reference_files_byte_for_byte(io, fx, tmp_path):
    J = np.array(fx["jint8"]["J"], np.int8)
    io.write_jint8(tmp_path / "a.jint8", J)
    io.write_bias(tmp_path / "a.bias", np.array(fx["bias"]["b"], np.int32))
    assert (tmp_path / "a.jint8").read_bytes() == (GOLDEN / fx["jint8"]["file"]).read_bytes()
    assert (tmp_path / "a.bias").read_bytes() == (GOLDEN / fx["bias"]["file"]).read_bytes()


def test_scores_equal_reference_scores(io, fx):
    J = np.array(fx["jint8"]["J"], np.int8)
    b = np.array(fx["bias"]["b"], np.int32)
    spins = np.array([s["spins"] for s in fx["scores"]], np.uint64)
    fields = np.array([s["fields"] for s in fx["scores"]], np.int32)
    r = io.score(J, spins, bias=b, fields=fields)
    assert r["cut"].tolist() == [s["cut"] for s in fx["scores"]]
    assert r["energy"].tolist() == [s["energy"] for s in fx["scores"]]
    assert r["mismatches"].tolist() == [s["mismatches"] for s in fx["scores"]] == [3, 3, 3, 3]


def test_gset_hand_checked(io, fx):
    """Edges (w): 1-2 (1), 2-3 (-1), 3-4 (2), 4-5 (1), 5-6 (-3), 6-1 (1), 1-4 (5); spins of vertices 1..6 =
    + - + + - +. Cut edges 1-2, 2-3, 4-5, 5-6: cut = 1 - 1 + 1 - 3 = -2. Energy = sum w s_i s_j = -1 + 1 + 2 - 1 + 3
    + 1 + 5 = 10. sumw = 6."""
    J = io.load("gset", GOLDEN / fx["gset"]["file"])
    assert np.array_equal(J, np.array(fx["gset"]["J"], np.int8))
    assert J[0, 3] == J[3, 0] == -5 and J[4, 5] == 3 and J[0, 2] == 0
    st = io.stats(J)
    assert st["sumw"] == 6 == fx["gset"]["stats"]["sumw"] and st["nnz"] == 14
    r = io.score(J, np.array([0b101101], np.uint64))
    assert int(r["cut"][0]) == -2 == fx["gset"]["cut"]
    assert int(r["energy"][0]) == 10 == fx["gset"]["energy"]


# ------------------------------------------------------------------------------------ against the reference's own layer
@needs_ref
@pytest.mark.parametrize("n,kbits,pzero,bmax", [(300, 4, 0.3, 50), (2000, 2, 0.0, 0), (777, 8, 0.9, 300)])
def test_round_trips_through_the_reference(io, tmp_path, n, kbits, pzero, bmax):
    J, b = O.gen_ising_random(900 + n, n, kbits=kbits, pzero=pzero, bmax=bmax)
    J, b = J[:n, :n].copy(), b[:n].copy()
    # ours -> reference reader
    io.write_jint8(tmp_path / "o.jint8", J)
    Jr, st = O.ref_io_load("jint8", tmp_path / "o.jint8")
    assert np.array_equal(Jr, J)
    mine = io.stats(J)
    for k in ("sumw", "maxabs", "hmax", "nnz"):
        assert mine[k] == st[k], k
    assert mine["mean_row_sq"] == st["mean_row_sq"]
    io.write_gset(tmp_path / "o.gset", J)
    assert np.array_equal(O.ref_io_load("gset", tmp_path / "o.gset")[0], J)
    io.write_bias(tmp_path / "o.bias", b)
    assert np.array_equal(O.ref_io_load_bias(tmp_path / "o.bias", n), b)
    # reference writer -> ours, byte-identical files
    O.ref_io_write_jint8(J, tmp_path / "r.jint8")
    O.ref_io_write_bias(b, tmp_path / "r.bias")
    assert (tmp_path / "r.jint8").read_bytes() == (tmp_path / "o.jint8").read_bytes()
    assert (tmp_path / "r.bias").read_bytes() == (tmp_path / "o.bias").read_bytes()
    assert np.array_equal(io.load("jint8", tmp_path / "r.jint8"), J)
    if kbits == 2 and pzero == 0.0:
        # +-1 complete graph: SNOWGPU1 (the reference reads it only at n = 2000)
        io.write_graph(tmp_path / "o.bin", J)
        Jg, stg = O.ref_io_load("graph", tmp_path / "o.bin")
        assert np.array_equal(Jg, J) and stg["sumw"] == mine["sumw"]
        assert np.array_equal(io.load(None, tmp_path / "o.bin"), J)


@needs_ref
@pytest.mark.parametrize("n", [37, 300, 2000])
def test_scores_equal_reference_on_random_replicas(io, n):
    J, b = O.gen_ising_random(4000 + n, n, kbits=4, pzero=0.5, bmax=100)
    J, b = J[:n, :n].copy(), b[:n].copy()
    rng = np.random.default_rng(n)
    R = 6
    words = (n + 63) // 64
    spins = rng.integers(0, 2**63, size=(R, words), dtype=np.int64).astype(np.uint64)
    spins[0] = 0                        # all -1
    spins[1] = np.uint64(2**64 - 1)     # all +1
    fields = np.zeros((R, n), np.int32)
    for r in range(R):
        fields[r] = O.fields(J, n, spins[r], bias=b)[:n]
    bad = rng.choice(n, size=min(n, 9), replace=False)
    fields[2, bad] += 1
    got = io.score(J, spins, bias=b, fields=fields, nthreads=3)
    for r in range(R):
        ref = O.ref_io_score(J, spins[r], bias=b, fields=fields[r])
        assert int(got["cut"][r]) == ref["cut"] and int(got["energy"][r]) == ref["energy"], r
        assert int(got["mismatches"][r]) == ref["mismatches"] == (len(bad) if r == 2 else 0), r


def _jint8_bytes(n, body, rsv=0, magic=b"SCAJINT8"):
    return magic + struct.pack("<II", n, rsv) + bytes(body)


def _bad_files():
    J = np.array([[0, 1, -1], [1, 0, 2], [-1, 2, 0]], np.int8)
    asym = J.copy(); asym[0, 1] = 3
    diag = J.copy(); diag[1, 1] = 1
    body = J.tobytes()
    b = struct.pack("<3i", 1, -2, 3)
    return [
        ("jint8", "bad_magic", _jint8_bytes(3, body, magic=b"SCAJINT9")),
        ("jint8", "reserved", _jint8_bytes(3, body, rsv=1)),
        ("jint8", "n_zero", _jint8_bytes(0, b"")),
        ("jint8", "short", _jint8_bytes(3, body[:-1])),
        ("jint8", "trailing", _jint8_bytes(3, body + b"\0")),
        ("jint8", "asymmetric", _jint8_bytes(3, asym.tobytes())),
        ("jint8", "diagonal", _jint8_bytes(3, diag.tobytes())),
        ("gset", "header", b"x y\n"),
        ("gset", "self_loop", b"3 1\n2 2 1\n"),
        ("gset", "out_of_range", b"3 1\n1 4 1\n"),
        ("gset", "duplicate", b"3 2\n1 2 1\n2 1 1\n"),
        ("gset", "weight", b"3 1\n1 2 128\n"),
        ("gset", "short", b"3 2\n1 2 1\n"),
        ("gset", "trailing_integer", b"3 1\n1 2 1\n7\n"),
        ("bias", "bad_magic", b"SCABIAS2" + struct.pack("<II", 3, 0) + b),
        ("bias", "length", b"SCABIAS1" + struct.pack("<II", 4, 0) + b),
        ("bias", "short", b"SCABIAS1" + struct.pack("<II", 3, 0) + b[:-1]),
        ("bias", "trailing", b"SCABIAS1" + struct.pack("<II", 3, 0) + b + b"\0"),
        ("bias", "range", b"SCABIAS1" + struct.pack("<II", 3, 0) + struct.pack("<3i", 1, 40000, 0)),
    ]


@pytest.mark.parametrize("kind,name,data", _bad_files(), ids=[f"{k}-{n}" for k, n, _ in _bad_files()])
def test_malformed_files_are_refused(io, tmp_path, kind, name, data):
    from anechoic_amd import _native as N
    p = tmp_path / f"{name}.{kind}"
    p.write_bytes(data)
    with pytest.raises(N.AncError):
        io.load_bias(p, 3) if kind == "bias" else io.load(kind, p)
    if O.ref_io_available():
        with pytest.raises(O.RefIOError):
            O.ref_io_load_bias(p, 3) if kind == "bias" else O.ref_io_load(kind, p)


def test_trailing_text_after_gset_edges_is_accepted(io, tmp_path):
    """A non-numeric tail is not an extra integer: the reference reader (jmat.hpp:82) and this one both accept it."""
    p = tmp_path / "tail.gset"
    p.write_bytes(b"3 1\n1 2 1\n# end\n")
    J = io.load("gset", p)
    assert J[0, 1] == J[1, 0] == -1
    if O.ref_io_available():
        assert np.array_equal(O.ref_io_load("gset", p)[0], J)


def test_writer_argument_checks(io, tmp_path):
    from anechoic_amd import _native as N
    J = np.array([[0, 2], [2, 0]], np.int8)
    with pytest.raises(N.AncError, match="SNOWGPU1"):
        io.write_graph(tmp_path / "x.bin", J)
    with pytest.raises(N.AncError, match="symmetric"):
        io.write_jint8(tmp_path / "x.jint8", np.array([[0, 1], [2, 0]], np.int8))
    with pytest.raises(N.AncError, match="32767"):
        io.write_bias(tmp_path / "x.bias", np.array([0, -40000], np.int32))
    with pytest.raises(N.AncError, match="cannot open"):
        io.load("jint8", tmp_path / "missing.jint8")


# --------------------------------------------------------------------------------------------- golden anneal records
def test_identities_on_golden_anneal_records(io, golden_trials):
    """energy = -(sum_sh + sum_bs)/2 and cut = (sumw + sum_sh/2)/2 (sca_ref.hpp:153) for the reference's own anneal
    results, with the cut and energy recomputed from J and the returned spins."""
    from anechoic_amd import instances as I
    checked = 0
    for name, rec in golden_trials.items():
        inst = rec["instance"]
        n = inst["n"]
        if inst["generator"] == "ising_random":
            J, b = I.ising_random(inst["seed"], n, kbits=inst["kbits"], pzero=inst["pzero"], bmax=inst.get("bmax", 0))
        elif inst["generator"] == "er":
            J, b = I.erdos_renyi(inst["seed"], n, inst["density"])
        else:
            J, b = I.torus(inst["seed"], *I.TORUS_SHAPES[n])
        st = io.stats(J)
        assert "sumw" not in inst or st["sumw"] == inst["sumw"], name
        res = rec["results"]
        spins = np.array([hex_to_words(r["spins"]) for r in res], np.uint64)
        sc = io.score(J, spins, bias=b)
        for r, row in enumerate(res):
            assert int(sc["energy"][r]) * 2 == -(row["sum_sh"] + row["sum_bs"]), (name, r)
            if not np.any(b):
                assert int(sc["cut"][r]) * 4 == 2 * st["sumw"] + row["sum_sh"], (name, r)
            checked += 1
    assert checked >= 25
