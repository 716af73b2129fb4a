"""ORACLE -- test infrastructure only.

ctypes bindings for the plain-C restatement of the reference golden model (oracle/golden.c -> oracle/liboracle.so)
and, in the development container only, for the reference's own headers behind C drivers (oracle/_ref/libref.so; the
instance-file layer jmat.hpp in oracle/_ref/libref_io.so).
Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may import this module, and only as the checker
(or, for the CPU baseline, as the thing timed on the host); the product package never imports it.
"""
import ctypes
import os
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "liboracle.so"
REF_LIB = HERE / "_ref" / "libref.so"
REF_GEN_ISING = HERE / "_ref" / "gen_ising"
REF_IO_LIB = HERE / "_ref" / "libref_io.so"
LANES = 256

_c_i8p = ctypes.POINTER(ctypes.c_int8)
_c_i32p = ctypes.POINTER(ctypes.c_int32)
_c_i64p = ctypes.POINTER(ctypes.c_int64)
_c_u32p = ctypes.POINTER(ctypes.c_uint32)
_c_u64p = ctypes.POINTER(ctypes.c_uint64)
_c_dp = ctypes.POINTER(ctypes.c_double)


def _ptr(a, ct):
    if a is None:
        return None
    return a.ctypes.data_as(ct)


class _Sched(ctypes.Structure):
    _fields_ = [("S", ctypes.c_int), ("shape", ctypes.c_int), ("t0", ctypes.c_double), ("t1", ctypes.c_double),
                ("q", ctypes.c_double), ("lam", ctypes.c_double), ("jv", ctypes.c_double), ("kappa", ctypes.c_double),
                ("scale", ctypes.c_double), ("ramp", ctypes.c_int), ("Texp", _c_dp)]


_lib = None
_ref = None


def build(lib_path=None, oflags=None):
    """Compile liboracle.so (make), optionally with other optimisation flags into another path."""
    if lib_path is None:
        subprocess.run(["make", "-s", "-C", str(HERE), "all"], check=True)
        return LIB
    flags = oflags or "-O3 -march=native"
    cmd = f"gcc {flags} -std=c11 -fPIC -shared -o {lib_path} {HERE / 'golden.c'} -lm -lpthread"
    subprocess.run(cmd, shell=True, check=True)
    return Path(lib_path)


def lib(path=None):
    global _lib
    if path is not None:
        return _bind(ctypes.CDLL(str(path)))
    if _lib is None:
        if not LIB.exists():
            build()
        _lib = _bind(ctypes.CDLL(str(LIB)))
    return _lib


def _bind(L):
    L.orc_np.argtypes = [ctypes.c_int]; L.orc_np.restype = ctypes.c_int
    L.orc_splitmix.argtypes = [_c_u64p]; L.orc_splitmix.restype = ctypes.c_uint64
    L.orc_rng_seed.argtypes = [_c_u32p, ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint32]; L.orc_rng_seed.restype = None
    L.orc_rng_next.argtypes = [_c_u32p]; L.orc_rng_next.restype = ctypes.c_uint32
    L.orc_tables.argtypes = [ctypes.POINTER(_Sched), _c_i32p, _c_i32p, _c_i64p, _c_i32p]; L.orc_tables.restype = ctypes.c_int
    L.orc_run_trial.argtypes = [ctypes.c_int, _c_i8p, ctypes.c_int64, _c_i32p, ctypes.c_int, _c_i32p, _c_i32p, _c_i64p, _c_i32p,
                                ctypes.c_uint64, ctypes.c_uint32, _c_u64p, _c_i64p, _c_i64p, _c_i64p, _c_i32p, _c_i32p]
    L.orc_run_trial.restype = ctypes.c_int
    L.orc_run_batch.argtypes = [ctypes.c_int, _c_i8p, ctypes.c_int64, _c_i32p, ctypes.c_int, _c_i32p, _c_i32p, _c_i64p, _c_i32p,
                                ctypes.c_uint64, _c_u32p, ctypes.c_int, ctypes.c_int, _c_u64p, _c_i64p, _c_i64p, _c_i64p, _c_i32p]
    L.orc_run_batch.restype = ctypes.c_int
    L.orc_gen_random_k.argtypes = [ctypes.c_uint64, ctypes.c_double, ctypes.c_int, ctypes.c_int, ctypes.c_int, _c_i8p,
                                   ctypes.c_int64, _c_i32p]
    L.orc_gen_random_k.restype = None
    L.orc_gen_ising_random.argtypes = [ctypes.c_uint64, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_int64, _c_i8p,
                                       ctypes.c_int64, _c_i32p]
    L.orc_gen_ising_random.restype = None
    L.orc_gen_mbb_bias.argtypes = [ctypes.c_uint64, ctypes.c_int, ctypes.c_int, _c_i32p]; L.orc_gen_mbb_bias.restype = None
    L.orc_fingerprint.argtypes = [_c_i8p, ctypes.c_int64, ctypes.c_int]; L.orc_fingerprint.restype = ctypes.c_uint64
    L.orc_j2_scale.argtypes = [_c_i8p, ctypes.c_int64, ctypes.c_int, ctypes.c_int]; L.orc_j2_scale.restype = ctypes.c_double
    L.orc_sumw.argtypes = [_c_i8p, ctypes.c_int64, ctypes.c_int]; L.orc_sumw.restype = ctypes.c_int64
    L.orc_fields.argtypes = [_c_i8p, ctypes.c_int64, ctypes.c_int, _c_i32p, _c_u64p, _c_i64p]; L.orc_fields.restype = None
    return L


def np_of(n):
    return lib().orc_np(int(n))


# ------------------------------------------------------------------------------------------------------------ RNG
def rng_stream(seed, trial, lane, count):
    s = np.zeros(4, np.uint32)
    L = lib()
    L.orc_rng_seed(_ptr(s, _c_u32p), np.uint64(seed), np.uint32(trial), np.uint32(lane))
    return np.array([L.orc_rng_next(_ptr(s, _c_u32p)) for _ in range(count)], np.uint32)


def splitmix_seq(x, count):
    st = np.array([x], np.uint64)
    return np.array([lib().orc_splitmix(_ptr(st, _c_u64p)) for _ in range(count)], np.uint64)


# ------------------------------------------------------------------------------------------------------------ tables
def sched_struct(sch, T=None):
    shape = {"geom": 0, "linear": 1, "explicit": 2}[sch.get("shape", "geom")]
    c = _Sched(int(sch["S"]), shape, float(sch.get("t0", 0.0)), float(sch.get("t1", 0.0)), float(sch["q"]),
               float(sch.get("lam", 0.0)), float(sch.get("jv", 0.0)), float(sch.get("kappa", 0.0)),
               float(sch.get("scale", 1.0)), int(bool(sch.get("ramp", False))), None)
    if shape == 2:
        T = np.ascontiguousarray(sch["T"], np.float64)
        c.Texp = _ptr(T, _c_dp)
    return c, T


def tables(sch):
    """sch: dict(S, q, t0, t1, lam, jv, kappa, scale, ramp, shape='geom'|'linear'|'explicit', T=...)."""
    S = int(sch["S"])
    c, keep = sched_struct(sch)
    out = dict(fourT=np.zeros(S, np.int32), q=np.zeros(S, np.int32), kcorr=np.zeros(S, np.int64), kconst=np.zeros(S, np.int32))
    rc = lib().orc_tables(ctypes.byref(c), _ptr(out["fourT"], _c_i32p), _ptr(out["q"], _c_i32p), _ptr(out["kcorr"], _c_i64p),
                          _ptr(out["kconst"], _c_i32p))
    del keep
    if rc:
        raise ValueError("orc_tables failed")
    return out


# ------------------------------------------------------------------------------------------------------------ anneals
def _prep_J(J):
    J = np.ascontiguousarray(J, np.int8)
    return J, int(J.shape[1])


def run_trial(J, n, tb, seed, trial, bias=None, trace=False, want_h=False, L=None):
    L = L or lib()
    J, ld = _prep_J(J)
    S = len(tb["fourT"])
    np_ = L.orc_np(n)
    spins = np.zeros(np_ // 64, np.uint64)
    sh = ctypes.c_int64(); bs = ctypes.c_int64(); fl = ctypes.c_int64()
    nlin = np.zeros(S, np.int32) if trace else None
    h = np.zeros(n, np.int32) if want_h else None
    b = None if bias is None else np.ascontiguousarray(bias, np.int32)
    rc = L.orc_run_trial(int(n), _ptr(J, _c_i8p), ld, _ptr(b, _c_i32p), S, _ptr(tb["fourT"], _c_i32p), _ptr(tb["q"], _c_i32p),
                         _ptr(tb["kcorr"], _c_i64p), _ptr(tb["kconst"], _c_i32p), np.uint64(seed), np.uint32(trial),
                         _ptr(spins, _c_u64p), ctypes.byref(sh), ctypes.byref(bs), ctypes.byref(fl), _ptr(nlin, _c_i32p),
                         _ptr(h, _c_i32p))
    if rc:
        raise ValueError("orc_run_trial failed")
    out = dict(spins=spins, sum_sh=sh.value, sum_bs=bs.value, flips=fl.value)
    if trace:
        out["nlin"] = nlin
    if want_h:
        out["h"] = h
    return out


def run_batch(J, n, tb, seed, trials, bias=None, trace=False, nthreads=0, L=None):
    L = L or lib()
    J, ld = _prep_J(J)
    S = len(tb["fourT"])
    trials = np.ascontiguousarray(trials, np.uint32)
    R = len(trials)
    words = L.orc_np(n) // 64
    spins = np.zeros((R, words), np.uint64)
    sh = np.zeros(R, np.int64); bs = np.zeros(R, np.int64); fl = np.zeros(R, np.int64)
    nlin = np.zeros((R, S), np.int32) if trace else None
    b = None if bias is None else np.ascontiguousarray(bias, np.int32)
    rc = L.orc_run_batch(int(n), _ptr(J, _c_i8p), ld, _ptr(b, _c_i32p), S, _ptr(tb["fourT"], _c_i32p), _ptr(tb["q"], _c_i32p),
                         _ptr(tb["kcorr"], _c_i64p), _ptr(tb["kconst"], _c_i32p), np.uint64(seed), _ptr(trials, _c_u32p), R,
                         int(nthreads), _ptr(spins, _c_u64p), _ptr(sh, _c_i64p), _ptr(bs, _c_i64p), _ptr(fl, _c_i64p),
                         _ptr(nlin, _c_i32p))
    if rc:
        raise ValueError("orc_run_batch failed")
    out = dict(spins=spins, sum_sh=sh, sum_bs=bs, flips=fl)
    if trace:
        out["nlin"] = nlin
    return out


# ------------------------------------------------------------------------------------------------------------ generators
def gen_random_k(seed, density, maxabs, n, bmax=0, ld=2048, rows=None):
    rows = rows or max(ld, n)
    J = np.zeros((rows, ld), np.int8)
    b = np.zeros(n, np.int32)
    lib().orc_gen_random_k(np.uint64(seed), float(density), int(maxabs), int(n), int(bmax), _ptr(J, _c_i8p), ld, _ptr(b, _c_i32p))
    return J, b


def gen_ising_random(seed, n, kbits=2, pzero=0.0, bmax=0, ld=None):
    ld = ld or max(2048, n)
    J = np.zeros((max(ld, n), ld), np.int8)
    b = np.zeros(n, np.int32)
    lib().orc_gen_ising_random(np.uint64(seed), int(n), int(kbits), float(pzero), int(bmax), _ptr(J, _c_i8p), ld, _ptr(b, _c_i32p))
    return J, b


def gen_mbb_bias(fingerprint, maxabs, n):
    b = np.zeros(n, np.int32)
    lib().orc_gen_mbb_bias(np.uint64(fingerprint), int(maxabs), int(n), _ptr(b, _c_i32p))
    return b


def fingerprint(J, np_=2048):
    J, ld = _prep_J(J)
    return int(lib().orc_fingerprint(_ptr(J, _c_i8p), ld, int(np_)))


def j2_scale(J, nsum, div):
    J, ld = _prep_J(J)
    return float(lib().orc_j2_scale(_ptr(J, _c_i8p), ld, int(nsum), int(div)))


def sumw(J, n):
    J, ld = _prep_J(J)
    return int(lib().orc_sumw(_ptr(J, _c_i8p), ld, int(n)))


def fields(J, n, spins, bias=None):
    J, ld = _prep_J(J)
    h = np.zeros(n, np.int64)
    b = None if bias is None else np.ascontiguousarray(bias, np.int32)
    sp = np.ascontiguousarray(spins, np.uint64)
    lib().orc_fields(_ptr(J, _c_i8p), ld, int(n), _ptr(b, _c_i32p), _ptr(sp, _c_u64p), _ptr(h, _c_i64p))
    return h


def spins_pm(words, n):
    """Packed words (bit 1 = +1) -> int8 array of +-1 for the first n slots."""
    w = np.asarray(words, np.uint64)
    bits = np.unpackbits(w.view(np.uint8), bitorder="little")[:n]
    return np.where(bits == 1, 1, -1).astype(np.int8)


# ------------------------------------------------------------------------------------------------------------ reference driver
def ref_available():
    return REF_LIB.exists()


def ref():
    global _ref
    if _ref is None:
        R = ctypes.CDLL(str(REF_LIB))
        R.ref_np.restype = ctypes.c_int
        R.ref_lanes.restype = ctypes.c_int
        R.ref_mb_tables.argtypes = [ctypes.c_int] + [ctypes.c_double] * 7 + [ctypes.c_int, _c_i32p, _c_i32p, _c_i64p, _c_i32p]
        R.ref_build_tables.argtypes = [ctypes.c_int] + [ctypes.c_double] * 6 + [ctypes.c_int, _c_i32p, _c_i32p, _c_i64p, _c_i32p]
        R.ref_make_tables.argtypes = [ctypes.c_int, _c_dp, _c_dp, _c_dp, ctypes.c_double, _c_i32p, _c_i32p, _c_i64p, _c_i32p]
        R.ref_run_trial_bias.argtypes = [ctypes.c_int, _c_i8p, _c_i32p, ctypes.c_int, _c_i32p, _c_i32p, _c_i64p, _c_i32p,
                                         ctypes.c_uint64, ctypes.c_uint32, _c_u64p, _c_i64p, _c_i64p, _c_i64p, _c_i32p]
        R.ref_run_trial.argtypes = [_c_i8p, ctypes.c_int64, ctypes.c_int, _c_i32p, _c_i32p, _c_i64p, _c_i32p, ctypes.c_uint64,
                                    ctypes.c_uint32, _c_u64p, _c_i64p, _c_i64p, _c_i32p]
        R.ref_mb_random_k.argtypes = [ctypes.c_uint64, ctypes.c_double, ctypes.c_int, _c_i8p, _c_i64p, _c_u64p, _c_dp]
        R.ref_mb3_random_k.argtypes = [ctypes.c_uint64, ctypes.c_double, ctypes.c_int, ctypes.c_int, ctypes.c_int, _c_i8p, _c_i32p,
                                       _c_i64p, _c_u64p, _c_dp]
        _ref = R
    return _ref


def _tab_arrays(S):
    return dict(fourT=np.zeros(S, np.int32), q=np.zeros(S, np.int32), kcorr=np.zeros(S, np.int64), kconst=np.zeros(S, np.int32))


def _tab_ptrs(o):
    return (_ptr(o["fourT"], _c_i32p), _ptr(o["q"], _c_i32p), _ptr(o["kcorr"], _c_i64p), _ptr(o["kconst"], _c_i32p))


def ref_mb_tables(sch):
    S = int(sch["S"]); o = _tab_arrays(S)
    ref().ref_mb_tables(S, float(sch["t0"]), float(sch["t1"]), float(sch["q"]), float(sch.get("lam", 0.0)),
                        float(sch.get("jv", 0.0)), float(sch.get("kappa", 0.0)), float(sch.get("scale", 1.0)),
                        int(bool(sch.get("ramp", False))), *_tab_ptrs(o))
    return o


def ref_build_tables(sch):
    S = int(sch["S"]); o = _tab_arrays(S)
    ref().ref_build_tables(S, float(sch["t0"]), float(sch["t1"]), float(sch["q"]), float(sch.get("lam", 0.0)),
                           float(sch.get("jv", 0.0)), float(sch.get("kappa", 0.0)), int(bool(sch.get("ramp", False))),
                           *_tab_ptrs(o))
    return o


def ref_make_tables(T, q, lam, jv=0.0):
    T = np.ascontiguousarray(T, np.float64); q = np.ascontiguousarray(q, np.float64); lam = np.ascontiguousarray(lam, np.float64)
    S = len(T); o = _tab_arrays(S)
    ref().ref_make_tables(S, _ptr(T, _c_dp), _ptr(q, _c_dp), _ptr(lam, _c_dp), float(jv), *_tab_ptrs(o))
    return o


def ref_run_trial_bias(n, dense2048, tb, seed, trial, bias=None, trace=False):
    R = ref()
    J = np.ascontiguousarray(dense2048, np.int8)
    assert J.shape == (2048, 2048)
    S = len(tb["fourT"])
    spins = np.zeros(32, np.uint64)
    sh = ctypes.c_int64(); bs = ctypes.c_int64(); fl = ctypes.c_int64()
    nlin = np.zeros(S, np.int32) if trace else None
    b = None
    if bias is not None:
        b = np.zeros(2048, np.int32); b[:len(bias)] = bias
    rc = R.ref_run_trial_bias(int(n), _ptr(J, _c_i8p), _ptr(b, _c_i32p), S, *_tab_ptrs(tb), np.uint64(seed), np.uint32(trial),
                              _ptr(spins, _c_u64p), ctypes.byref(sh), ctypes.byref(bs), ctypes.byref(fl), _ptr(nlin, _c_i32p))
    if rc:
        raise ValueError("ref_run_trial_bias failed")
    out = dict(spins=spins, sum_sh=sh.value, sum_bs=bs.value, flips=fl.value)
    if trace:
        out["nlin"] = nlin
    return out


def ref_mb3_random_k(seed, density, maxabs, n, bmax):
    J = np.zeros((2048, 2048), np.int8); b = np.zeros(2048, np.int32)
    sw = ctypes.c_int64(); fp = ctypes.c_uint64(); j2 = ctypes.c_double()
    mx = ref().ref_mb3_random_k(np.uint64(seed), float(density), int(maxabs), int(n), int(bmax), _ptr(J, _c_i8p), _ptr(b, _c_i32p),
                                ctypes.byref(sw), ctypes.byref(fp), ctypes.byref(j2))
    return dict(J=J, bias=b, sumw=sw.value, fingerprint=fp.value, j2=j2.value, maxabs=mx)


def ref_mb_random_k(seed, density, maxabs):
    J = np.zeros((2048, 2048), np.int8)
    sw = ctypes.c_int64(); fp = ctypes.c_uint64(); j2 = ctypes.c_double()
    mx = ref().ref_mb_random_k(np.uint64(seed), float(density), int(maxabs), _ptr(J, _c_i8p), ctypes.byref(sw), ctypes.byref(fp),
                               ctypes.byref(j2))
    return dict(J=J, sumw=sw.value, fingerprint=fp.value, j2=j2.value, maxabs=mx)


def ref_gen_ising_random(seed, n, kbits, pzero, bmax, workdir):
    """Runs the reference's own gen_ising binary (built from its source in oracle/_ref) and reads its SCAJINT8/SCABIAS1."""
    prefix = os.path.join(str(workdir), f"gi_{seed}_{n}_{kbits}")
    subprocess.run([str(REF_GEN_ISING), "random", "--n", str(n), "--kbits", str(kbits), "--pzero", repr(float(pzero)),
                    "--bmax", str(bmax), "--seed", str(seed), "--out", prefix], check=True, capture_output=True)
    raw = Path(prefix + ".jint8").read_bytes()
    assert raw[:8] == b"SCAJINT8"
    N = int(np.frombuffer(raw[8:12], np.uint32)[0])
    J = np.frombuffer(raw[16:16 + N * N], np.int8).reshape(N, N).copy()
    rb = Path(prefix + ".bias").read_bytes()
    b = np.frombuffer(rb[16:16 + 4 * N], np.int32).copy()
    return J, b


# ------------------------------------------------------------------------------------------- reference instance-file layer
_ref_io = None
REF_NP = 2048


def ref_io_available():
    return REF_IO_LIB.exists()


def ref_io():
    global _ref_io
    if _ref_io is None:
        R = ctypes.CDLL(str(REF_IO_LIB))
        R.ref_io_error.restype = ctypes.c_char_p
        R.ref_io_load.argtypes = [ctypes.c_int, ctypes.c_char_p, _c_i8p, _c_i32p, _c_i64p, _c_dp]
        R.ref_io_write_jint8.argtypes = [ctypes.c_int, _c_i8p, ctypes.c_char_p]
        R.ref_io_load_bias.argtypes = [ctypes.c_char_p, ctypes.c_int, _c_i32p]
        R.ref_io_write_bias.argtypes = [ctypes.c_int, _c_i32p, ctypes.c_char_p]
        R.ref_io_score.argtypes = [ctypes.c_int, _c_i8p, _c_i32p, _c_u64p, _c_i32p, _c_i64p]
        _ref_io = R
    return _ref_io


class RefIOError(RuntimeError):
    pass


def _ref_io_check(rc):
    if rc:
        raise RefIOError(ref_io().ref_io_error().decode(errors="replace"))


def _pad(J):
    n = J.shape[0]
    D = np.zeros((REF_NP, REF_NP), np.int8)
    D[:n, :n] = J
    return D


def ref_io_load(kind, path):
    """jmat::load(kind, path): (J n x n, stats dict of jmat::finish)."""
    D = np.zeros((REF_NP, REF_NP), np.int8)
    n = ctypes.c_int32(); st = np.zeros(4, np.int64); sq = ctypes.c_double()
    _ref_io_check(ref_io().ref_io_load({"jint8": 0, "graph": 1, "gset": 2}[kind], os.fsencode(str(path)), _ptr(D, _c_i8p),
                                       ctypes.byref(n), _ptr(st, _c_i64p), ctypes.byref(sq)))
    k = n.value
    return D[:k, :k].copy(), dict(sumw=int(st[0]), maxabs=int(st[1]), hmax=int(st[2]), nnz=int(st[3]), mean_row_sq=sq.value)


def ref_io_write_jint8(J, path):
    D = _pad(np.asarray(J, np.int8))
    _ref_io_check(ref_io().ref_io_write_jint8(J.shape[0], _ptr(D, _c_i8p), os.fsencode(str(path))))


def ref_io_load_bias(path, n):
    b = np.zeros(REF_NP, np.int32)
    _ref_io_check(ref_io().ref_io_load_bias(os.fsencode(str(path)), int(n), _ptr(b, _c_i32p)))
    return b[:n].copy()


def ref_io_write_bias(b, path):
    b = np.ascontiguousarray(b, np.int32)
    _ref_io_check(ref_io().ref_io_write_bias(len(b), _ptr(b, _c_i32p), os.fsencode(str(path))))


def ref_io_score(J, spins, bias=None, fields=None):
    """jmat scalar_cut, energy, field_mismatches of one replica (spins: packed words, n <= 2048)."""
    n = J.shape[0]
    D = _pad(np.asarray(J, np.int8))
    b = np.zeros(REF_NP, np.int32)
    if bias is not None:
        b[:n] = bias
    w = np.zeros(REF_NP // 64, np.uint64)
    sp = np.asarray(spins).view(np.uint64)
    w[:len(sp[:REF_NP // 64])] = sp[:REF_NP // 64]
    h = None
    if fields is not None:
        h = np.zeros(REF_NP, np.int32); h[:n] = np.asarray(fields)[:n]
    out = np.zeros(3, np.int64)
    _ref_io_check(ref_io().ref_io_score(n, _ptr(D, _c_i8p), _ptr(b, _c_i32p), _ptr(w, _c_u64p), _ptr(h, _c_i32p),
                                        _ptr(out, _c_i64p)))
    return dict(cut=int(out[0]), energy=int(out[1]), mismatches=int(out[2]))


def _read_gen_ising_files(prefix):
    raw = Path(prefix + ".jint8").read_bytes()
    assert raw[:8] == b"SCAJINT8"
    N = int(np.frombuffer(raw[8:12], np.uint32)[0])
    J = np.frombuffer(raw[16:16 + N * N], np.int8).reshape(N, N).copy()
    rb = Path(prefix + ".bias").read_bytes()
    b = np.frombuffer(rb[16:16 + 4 * N], np.int32).copy()
    import json
    meta = json.loads(Path(prefix + ".json").read_text())
    return J, b, meta


def ref_gen_ising_gp(seed, n, deg, k, A, B, workdir, brute=False):
    """The reference's `gen_ising gp` (built from its source in oracle/_ref): J, bias, metadata (C, ...) and edges."""
    prefix = os.path.join(str(workdir), f"gp_{seed}_{n}_{deg}_{k}")
    cmd = [str(REF_GEN_ISING), "gp", "--n", str(n), "--deg", str(deg), "--k", str(k), "--A", str(A), "--B", str(B),
           "--seed", str(seed), "--out", prefix] + (["--brute", "1"] if brute else [])
    subprocess.run(cmd, check=True, capture_output=True)
    J, b, meta = _read_gen_ising_files(prefix)
    edges = np.loadtxt(prefix + ".edges", dtype=np.int64, ndmin=2)
    return J, b, meta, edges


def ref_gen_ising_tsp(seed, cities, wmax, A, B, workdir, brute=False):
    """The reference's `gen_ising tsp`: J, bias and metadata (C, W, ...)."""
    prefix = os.path.join(str(workdir), f"tsp_{seed}_{cities}_{wmax}")
    cmd = [str(REF_GEN_ISING), "tsp", "--n", str(cities), "--wmax", str(wmax), "--A", str(A), "--B", str(B),
           "--seed", str(seed), "--out", prefix] + (["--brute", "1"] if brute else [])
    subprocess.run(cmd, check=True, capture_output=True)
    return _read_gen_ising_files(prefix)
