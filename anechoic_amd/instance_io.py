"""Instance files and independent scorers (host side; the work is done by libanechoic_mi355x's C++ code).

Mirrors the reference's instance layer (gpu/multibit_bias_20261008/src/jmat.hpp):
* ``load(kind, path)`` as ``jmat::load`` (:115-121) with kinds "graph" (SNOWGPU1, the K2000.bin format of
  fpga/v80_sca/src/sca_ref.hpp:31-49), "gset" (edge-list text, J = -w) and "jint8" (SCAJINT8);
* ``write_jint8`` / ``write_graph`` / ``write_gset`` and ``load_bias`` / ``write_bias`` (jmat.hpp:105-113, :164-186);
* ``stats(J)`` as ``jmat::finish`` (:35-55): sumw, maxabs, hmax (field bound), mean_row_sq, nnz;
* ``score(J, spins, bias, fields)`` as ``scalar_cut``, ``energy`` and ``field_mismatches`` (:130-162) for a batch of
  replicas.
Errors raise ``anechoic_amd._native.AncError`` with the library's message (the reference throws runtime_error).
Matrices are numpy int8 (n, n); spins are the engine's packed 64-bit words (bit i = 1 iff s_i = +1).
"""
import ctypes
import os

import numpy as np

from . import _native as N


def _path(p):
    return os.fsencode(os.fspath(p))


def _fmt(kind):
    if kind not in N.FORMATS:
        raise ValueError(f"unknown instance kind {kind!r} (one of {sorted(N.FORMATS)})")
    return N.FORMATS[kind]


def detect(path):
    """Instance kind of a file from its first bytes: "jint8", "graph" (SNOWGPU1) or "gset" (text)."""
    with open(path, "rb") as f:
        head = f.read(8)
    return {b"SCAJINT8": "jint8", b"SNOWGPU1": "graph"}.get(head, "gset")


def load(kind, path):
    """J (n x n int8) from a file of the given kind ("graph", "gset", "jint8", or None to detect it)."""
    kind = kind or detect(path)
    L, fmt = N.lib(), _fmt(kind)
    n = ctypes.c_int32(0)
    N.check(L.anc_read_instance(_path(path), fmt, ctypes.byref(n), None, 0))
    J = np.empty((n.value, n.value), dtype=np.int8)
    N.check(L.anc_read_instance(_path(path), fmt, ctypes.byref(n), J.ctypes.data, n.value))
    return J


def _write(kind, path, J):
    J = np.ascontiguousarray(J, dtype=np.int8)
    if J.ndim != 2 or J.shape[0] != J.shape[1]:
        raise ValueError("J must be square")
    N.check(N.lib().anc_write_instance(_path(path), _fmt(kind), J.shape[0], J.ctypes.data, J.shape[1]))


def write_jint8(path, J):
    _write("jint8", path, J)


def write_graph(path, J):
    """SNOWGPU1 (J in {-1, +1} off the diagonal)."""
    _write("graph", path, J)


def write_gset(path, J):
    _write("gset", path, J)


def load_bias(path, n):
    b = np.empty(n, dtype=np.int32)
    N.check(N.lib().anc_read_bias(_path(path), n, b.ctypes.data))
    return b


def write_bias(path, b):
    b = np.ascontiguousarray(b, dtype=np.int32)
    N.check(N.lib().anc_write_bias(_path(path), b.shape[0], b.ctypes.data))


def stats(J):
    J = np.ascontiguousarray(J, dtype=np.int8)
    st = N.JStats()
    N.check(N.lib().anc_instance_stats(J.shape[0], J.ctypes.data, J.shape[1], ctypes.byref(st)))
    return dict(sumw=st.sumw, maxabs=st.maxabs, hmax=st.hmax, mean_row_sq=st.mean_row_sq, nnz=st.nnz)


def score(J, spins, bias=None, fields=None, nthreads=0):
    """Independent cut, energy and (with fields) field-mismatch counts of every replica.

    spins: (R, words) uint64 / int64 packed words (or one replica as a 1-D array); fields: (R, >= n) int32."""
    J = np.ascontiguousarray(J, dtype=np.int8)
    n = J.shape[0]
    sp = np.ascontiguousarray(np.atleast_2d(np.asarray(spins)).view(np.uint64))
    R = sp.shape[0]
    b = None if bias is None else np.ascontiguousarray(bias, dtype=np.int32)
    if b is not None and b.shape[0] < n:
        raise ValueError("bias shorter than n")
    cut, energy = np.zeros(R, np.int64), np.zeros(R, np.int64)
    mis = None
    f = None
    if fields is not None:
        f = np.ascontiguousarray(np.atleast_2d(np.asarray(fields)), dtype=np.int32)
        if f.shape[0] != R:
            raise ValueError("fields and spins hold different replica counts")
        mis = np.zeros(R, np.int64)
    N.check(N.lib().anc_score(n, J.ctypes.data, J.shape[1], None if b is None else b.ctypes.data, R, sp.ctypes.data,
                              sp.shape[1], None if f is None else f.ctypes.data, 0 if f is None else f.shape[1],
                              cut.ctypes.data, energy.ctypes.data, None if mis is None else mis.ctypes.data, nthreads))
    out = dict(cut=cut, energy=energy)
    if mis is not None:
        out["mismatches"] = mis
    return out
