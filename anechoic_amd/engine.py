"""Python host surface of the MI355X SCA/Onsager engine.

Mirrors the reference's engine call surfaces:
  * research/reaim_benchmarks_20261008/hwmodel.py:179-195 ``run(P, sch, seed, trials, trace)`` -> dict(spins (B, n) int8,
    flips, sum_sh, sum_bs, nlin_total[, nlin_trace_last]) with ``sch = dict(t0, t1, S, q, lam, jv, kappa, ramp)``;
  * hwmodel.py:32-52 / fpga/v80_sca/src/mb_common.hpp:141-156 ``tables(sch)`` (fourT, q, kcorr, kconst), here built
    by the library's C++ table builder so that every double rounding matches the reference;
  * fpga/v80_sca/src/sca_ref_bias.hpp:25 ``run_trial_bias(n, dense, bias, tables, seed, trial, trace)``.
The couplings are a dense int8 matrix (numpy array or torch tensor, n x n or padded), optionally with an int32 bias.
Every anneal runs in the HIP library (libanechoic_mi355x.so) on the GPU; torch only holds the device buffers.
"""
import ctypes
import warnings

import numpy as np
import torch

from . import _native as N

_SCHED_DEFAULTS = dict(t0=30.0, t1=5.0, q=4.0, lam=0.0, jv=0.0, kappa=0.0, scale=1.0, ramp=False, shape="geom")


def _sched_struct(sch):
    s = dict(_SCHED_DEFAULTS)
    s.update(sch)
    if "S" not in s:
        raise ValueError("schedule needs S (steps)")
    shape = N.SHAPES[s["shape"]]
    keep = None
    c = N.Sched(int(s["S"]), shape, float(s["t0"]), float(s["t1"]), float(s["q"]), float(s["lam"]), float(s["jv"]),
                float(s["kappa"]), float(s["scale"]), int(bool(s["ramp"])), None)
    if shape == 2:
        keep = np.ascontiguousarray(s["T"], np.float64)
        if len(keep) != c.S:
            raise ValueError("explicit T must have S entries")
        c.T = keep.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    return c, keep


def tables(sch):
    """Schedule dict -> fixed-point tables dict(fourT int32, q int32, kcorr int64, kconst int32) (mb::tables rules).

    Keys: S, t0, t1, q, lam, jv, kappa, scale (correction scale gamma), ramp, shape ('geom' | 'linear' | 'explicit'),
    T (explicit temperatures)."""
    c, keep = _sched_struct(sch)
    S = c.S
    out = dict(fourT=np.zeros(S, np.int32), q=np.zeros(S, np.int32), kcorr=np.zeros(S, np.int64), kconst=np.zeros(S, np.int32))
    N.check(N.lib().anc_make_tables(ctypes.byref(c), out["fourT"].ctypes.data_as(N._c_i32p), out["q"].ctypes.data_as(N._c_i32p),
                                    out["kcorr"].ctypes.data_as(N._c_i64p), out["kconst"].ctypes.data_as(N._c_i32p)))
    del keep
    return out


def np_of(n):
    return N.lib().anc_np(int(n))


def _stream_ptr(stream):
    if stream is None:
        stream = torch.cuda.current_stream()
    return ctypes.c_void_p(stream.cuda_stream)


class Engine:
    """A context on one GPU: tables, registered instances and the exchange workspace of the team kernels.

    fallback=True opts in to recovery from an exchange time-out: the library then switches the context to the
    device-scope exchange and run()/run_multi() repeat the run once, with a warning. By default a time-out raises."""

    def __init__(self, device=0, fallback=False):
        if not torch.cuda.is_available():
            raise RuntimeError("anechoic_amd needs a ROCm GPU (torch.cuda.is_available() is False)")
        self.device = torch.device("cuda", device)
        self._ctx = N.lib().anc_create(int(device))
        if not self._ctx:
            raise N.AncError(-4, N.lib().anc_last_error().decode())
        self._inst = []      # (n, J tensor, bias tensor) kept alive while registered
        self.S = 0
        self.fallback = bool(fallback)
        if self.fallback:
            self.set_option("fallback", 1)

    def close(self):
        if self._ctx:
            N.lib().anc_destroy(self._ctx)
            self._ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ------------------------------------------------------------------------------------------------ set-up
    def set_schedule(self, sch):
        return self.set_tables(tables(sch))

    def set_tables(self, tb):
        ft = np.ascontiguousarray(tb["fourT"], np.int32); q = np.ascontiguousarray(tb["q"], np.int32)
        kc = np.ascontiguousarray(tb["kcorr"], np.int64); kk = np.ascontiguousarray(tb["kconst"], np.int32)
        S = len(ft)
        N.check(N.lib().anc_set_tables(self._ctx, S, ft.ctypes.data_as(N._c_i32p), q.ctypes.data_as(N._c_i32p),
                                       kc.ctypes.data_as(N._c_i64p), kk.ctypes.data_as(N._c_i32p)))
        self.S = S
        self.tables = dict(fourT=ft, q=q, kcorr=kc, kconst=kk)
        return self

    def add_instance(self, J, bias=None, n=None, stream=None):
        """J: (rows, ld) int8 (numpy or torch), n x n used; bias: n int32 or None. Returns the instance id."""
        Jt = torch.as_tensor(J)
        if Jt.dtype != torch.int8 or Jt.dim() != 2:
            raise ValueError("J must be a 2-D int8 matrix")
        n = int(n if n is not None else Jt.shape[0])
        if Jt.shape[0] < n or Jt.shape[1] < n:
            raise ValueError("J smaller than n x n")
        Jt = Jt.to(self.device).contiguous()
        bt = None
        if bias is not None:
            bt = torch.as_tensor(bias).to(device=self.device, dtype=torch.int32).contiguous()
            if bt.numel() < n:
                raise ValueError("bias needs n entries")
        inst = N.check(N.lib().anc_add_instance(self._ctx, n, ctypes.c_void_p(Jt.data_ptr()), Jt.shape[1],
                                                ctypes.c_void_p(bt.data_ptr()) if bt is not None else None, _stream_ptr(stream)))
        self._inst.append((n, Jt, bt))
        assert inst == len(self._inst) - 1
        return inst

    def clear_instances(self):
        """Unregisters every instance (and frees the context's captured runs, which reference them)."""
        N.check(N.lib().anc_clear_instances(self._ctx))
        self._inst = []

    def bound(self, inst):
        return N.check(N.lib().anc_instance_bound(self._ctx, int(inst)))

    def set_option(self, key, value):
        N.check(N.lib().anc_set_option(self._ctx, key.encode(), float(value)))

    # ------------------------------------------------------------------------------------------------ runs
    def alloc_out(self, inst, R, trace=False, fields=False):
        n = self._inst[inst][0]
        words = np_of(n) // 64
        dev = self.device
        o = dict(spins=torch.empty((R, words), dtype=torch.int64, device=dev),
                 sum_sh=torch.empty(R, dtype=torch.int64, device=dev), sum_bs=torch.empty(R, dtype=torch.int64, device=dev),
                 flips=torch.empty(R, dtype=torch.int64, device=dev))
        if trace:
            o["nlin"] = torch.empty((R, self.S), dtype=torch.int32, device=dev)
        if fields:
            o["fields"] = torch.empty((R, n), dtype=torch.int32, device=dev)
        return o

    @staticmethod
    def _out_struct(o):
        return N.Out(o["spins"].data_ptr(), o["spins"].shape[1], o["sum_sh"].data_ptr(), o["sum_bs"].data_ptr(),
                     o["flips"].data_ptr(), o["nlin"].data_ptr() if "nlin" in o else None,
                     o["fields"].data_ptr() if "fields" in o else None, o["fields"].shape[1] if "fields" in o else 0)

    def launch(self, inst, seed, trials, out, kernel="batched", stream=None):
        """Enqueues the anneals on the stream without synchronising. trials: device uint32 tensor (batched) or int (plain)."""
        st = _stream_ptr(stream)
        os_ = self._out_struct(out)
        if kernel == "plain":
            N.check(N.lib().anc_run_plain(self._ctx, int(inst), ctypes.c_uint64(int(seed) & (2**64 - 1)), int(trials),
                                          ctypes.byref(os_), st))
        elif kernel in ("batched", "batched_noalloc"):
            R = out["sum_sh"].shape[0]
            tp = None
            if trials is not None:
                assert trials.dtype == torch.int32 and trials.is_cuda and trials.numel() == R
                tp = ctypes.c_void_p(trials.data_ptr())
            fn = N.lib().anc_run_batched if kernel == "batched" else N.lib().anc_run_batched_noalloc
            N.check(fn(self._ctx, int(inst), ctypes.c_uint64(int(seed) & (2**64 - 1)), tp, R, ctypes.byref(os_), st))
        else:
            raise ValueError(f"unknown kernel {kernel}")

    def launch_multi(self, groups, seed, trials, out, stream=None):
        """Enqueues an uneven multi-instance batch: groups = [(instance id, replicas), ...]; output rows follow the
        group order; trials: device int32 tensor (one id per row) or None (row index)."""
        inst = np.array([g[0] for g in groups], np.int32)
        cnt = np.array([g[1] for g in groups], np.int32)
        tp = None
        if trials is not None:
            assert trials.dtype == torch.int32 and trials.is_cuda and trials.numel() == int(cnt.sum())
            tp = ctypes.c_void_p(trials.data_ptr())
        os_ = self._out_struct(out)
        N.check(N.lib().anc_run_multi(self._ctx, len(groups), inst.ctypes.data_as(N._c_i32p), cnt.ctypes.data_as(N._c_i32p),
                                      ctypes.c_uint64(int(seed) & (2**64 - 1)), tp, ctypes.byref(os_), _stream_ptr(stream)))

    def capture_multi(self, groups, seed, trials, out):
        """Captures the launch of launch_multi(groups, seed, trials, out) as a hipGraph owned by the context; returns a
        CapturedRun whose launch() replays it. The trial-id tensor and the output tensors are referenced, not copied:
        change the trial ids in place between replays to anneal fresh replicas."""
        inst = np.array([g[0] for g in groups], np.int32)
        cnt = np.array([g[1] for g in groups], np.int32)
        tp = None
        if trials is not None:
            assert trials.dtype == torch.int32 and trials.is_cuda and trials.numel() == int(cnt.sum())
            tp = ctypes.c_void_p(trials.data_ptr())
        os_ = self._out_struct(out)
        gid = N.check(N.lib().anc_graph_multi(self._ctx, len(groups), inst.ctypes.data_as(N._c_i32p),
                                              cnt.ctypes.data_as(N._c_i32p), ctypes.c_uint64(int(seed) & (2**64 - 1)), tp,
                                              ctypes.byref(os_)))
        return CapturedRun(self, gid, trials, out)

    def get_option(self, key):
        v = N.lib().anc_get_option(self._ctx, key.encode())
        if v < 0:
            raise N.AncError(int(v), N.lib().anc_last_error().decode())
        return v

    def alloc_out_rows(self, R, words, n_fields=0, trace=False):
        dev = self.device
        o = dict(spins=torch.empty((R, words), dtype=torch.int64, device=dev),
                 sum_sh=torch.empty(R, dtype=torch.int64, device=dev), sum_bs=torch.empty(R, dtype=torch.int64, device=dev),
                 flips=torch.empty(R, dtype=torch.int64, device=dev))
        if trace:
            o["nlin"] = torch.empty((R, self.S), dtype=torch.int32, device=dev)
        if n_fields:
            o["fields"] = torch.empty((R, n_fields), dtype=torch.int32, device=dev)
        return o

    def _retry(self, fn):
        """fn() once; with the fallback opt-in, once more after an exchange time-out (logged as a warning)."""
        try:
            return fn()
        except N.AncError as e:
            if not (self.fallback and "timed out" in str(e)):
                raise
            warnings.warn(f"anechoic_amd: {e}; repeating the run on the device-scope exchange", RuntimeWarning)
            return fn()

    def run_multi(self, groups, seed, trials=None, trace=False, fields=False):
        """Uneven multi-instance batch; returns host numpy results (rows in group order)."""
        return self._retry(lambda: self._run_multi(groups, seed, trials, trace, fields))

    def _run_multi(self, groups, seed, trials, trace, fields):
        R = int(sum(g[1] for g in groups))
        words = max(np_of(self._inst[g[0]][0]) for g in groups) // 64
        nf = max(self._inst[g[0]][0] for g in groups) if fields else 0
        out = self.alloc_out_rows(R, words, nf, trace)
        tt = None
        if trials is not None:
            tt = torch.from_numpy(np.asarray(trials, np.uint32).view(np.int32)).to(self.device)
        self.launch_multi(groups, seed, tt, out)
        self.sync()
        return {k: v.cpu().numpy() for k, v in out.items()}

    def sync(self, stream=None):
        N.check(N.lib().anc_sync_status(self._ctx, _stream_ptr(stream)))

    def run(self, inst, seed, trials, trace=False, fields=False, kernel="batched"):
        """Runs anneals (trial ids `trials`) of instance `inst`; returns host numpy results.

        kernel='plain' launches the one-anneal kernel once per trial; 'batched' runs them all in team launches
        ('batched_noalloc': the same through anc_run_batched_noalloc, which allocates nothing). An
        exchange time-out raises AncError (repeated once on the device-scope exchange only with fallback=True)."""
        return self._retry(lambda: self._run(inst, seed, trials, trace, fields, kernel))

    def _run(self, inst, seed, trials, trace, fields, kernel):
        trials = np.asarray(trials, np.uint32).reshape(-1)
        R = len(trials)
        out = self.alloc_out(inst, R, trace=trace, fields=fields)
        if kernel == "plain":
            for r, tr in enumerate(trials):
                sub = {k: v[r:r + 1] for k, v in out.items()}
                self.launch(inst, seed, int(tr), sub, kernel="plain")
        else:
            tt = torch.from_numpy(trials.view(np.int32)).to(self.device)
            self.launch(inst, seed, tt, out, kernel=kernel)
        self.sync()
        return {k: v.cpu().numpy() for k, v in out.items()}


class CapturedRun:
    """A multi-instance run captured as a hipGraph (Engine.capture_multi); keeps its trial-id and output tensors alive."""

    def __init__(self, eng, gid, trials, out):
        self.eng, self.gid, self.trials, self.out = eng, gid, trials, out

    def launch(self, stream=None):
        N.check(N.lib().anc_graph_launch(self.eng._ctx, self.gid, _stream_ptr(stream)))

    def destroy(self):
        if self.gid is not None and self.eng._ctx:
            N.check(N.lib().anc_graph_destroy(self.eng._ctx, self.gid))
        self.gid = None


def spins_pm(words, n):
    """Packed spin words (bit i = 1 iff s_i = +1) -> (B, n) int8 array of +-1."""
    w = np.ascontiguousarray(np.atleast_2d(words)).view(np.uint64)
    bits = np.unpackbits(w.view(np.uint8), axis=1, bitorder="little")[:, :n]
    return np.where(bits == 1, 1, -1).astype(np.int8)


def run(J, sch, seed, trials, bias=None, trace=False, kernel="batched", engine=None, n=None):
    """hwmodel.run-shaped entry point: anneals `trials` of (J, bias) under schedule `sch` on the GPU.

    Returns dict(spins (B, n) int8, flips (B,), sum_sh (B,), sum_bs (B,), nlin_total (B,), nlin_trace_last (S,) if trace,
    words (B, NP/64) packed spins)."""
    eng = engine or Engine()
    eng.set_schedule(sch)
    inst = eng.add_instance(J, bias=bias, n=n)
    n = eng._inst[inst][0]
    res = eng.run(inst, seed, trials, trace=True, kernel=kernel)
    out = dict(spins=spins_pm(res["spins"], n), flips=res["flips"], sum_sh=res["sum_sh"], sum_bs=res["sum_bs"],
               nlin_total=res["nlin"].astype(np.int64).sum(axis=1), words=res["spins"])
    if trace:
        out["nlin_trace_last"] = res["nlin"][-1].astype(np.int64)
        out["nlin"] = res["nlin"]
    return out
