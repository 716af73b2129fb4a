"""ctypes binding of libanechoic_mi355x.so (the C ABI declared in include/anechoic.h)."""
import ctypes
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "_lib" / "libanechoic_mi355x.so"

ANC_OK = 0
ERRORS = {-1: "ARG", -2: "INSTANCE", -3: "RANGE", -4: "HIP", -5: "UNSUPPORTED", -6: "STATE", -7: "IO"}
FORMATS = {"jint8": 0, "snowgpu1": 1, "graph": 1, "gset": 2}
SHAPES = {"geom": 0, "geometric": 0, "linear": 1, "explicit": 2}

# Every entry point the header declares, with its ctypes signature.
_c_i32p = ctypes.POINTER(ctypes.c_int32)
_c_i64p = ctypes.POINTER(ctypes.c_int64)
_vp = ctypes.c_void_p


class Sched(ctypes.Structure):
    _fields_ = [("S", ctypes.c_int32), ("shape", ctypes.c_int32), ("t0", ctypes.c_double), ("t1", ctypes.c_double),
                ("q", ctypes.c_double), ("lam", ctypes.c_double), ("jv", ctypes.c_double), ("kappa", ctypes.c_double),
                ("scale", ctypes.c_double), ("ramp", ctypes.c_int32), ("T", ctypes.POINTER(ctypes.c_double))]


class JStats(ctypes.Structure):
    _fields_ = [("sumw", ctypes.c_int64), ("maxabs", ctypes.c_int32), ("hmax", ctypes.c_int64), ("mean_row_sq", ctypes.c_double),
                ("nnz", ctypes.c_int64)]


class Out(ctypes.Structure):
    _fields_ = [("spins", _vp), ("spin_stride", ctypes.c_int64), ("sum_sh", _vp), ("sum_bs", _vp), ("flips", _vp),
                ("nlin", _vp), ("fields", _vp), ("field_stride", ctypes.c_int64)]


SIGNATURES = {
    "anc_make_tables": (ctypes.c_int, [ctypes.POINTER(Sched), _c_i32p, _c_i32p, _c_i64p, _c_i32p]),
    "anc_np": (ctypes.c_int, [ctypes.c_int]),
    "anc_last_error": (ctypes.c_char_p, []),
    "anc_version": (ctypes.c_int, []),
    "anc_create": (_vp, [ctypes.c_int]),
    "anc_destroy": (None, [_vp]),
    "anc_set_tables": (ctypes.c_int, [_vp, ctypes.c_int, _c_i32p, _c_i32p, _c_i64p, _c_i32p]),
    "anc_add_instance": (ctypes.c_int, [_vp, ctypes.c_int, _vp, ctypes.c_int64, _vp, _vp]),
    "anc_clear_instances": (ctypes.c_int, [_vp]),
    "anc_instance_bound": (ctypes.c_int64, [_vp, ctypes.c_int]),
    "anc_run_plain": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_uint64, ctypes.c_uint32, ctypes.POINTER(Out), _vp]),
    "anc_run_batched": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_uint64, _vp, ctypes.c_int, ctypes.POINTER(Out), _vp]),
    "anc_run_batched_noalloc": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_uint64, _vp, ctypes.c_int, ctypes.POINTER(Out), _vp]),
    "anc_run_multi": (ctypes.c_int, [_vp, ctypes.c_int, _c_i32p, _c_i32p, ctypes.c_uint64, _vp, ctypes.POINTER(Out), _vp]),
    "anc_gen_ising": (ctypes.c_int, [ctypes.c_uint64, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_int, _vp, ctypes.c_int64, _vp, _vp]),
    "anc_graph_multi": (ctypes.c_int, [_vp, ctypes.c_int, _c_i32p, _c_i32p, ctypes.c_uint64, _vp, ctypes.POINTER(Out)]),
    "anc_graph_launch": (ctypes.c_int, [_vp, ctypes.c_int, _vp]),
    "anc_graph_destroy": (ctypes.c_int, [_vp, ctypes.c_int]),
    "anc_sync_status": (ctypes.c_int, [_vp, _vp]),
    "anc_set_option": (ctypes.c_int, [_vp, ctypes.c_char_p, ctypes.c_double]),
    "anc_get_option": (ctypes.c_double, [_vp, ctypes.c_char_p]),
    "anc_get_profile": (ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_longlong), ctypes.c_int]),
    "anc_mf_workspace": (ctypes.c_size_t, [ctypes.c_int, ctypes.c_int, ctypes.c_int]),
    "anc_mf_forward": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, _vp, ctypes.c_int, ctypes.c_int64, ctypes.c_int64, ctypes.c_float,
                                      ctypes.c_int, _vp, _vp, ctypes.c_int, _vp, ctypes.c_int, _vp, _vp, _vp, _vp]),
    "anc_mf_backward": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, _vp, ctypes.c_int, ctypes.c_int64, ctypes.c_int64, ctypes.c_float,
                                       ctypes.c_int, _vp, _vp, ctypes.c_int, ctypes.c_int, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),  # dT, dcoef: double*
    "anc_mf_set_option": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_double]),
    "anc_mf_get_option": (ctypes.c_double, [ctypes.c_char_p]),
    "anc_mf_status": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, _vp]),
    "anc_mf_gemv": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, _vp, ctypes.c_int, ctypes.c_int64, ctypes.c_int64, _vp, _vp, _vp]),
    "anc_mf_pack_j4": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, _vp, ctypes.c_int64, ctypes.c_int64, _vp, ctypes.c_int64,
                                      ctypes.c_int64, _vp, _vp]),
    "anc_mf_pack_j2": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, _vp, ctypes.c_int64, ctypes.c_int64, _vp, ctypes.c_int64,
                                      ctypes.c_int64, _vp, _vp]),
    "anc_read_instance": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_int, _c_i32p, _vp, ctypes.c_int64]),
    "anc_write_instance": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_int, ctypes.c_int32, _vp, ctypes.c_int64]),
    "anc_read_bias": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_int32, _vp]),
    "anc_write_bias": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_int32, _vp]),
    "anc_instance_stats": (ctypes.c_int, [ctypes.c_int32, _vp, ctypes.c_int64, ctypes.POINTER(JStats)]),
    "anc_gen_gp": (ctypes.c_int, [ctypes.c_uint64, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, ctypes.c_int64, ctypes.c_int64,
                                  _vp, ctypes.c_int64, _vp, _vp, _c_i64p]),
    "anc_gen_tsp": (ctypes.c_int, [ctypes.c_uint64, ctypes.c_int32, ctypes.c_int32, ctypes.c_int64, ctypes.c_int64, _vp,
                                   ctypes.c_int64, _vp, _vp, _c_i64p]),
    "anc_score": (ctypes.c_int, [ctypes.c_int32, _vp, ctypes.c_int64, _vp, ctypes.c_int32, _vp, ctypes.c_int64, _vp, ctypes.c_int64,
                                 _vp, _vp, _vp, ctypes.c_int]),
}

_lib = None


class AncError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"anechoic error {code} ({ERRORS.get(code, '?')}): {msg}")
        self.code = code


def lib():
    """Loads the HIP library; raises if it was not built (there is no fallback path)."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(f"{LIB_PATH} is missing: build it with anechoic_amd.build.build() (hipcc, gfx950)")
        L = ctypes.CDLL(str(LIB_PATH))
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(rc):
    if rc < 0:
        raise AncError(rc, lib().anc_last_error().decode(errors="replace"))
    return rc
