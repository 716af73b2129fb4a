"""Builds the HIP library libanechoic_mi355x.so in-tree (anechoic_amd/_lib) with hipcc for gfx950.

Every source compiles to its own object in parallel (a temporary directory), then hipcc links the shared library.
"""
import os
import shutil
import subprocess
import tempfile
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

HERE = Path(__file__).resolve().parent
CSRC = HERE / "csrc"
LIBDIR = HERE / "_lib"
LIB = LIBDIR / "libanechoic_mi355x.so"
SOURCES = ["anc_engine.hip", "anc_team_k32.hip", "anc_team_flip.hip", "anc_team_crumb.hip", "anc_team16.hip", "anc_team_k64.hip", "anc_team_mixed.hip", "anc_large.hip",
           "anc_large_k16.hip", "anc_large_k32.hip", "anc_large_k64.hip", "anc_large_c16.hip", "anc_large_c32.hip",
           "anc_large_c64.hip", "anc_host.hip", "anc_meanfield.hip", "anc_mf_team.hip", "anc_mf_host.hip", "anc_tables.cpp", "anc_io.cpp",
           "anc_encode.cpp"]
# -packed-fp32-ops off: packed FP32 instructions are not generated (hipcc's host pass warns that it does not know the
# feature, which is harmless). The raised pragma-unroll threshold makes `#pragma unroll` loops over the accumulator
# tiles unroll fully (a loop left rolled would index the accumulators dynamically and move them to scratch memory).
HIPFLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-Wno-unused-result", "-Wno-unused-value",
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-mllvm", "-pragma-unroll-threshold=1000000"]


def hipcc():
    for c in (os.environ.get("HIPCC"), shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if c and Path(c).exists():
            return c
    raise RuntimeError("hipcc not found")


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"{' '.join(cmd[:2])} ... failed:\n" + r.stderr[-4000:])
    return r.stderr


def build(verbose=False, jobs=None):
    LIBDIR.mkdir(exist_ok=True)
    cc = hipcc()
    jobs = jobs or max(1, min(len(SOURCES), os.cpu_count() or 1))
    with tempfile.TemporaryDirectory(prefix="anc_build_") as tmpdir:
        objs = [Path(tmpdir) / (Path(s).stem + ".o") for s in SOURCES]
        cmds = [[cc, *HIPFLAGS, "-c", str(CSRC / s), "-o", str(o)] for s, o in zip(SOURCES, objs)]
        with ThreadPoolExecutor(jobs) as ex:
            logs = list(ex.map(_run, cmds))
        tmp = LIB.with_suffix(".so.tmp")
        logs.append(_run([cc, "--offload-arch=gfx950", "-shared", "-fPIC", *map(str, objs), "-o", str(tmp)]))
    if verbose:
        for lg in logs:
            if lg:
                print(lg[-2000:])
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
