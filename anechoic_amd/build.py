"""Builds the HIP library libanechoic_mi355x.so in-tree (anechoic_amd/_lib) with hipcc for gfx950."""
import os
import shutil
import subprocess
from pathlib import Path

HERE = Path(__file__).resolve().parent
CSRC = HERE / "csrc"
LIBDIR = HERE / "_lib"
LIB = LIBDIR / "libanechoic_mi355x.so"
SOURCES = ["anc_engine.hip", "anc_host.hip", "anc_meanfield.hip", "anc_mf_host.hip", "anc_tables.cpp"]
# -packed-fp32-ops off: packed FP32 instructions are not generated (hipcc's host pass warns that it does not know the
# feature, which is harmless).
HIPFLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared", "-Wno-unused-result", "-Wno-unused-value",
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def hipcc():
    for c in (os.environ.get("HIPCC"), shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if c and Path(c).exists():
            return c
    raise RuntimeError("hipcc not found")


def build(verbose=False):
    LIBDIR.mkdir(exist_ok=True)
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [hipcc(), *HIPFLAGS, *[str(CSRC / s) for s in SOURCES], "-o", str(tmp)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed:\n" + r.stderr[-4000:])
    if verbose and r.stderr:
        print(r.stderr[-2000:])
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
