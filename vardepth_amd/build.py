"""Compile the HIP engine in-tree for gfx950.

    python -m vardepth_amd.build        # -> vardepth_amd/libvdhmm.so, libvdhmm_core.so

hipcc cross-compiles without a GPU.  Every HIP source is built with packed
FP32 instructions disabled (-target-feature -packed-fp32-ops).
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
LIB = PKG / "libvdhmm.so"
CORE_LIB = PKG / "libvdhmm_core.so"
ARCH = os.environ.get("VD_OFFLOAD_ARCH", "gfx950")

HIP_FLAGS = ["-O3", "-std=c++17", "-fPIC", "--offload-arch=" + ARCH,
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-I" + str(ROOT / "include"), "-I" + str(CSRC)]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found")


def _run(cmd):
    proc = subprocess.run(cmd, capture_output=True, text=True)
    noise = "'-packed-fp32-ops' is not a recognized feature"
    err = "\n".join(l for l in proc.stderr.splitlines() if noise not in l)
    if proc.returncode != 0:
        raise RuntimeError("build failed: %s\n%s" % (" ".join(cmd), err))
    if err.strip():
        print(err, file=sys.stderr)


def build(verbose: bool = False) -> Path:
    """Compile libvdhmm.so (kernels + engine + C ABI) and libvdhmm_core.so."""
    hc = hipcc()
    objs = []
    tmp = PKG / "_build"
    tmp.mkdir(exist_ok=True)
    for src, lang in ((CSRC / "vdhmm_kernels.hip", "hip"), (CSRC / "vdhmm_grad.hip", "hip"),
                      (CSRC / "vdhmm_plain.hip", "hip"), (CSRC / "vdhmm_genotype.hip", "hip"),
                      (CSRC / "vdhmm_noalloc.hip", "hip"), (CSRC / "vdhmm_refdesign.hip", "hip"),
                      (CSRC / "vdhmm_engine.cpp", "c++")):
        obj = tmp / (src.stem + ".o")
        cmd = [hc] + HIP_FLAGS + (["-x", "hip"] if lang == "hip" else []) + ["-c", str(src), "-o", str(obj)]
        if verbose:
            print(" ".join(cmd))
        _run(cmd)
        objs.append(str(obj))
    out_tmp = LIB.with_suffix(".so.tmp")
    _run([hc, "-shared", "-fPIC", "--offload-arch=" + ARCH] + objs + ["-o", str(out_tmp)])
    out_tmp.replace(LIB)
    core_tmp = CORE_LIB.with_suffix(".so.tmp")
    _run([hc, "-O2", "-std=c++17", "-fPIC", "-shared", "-I" + str(ROOT / "include"),
          str(CSRC / "vdhmm_core_alias.cpp"), "-L" + str(PKG), "-lvdhmm",
          "-Wl,-rpath,$ORIGIN", "-o", str(core_tmp)])
    core_tmp.replace(CORE_LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
