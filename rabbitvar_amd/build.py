"""Builds the native parts of rabbitvar_amd in-tree.

  rabbitvar_amd/_build/librabbitvar_amd.so   HIP kernels for gfx950 + host engine + C ABI
  rabbitvar_amd/_build/rvsynth                seeded synthetic input generator

Run:  python -m rabbitvar_amd.build      (or import and call build())
"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "_build")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

# -packed-fp32-ops: keep the compiler from emitting packed FP32 ops; -ffp-contract=off keeps the
# FP64 arithmetic that feeds printed values identical to the reference's non-contracted x86 build.
HIP_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", "--offload-arch=gfx950", "-ffp-contract=off", "-fno-fast-math",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]

LIB_SOURCES = ["rv_pileup.hip", "rv_score.hip", "rv_order.hip", "rv_slotorder.hip", "rv_realign.hip", "rv_host.cpp", "rv_engine.cpp", "rv_region.cpp",
               "rv_realign.cpp", "rv_format.cpp", "rv_order.cpp", "rv_somatic.cpp", "rv_screen.hip", "rv_mlp.cpp", "rv_bam.cpp", "rv_capi.cpp"]


def _run(cmd):
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build(verbose=True, jobs=None):
    from concurrent.futures import ThreadPoolExecutor

    os.makedirs(OUT, exist_ok=True)
    inc = ["-I", CSRC, "-I", os.path.join(ROOT, "include")]
    cmds, objs = [], []
    for src in LIB_SOURCES:
        path = os.path.join(CSRC, src)
        obj = os.path.join(OUT, src.replace(".", "_") + ".o")
        cmds.append([HIPCC, "-x", "hip"] + HIP_FLAGS + inc + ["-c", path, "-o", obj])
        objs.append(obj)
    jobs = jobs or min(len(cmds), max(1, min(8, os.cpu_count() or 1)))
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        list(ex.map(_run, cmds))
    lib = os.path.join(OUT, "librabbitvar_amd.so")
    _run([HIPCC] + HIP_FLAGS + ["-shared", "-o", lib + ".tmp"] + objs + ["-lz"])
    os.replace(lib + ".tmp", lib)
    gxx = shutil.which("g++") or "g++"
    synth = os.path.join(OUT, "rvsynth")
    _run([gxx, "-O2", "-std=c++17", "-pthread", "-o", synth + ".tmp", os.path.join(HERE, "synth", "rv_synth.cpp"), "-lz"])
    os.replace(synth + ".tmp", synth)
    return lib


if __name__ == "__main__":
    build()
    sys.exit(0)
