"""Build the MI355X back-end in-tree.

    python -m fletcher_petrobras_amd.build

produces
    fletcher_petrobras_amd/libfletcher_amd.so   C-ABI library: driver.h + fl_hip.h +
                                                fl_model.h + fl_run.h (HIP code for gfx950)
    fletcher_petrobras_amd/bin/fletcher         reference-grammar CLI linked against it

Host C is compiled by gcc (-fopenmp for the model/precompute loops); device code by
hipcc for gfx950 with packed-FP32 instructions disabled.
"""
import os
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
INC = os.path.join(ROOT, "include")
BUILD = os.path.join(PKG, "_build")
LIB = os.path.join(PKG, "libfletcher_amd.so")
BIN = os.path.join(PKG, "bin", "fletcher")
ARCH = os.environ.get("FL_OFFLOAD_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

HOST_SRC = ["fl_model.c", "fl_synth.c", "fl_run.c"]
HIP_SRC = ["fl_driver.hip"]
HIP_FLAGS = [
    "--offload-arch=" + ARCH, "-O3", "-fPIC", "-std=c++17",
    # FMAs are written explicitly in the step kernels: no implicit contraction, so every
    # copy of an expression rounds the same way
    "-ffp-contract=off",
    # keep packed FP32 (v_pk_*_f32) out of the generated code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def _run(cmd, quiet=False):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l
                    and "not a recognized feature" not in l)
    if r.returncode != 0:
        sys.stderr.write(out + "\n")
        raise RuntimeError("build step failed: " + " ".join(cmd))
    if out.strip() and not quiet:
        sys.stderr.write(out + "\n")


def build(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    os.makedirs(os.path.dirname(BIN), exist_ok=True)
    objs = []
    for src in HOST_SRC:
        o = os.path.join(BUILD, src.replace(".c", ".o"))
        _run(["gcc", "-O2", "-fPIC", "-fopenmp", "-std=gnu11", "-Wall", "-Wno-unknown-pragmas",
              "-I" + INC, "-c", os.path.join(PKG, "host", src), "-o", o], quiet=not verbose)
        objs.append(o)
    for src in HIP_SRC:
        o = os.path.join(BUILD, src.replace(".hip", ".o"))
        _run([HIPCC] + HIP_FLAGS + ["-I" + INC, "-c", os.path.join(PKG, "csrc", src), "-o", o],
             quiet=not verbose)
        objs.append(o)
    tmp = LIB + ".tmp"
    _run([HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp] + objs + ["-lgomp"],
         quiet=not verbose)
    os.replace(tmp, LIB)
    _run(["gcc", "-O2", "-std=gnu11", "-I" + INC, os.path.join(PKG, "host", "fl_main.c"),
          "-L" + PKG, "-lfletcher_amd", "-Wl,-rpath,$ORIGIN/..", "-o", BIN], quiet=not verbose)
    # HBM copy calibration used by the roofline / PMC profiling
    _run([HIPCC] + HIP_FLAGS + [os.path.join(PKG, "csrc", "fl_stream.hip"), "-o",
                                os.path.join(PKG, "bin", "fl_stream")], quiet=not verbose)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
