"""Build the MI355X back-end in-tree.

    python -m fletcher_petrobras_amd.build

produces
    fletcher_petrobras_amd/libfletcher_amd.so   C-ABI library: driver.h + fl_hip.h +
                                                fl_model.h + fl_run.h (HIP code for gfx950)
    fletcher_petrobras_amd/bin/fletcher         reference-grammar CLI linked against it

Host C is compiled by gcc (-fopenmp for the model/precompute loops); device code by
hipcc for gfx950 with packed-FP32 instructions disabled.
"""
import concurrent.futures
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
# (source, object, extra flags): the driver, and the step kernels of each tile height for the
# single grid and for batched shots in translation units of their own, compiled in parallel
HIP_UNITS = [("fl_driver.hip", "fl_driver.o", [])] + [
    ("fl_inst.hip", "fl_inst_%s%d.o" % (part, ty), ["-DFL_TY=%d" % ty, "-DFL_PART_%s" % part.upper()])
    for part in ("grid", "batch") for ty in (4, 8, 12, 16)]
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
    jobs = []
    for src, obj, extra in HIP_UNITS:
        o = os.path.join(BUILD, obj)
        jobs.append([HIPCC] + HIP_FLAGS + extra + ["-I" + INC, "-c", os.path.join(PKG, "csrc", src), "-o", o])
        objs.append(o)
    workers = max(1, min(len(jobs), int(os.environ.get("MAX_JOBS", os.cpu_count() or 1))))
    with concurrent.futures.ThreadPoolExecutor(workers) as ex:
        for f in [ex.submit(_run, j, not verbose) for j in jobs]:
            f.result()
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
