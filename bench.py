"""Headline benchmark: MSamples/s of the FP32 TTI drop-in at 344^3 on one MI355X.

    python bench.py [--steps K] [--warmup W] [--gpus 1]

Workload (synthetic, seeded): interior 344^3 with a 16-cell absorbing zone and the
4-cell halo (array 384^3), dx=dy=dz=12.5 m; vpz rising 1500->4500 m/s with depth
times +-10% Gaussian blobs, smooth eps in [0,0.3], delta in [0,0.15] (<= eps), theta in
[0,pi/3], phi in [0,2pi); the reference random-velocity boundary; dt from Courant 0.4
on max vpz*sqrt(1+2 eps); Ricker source at (sx/2, sy/2, 28).

A step is one reference time step: DRIVER_InsertSource + DRIVER_Propagate (the kernel
updates all (sx-8)(sy-8)(sz-8) samples).  Fields and coefficients are resident in HBM
before the timed region.  MSamples/s = 1e-6 * samples * K / seconds (model.c:44-45,117).

roofline: algorithmic bytes of the step kernel (64 B per sample: pc, qc read, pp, qp
read and written, 10 FP32 coefficients) over its mean duration from HIP events recorded
around every launch on the library's stream during the timed region.
cpu_baseline: the reference C/OpenMP program (oracle/_ref, built from the reference
sources) timed on a few steps of the same grid on the host cores.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def log(msg):
    print("[bench] " + msg, file=sys.stderr, flush=True)


def traffic_from_profiles(tag):
    """HBM bytes per dispatch measured by rocprofv3 --pmc (FETCH_SIZE, WRITE_SIZE), scaled by
    the copy-kernel calibration; recorded in profiles/traffic.json by the profiling run."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        with open(path) as f:
            t = json.load(f)
        return t.get(tag, {}).get("bytes_per_dispatch")
    except (OSError, ValueError):
        return None


def cpu_baseline(model, shape, dt, isrc, steps, threads):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    if not O.ref_available("original"):
        return None
    R = O.Ref("original")
    R.set_threads(threads)
    R.setup(shape, 4, (12.5, 12.5, 12.5), dt, model)
    R.steps(0, 1, isrc)  # first touch
    wall = R.steps(1, steps, isrc)
    samples = (shape[0] - 8) * (shape[1] - 8) * (shape[2] - 8)
    R.L.ref_free()
    return {"value": round(1e-6 * samples * steps / wall, 2), "unit": "MSamples/s", "cores": threads,
            "kind": "reference",
            "sample": "%d steps of the reference original/ OpenMP Propagate (direct DerCross) on the "
                      "same 384^3 grid and model, Propagate+swap wall time as model.c:82-88" % steps}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--nx", type=int, default=344)
    ap.add_argument("--absorb", type=int, default=16)
    ap.add_argument("--variant", default="fused", choices=["fused", "direct", "twopass"])
    ap.add_argument("--precision", default="fp32", choices=["fp32", "bf16", "fp16", "fp16r"])
    ap.add_argument("--tiles", default="auto", help="ty,zt or 'auto' (autotuner)")
    ap.add_argument("--cpu-steps", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: the Fletcher hot path here is single-GPU "
              "(multi-GPU halo exchange is out of scope); run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import fletcher_petrobras_amd as F

    n = args.nx
    t0 = time.time()
    model = F.model_gradient((n, n, n), args.absorb, seed=344)
    shape = model[0].shape
    dt = F.cfl_dt(model, 12.5, 0.4)
    isrc = F.index(shape, shape[2] // 2, shape[1] // 2, 4 + args.absorb + 8)
    samples = (shape[0] - 8) * (shape[1] - 8) * (shape[2] - 8)
    log("model %s built in %.1f s, dt=%.6g" % (shape, time.time() - t0, dt))

    d = F.Driver(variant=args.variant, precision=args.precision, sync=False)
    t0 = time.time()
    d.initialize(shape, (12.5, 12.5, 12.5), dt, model)
    log("initialize (precompute + upload) %.1f s, device %s" % (time.time() - t0, F.device_name()))
    if args.tiles == "auto":
        ty, zt = d.autotune(3)
    else:
        ty, zt = (int(v) for v in args.tiles.split(","))
        d.set_tiles(ty, zt)
    log("tiles ty=%d zt=%d" % (ty, zt))
    bpc = d.bytes_per_cell()

    # warm-up steps it = 1..W
    for it in range(1, args.warmup + 1):
        d.insert_source(it - 1, isrc, F.source(dt, it - 1))
        d.propagate(it)
    d.synchronize()
    d.kernel_times()  # clear
    # timed region: steps W+1 .. W+K, per-launch HIP events on the library stream
    src = F.sources(dt, args.steps, args.warmup)
    d.set_timing(True)
    d.synchronize()
    d.region_begin()
    w0 = time.time()
    for k in range(args.steps):
        it = args.warmup + 1 + k
        d.insert_source(it - 1, isrc, float(src[k]))
        d.propagate(it)
    region_ms = d.region_end()
    wall = time.time() - w0
    d.set_timing(False)
    kms, kcount = d.kernel_times()
    pc = d.field("pc")
    finite = bool(np.isfinite(pc).all()) and float(np.abs(pc).max()) > 0
    d.finalize()
    sec = region_ms * 1e-3
    value = 1e-6 * samples * args.steps / sec
    avg_ms = kms / max(kcount, 1)
    achieved = bpc * samples / (avg_ms * 1e-3) / 1e9
    log("region %.1f ms (host wall %.1f ms), kernel avg %.4f ms over %d, %.0f GB/s, finite=%s"
        % (region_ms, wall * 1e3, avg_ms, kcount, achieved, finite))
    if not finite:
        print("bench: the wavefield is not finite; refusing to report", file=sys.stderr)
        sys.exit(1)

    cpu = None
    if not args.no_cpu:
        threads = min(os.cpu_count() or 1, 16)
        t0 = time.time()
        cpu = cpu_baseline(model, shape, dt, isrc, args.cpu_steps, threads)
        log("cpu baseline %s (%.1f s)" % (cpu, time.time() - t0))

    tag = "%s_%s_%d" % (args.variant, args.precision, n)
    line = {
        "metric": "MSamples/s and % HBM peak, TTI 344^3: FP32 drop-in vs mixed-prec batched",
        "value": round(value, 1),
        "unit": "MSamples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(region_ms / args.steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if args.precision == "fp32" else "f32 fields, %s coefficients" % args.precision,
        "data": "synthetic seeded heterogeneous TTI model (vpz gradient + blobs, smooth eps/delta/theta/phi, "
                "reference random boundary)",
        "config": {"workload": "TTI %d^3 interior, absorb %d, array %dx%dx%d, dx=12.5 m, dt=%.6g s, "
                               "%s variant, %s coefficients, tiles %dx64x%d"
                               % (n, args.absorb, shape[2], shape[1], shape[0], dt, args.variant,
                                  args.precision, ty, zt),
                   "samples_per_step": samples, "bytes_per_sample": bpc},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic_from_profiles(tag)},
        "pct_hbm_peak": round(100.0 * achieved / HBM_PEAK_GBS, 2),
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
