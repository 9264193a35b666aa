"""Headline benchmark: MSamples/s of the FP32 TTI drop-in at 344^3 on one MI355X, with the
mixed-precision and batched forms measured beside it.

    python bench.py [--steps K] [--warmup W] [--gpus 1] [--quick]

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
cpu_baseline: the reference C/OpenMP back-ends original, der1der1, der1der1hm and
der1der1lm (oracle/_ref, built from the reference sources) timed on a few steps of the same
grid on the host cores; the headline object is the fastest, every variant is listed with
the CPU model and core counts.

extras (skipped with --quick), each timed over its whole run with inputs resident; runs in
graphs replay a graph captured by an untimed first run:
  twopass   the der1der1 two-pass variant on the same grid (104 algorithmic B/sample)
  fp32_graph / mixed   FP32, the FP32 in-kernel recompute (fp32r: FP32 arithmetic, 24-byte
            records), BF16 / FP16 coefficients, the FP16 in-kernel recompute and its
            block floating-point variant (q16r) on the same grid, receivers on plane z=28
            every 4 steps, 50-step graph chunks
  rel_l2_vs_fp64   relative L2 (in FP64) of the receiver traces and the final pc against
            the device FP64 reference (fl_ref64_*: the der1der1 loop in double, pinned
            bitwise to the FP64 oracle by the tests) on the same grid and steps
  batched   32 shots on uneven (128..256)^2 x 320 apertures cut from one seeded
            512x512x320 model (absorb 16), 800 steps in 50-step graphs, autotuned
            tiles, FP32, BF16, FP16-recompute and block floating-point-recompute coefficients;
            receivers every 4 steps, their relative L2 against the FP64 device reference
            for the first two shots
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
D = (12.5, 12.5, 12.5)


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


def host_cpu():
    """CPU model name, logical CPUs of the host and CPUs this process may run on."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        allowed = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        allowed = os.cpu_count() or 1
    return model, os.cpu_count() or 1, allowed


CPU_VARIANTS = (("original", "direct 64-point DerCross"),
                ("der1der1", "first-of-first, four first-derivative volumes"),
                ("der1der1hm", "first-of-first, derivative planes formed in z order"),
                ("der1der1lm", "first-of-first, 9-plane derivative ring"))


def cpu_baseline(model, shape, dt, isrc, steps, threads):
    """The reference C/OpenMP back-ends (oracle/_ref, built from the reference sources) timed
    on `steps` steps of the same grid and model: Propagate + swap wall time, as
    model.c:82-88.  The headline object is the fastest of them; every variant is listed."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    samples = (shape[0] - 8) * (shape[1] - 8) * (shape[2] - 8)
    cpu_model, nproc, allowed = host_cpu()
    rows = {}
    for v, what in CPU_VARIANTS:
        if not O.ref_available(v):
            continue
        R = O.Ref(v)
        R.set_threads(threads)
        R.setup(shape, 4, D, dt, model)
        R.steps(0, 1, isrc)  # first touch of the arrays
        wall = R.steps(1, steps, isrc)
        R.L.ref_free()
        rows[v] = {"value": round(1e-6 * samples * steps / wall, 2), "seconds": round(wall, 2), "form": what}
        log("cpu baseline %s: %.1f MS/s (%d steps, %.1f s, %d threads)" % (v, rows[v]["value"], steps, wall, threads))
    if not rows:
        return None
    best = max(rows, key=lambda k: rows[k]["value"])
    return {"value": rows[best]["value"], "unit": "MSamples/s", "cores": threads, "kind": "reference",
            "variant": best,
            "sample": "%d steps of each reference OpenMP back-end (%s) on the same %dx%dx%d grid and model, "
                      "Propagate+swap wall time as model.c:82-88; value = the fastest (%s)"
                      % (steps, ", ".join(rows), shape[2], shape[1], shape[0], best),
            "variants": rows, "cpu_model": cpu_model, "nproc": nproc, "cpus_allowed": allowed}


def tiles_for(d, spec, tune_steps=10):
    """(ty, zt, stream hints): the autotuner's choice for 'auto' (tune_steps steps a
    candidate), else 'ty,zt[,hints]'."""
    if spec == "auto":
        ty, zt = d.autotune(tune_steps)
        return ty, zt, int(d.stream_hints())
    v = [int(x) for x in spec.split(",")]
    d.set_tiles(v[0], v[1])
    d.set_stream_hints(v[2] if len(v) > 2 else 0)
    ty, zt = d.tiles()  # zt 0 resolved to the grid's default chunk depth
    return ty, zt, int(d.stream_hints())


def dropin(F, np, model, shape, dt, isrc, args, d3=D):
    """The reference loop through the drop-in entry points, per-launch HIP events."""
    samples = (shape[0] - 8) * (shape[1] - 8) * (shape[2] - 8)
    d = F.Driver(variant=args.variant, precision=args.precision, sync=False)
    t0 = time.time()
    d.initialize(shape, d3, dt, model)
    log("initialize (precompute + upload) %.1f s, device %s" % (time.time() - t0, F.device_name()))
    ty, zt, hints = tiles_for(d, args.tiles)
    log("tiles ty=%d zt=%d stream hints %d" % (ty, zt, hints))
    bpc = d.bytes_per_cell()
    for it in range(1, args.warmup + 1):
        d.insert_source(it - 1, isrc, F.source(dt, it - 1))
        d.propagate(it)
    d.synchronize()
    d.kernel_times()  # clear
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
    avg_ms = kms / max(kcount, 1)
    r = {"value": 1e-6 * samples * args.steps / (region_ms * 1e-3), "region_ms": region_ms,
         "avg_kernel_ms": avg_ms, "launches": kcount, "bpc": bpc, "tiles": (ty, zt), "hints": hints,
         "achieved": bpc * samples / (avg_ms * 1e-3) / 1e9, "finite": finite, "pc": pc}
    log("drop-in %s/%s: region %.1f ms (host wall %.1f ms), kernel avg %.4f ms over %d, %.0f GB/s, "
        "%.0f MS/s, finite=%s" % (args.variant, args.precision, region_ms, wall * 1e3, avg_ms, kcount,
                                  r["achieved"], r["value"], finite))
    return r


def rel_l2(np, a, b):
    """Relative L2 of a against b, accumulated in FP64."""
    a = np.asarray(a, np.float64).ravel()
    b = np.asarray(b, np.float64).ravel()
    return float(np.sqrt(np.dot(a - b, a - b) / np.dot(b, b)))


def fp64_reference(F, np, model, shape, dt, isrc, steps, extra, rec_iz, d3=D):
    """The device FP64 reference (der1der1 in double, exact weights): traces every 4 steps
    and pc after `steps` steps, then pc after `steps + extra`."""
    t0 = time.time()
    r = F.Ref64(shape, d3, dt, model)
    try:
        src = F.sources(dt, steps + extra)
        r.run(steps, isrc, src=src[:steps], rec_iz=rec_iz, rec_every=4, max_traces=(steps + 3) // 4)
        out = {"traces": r.traces(), "pc_steps": r.field("pc")}
        if extra:
            r.run(extra, isrc, it0=steps, src=src[steps:])
        out["pc_total"] = r.field("pc")
    finally:
        r.destroy()
    log("FP64 reference: %d + %d steps in %.1f s" % (steps, extra, time.time() - t0))
    return out


def pin_fp64_reference(F, np, model, shape, dt, isrc, steps, rec_iz, d3=D):
    """The device FP64 reference against the FP64 oracle on the host over the first `steps`
    steps of the same grid and model: pc and traces equal bit for bit."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    t0 = time.time()
    src = F.sources(dt, steps)
    ro = O.run("der1der1", shape, 4, d3, dt, steps, isrc, O.precompute(model, "f64"), fp="f64", rec_iz=rec_iz,
               rec_every=2, src=src)
    r = F.Ref64(shape, d3, dt, model)
    try:
        r.run(steps, isrc, src=src, rec_iz=rec_iz, rec_every=2, max_traces=steps // 2)
        same = bool(np.array_equal(r.field("pc"), ro["pc"]) and np.array_equal(r.traces(), ro["traces"]))
    finally:
        r.destroy()
    log("FP64 oracle on %s, %d steps: %.1f s; device FP64 reference equal bit for bit: %s"
        % (shape, steps, time.time() - t0, same))
    return same


def graph_run(F, np, model, shape, dt, isrc, variant, precision, steps, rec_iz, tiles, ref=None, d3=D):
    """steps time steps from rest in 50-step graph chunks, receivers every 4 steps; the timed
    run replays the graph that a first, untimed run of up to 50 steps captured and
    instantiated (the fields are zeroed in between), so the region holds only the run."""
    samples = (shape[0] - 8) * (shape[1] - 8) * (shape[2] - 8)
    d = F.Driver(variant=variant, precision=precision, sync=False)
    d.initialize(shape, d3, dt, model)
    ty, zt, hints = tiles_for(d, tiles)
    bpc = d.bytes_per_cell()
    d.record_traces(rec_iz, 4, (steps + 3) // 4)
    src = F.sources(dt, steps)
    d.run(min(steps, 50), isrc, graph_chunk=50, src=src[:min(steps, 50)])  # warm-up, graph capture
    d.reset_fields()
    d.synchronize()
    d.region_begin()
    d.run(steps, isrc, graph_chunk=50, src=src)
    region_ms = d.region_end()
    tr = d.traces()
    pc = d.field("pc")
    d.finalize()
    finite = bool(np.isfinite(pc).all() and np.isfinite(tr).all()) and float(np.abs(tr).max()) > 0
    value = 1e-6 * samples * steps / (region_ms * 1e-3)
    r = {"value": round(value, 1), "ms_per_step": round(region_ms / steps, 4),
         "bytes_per_sample": bpc, "achieved_GBs": round(bpc * value * 1e-3, 1),
         "frac": round(bpc * value * 1e-3 / HBM_PEAK_GBS, 4), "tiles": [ty, zt], "stream_hints": hints, "finite": finite}
    if ref is not None:
        r["rel_l2_vs_fp64"] = {"traces": round(rel_l2(np, tr, ref["traces"]), 9),
                               "pc": round(rel_l2(np, pc, ref["pc_steps"]), 9)}
    log("%s/%s graph run: %s" % (variant, precision, r))
    return r


def batched(F, np, precisions, steps, nshots, tiles=None, persistent=0):
    """Batched shots on apertures cut from one 512x512x320 parent model, records computed on
    the device; tiles (ty, zt), or None for the autotuner's choice per precision."""
    from fletcher_petrobras_amd import shots as S
    absorb = 16
    t0 = time.time()
    parent = F.model_gradient((512, 512, 320), absorb, seed=512)
    interior = S.parent_interior(parent, absorb)
    sh = S.make_shots(interior, nshots, absorb, 128, 256, src_depth=4 + absorb + 8,
                      rec_depth=4 + absorb + 8, seed=32)
    del parent, interior
    dt = min(F.cfl_dt(s["model"], 12.5, 0.4) for s in sh)
    samples = sum((s["shape"][0] - 8) * (s["shape"][1] - 8) * (s["shape"][2] - 8) for s in sh)
    log("batched: %d shots, %.3g samples/step, dt=%.6g, setup %.1f s"
        % (len(sh), samples, dt, time.time() - t0))
    out = {"shots": len(sh), "samples_per_step": samples, "steps": steps, "graph_chunk": 50,
           "dt": dt}
    src = F.sources(dt, steps)
    # FP64 device reference traces of the first two shots (their own sub-grids, same
    # source schedule): the batched forms' receiver error
    nref = min(2, len(sh))
    ref_tr = []
    for k in range(nref):
        r = F.Ref64(sh[k]["shape"], D, dt, sh[k]["model"])
        try:
            r.run(steps, sh[k]["isource"], src=src, rec_iz=sh[k]["rec_iz"], rec_every=4,
                  max_traces=(steps + 3) // 4)
            ref_tr.append(r.traces())
        finally:
            r.destroy()
    for prec in precisions:
        t0 = time.time()
        F.lib().fl_set_device_precompute(1)  # the shots' records computed on the device
        try:
            b = F.Batch(sh, D, dt, precision=prec)
        finally:
            F.lib().fl_set_device_precompute(0)
        t_up = time.time() - t0
        try:
            b.set_persistent(persistent)
            if tiles is None:
                b.autotune(6)
            else:
                b.set_tiles(*tiles)
            tl = b.tiles()
            t_tune = time.time() - t0 - t_up
            # untimed first run of up to 50 steps: work list upload, graph capture
            b.run(min(steps, 50), graph_chunk=50, rec_every=4, max_traces=(steps + 3) // 4,
                  src=src[:min(steps, 50)])
            b.reset()
            b.run(steps, graph_chunk=50, rec_every=4, max_traces=(steps + 3) // 4, src=src)
            ms = b.last_ms()
            tr = b.traces(0)
            finite = bool(np.isfinite(tr).all()) and float(np.abs(tr).max()) > 0
            err = [rel_l2(np, b.traces(k), ref_tr[k]) for k in range(nref)]
        finally:
            b.destroy()
        value = 1e-6 * samples * steps / (ms * 1e-3)
        out[prec] = {"value": round(value, 1), "ms_per_step": round(ms / steps, 4), "tiles": list(tl),
                     "finite": finite,
                     "traces_rel_l2_vs_fp64": {"shot%d" % k: round(e, 9) for k, e in enumerate(err)}}
        log("batched %s: %.0f MS/s (%.1f ms for %d steps; %.1f s in all: upload %.1f s, tuning %.1f s), "
            "traces vs FP64 %s" % (prec, value, ms, steps, time.time() - t0, t_up, t_tune, ["%.2e" % e for e in err]))
    return out


def window_model(F, np, stream, corner, w, absorb):
    """A 256^3 (w^3) window of the streamed model with its own absorbing zone: the model's
    cells around the window interior (which lies inside the model's interior), then the
    reference random boundary of the window grid (seed 1) over its absorbing zone."""
    bord = 4
    side = w + 2 * (bord + absorb)
    z0, y0, x0 = (c - bord - absorb for c in corner)
    sz, sy, sx = stream.shape
    out = [np.empty((side, side, side), np.float32) for _ in range(6)]
    slab = 16
    for zs in range(0, z0 + side, slab):  # the stream is closed after this partial pass
        n = min(slab, sz - zs)
        planes = stream.fill(zs, n)
        lo, hi = max(zs, z0), min(zs + n, z0 + side)
        if lo < hi:
            for k in range(6):
                out[k][lo - z0:hi - z0] = planes[k][lo - zs:hi - zs, y0:y0 + side, x0:x0 + side]
    F.random_boundary(out, (w, w, w), absorb)
    return out


def large(F, np, args, tune_steps=3):
    """The 1024x1024x768 layered model (array 1096x1096x840, absorb 32, dx=dy=dz=10 m, Courant
    0.4): produced in z slabs and initialised on the device (fl_initialize_streamed), so no
    whole model arrays exist on the host.  FP32 drop-in through the per-step API, and 50-step
    graph runs of FP32 recompute, FP16 recompute and block floating point; then the error of
    each against FP64 on a 256^3 sub-window of the model with its own absorbing zone: the
    device FP64 reference over all steps, and the FP64 oracle on the host over the first 40
    steps (which the device reference must equal bit for bit).  Returns the result line."""
    d3 = (10.0, 10.0, 10.0)
    absorb = 32
    t0 = time.time()
    stream = F.ModelStream("layered", (1024, 1024, 768), absorb, 1024)
    shape = stream.shape
    dt = stream.cfl_dt(10.0, 0.4)
    rec_iz = 4 + absorb + 8
    isrc = F.index(shape, shape[2] // 2, shape[1] // 2, rec_iz)
    samples = (shape[0] - 8) * (shape[1] - 8) * (shape[2] - 8)
    steps = args.large_steps
    log("large model %s streamed (two passes) in %.1f s, dt=%.6g" % (shape, time.time() - t0, dt))

    # FP32 drop-in, per-step API
    d = F.Driver(variant="fused", precision="fp32", sync=False)
    t0 = time.time()
    d.initialize_streamed(stream, d3, dt)
    log("streamed initialisation %.1f s" % (time.time() - t0))
    ty, zt, hints = tiles_for(d, args.tiles, tune_steps)
    bpc = d.bytes_per_cell()
    src = F.sources(dt, max(steps, 5))
    for it in range(1, 6):
        d.insert_source(it - 1, isrc, float(src[it - 1]))
        d.propagate(it)
    d.reset_fields()
    d.synchronize()
    d.kernel_times()
    d.set_timing(True)
    d.region_begin()
    for it in range(1, steps + 1):
        d.insert_source(it - 1, isrc, float(src[it - 1]))
        d.propagate(it)
    region_ms = d.region_end()
    d.set_timing(False)
    kms, kcount = d.kernel_times()
    pc = d.field("pc")
    finite = bool(np.isfinite(pc).all()) and float(np.abs(pc).max()) > 0
    del pc
    d.finalize()
    avg = kms / max(kcount, 1)
    value = 1e-6 * samples * steps / (region_ms * 1e-3)
    achieved = bpc * samples / (avg * 1e-3) / 1e9
    log("large drop-in: %.0f MS/s, kernel %.3f ms, %.0f GB/s, finite=%s" % (value, avg, achieved, finite))
    if args.large_dropin_only:  # for profiling the FP32 step on this grid
        stream.close()
        return {"metric": "large drop-in only", "value": round(value, 1), "kernel_ms": round(avg, 4),
                "achieved_GBs": round(achieved, 1), "tiles": [ty, zt], "stream_hints": hints, "finite": finite}
    graphs = {}
    for prec in ("fp32r", "fp16r", "q16r"):
        g = F.Driver(variant="fused", precision=prec, sync=False)
        g.initialize_streamed(stream, d3, dt)
        gty, gzt, gh = tiles_for(g, args.tiles, tune_steps)
        g.run(min(steps, 50), isrc, graph_chunk=50, src=src[:min(steps, 50)])
        g.reset_fields()
        g.synchronize()
        g.region_begin()
        g.run(steps, isrc, graph_chunk=50, src=src)
        ms = g.region_end()
        gb = g.bytes_per_cell()
        g.finalize()
        v = 1e-6 * samples * steps / (ms * 1e-3)
        graphs[prec] = {"value": round(v, 1), "ms_per_step": round(ms / steps, 4), "bytes_per_sample": gb,
                        "frac": round(gb * v * 1e-3 / HBM_PEAK_GBS, 4), "tiles": [gty, gzt], "stream_hints": gh,
                        "vs_fp32_dropin": round(v / value, 3)}
        log("large %s graph run: %s" % (prec, graphs[prec]))

    # FP64 check on a 256^3 window around the source column, below the absorbing zone
    w = 256
    corner = (4 + absorb + 8, shape[1] // 2 - w // 2, shape[2] // 2 - w // 2)
    t0 = time.time()
    wm = window_model(F, np, stream, corner, w, absorb)
    stream.close()
    wshape = wm[0].shape
    wsrc = F.index(wshape, wshape[2] // 2, wshape[1] // 2, rec_iz)
    log("256^3 window model %s built in %.1f s" % (wshape, time.time() - t0))
    ref = fp64_reference(F, np, wm, wshape, dt, wsrc, steps, 0, rec_iz, d3)
    nchk = min(40, steps)
    ref_pin = pin_fp64_reference(F, np, wm, wshape, dt, wsrc, nchk, rec_iz, d3)
    window = {"window": "256^3 interior at (z, y, x) = %s of the model, absorb 32 of its own, array %s, %d steps"
                        % (corner, "x".join(map(str, wshape[::-1])), steps),
              "fp64_reference_equals_oracle_%d_steps" % nchk: ref_pin, "rel_l2_vs_fp64": {}}
    for prec in ("fp32", "fp32r", "fp16r", "q16r"):
        r = graph_run(F, np, wm, wshape, dt, wsrc, "fused", prec, steps, rec_iz, args.tiles, ref, d3)
        window["rel_l2_vs_fp64"][prec] = r["rel_l2_vs_fp64"]
    line = {
        "metric": "MSamples/s and % HBM peak, TTI 1024x1024x768: FP32 drop-in vs FP16 recompute",
        "value": round(value, 1), "unit": "MSamples/s", "n_gpus": 1, "steps": steps, "warmup": 5,
        "ms_per_step": round(region_ms / steps, 4), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32",
        "data": "synthetic seeded layered TTI model (12 dipping layers, smooth theta/phi, reference random boundary), "
                "streamed to the device in z slabs",
        "config": {"workload": "TTI 1024x1024x768 interior, absorb 32, array %dx%dx%d, dx=10 m, dt=%.6g s, "
                               "fused variant, fp32 coefficients, tiles 32x%dx%d, stream hints %s"
                               % (shape[2], shape[1], shape[0], dt, ty, zt, "on" if hints else "off"),
                   "samples_per_step": samples, "bytes_per_sample": bpc},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None},
        "finite": finite,
        "graph_runs": graphs,
        "fp64_window": window,
    }
    return line


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--nx", type=int, default=344)
    ap.add_argument("--absorb", type=int, default=16)
    ap.add_argument("--variant", default="fused", choices=["fused", "direct", "twopass"])
    ap.add_argument("--precision", default="fp32", choices=["fp32", "bf16", "fp16", "fp16r", "fp32r", "q16r"])
    ap.add_argument("--tiles", default="auto", help="ty,zt[,stream hints 0/1] or 'auto' (autotuner)")
    ap.add_argument("--cpu-steps", type=int, default=30)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--quick", action="store_true", help="headline line only, no extras")
    ap.add_argument("--shots", type=int, default=32)
    ap.add_argument("--batch-steps", type=int, default=800)
    ap.add_argument("--large", action="store_true",
                    help="only the 1024x1024x768 layered model (FP32 + FP16 recompute vs FP64)")
    ap.add_argument("--large-steps", type=int, default=500)
    ap.add_argument("--no-large", action="store_true", help="skip the 1024x1024x768 section of the full run")
    ap.add_argument("--large-dropin-only", action="store_true",
                    help="with --large: only the FP32 drop-in on the 1024x1024x768 grid (for profiling)")
    ap.add_argument("--batched", default="",
                    help="only the batched shots, in these precisions (comma list; for profiling)")
    ap.add_argument("--batch-tiles", default="", help="ty,zt of the batched runs (default: autotuned)")
    ap.add_argument("--batch-persistent", type=int, default=0,
                    help="launch form of the batched step: 0 one workgroup per item, 1 persistent (resident "
                         "workgroups), >1 persistent over at most that many workgroups")
    ap.add_argument("--lane-pairs", type=int, default=-1, choices=[-1, 0, 1],
                    help="lane layout of the fp16r / q16r step: 1 both fields per lane, 0 one field per lane, "
                         "-1 the library's default")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: the Fletcher hot path here is single-GPU "
              "(multi-GPU halo exchange is out of scope); run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import fletcher_petrobras_amd as F

    if args.lane_pairs >= 0:
        F._check(F.lib().fl_set_lane_pairs(args.lane_pairs))
    if args.large:
        print(json.dumps(large(F, np, args)), flush=True)
        return
    if args.batched:
        precs = tuple(args.batched.split(","))
        bt = tuple(int(v) for v in args.batch_tiles.split(",")) if args.batch_tiles else None
        print(json.dumps({"batched": batched(F, np, precs, args.batch_steps, args.shots, bt, args.batch_persistent)}),
              flush=True)
        return
    n = args.nx
    t0 = time.time()
    model = F.model_gradient((n, n, n), args.absorb, seed=344)
    shape = model[0].shape
    dt = F.cfl_dt(model, 12.5, 0.4)
    rec_iz = 4 + args.absorb + 8
    isrc = F.index(shape, shape[2] // 2, shape[1] // 2, rec_iz)
    samples = (shape[0] - 8) * (shape[1] - 8) * (shape[2] - 8)
    log("model %s built in %.1f s, dt=%.6g" % (shape, time.time() - t0, dt))

    main_r = dropin(F, np, model, shape, dt, isrc, args)
    if not main_r["finite"]:
        print("bench: the wavefield is not finite; refusing to report", file=sys.stderr)
        sys.exit(1)
    value = main_r["value"]

    pc_main = main_r.pop("pc")
    extras = None
    if not args.quick:
        ref = fp64_reference(F, np, model, shape, dt, isrc, args.steps, args.warmup, rec_iz)
        main_err = rel_l2(np, pc_main, ref["pc_total"])
        del pc_main
        extras = {"fp32_dropin_rel_l2_vs_fp64": {"pc": round(main_err, 9)},
                  "fp64_reference_equals_oracle_10_steps": pin_fp64_reference(F, np, model, shape, dt, isrc, 10,
                                                                               rec_iz),
                  "twopass": graph_run(F, np, model, shape, dt, isrc, "twopass", "fp32", args.steps,
                                       rec_iz, "auto", ref),
                  "fp32_graph": graph_run(F, np, model, shape, dt, isrc, "fused", "fp32", args.steps,
                                          rec_iz, "auto", ref),
                  "mixed": {}}
        for prec in ("fp32r", "bf16", "fp16", "fp16r", "q16r"):
            r = graph_run(F, np, model, shape, dt, isrc, "fused", prec, args.steps, rec_iz, "auto", ref)
            r["vs_fp32_dropin"] = round(r["value"] / value, 3)
            extras["mixed"][prec] = r
        del ref
        bt = batched(F, np, ("fp32", "bf16", "fp16r", "q16r"), args.batch_steps, args.shots)
        for prec in ("fp32", "bf16", "fp16r", "q16r"):
            bt[prec]["vs_fp32_dropin"] = round(bt[prec]["value"] / value, 3)
        extras["batched"] = bt
        if not args.no_large:
            t0 = time.time()
            lg = large(F, np, args)
            extras["large_1024x1024x768"] = {k: lg[k] for k in ("metric", "value", "unit", "steps", "ms_per_step",
                                                                  "config", "roofline", "finite", "graph_runs",
                                                                  "fp64_window")}
            log("large model section %.1f s" % (time.time() - t0))

    cpu = None
    if not args.no_cpu:
        # every CPU this process may run on, or the OpenMP thread count the host assigns to
        # the job when it sets one
        threads = int(os.environ.get("OMP_NUM_THREADS") or 0) or host_cpu()[2]
        t0 = time.time()
        cpu = cpu_baseline(model, shape, dt, isrc, args.cpu_steps, threads)
        log("cpu baseline %s (%.1f s)" % (cpu, time.time() - t0))

    ty, zt = main_r["tiles"]
    bpc = main_r["bpc"]
    achieved = main_r["achieved"]
    tag = "%s_%s_%d" % (args.variant, args.precision, n)
    line = {
        "metric": "MSamples/s and % HBM peak, TTI 344^3: FP32 drop-in vs mixed-prec batched",
        "value": round(value, 1),
        "unit": "MSamples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(main_r["region_ms"] / args.steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if args.precision in ("fp32", "fp32r") else "f32 fields, %s coefficients" % args.precision,
        "data": "synthetic seeded heterogeneous TTI model (vpz gradient + blobs, smooth eps/delta/theta/phi, "
                "reference random boundary)",
        "config": {"workload": "TTI %d^3 interior, absorb %d, array %dx%dx%d, dx=12.5 m, dt=%.6g s, "
                               "%s variant, %s coefficients, %s"
                               % (n, args.absorb, shape[2], shape[1], shape[0], dt, args.variant, args.precision,
                                  "64x4-cell tiles, 10-plane LDS ring" if args.variant == "direct" else
                                  "tiles 32x%dx%d, stream hints %s" % (ty, zt, "on" if main_r["hints"] else "off")),
                   "samples_per_step": samples, "bytes_per_sample": bpc},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic_from_profiles(tag)},
        "pct_hbm_peak": round(100.0 * achieved / HBM_PEAK_GBS, 2),
        "cpu_baseline": cpu,
    }
    if extras is not None:
        line["extras"] = extras
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
