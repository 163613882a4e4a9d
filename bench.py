"""Benchmark: Msamples/s of the wavefront path tracer on the 1080p env-lit workload.

One step = one complete 16-spp, depth-12 render of the env-lit scene (500
displaced spheres, 1,000,002 triangles, 2048x1024 HDR sky with binary64-built
CDFs) at 1920x1080, i.e. 33,177,600 path samples, from accumulation reset to
the gamma-packed framebuffer.  Scene generation, BVH build and CDF build happen
before the timed region (inputs resident in HBM).

The JSON line also carries
  roofline       : for the dominant kernel (the BVH4 closest-hit pass), the HBM
                   bytes one launch moves (rocprofv3 FETCH_SIZE + WRITE_SIZE,
                   copy-calibrated, from profiles/) over its average launch
                   duration measured here with HIP events on the library's
                   stream, against MI355X's 8 TB/s; beside it the bytes the
                   traversal requests and its measured lane utilisation;
  cpu_baseline   : the CPU oracle (same estimator, binary32, OpenMP over all
                   host threads) on a bounded random subset of the frame's pixels;
  time_to_relmse : relMSE <= 0.01 against a 65,536-spp binary64 reference
                   rendered in this run: plain spp (doubling up to 16,384 spp;
                   a plain side that does not get there is a lower bound) vs
                   4 spp + the learned denoiser.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Msamples/s at 1080p; time to relMSE≤0.01 (4spp+denoiser vs plain spp)"
HBM_PEAK_GBPS = 8000.0
# Algorithmic bytes of one extend ray: path-state read (origin 16 B + direction
# 16 B) + hit record write 16 B; per BVH2 node visited 64 B (BVH4: 128 B); per
# triangle tested 48 B (three float4 vertices).
EXTEND_RAY_BYTES = 48
NODE_BYTES = 64
TRI_BYTES = 48


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def profile_value(name, key):
    """A `key: value` line of a committed rocprofv3 summary under profiles/."""
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return None
    with open(path) as fh:
        for line in fh:
            if line.startswith(key + ":"):
                return float(line.split(":", 1)[1])
    return None


def cpu_baseline(sd, seconds):
    """The oracle (binned-SAH BVH2, nearer child first) on the host's cores, on
    a bounded sample of the workload's pixel-samples."""
    os.environ.setdefault("OMP_PROC_BIND", "close")     # read when the OpenMP runtime starts
    os.environ.setdefault("OMP_PLACES", "cores")
    import oracle
    path, build = oracle.baseline_library()
    info = oracle.cpu_info()
    with oracle.using_library(path) as L:
        ref = oracle.OracleScene(sd, env_cdf=None)
        threads = L.ora_num_threads()
        W, H = sd.width, sd.height
        rs = np.random.default_rng(1234)
        n = 65536
        t0 = time.perf_counter()
        ref.trace_samples(rs.integers(0, W * H, n).astype(np.int32), W, H, 0, sd.max_depth)
        dt = max(time.perf_counter() - t0, 1e-3)
        n2 = int(min(60_000_000, max(n, n * seconds / dt)))
        pix = rs.integers(0, W * H, n2).astype(np.int32)
        t0 = time.perf_counter()
        ref.trace_samples(pix, W, H, 0, sd.max_depth)
        dt2 = time.perf_counter() - t0
        del ref
    return {"value": round(n2 / dt2 / 1e6, 4), "unit": "Msamples/s", "cores": int(threads), "kind": "port",
            "sample": f"{n2} random pixels of frame 0 (1 spp each) of the same 1920x1080 depth-12 workload, "
                      f"{dt2:.1f} s on {threads} OpenMP threads",
            "build": build, "bvh": "binned-SAH BVH2 (16 bins), nearer child first",
            "cpu": {"model": info["model"], "logical_cpus": info["logical_cpus"],
                    "physical_cores": info["physical_cores"],
                    "smt": round(info["logical_cpus"] / max(1, info["physical_cores"]), 2),
                    "omp": {k: os.environ.get(k) for k in ("OMP_NUM_THREADS", "OMP_PROC_BIND", "OMP_PLACES")}}}


def relmse(x, ref):
    """mean over pixels and RGB of (x - r)^2 / (r^2 + 1e-2)"""
    x = np.asarray(x, np.float64)
    return float(np.mean((x - ref) ** 2 / (ref * ref + 1e-2)))


def relmse_by_luminance(x, ref):
    """share of the relMSE sum from pixels binned by reference luminance, and
    each bin's energy bias sum(x) / sum(ref) - 1"""
    e = ((np.asarray(x, np.float64) - ref) ** 2 / (ref * ref + 1e-2)).mean(axis=-1)
    lum = ref.mean(axis=-1)
    edges = [0.0, 0.01, 0.03, 0.1, 0.3, 1.0, 3.0, np.inf]
    total = e.sum()
    out = {}
    for lo, hi in zip(edges[:-1], edges[1:]):
        m = (lum >= lo) & (lum < hi)
        xs, rs = np.asarray(x, np.float64)[m].sum(), ref[m].sum()
        out[f"{lo:g}-{hi:g}"] = {"px_frac": round(float(m.mean()), 4), "err_share": round(float(e[m].sum() / total), 4),
                                 "bias": round(float(xs / rs - 1.0), 4) if rs > 0 else None}
    return out


def denoiser_bench(args):
    """Inference time of the U-Net on the 1080p frame (4 spp + AOVs) and the
    time of one AdamW training step (batch 8 of 256x256), with MFMA rooflines."""
    from bluetruthengine_amd import denoiser, renderer, scenes
    sd = scenes.envlit_c3(spp=4)
    r = renderer.Renderer(sd)
    r.enable_aov(True)
    r.build_accel()
    r.render(4)
    if os.path.exists(denoiser.WEIGHTS_PATH):
        denoiser.load()
    else:
        denoiser.init(0)
    flips = denoiser.get_flips()

    def timed_denoise(n_flips):
        denoiser.set_flips(n_flips)
        for _ in range(2):
            denoiser.denoise()
        times = []
        for _ in range(args.steps):
            t0 = time.perf_counter()
            denoiser.denoise()
            times.append(time.perf_counter() - t0)
        return float(np.median(times)) * 1e3
    single_ms = timed_denoise(1)
    inf_ms = timed_denoise(flips)       # the default: the flip ensemble as one batch
    hp, wp = (sd.height + 7) // 8 * 8, (sd.width + 7) // 8 * 8
    fwd_flops = denoiser.flops_per_pixel() * hp * wp * flips
    # training step: batch 8 of 256x256 captured pairs
    tsd = scenes.small_envlit(width=256, height=256)
    t = renderer.Renderer(tsd)
    t.enable_aov(True)
    t.render(4)
    denoiser.dataset_reserve(8, 256, 256)
    for i in range(8):
        denoiser.capture(i, 0)
        denoiser.capture(i, 1)
    denoiser.train(3, 8)
    t0 = time.perf_counter()
    n_train = 20
    denoiser.train(n_train, 8)
    train_ms = (time.perf_counter() - t0) / n_train * 1e3
    train_flops = 3.0 * denoiser.flops_per_pixel() * 8 * 256 * 256   # forward + input grads + weight grads
    peak = 2500.0
    out = {"metric": "denoiser", "inference_ms_1080p": round(inf_ms, 3), "flips": flips,
           "single_pass_ms_1080p": round(single_ms, 3), "frame": f"{sd.width}x{sd.height}",
           "forward_tflop": round(fwd_flops / 1e12, 4),
           "roofline": {"bound": "mfma", "achieved": round(fwd_flops / (inf_ms * 1e-3) / 1e12, 2), "peak": peak,
                        "unit": "TFLOP/s", "frac": round(fwd_flops / (inf_ms * 1e-3) / 1e12 / peak, 4),
                        "kernel": f"U-Net forward over the {flips}-image flip ensemble (15 bf16 MFMA convs + "
                                  "pooling, features, output)"},
           "train_step_ms_b8_256": round(train_ms, 3),
           "train_tflops": round(train_flops / (train_ms * 1e-3) / 1e12, 2), "dtype": "bf16 (fp32 accumulate)"}
    print(json.dumps(out), flush=True)


def measure_relmse(r, sd, ref_spp, plain_max_spp, weight_files=None, windows=8, ref_cache="", skip_plain=False):
    """Time to relMSE <= 0.01 on the 1080p env-lit frame: plain path tracing
    (smallest spp that reaches it, if it does by plain_max_spp) against 4 spp
    + the learned denoiser.  The reference is the same estimator at ref_spp
    samples per pixel summed in binary64, from frames 2^24 on (a sample stream
    independent of the test renders' frames 0..).  r is a Renderer of sd (its
    launch shape already chosen); its accumulation is reset afterwards."""
    from bluetruthengine_amd import denoiser
    W, H = sd.width, sd.height
    t_all = time.perf_counter()
    r.enable_aov(False)
    r.enable_accum64(True)
    r.reset()
    r.set_frame_index(1 << 24)
    ref = None
    if ref_cache and os.path.exists(ref_cache):
        with np.load(ref_cache) as z:
            if int(z["spp"]) == ref_spp and z["ref"].shape == (H, W, 3):
                ref = z["ref"].astype(np.float64)
                log(f"reference: {ref_spp} spp read from {ref_cache}")
    if ref is None:
        t0 = time.perf_counter()
        done = 0
        while done < ref_spp:
            n = min(256, ref_spp - done)
            r.render(n)
            done += n
            if done % 4096 == 0 or done == ref_spp:
                log(f"reference: {done}/{ref_spp} spp, {time.perf_counter() - t0:.0f} s")
        ref = r.accum64() / float(ref_spp)
        if ref_cache:
            np.savez(ref_cache, ref=ref.astype(np.float32), spp=np.int64(ref_spp))
            ref = ref.astype(np.float32).astype(np.float64)   # same values as a later read of the file
    ref_s = time.perf_counter() - t_all
    r.enable_accum64(False)

    # One progressive pass (a render of n spp equals the first n frames of a
    # longer one): relMSE at powers of two until it reaches 0.01 or
    # plain_max_spp; when reached, a second pass re-accumulates to the power
    # of two below and checks every 1/64 of the final interval.
    curve = {}
    reached = False
    plain_spp = None
    if not skip_plain:
        r.reset()
        spp, done = 1, 0
        while True:
            r.render(spp - done)
            done = spp
            curve[spp] = relmse(r.accum(), ref)
            log(f"plain {spp} spp: relMSE {curve[spp]:.5f}")
            if curve[spp] <= 0.01 or spp * 2 > plain_max_spp:
                break
            spp *= 2
        plain_spp = spp
        reached = curve[spp] <= 0.01
        if reached and spp > 1:
            lo = spp // 2
            step = max(1, (spp - lo) // 64)
            r.reset()
            r.render(lo)
            done = lo
            while done + step < spp:
                r.render(step)
                done += step
                curve[done] = relmse(r.accum(), ref)
                if curve[done] <= 0.01:
                    plain_spp = done
                    break
        log(f"plain: {plain_spp} spp, relMSE {curve[plain_spp]:.5f} ({'reached' if reached else 'not reached'})")

    # render time is linear in spp (passes of frames in flight): time a
    # bounded render and scale it
    plain_ms = None
    if plain_spp:
        timed_spp = min(plain_spp, 1024)

        def timed_plain():
            r.reset()
            t = time.perf_counter()
            r.render(timed_spp)
            return time.perf_counter() - t
        timed_plain()
        plain_ms = min(timed_plain() for _ in range(2)) * 1e3 * plain_spp / timed_spp

    # 4 spp + denoiser
    r.enable_aov(True)
    weight_files = weight_files or [denoiser.WEIGHTS_PATH]

    def timed_dn(spp):
        r.reset()
        t = time.perf_counter()
        r.render(spp)
        denoiser.denoise()
        return time.perf_counter() - t
    curves, noisy_curve, dn4 = {}, {}, {}
    for wf in weight_files:
        denoiser.load(wf)
        r.reset()
        done = 0
        cur = curves[wf] = {}
        for spp in (4, 8, 16, 32, 64):
            r.render(spp - done)
            done = spp
            denoiser.denoise()
            img = denoiser.denoised(W, H)
            if spp == 4:
                dn4[wf] = img
            cur[spp] = relmse(img, ref)
            noisy_curve[spp] = relmse(r.accum(), ref)
            log(f"{os.path.basename(wf)}: denoised {spp} spp: relMSE {cur[spp]:.5f} (input {noisy_curve[spp]:.5f})")
    compared = {wf: curves[wf][4] for wf in weight_files}
    best = min(compared, key=compared.get)
    dn_curve = curves[best]
    denoiser.load(best)
    timed_dn(4)
    dn_ms = min(timed_dn(4) for _ in range(3)) * 1e3
    dn_ok = dn_curve[4] <= 0.01
    out = {"reference_spp": ref_spp, "reference_s": round(ref_s, 1),
           "plain": None if skip_plain else {
               "reached": reached, "spp": plain_spp, "relmse": round(curve[plain_spp], 5), "ms": round(plain_ms, 2),
               "ms_is_lower_bound": not reached, "max_spp_tried": plain_max_spp,
               "ms_from": f"{min(plain_spp, 1024)}-spp render x {plain_spp / min(plain_spp, 1024):g}"},
           "denoised_4spp": {"relmse": round(dn_curve[4], 5), "ms": round(dn_ms, 2), "reached": dn_ok,
                             "input_relmse": round(noisy_curve[4], 5), "weights": os.path.basename(best),
                             "flips": denoiser.get_flips()},
           "target_speedup": 5.0,
           "plain_curve": {str(k): round(v, 5) for k, v in sorted(curve.items())},
           "denoised_curve": {str(k): round(v, 5) for k, v in sorted(dn_curve.items())},
           "input_curve": {str(k): round(v, 5) for k, v in sorted(noisy_curve.items())}}
    if plain_ms and dn_ok:
        out["speedup_4spp" if reached else "speedup_4spp_lower_bound"] = round(plain_ms / dn_ms, 2)
    if len(weight_files) > 1:
        out["denoised_4spp"]["weights_compared"] = {os.path.basename(k): round(v, 5) for k, v in compared.items()}
        out["denoised_curves"] = {os.path.basename(w): {str(k): round(v, 5) for k, v in sorted(c.items())}
                                  for w, c in curves.items()}
    out["denoised_4spp"]["error_by_reference_luminance"] = relmse_by_luminance(dn4[best], ref)
    if windows > 0:
        out["denoised_4spp"]["windows"] = denoise_windows(r, denoiser, ref, windows, W, H)
    # relMSE(N) = a / N + b for independent samples: b is the reference's own
    # noise, a / 0.01 the spp a noise-free reference would need
    ks = sorted(k for k in curve if k >= 1024 and (k & (k - 1)) == 0)
    if len(ks) >= 2 and plain_ms:
        n1, n2 = ks[-2], ks[-1]
        a = (curve[n1] - curve[n2]) / (1.0 / n1 - 1.0 / n2)
        b = curve[n2] - a / n2
        ms_per_spp = plain_ms / plain_spp
        fit = {"a": round(a, 3), "reference_noise_b": round(b, 5), "from_spp": [n1, n2]}
        if b < 0.01:
            need = a / (0.01 - b)
            fit["plain_spp_needed_vs_this_reference"] = int(np.ceil(need))
            fit["plain_ms_needed_vs_this_reference"] = round(need * ms_per_spp, 1)
        fit["plain_spp_needed_noise_free_reference"] = int(np.ceil(a / 0.01))
        fit["plain_ms_needed_noise_free_reference"] = round(a / 0.01 * ms_per_spp, 1)
        if dn_ok:
            fit["speedup_4spp_vs_extrapolated_plain"] = round(a / 0.01 * ms_per_spp / dn_ms, 1)
        out["extrapolation"] = fit
    r.enable_aov(False)
    r.reset()
    out["wall_s"] = round(time.perf_counter() - t_all, 1)
    return out


def relmse_bench(args):
    """--relmse: the time-to-relMSE measurement alone, with a larger plain
    budget and optionally several weight files."""
    from bluetruthengine_amd import renderer, scenes
    sd = scenes.envlit_c3(spp=16)
    r = renderer.Renderer(sd)
    r.build_accel()
    r.autotune(16)
    weight_files = [w for w in args.weights.split(",") if w] or None
    out = measure_relmse(r, sd, args.ref_spp, args.plain_max_spp or args.ref_spp // 2, weight_files, args.windows,
                         args.ref_cache, args.skip_plain)
    out["metric"] = "time to relMSE<=0.01, 1080p env-lit"
    print(json.dumps(out), flush=True)


def denoise_windows(r, denoiser, ref, n, W, H, spp=4):
    """relMSE of 4 spp + the denoiser over n independent renders of the frame
    (frames 4k..4k+3 of one accumulation: the window's radiance from the
    binary64 sums, its AOV means from the running means), each denoised from
    caller buffers (bte_denoise_image)."""
    r.reset()
    r.enable_accum64(True)
    prev = None
    errs, noisy = [], []
    for k in range(n):
        r.render(spp)
        s64 = r.accum64().astype(np.float64)
        a, m = (v.astype(np.float64) for v in r.aov())
        f = spp * (k + 1)
        if prev is None:
            rgb, alb, nrm = s64 / spp, a, m
        else:
            ps, pa, pm = prev
            rgb, alb, nrm = (s64 - ps) / spp, (a * f - pa * (f - spp)) / spp, (m * f - pm * (f - spp)) / spp
        prev = (s64, a, m)
        img = denoiser.denoise_image(rgb.astype(np.float32), alb.astype(np.float32), nrm.astype(np.float32))
        errs.append(relmse(img, ref))
        noisy.append(relmse(rgb, ref))
        log(f"window {k} (frames {f - spp}..{f - 1}): denoised relMSE {errs[-1]:.5f} (input {noisy[-1]:.3f})")
    r.enable_accum64(False)
    return {"n": n, "spp": spp, "relmse": [round(e, 5) for e in errs], "mean": round(float(np.mean(errs)), 5),
            "max": round(float(np.max(errs)), 5), "reaching_0.01": int(sum(e <= 0.01 for e in errs)),
            "input_relmse": [round(e, 3) for e in noisy]}


def serve_scenes(n):
    """The 64 uneven serving scenes (scenes.serve_c5)."""
    from bluetruthengine_amd import scenes
    t0 = time.perf_counter()
    out = scenes.serve_c5(n)
    log(f"generated {n} scenes in {time.perf_counter() - t0:.1f} s")
    return out


def serve_bench(args):
    """Serving of the 64 uneven scenes at 8 spp, depth 10, each frame followed
    by the denoiser, two ways: all scenes as one batch (every bounce of every
    scene in one persistent launch, then the U-Net over every slot), and one
    scene after another.  Scene generation, upload and BVH build are outside
    the timed regions; a first untimed round captures the graphs."""
    from bluetruthengine_amd import denoiser, renderer
    renderer.set_accel_builder(args.accel_builder)
    sds = serve_scenes(args.serve_scenes)
    have_weights = os.path.exists(denoiser.WEIGHTS_PATH)
    samples = sum(sd.width * sd.height * 8 for sd in sds)

    def load_weights():
        if have_weights:
            denoiser.load()
        else:
            denoiser.init(0)
        if args.flips:
            denoiser.set_flips(args.flips)

    # one scene after another
    cold, warm, dn_ms = [], [], []
    t_all = 0.0
    for k, sd in enumerate([] if args.skip_sequential else sds):
        r = renderer.Renderer(sd)
        r.enable_aov(True)
        r.build_accel()
        load_weights()
        t0 = time.perf_counter()
        r.render(8)
        denoiser.denoise()
        cold.append((time.perf_counter() - t0) * 1e3)
        r.reset()
        t0 = time.perf_counter()
        r.render(8)
        t1 = time.perf_counter()
        denoiser.denoise()
        t2 = time.perf_counter()
        warm.append((t2 - t0) * 1e3)
        dn_ms.append((t2 - t1) * 1e3)
        t_all += t2 - t0
        if k % 8 == 7:
            log(f"sequential: scene {k}: {sd.width}x{sd.height}, {sd.triangles.shape[0]} tris: first frame "
                f"{cold[-1]:.1f} ms, frame {warm[-1]:.1f} ms (denoise {dn_ms[-1]:.2f} ms)")
    sequential = None if args.skip_sequential else {"frame_p50_ms": round(float(np.percentile(warm, 50)), 2),
                  "frame_p99_ms": round(float(np.percentile(warm, 99)), 2),
                  "first_frame_p50_ms": round(float(np.percentile(cold, 50)), 2),
                  "denoise_p50_ms": round(float(np.percentile(dn_ms, 50)), 2),
                  "all_scenes_ms": round(t_all * 1e3, 1),
                  "msamples_per_s": round(samples / t_all / 1e6, 1)}

    # one batch
    t0 = time.perf_counter()
    b = renderer.Batch(max_depth=10, aov=True, graphs=0 if args.no_graphs else 1,
                       frames_in_flight=args.frames_in_flight, memory_budget_mb=args.memory_budget_mb,
                       tail_bounce=args.tail_bounce)
    build_ms = []
    for k, sd in enumerate(sds):
        b.add(sd)
        build_ms.append(renderer.accel_stats()["buildMs"])
        if k % 8 == 7:
            log(f"batch: {k + 1}/{len(sds)} scenes uploaded, BVHs built ({time.perf_counter() - t0:.0f} s)")
    load_weights()
    tuned = None
    if not args.no_autotune:
        ms, cfg = b.autotune(8)
        tuned = {k: cfg[k] for k in ("traversalStack", "workChunk", "traversalBlocksPerCU", "bvhWidth", "tailBounce")}
        log(f"batch autotune: best {ms:.1f} ms with {tuned}")
    rounds = []
    for it in range(1 + max(1, args.steps)):
        b.reset()
        t0 = time.perf_counter()
        b.render(8)
        t1 = time.perf_counter()
        b.denoise()
        t2 = time.perf_counter()
        if it > 0:
            rounds.append(((t2 - t0) * 1e3, (t1 - t0) * 1e3, (t2 - t1) * 1e3))
        log(f"batch round {it}: {len(sds)} frames in {(t2 - t0) * 1e3:.1f} ms "
            f"(trace {(t1 - t0) * 1e3:.1f}, denoise {(t2 - t1) * 1e3:.1f})" + (" [warm-up]" if it == 0 else ""))
    st = b.stats()
    lat = [x[0] for x in rounds]
    batch = {"frame_p50_ms": round(float(np.percentile(lat, 50)), 2),
             "frame_p99_ms": round(float(np.percentile(lat, 99)), 2),
             "trace_ms": round(float(np.median([x[1] for x in rounds])), 2),
             "denoise_ms": round(float(np.median([x[2] for x in rounds])), 2),
             "msamples_per_s": round(samples / (np.median(lat) * 1e-3) / 1e6, 1),
             "frames_in_flight": st.framesInFlight, "rounds": len(rounds), "autotuned": tuned,
             "latency_note": "every frame of the batch completes when the batch does"}
    out = {"metric": "batched serving: per-frame latency, 8 spp + denoiser", "scenes": len(sds),
           "pixels": int(sum(sd.width * sd.height for sd in sds)),
           "triangles": int(sum(sd.triangles.shape[0] for sd in sds)),
           "p50_ms": batch["frame_p50_ms"], "p99_ms": batch["frame_p99_ms"],
           "batch": batch, "sequential": sequential, "denoiser_flips": denoiser.get_flips(),
           "bvh_build_ms": {"builder": "gpu" if args.accel_builder == 1 else "host",
                            "all_scenes": round(float(np.sum(build_ms)), 1),
                            "max_scene": round(float(np.max(build_ms)), 1)},
           "weights": "trained" if have_weights else "untrained (random init)"}
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--spp", type=int, default=16)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-instrumented", action="store_true")
    ap.add_argument("--traversal-stack", type=int, default=0, help="LDS stack entries per lane (8/16/32)")
    ap.add_argument("--work-chunk", type=int, default=0, help="64-slot rows per work grab (1/2/4/8)")
    ap.add_argument("--blocks-per-cu", type=int, default=0, help="resident traversal blocks per CU (0 = max)")
    ap.add_argument("--bvh-width", type=int, default=4, help="2, 4 or 8: BVH traversed by extend/connect")
    ap.add_argument("--tail-bounce", type=int, default=None,
                    help="bounce from which one fused pass finishes the paths (0 = off; default: the library's)")
    ap.add_argument("--no-graphs", action="store_true", help="plain launches instead of captured HIP graphs")
    ap.add_argument("--accel-builder", type=int, default=1, help="BVH builder: 1 on the GPU (default), 0 on the host")
    ap.add_argument("--frames-in-flight", type=int, default=0,
                    help="frames traced per wavefront pass (0 = as many as the memory budget allows)")
    ap.add_argument("--memory-budget-mb", type=int, default=0, help="path-state budget (0 = 32768 MB)")
    ap.add_argument("--skip-sequential", action="store_true", help="--serve: only the batched mode")
    ap.add_argument("--no-autotune", action="store_true", help="skip the launch-shape autotuner")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB (3 times) to calibrate rocprof byte counters")
    ap.add_argument("--denoiser", action="store_true", help="only time the U-Net denoiser (inference, training)")
    ap.add_argument("--relmse", action="store_true",
                    help="only measure time to relMSE<=0.01: plain spp vs 4 spp + denoiser")
    ap.add_argument("--ref-spp", type=int, default=65536, help="samples per pixel of the relMSE reference")
    ap.add_argument("--plain-max-spp", type=int, default=0,
                    help="largest plain spp tried for relMSE <= 0.01 (0: 16384 in the default run, ref-spp/2 with --relmse)")
    ap.add_argument("--no-relmse", action="store_true", help="skip the time-to-relMSE measurement")
    ap.add_argument("--skip-plain", action="store_true", help="--relmse: only the denoised side")
    ap.add_argument("--windows", type=int, default=8,
                    help="--relmse: also denoise this many independent 4-spp renders (0 = off)")
    ap.add_argument("--ref-cache", default="", metavar="NPZ",
                    help="--relmse: read the reference image from this file if it holds --ref-spp samples, "
                         "else render it and write it there (float32)")
    ap.add_argument("--serve", action="store_true", help="only run the 64-scene serving benchmark")
    ap.add_argument("--serve-scenes", type=int, default=64)
    ap.add_argument("--flips", type=int, default=0, help="--serve: denoiser flip ensemble 1/2/4 (0 = library default)")
    ap.add_argument("--weights", default="", help="--relmse: comma-separated denoiser weight files to compare "
                                                  "(default: the shipped weights)")
    args = ap.parse_args()
    if args.copy_calibration:
        import torch
        n = args.copy_calibration * 1024 * 1024 // 4
        a = torch.ones(n, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": n * 4, "copy_bytes_written": n * 4, "repeats": 3}), flush=True)
        return
    if args.denoiser:
        return denoiser_bench(args)
    if args.relmse:
        return relmse_bench(args)
    if args.serve:
        return serve_bench(args)
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU only (asked for --gpus {args.gpus}); "
              "the path shards by sample index across GPUs with one final reduce, which is not run here",
              file=sys.stderr)
        sys.exit(2)

    from bluetruthengine_amd import renderer, scenes
    renderer.set_accel_builder(args.accel_builder)
    t0 = time.perf_counter()
    sd = scenes.envlit_c3(spp=args.spp)
    log(f"scene: {sd.triangles.shape[0]} triangles, {sd.bsdfs.shape[0]} bsdfs, generated in {time.perf_counter() - t0:.1f} s")
    t0 = time.perf_counter()
    r = renderer.Renderer(sd, frames_in_flight=args.frames_in_flight, memory_budget_mb=args.memory_budget_mb,
                          traversal_stack=args.traversal_stack, graphs=0 if args.no_graphs else 1,
                          work_chunk=args.work_chunk, traversal_blocks_per_cu=args.blocks_per_cu,
                          bvh_width=args.bvh_width, tail_bounce=args.tail_bounce)
    r.build_accel()
    accel = r.accel_stats()
    log(f"upload + env CDF + BVH build: {time.perf_counter() - t0:.1f} s (BVH: {accel})")
    W, H, spp = sd.width, sd.height, sd.spp
    if not args.no_autotune:
        t0 = time.perf_counter()
        best_ms, cfg = r.autotune(spp)
        log(f"autotune ({time.perf_counter() - t0:.1f} s): best {best_ms:.2f} ms with {cfg}")
    cfg = r.config()

    def step():
        r.reset()
        r.render(spp)

    for i in range(args.warmup):
        step()
        log(f"warmup {i} done: {r.stats().renderMs:.1f} ms")
    times = []
    t_all = time.perf_counter()
    for i in range(args.steps):
        ts = time.perf_counter()
        step()
        times.append(time.perf_counter() - ts)
        log(f"step {i}: {times[-1] * 1e3:.1f} ms (gpu {r.stats().renderMs:.1f} ms)")
    elapsed = time.perf_counter() - t_all
    st = r.stats().as_dict()
    samples = W * H * spp
    value = samples * args.steps / elapsed / 1e6
    rays = st["extensionRays"] + st["shadowRays"]

    roofline = None
    kernel = {}
    if not args.no_instrumented:
        r.set_instrumentation(1)
        step()
        s1 = r.stats()                 # kernel split of the configuration measured above
        tail = r.set_tail_bounce(0)
        step()
        s0 = r.stats()                 # every bounce a per-bounce launch: the extension kernel's roofline
        r.set_tail_bounce(tail)
        r.set_instrumentation(2)
        step()
        s2 = r.stats()                 # node / triangle counts (per-bounce passes)
        r.set_instrumentation(0)
        launches = max(1, s0.extendLaunches)
        avg_ms = s0.extendMs / launches
        node_bytes = NODE_BYTES * (cfg["bvhWidth"] // 2 if cfg["bvhWidth"] in (4, 8) else 1)
        request_total = EXTEND_RAY_BYTES * s2.extensionRays + node_bytes * s2.extendNodes + TRI_BYTES * s2.extendTris
        request_per_launch = request_total / launches
        # bytes that cross HBM: rocprofv3 FETCH_SIZE + WRITE_SIZE of this kernel
        # on this workload, bounce by bounce, calibrated with a copy of known
        # size, over the same launches (every bounce, no tail pass)
        per_bounce = [profile_value("extend_c3_hbm_traffic.txt", f"extend_bytes_bounce_{b}") for b in range(launches)]
        traffic = sum(per_bounce) / launches if all(v is not None for v in per_bounce) else None
        achieved = traffic / (avg_ms * 1e-3) / 1e9 if traffic else None
        roofline = {"bound": "hbm", "achieved": round(achieved, 1) if achieved else None, "peak": HBM_PEAK_GBPS,
                    "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBPS, 4) if achieved else None,
                    "traffic": int(traffic) if traffic else None,
                    "traffic_source": "profiles/extend_c3_hbm_traffic.txt (rocprofv3 FETCH_SIZE + WRITE_SIZE per "
                                      "launch, copy-calibrated; every bounce a launch, no tail pass)",
                    "kernel": ("k_extend_pool (BVH4 closest hit; camera rays of bounce 0 by k_extend_packet)"
                               if cfg["bvhWidth"] == 4 else
                               f"k_extend_pool (BVH{cfg['bvhWidth']} closest hit)" if cfg["bvhWidth"] == 8
                               else "k_extend (BVH2 closest hit)"), "launches": launches,
                    "avg_launch_ms": round(avg_ms, 4),
                    "request_bytes_per_launch": int(request_per_launch),
                    "request_GBps": round(request_per_launch / (avg_ms * 1e-3) / 1e9, 1),
                    "request_note": f"bytes the traversal asks the memory system for ({node_bytes} B per node visit, "
                                    f"{TRI_BYTES} B per triangle test, {EXTEND_RAY_BYTES} B ray + hit); the L2 and "
                                    "MALL serve all but the traffic above",
                    "l2_hit_rate": profile_value("extend_c3_hbm_traffic.txt", "extend_l2_hit_rate"),
                    "valu_utilization_pct": profile_value("c3_occupancy_divergence.md", "extend_valu_utilization"),
                    "waves_per_cu": profile_value("c3_occupancy_divergence.md", "extend_waves_per_cu"),
                    "limiter": "vector-ALU issue of divergent traversal waves (lanes busy per VALU instruction "
                               "above), not HBM bandwidth"}
        kernel = {"extend_ms": round(s1.extendMs, 2), "shade_ms": round(s1.shadeMs, 2),
                  "connect_ms": round(s1.connectMs, 2), "tail_ms": round(s1.tailMs, 2),
                  "tail_from_bounce": cfg["tailBounce"] if s1.tailLaunches else None,
                  "nodes_per_ray": round(s2.extendNodes / max(1, s2.extensionRays), 2),
                  "tris_per_ray": round(s2.extendTris / max(1, s2.extensionRays), 2),
                  "shadow_nodes_per_ray": round(s2.connectNodes / max(1, s2.shadowRays), 2)}

    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(sd, args.cpu_seconds)
        log(f"cpu baseline: {cpu}")

    # second half of the metric: time to relMSE <= 0.01, measured in this run
    relmse_rec = None
    if not args.no_relmse:
        relmse_rec = measure_relmse(r, sd, args.ref_spp, args.plain_max_spp or 16384, windows=args.windows)
        log(f"time to relMSE: {json.dumps({k: relmse_rec.get(k) for k in ('plain', 'speedup_4spp', 'speedup_4spp_lower_bound')})}")
    out = {
        "metric": METRIC, "value": round(value, 3), "unit": "Msamples/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(elapsed / args.steps * 1e3, 2), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"env-lit 1920x1080, {spp} spp, depth 12, 500 displaced spheres "
                               f"({sd.triangles.shape[0]} tris), 2048x1024 HDR sky, NEE+MIS",
                   "width": W, "height": H, "spp": spp, "max_depth": sd.max_depth,
                   "frames_in_flight": st["framesInFlight"], "traversal_stack": cfg["traversalStack"],
                   "work_chunk_rows": cfg["workChunk"], "traversal_blocks_per_cu": cfg["traversalBlocksPerCU"],
                   "bvh_width": cfg["bvhWidth"], "tail_bounce": cfg["tailBounce"],
                   "hip_graphs": bool(cfg["graphs"]), "autotuned": not args.no_autotune},
        "mrays_per_s": round(rays * args.steps / elapsed / 1e6, 1) if rays else None,
        "rays_per_step": {"extension": st["extensionRays"], "shadow": st["shadowRays"], "hits": st["hits"],
                          "extension_per_bounce": st["raysPerBounce"]},
        "roofline": roofline, "kernels": kernel, "cpu_baseline": cpu, "time_to_relmse": relmse_rec,
        "bvh_build": {"builder": "gpu" if accel["builder"] == 1 else "host", "ms": round(accel["buildMs"], 1),
                      "triangles": accel["triangles"], "bvh2_nodes": accel["nodes2"], "bvh4_nodes": accel["nodes4"],
                      "depth": accel["depth2"]},
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
