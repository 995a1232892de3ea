#!/usr/bin/env python3
"""Headline benchmark: paths/s of the path tracer at 1024^2 on one MI355X.

Workload (the configuration the metric is quoted on): the GGX material grid --
7x7 icospheres of 5,120 triangles on a ground plane (250,882 triangles), GGX
alpha in [0.02, 1] across columns, rows alternating conductor / opaque
dielectric, an importance-sampled 2048x1024 sky with a 5e4 sun, f/2.8 thin
lens, 1024x1024, 256 spp, max depth 16, NEE + MIS, seed 1 -- generated in the
reference's own input formats and loaded through the C ABI.

A step is one complete 256-spp render into device-resident films (scene, BVH
and films already in HBM; parse + BVH build + upload are reported separately as
load_s).  value = width * height * spp * steps / timed wall time.

The roofline object covers the dominant kernel (pt_extend, the closest-hit
traversal): its algorithmic bytes per launch are the canonical per-ray bytes
(64 B per BVH2 node visit + 48 B per triangle test, counted by the CPU
oracle's binned-SAH BVH2 over 2^16 paths) plus the 52 B of queue/ray/hit
records each ray moves, times the rays one launch processes; divided by the
launch's average duration from HIP events on the library's stream.

cpu_baseline: the oracle (the project's C++ restatement of the reference path
tracer) timed on the host on 48 passes (48 spp) of the same 1024x1024 workload.

roofline.traffic: HBM bytes per pt_extend launch measured by rocprofv3
FETCH_SIZE and WRITE_SIZE runs, scaled by the factor the same counters read
for a plain copy of known size (--copy-calibration); read from the committed
summary in profiles/ when present.

Other workloads (not the headline line): --workload sppm times photon-mapping
passes on the caustic box (prism + 2M-triangle torus, 1920x1080);
--workload batch renders 32 random-cuboid jobs in one batch (autotuned);
--workload guide measures the learned guide's equal-time relMSE gain.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
TRAFFIC_SUMMARY = os.path.join(ROOT, "profiles", "pt_extend_ggx_grid_1024_traffic.md")
STATE_BYTES_PER_RAY = 52  # queue index 4 + ray origin/direction 32 + hit record 16


def log(msg):
    print("[bench] " + msg, file=sys.stderr, flush=True)


class Hip:
    """Minimal HIP runtime access for device films (no framework needed)."""

    def __init__(self):
        self.lib = C.CDLL("libamdhip64.so")
        self.lib.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.lib.hipMemset.argtypes = [C.c_void_p, C.c_int, C.c_size_t]
        self.lib.hipFree.argtypes = [C.c_void_p]

    def malloc(self, nbytes):
        p = C.c_void_p()
        rc = self.lib.hipMalloc(C.byref(p), nbytes)
        if rc != 0:
            raise RuntimeError("hipMalloc failed (%d)" % rc)
        return p.value

    def memset(self, p, nbytes):
        if self.lib.hipMemset(C.c_void_p(p), 0, nbytes) != 0:
            raise RuntimeError("hipMemset failed")

    def sync(self):
        if self.lib.hipDeviceSynchronize() != 0:
            raise RuntimeError("hipDeviceSynchronize failed")

    def free(self, p):
        self.lib.hipFree(C.c_void_p(p))


def host_cpus():
    """What the host offers this process: the CPU model (/proc/cpuinfo), the
    logical CPUs of the machine, those this process may run on (affinity),
    and the worker count the environment allots (OMP_NUM_THREADS, else the
    affinity count)."""
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
        affinity = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        affinity = os.cpu_count() or 1
    share = affinity
    env = os.environ.get("OMP_NUM_THREADS", "")
    if env.isdigit() and int(env) > 0:
        share = min(share, int(env))
    return {"model": model, "logical_cpus": os.cpu_count() or 1, "affinity_cpus": affinity, "share": share}


def cpu_baseline(job_path, seed, threads, spp):
    """Oracle timing on the host (bounded sample: `spp` passes of the 1024^2
    film) and the canonical BVH2 counts (2^16 paths: 128^2 film of the same
    camera, 4 spp)."""
    from oracle import pyoracle

    oj = pyoracle.OracleJob(job_path, seed=seed)
    info = oj.info()
    t0 = time.perf_counter()
    oj.render(0, spp, threads=threads)
    dt = time.perf_counter() - t0
    paths = info["width"] * info["height"] * spp
    oj.set(width=128, height=128)
    _, cnt = oj.render_counted(0, 4)
    hc = host_cpus()
    return {
        "value": paths / dt,
        "unit": "paths/s",
        "cores": threads,
        "kind": "port",
        "sample": "%d spp of the %dx%d film (%d paths) in %.2f s, %d host threads" % (
            spp, info["width"], info["height"], paths, dt, threads),
        "cpu_model": hc["model"],
        "host_logical_cpus": hc["logical_cpus"],
        "affinity_cpus": hc["affinity_cpus"],
    }, cnt


def sppm_cpu_baseline(job_path, args):
    """The oracle's photon mapper (a port of renderPhotonThread's pass) on a
    bounded sample: one pass of the same job at half the width and height
    (a quarter of the pixels, so also a quarter of the photon budget, which
    the reference sets per pixel: PhotonMapping.cpp:834-835), scaled to the
    full film as passes/s x 1/4.  Photons are emitted sequentially as in one
    reference thread's pass; camera paths run on the host threads."""
    from oracle import pyoracle

    t0 = time.perf_counter()
    o = pyoracle.OracleSppm(job_path, seed=3, alpha=args.sppm_alpha, store_fraction=args.sppm_store_fraction,
                            emit_fraction=args.sppm_emit_fraction, initial_radius=args.sppm_initial_radius,
                            width=960, height=540, fixed_point_flux=args.sppm_fixed_flux)
    load = time.perf_counter() - t0
    t0 = time.perf_counter()
    o.run_pass(0, 960, 540)
    dt = time.perf_counter() - t0
    hc = host_cpus()
    info = o.info()
    return {"value": 0.25 / dt, "unit": "passes/s", "cores": min(16, hc["logical_cpus"]), "kind": "port",
            "sample": "1 pass at 960x540 (a quarter of the 1920x1080 pixels and photon budget: %d photons "
                      "emitted, %d stored) in %.2f s after a %.1f s load; passes/s at 1920x1080 = 0.25 / that"
                      % (info["emitted"], info["stored"], dt, load),
            "cpu_model": hc["model"], "host_logical_cpus": hc["logical_cpus"]}


def measured_traffic():
    """HBM bytes per pt_extend launch from the committed rocprofv3 summary
    (line 'traffic_bytes_per_launch: N'), or None.  rocprofv3 counters need a
    run of their own, so this figure comes from the profiles/ summary of an
    earlier counter run of the same command, not from this run; the bench line
    says so beside it."""
    try:
        with open(TRAFFIC_SUMMARY) as f:
            for line in f:
                if line.startswith("traffic_bytes_per_launch:"):
                    return float(line.split(":", 1)[1])
    except OSError:
        return None
    return None


def copy_calibration(mib):
    """A plain device copy of `mib` MiB, five times (the known-size kernel the
    FETCH_SIZE / WRITE_SIZE counters are scaled by)."""
    import torch

    n = mib * (1 << 20) // 4
    src = torch.ones(n, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    for _ in range(5):
        dst.copy_(src)
    torch.cuda.synchronize()
    log("copy calibration: 5 copies of %d bytes" % (n * 4))


def bench_sppm(args):
    """Photon-mapping passes on the caustic box (dispersive prism, GGX gold
    torus of n_major x n_minor x 2 triangles, 1% ceiling light) at
    1920x1080; reference parameters unless overridden."""
    import hastyrenderer_amd as ha
    from hastyrenderer_amd import scenes

    scene_dir = args.scene_dir or os.path.join(tempfile.gettempdir(), "hasty_bench_caustic")
    job_path = os.path.join(scene_dir, "renderJob.json")
    if not os.path.exists(job_path):
        log("generating caustic scene in %s" % scene_dir)
        scenes.write_caustic_scene(scene_dir)
    t0 = time.perf_counter()
    scene = ha.Scene(job_path)
    load_s = time.perf_counter() - t0
    info = scene.info()
    W, H = info["width"], info["height"]
    log("scene loaded in %.2f s: %d triangles" % (load_s, info["num_triangles"]))
    sp = ha.Sppm(scene, 3, alpha=args.sppm_alpha, store_fraction=args.sppm_store_fraction,
                 emit_fraction=args.sppm_emit_fraction, initial_radius=args.sppm_initial_radius,
                 fixed_point_flux=args.sppm_fixed_flux)
    hip = Hip()
    nbytes = W * H * 3 * 4
    films = [hip.malloc(nbytes) for _ in range(3)]
    k = 0
    for _ in range(args.warmup):
        sp.pass_device(k, *films)
        k += 1
    # timed passes: no event timing; each pass's counters are read after it
    hip.sync()
    t_start = time.perf_counter()
    emitted = stored = overflows = 0
    for _ in range(args.steps):
        t = time.perf_counter()
        sp.pass_device(k, *films)
        i = sp.info()
        emitted += i["emitted"]
        stored += i["stored"]
        overflows += i["flux_overflows"]
        log("pass %d: %.3f s, emitted %d stored %d cell %.4g" % (k, time.perf_counter() - t, i["emitted"],
                                                              i["stored"], i["cell_size"]))
        k += 1
    hip.sync()
    elapsed = time.perf_counter() - t_start
    # per-kernel times and gather statistics from extra passes with HIP-event
    # timing on every launch (outside the timed region)
    prof_passes = max(1, min(3, args.steps))
    scene.reset_stats()
    scene.set_profiling(events=True, counters=False)
    queries = candidates = found = 0
    for _ in range(prof_passes):
        sp.pass_device(k, *films)
        i = sp.info()
        queries += i["gather_queries"]
        candidates += i["gather_candidates"]
        found += i["gather_found"]
        k += 1
    hip.sync()
    kstats = scene.kernel_stats()
    scene.set_profiling(events=False, counters=False)
    # roofline of the gather kernel: per gather 64 B of query record, 16 B
    # per candidate photon position read in the cells it visits, 32 B
    # (direction + flux) per photon within the radius
    roofline = None
    g_launches, g_ms = kstats.get("sppm_gather", (0, 0.0))
    if g_launches and g_ms > 0:
        gbytes = 64.0 * queries + 16.0 * candidates + 32.0 * found
        achieved = gbytes / (g_ms / 1e3) / 1e9
        roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None, "kernel": "sppm_gather",
                    "bytes_per_launch": gbytes / g_launches, "avg_launch_ms": round(g_ms / g_launches, 4),
                    "queries_per_pass": queries / prof_passes,
                    "candidates_per_query": candidates / max(1, queries), "found_per_query": found / max(1, queries)}
    base = None
    if not args.no_cpu_baseline:
        base = sppm_cpu_baseline(job_path, args)
        log("sppm cpu baseline %.4g passes/s at %dx%d" % (base["value"], W, H))
    line = {
        "metric": "photon-mapping passes/s at 1920x1080 (caustic box)",
        "value": args.steps / elapsed,
        "unit": "passes/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "sppm_caustic_box", "width": W, "height": H, "triangles": info["num_triangles"],
                   "alpha": args.sppm_alpha, "store_fraction": args.sppm_store_fraction,
                   "emit_fraction": args.sppm_emit_fraction, "initial_radius": args.sppm_initial_radius or "footprint",
                   "flux": "int64 fixed point" if args.sppm_fixed_flux else "fp32",
                   "parallelism": "single-gpu"},
        "photons_emitted_per_s": emitted / elapsed,
        "photons_stored_per_pass": stored / max(1, args.steps),
        "flux_overflows": overflows,
        "load_s": load_s,
        "kernel_ms_per_pass": {k2: round(v[1] / prof_passes, 3) for k2, v in kstats.items()},
        "kernel_ms_source": "%d extra passes with HIP-event timing per launch, after the timed passes" % prof_passes,
        "roofline": roofline,
        "cpu_baseline": base,
    }
    for f in films:
        hip.free(f)
    sp.close()
    scene.close()
    print(json.dumps(line), flush=True)


def bench_batch(args):
    """32 render jobs (random-cuboid scenes, resolutions log-uniform in
    128^2..1024^2, spp in 16..256, 1k..160k cuboids) rendered in one batch
    through one path pool, after the autotuner picked the launch shape; the
    same jobs rendered one after another are reported beside it."""
    import hastyrenderer_amd as ha
    from hastyrenderer_amd import scenes

    scene_dir = args.scene_dir or os.path.join(tempfile.gettempdir(), "hasty_bench_batch")
    marker = os.path.join(scene_dir, "jobs.json")
    if os.path.exists(marker):
        with open(marker) as f:
            paths = json.load(f)
    else:
        log("generating 32 cuboid jobs in %s" % scene_dir)
        paths = scenes.batch_jobs(scene_dir, n_jobs=32, seed=2)
        with open(marker, "w") as f:
            json.dump(paths, f)
    t0 = time.perf_counter()
    scs = []
    for i, pth in enumerate(paths):
        scs.append(ha.Scene(pth))
        log("loaded job %d: %d triangles" % (i, scs[-1].info()["num_triangles"]))
    load_s = time.perf_counter() - t0
    infos = [sc.info() for sc in scs]
    films = [ha.DeviceFilms(i["width"], i["height"]) for i in infos]
    jobs = [(sc, j, 0, i["num_samples"], f) for j, (sc, i, f) in enumerate(zip(scs, infos, films))]
    paths_total = sum(i["width"] * i["height"] * i["num_samples"] for i in infos)
    tuned = ha.autotune_batch([(sc, j, 0, 1, None) for sc, j, _, _, _ in jobs], trial_samples=1)
    log("autotuned launch: %s" % json.dumps(tuned))
    hip = Hip()
    for _ in range(args.warmup):
        ha.render_batch(jobs)
    hip.sync()
    t_start = time.perf_counter()
    for i in range(args.steps):
        t = time.perf_counter()
        for f in films:
            f.zero()
        ha.render_batch(jobs)
        log("batch step %d: %.3f s" % (i, time.perf_counter() - t))
    hip.sync()
    elapsed = time.perf_counter() - t_start
    t = time.perf_counter()
    for sc, seed, first, n, f in jobs:
        f.zero()
        sc.set_launch_ex(**{k: v for k, v in tuned.items()})
        sc.render_pt_device(seed, first, n, *f.ptrs)
    hip.sync()
    seq = time.perf_counter() - t
    line = {
        "metric": "paths/s over a batch of 32 render jobs (one path pool)",
        "value": paths_total * args.steps / elapsed,
        "unit": "paths/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "batch_32_cuboid_jobs", "jobs": len(jobs), "paths": paths_total,
                   "triangles": sum(i["num_triangles"] for i in infos),
                   "pixels": sum(i["width"] * i["height"] for i in infos), "launch": tuned,
                   "parallelism": "single-gpu"},
        "sequential_paths_per_s": paths_total / seq,
        "load_s": load_s,
    }
    for f in films:
        f.close()
    print(json.dumps(line), flush=True)


def bench_guide(args):
    """Equal-time relMSE, guided vs plain GPU tracer, on the metal-ring box
    (dispersive prism, GGX gold torus, 1% ceiling light) at 1024^2 with
    4096 spp against a 65536-spp reference by default (or the indirectly lit
    room at 512^2, 256 spp, 16384-spp reference), max depth 32.  The
    reference is a plain render of --guide-ref-spp samples; both
    methods render --guide-spp samples; the guided one trains online after
    groups of its first --guide-train-spp sample indices (one pass of Adam
    steps of --guide-minibatch samples per group of ~2^20 paths), then renders the rest
    with the trained guide (training time included).  Since relMSE falls as
    1 / spp, time to reach the plain render's error with the guide =
    T_guided * relMSE_guided / relMSE_plain; value = T_plain / that time."""
    import numpy as np

    import hastyrenderer_amd as ha
    from hastyrenderer_amd import scenes

    scene_dir = args.scene_dir or os.path.join(tempfile.gettempdir(),
                                               "hasty_bench_%s%d" % (args.guide_scene, args.guide_res))
    job_path = os.path.join(scene_dir, "renderJob.json")
    if not os.path.exists(job_path):
        log("generating %s scene in %s" % (args.guide_scene, scene_dir))
        if args.guide_scene == "ring":
            scenes.write_caustic_scene(scene_dir, width=args.guide_res, height=args.guide_res, passes=64, max_depth=32)
        else:  # Cornell room whose light faces the ceiling: lit indirectly
            scenes.write_cornell(scene_dir, width=args.guide_res, height=args.guide_res, spp=64, light_up=True,
                                 max_depth=32)
    hip = Hip()

    def render(sc, seed, n, chunk=0):
        """n samples into fresh films (in calls of `chunk` samples with a
        progress line each when chunk > 0); returns the mean image and the time."""
        info = sc.info()
        f = ha.DeviceFilms(info["width"], info["height"])
        hip.sync()
        t = time.perf_counter()
        step = chunk if chunk > 0 else n
        for first in range(0, n, step):
            sc.render_pt_device(seed, first, min(step, n - first), *f.ptrs)
            if chunk > 0:
                hip.sync()
                log("  %d / %d samples, %.1f s" % (min(first + step, n), n, time.perf_counter() - t))
        hip.sync()
        dt = time.perf_counter() - t
        c = f.download()[0] / n
        f.close()
        return c, dt

    def relmse(img, ref):
        return float(np.mean((img - ref) ** 2 / (ref ** 2 + 1e-2)))

    def relmse_trimmed(img, ref):
        """relMSE over the pixels outside the worst 0.1% (how much of the
        error sits in a few outlier pixels)."""
        e = np.sort(np.mean((img - ref) ** 2 / (ref ** 2 + 1e-2), axis=-1).ravel())
        return float(np.mean(e[: int(e.size * 0.999)]))

    plain = ha.Scene(job_path)
    info = plain.info()
    if args.guide_ref_file and os.path.exists(args.guide_ref_file):
        ref = np.load(args.guide_ref_file)
        log("reference: %s" % args.guide_ref_file)
    else:
        ref, tref = render(plain, 101, args.guide_ref_spp, chunk=1024)
        log("reference: %d spp in %.1f s" % (args.guide_ref_spp, tref))
        if args.guide_ref_file:
            np.save(args.guide_ref_file, ref)
    n = args.guide_spp
    chunk = 1024 if n > 1024 else 0  # progress lines for long renders
    pc, tp = render(plain, 5, n, chunk=chunk)
    ep = relmse(pc, ref)
    log("plain: %d spp in %.3f s, relMSE %.5f" % (n, tp, ep))
    gsc = ha.Scene(job_path)
    g = ha.Guide(gsc, seed=0)
    if args.guide_selection:
        g.set_selection(*args.guide_selection)
    ntrain = min(args.guide_train_spp, n)
    f = ha.DeviceFilms(info["width"], info["height"])
    hip.sync()
    t = time.perf_counter()
    gsc.set_guide_training(minibatch=args.guide_minibatch, weight_cap=args.guide_weight_cap)
    gsc.set_guide_photons(args.guide_photons, args.guide_photon_share)
    gsc.set_guide_depth(args.guide_max_depth, args.guide_max_throughput)
    gsc.set_guide(g, fraction=args.guide_fraction, train_lr=args.guide_lr)  # online training
    gsc.render_pt_device(5, 0, ntrain, *f.ptrs)
    gsc.set_guide(g, fraction=args.guide_fraction, train_lr=0.0)  # trained guide for the remaining samples
    step = chunk if chunk > 0 else n
    for first in range(ntrain, n, step):
        gsc.render_pt_device(5, first, min(step, n - first), *f.ptrs)
        if chunk > 0:
            hip.sync()
            log("  guided %d / %d samples, %.1f s" % (min(first + step, n), n, time.perf_counter() - t))
    hip.sync()
    tg = time.perf_counter() - t
    gc = f.download()[0] / n
    f.close()
    eg = relmse(gc, ref)
    if args.guide_save:
        np.save(args.guide_save + "_plain.npy", pc)
        np.save(args.guide_save + "_guided.npy", gc)
    loss, ns = gsc.guide_stats()
    log("guided (%d spp online training + %d spp): %.3f s, relMSE %.5f, last loss %.4f on %d samples" % (
        ntrain, n - ntrain, tg, eg, loss, ns))
    t_equal = tg * eg / ep
    line = {
        "metric": "equal-time relMSE gain of the learned guide over the plain GPU tracer",
        "value": tp / t_equal,
        "unit": "x (plain time / guided time to equal relMSE)",
        "n_gpus": 1,
        "steps": 1,
        "warmup": 0,
        "ms_per_step": tg * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 path tracing, bf16 MLP weights",
        "data": "synthetic",
        "config": {"workload": "metal_ring_box_guided" if args.guide_scene == "ring" else "indirect_room_guided",
                   "width": info["width"], "height": info["height"],
                   "max_depth": info["max_depth"], "triangles": info["num_triangles"], "spp": n,
                   "reference_spp": args.guide_ref_spp, "train_spp": ntrain,
                   "guide_fraction": (("learned in [%g, %g]" % tuple(args.guide_selection)) if args.guide_selection
                                      else args.guide_fraction),
                   "minibatch": args.guide_minibatch, "weight_cap": args.guide_weight_cap,
                   "photons_per_step": args.guide_photons, "photon_share": args.guide_photon_share,
                   "guided_depths": args.guide_max_depth, "guided_max_throughput": args.guide_max_throughput,
                   "adam_lr": args.guide_lr,
                   "parallelism": "single-gpu"},
        "plain": {"seconds": tp, "relmse": ep, "relmse_without_worst_0.1pct_pixels": relmse_trimmed(pc, ref)},
        "guided": {"seconds": tg, "relmse": eg, "equal_error_seconds": t_equal,
                   "relmse_without_worst_0.1pct_pixels": relmse_trimmed(gc, ref)},
    }
    # the same gain with the worst 0.1 % of pixels left out of both errors (a
    # few single-pixel outliers decide the untrimmed figure of a run)
    line["value_without_worst_0.1pct_pixels"] = tp / (tg * line["guided"]["relmse_without_worst_0.1pct_pixels"] /
                                                     line["plain"]["relmse_without_worst_0.1pct_pixels"])
    print(json.dumps(line), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--spp", type=int, default=256)
    ap.add_argument("--scene-dir", default=None)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=0,
                    help="host threads of the CPU baseline (0: the CPUs this process may use, capped by "
                         "OMP_NUM_THREADS when set)")
    ap.add_argument("--cpu-spp", type=int, default=48)
    ap.add_argument("--workload", choices=("ggx", "sppm", "batch", "guide"), default="ggx")
    ap.add_argument("--tune-spp", type=int, default=0,
                    help="samples per autotuner trial before timing (0: the library's default launch, which is "
                         "what the autotuner picks on this scene, without the run-to-run noise of short trials)")
    ap.add_argument("--launch", default="", metavar="SLOTS,TRACE,SHADE,CHUNK,GRAPH[,WIDTH]",
                    help="explicit launch shape (path slots, trace blocks, shade blocks, bounces per check, "
                         "hipGraph 0/1, BVH node width 4/8) instead of autotuning")
    ap.add_argument("--guide-spp", type=int, default=None,
                    help="samples of both renders (default: 4096 on the ring, 256 on the room)")
    ap.add_argument("--guide-res", type=int, default=None, help="image side (default: 1024 ring, 512 room)")
    ap.add_argument("--guide-train-spp", type=int, default=None,
                    help="sample indices rendered with online training (default: 48 ring, 16 room)")
    ap.add_argument("--guide-scene", choices=("ring", "room"), default="ring",
                    help="ring: the metal-ring caustic box; room: a Cornell room lit from behind its light")
    ap.add_argument("--guide-lr", type=float, default=1e-3, help="Adam learning rate of the online training")
    ap.add_argument("--guide-minibatch", type=int, default=None,
                    help="training samples per Adam step (default 65536 ring, 16384 room)")
    ap.add_argument("--guide-weight-cap", type=float, default=None,
                    help="cap of the normalised training weights (default 100 ring, 20 room)")
    ap.add_argument("--guide-photons", type=int, default=None,
                    help="caustic light paths per training step as extra training samples "
                         "(default 4194304 ring, 0 room: its light faces the ceiling, no caustics)")
    ap.add_argument("--guide-photon-share", type=float, default=0.5,
                    help="photon samples' weight sum relative to the camera samples'")
    ap.add_argument("--guide-max-depth", type=int, default=255, help="guide only vertices of depth below this")
    ap.add_argument("--guide-fraction", type=float, default=0.5, help="probability of sampling the guide at a vertex")
    ap.add_argument("--guide-selection", type=lambda v: tuple(float(x) for x in v.split(",") if float(x) > 0.0),
                    default=None, metavar="LO,HI",
                    help="learned guide/BSDF selection probability in [LO, HI], replacing --guide-fraction "
                         "(default 0.05,0.9 ring; 0: the fixed fraction, the room's default)")
    ap.add_argument("--guide-max-throughput", type=float, default=3.0e38,
                    help="guide a vertex only while the path throughput is at most this")
    ap.add_argument("--guide-save", default="", help="save the plain and guided images as PREFIX_plain.npy / _guided.npy")
    ap.add_argument("--guide-ref-file", default="",
                    help="reference image (.npy): loaded when it exists, else rendered and saved there")
    ap.add_argument("--guide-ref-spp", type=int, default=None,
                    help="reference samples (default: 65536 ring, 16384 room)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="also run a plain copy of MIB MiB (counter calibration)")
    ap.add_argument("--sppm-alpha", type=float, default=0.4)
    ap.add_argument("--sppm-store-fraction", type=float, default=0.5)
    ap.add_argument("--sppm-emit-fraction", type=float, default=10.0)
    ap.add_argument("--sppm-fixed-flux", action="store_true",
                    help="int64 fixed-point per-pixel flux (order-independent sums)")
    ap.add_argument("--sppm-initial-radius", type=float, default=0.0,
                    help="constant initial radius (0: 10x pixel footprint as the reference)")
    args = ap.parse_args()

    if args.cpu_threads <= 0:
        args.cpu_threads = host_cpus()["share"]
    if args.gpus != 1:
        print("bench.py measures one GPU only: this path tracer renders on a single MI355X "
              "(multi-GPU sharding is a design note, not a measured configuration); run with --gpus 1.",
              file=sys.stderr)
        sys.exit(2)

    if args.workload == "batch":
        bench_batch(args)
        return
    if args.workload == "guide":
        ring = args.guide_scene == "ring"
        if args.guide_minibatch is None:
            args.guide_minibatch = 65536 if ring else 16384
        if args.guide_weight_cap is None:
            args.guide_weight_cap = 100.0 if ring else 20.0
        if args.guide_photons is None:
            args.guide_photons = 4194304 if ring else 0
        if args.guide_selection is None:
            args.guide_selection = (0.05, 0.9) if ring else ()
        if args.guide_spp is None:
            args.guide_spp = 4096 if ring else 256
        if args.guide_train_spp is None:
            args.guide_train_spp = 48 if ring else 16
        if args.guide_res is None:
            args.guide_res = 1024 if ring else 512
        if args.guide_ref_spp is None:
            args.guide_ref_spp = 65536 if ring else 16384
        bench_guide(args)
        return
    if args.workload == "sppm":
        bench_sppm(args)
        if args.copy_calibration:
            copy_calibration(args.copy_calibration)
        return

    import hastyrenderer_amd as ha
    from hastyrenderer_amd import scenes

    scene_dir = args.scene_dir or os.path.join(tempfile.gettempdir(), "hasty_bench_ggx")
    job_path = os.path.join(scene_dir, "renderJob.json")
    if not os.path.exists(job_path):
        log("generating GGX grid scene in %s" % scene_dir)
        scenes.write_ggx_grid(scene_dir)
    seed = 1

    t0 = time.perf_counter()
    scene = ha.Scene(job_path)
    load_s = time.perf_counter() - t0
    info = scene.info()
    W, H = info["width"], info["height"]
    spp = args.spp
    log("scene loaded in %.2f s: %d triangles, BVH %d nodes depth %d, %.1f MB in HBM" % (
        load_s, info["num_triangles"], info["bvh_nodes"], info["bvh_depth"], info["device_bytes"] / 1e6))

    # launch shape: given (--launch), else what the autotuner picks from short
    # renders of this scene (untimed), else the default
    if args.launch:
        keys = ("path_slots", "trace_blocks", "shade_blocks", "chunk", "use_graph", "bvh_width")
        scene.set_launch_ex(**dict(zip(keys, (int(x) for x in args.launch.split(",")))))
        tuned = scene.get_launch()
    elif args.tune_spp > 0:
        tuned = scene.autotune(seed=seed, trial_samples=args.tune_spp)
    else:
        tuned = scene.get_launch()
    log("launch: %s" % json.dumps(tuned))
    hip = Hip()
    nbytes = W * H * 3 * 4
    films = [hip.malloc(nbytes) for _ in range(3)]

    def render():
        for f in films:
            hip.memset(f, nbytes)
        scene.render_pt_device(seed, 0, spp, *films)

    for i in range(args.warmup):
        t = time.perf_counter()
        render()
        log("warmup %d: %.3f s" % (i, time.perf_counter() - t))

    # timed steps: HIP events around the extension launches only (the
    # roofline's kernel), so that the events of the other launches add
    # nothing to the timed region
    scene.reset_stats()
    scene.set_profiling(events=True, counters=False, extend_only=True)
    hip.sync()
    t_start = time.perf_counter()
    for i in range(args.steps):
        t = time.perf_counter()
        render()
        log("step %d: %.3f s" % (i, time.perf_counter() - t))
    hip.sync()
    elapsed = time.perf_counter() - t_start
    kstats_timed = scene.kernel_stats()
    # per-kernel breakdown: one more, untimed render with every launch timed
    scene.set_profiling(events=True, counters=False)
    scene.reset_stats()
    render()
    kstats = scene.kernel_stats()
    scene.set_profiling(events=False, counters=True)
    scene.reset_stats()
    render()  # counted, untimed: rays per path, node visits, triangle tests
    counters = scene.counters()
    scene.set_profiling(events=False, counters=False)

    paths = W * H * spp
    value = paths * args.steps / elapsed
    rays_per_path = (counters["extension_rays"] + counters["shadow_rays"]) / paths
    log("paths/s %.4g, rays/path %.3f, kernels: %s" % (value, rays_per_path, json.dumps(kstats)))

    base = None
    canon = None
    if not args.no_cpu_baseline:
        log("cpu baseline: oracle, %d threads" % args.cpu_threads)
        base, canon = cpu_baseline(job_path, seed, args.cpu_threads, args.cpu_spp)
        log("cpu baseline %.4g paths/s" % base["value"])

    roofline = None
    launches, ext_ms = kstats_timed.get("pt_extend", (0, 0.0))
    if canon and launches:
        b_ray = (64.0 * canon["node_visits"] + 48.0 * canon["tri_tests"]) / canon["rays"]
        ext_rays_per_render = counters["extension_rays"]
        rays_per_launch = ext_rays_per_render / (launches / args.steps)
        bytes_per_launch = rays_per_launch * (b_ray + STATE_BYTES_PER_RAY)
        avg_s = ext_ms / launches / 1e3
        achieved = bytes_per_launch / avg_s / 1e9
        roofline = {
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": measured_traffic(),
            "traffic_source": "rocprofv3 FETCH_SIZE/WRITE_SIZE of an earlier run of this command, read from "
                              "profiles/pt_extend_ggx_grid_1024_traffic.md (not measured in this run)",
            "kernel": "pt_extend",
            "bytes_per_ray_canonical": round(b_ray, 1),
            "rays_per_launch": round(rays_per_launch, 1),
            "avg_launch_ms": round(avg_s * 1e3, 4),
        }

    line = {
        "metric": "paths/sec at 1024^2 (GGX sphere grid, 256 spp, max depth 16)",
        "value": value,
        "unit": "paths/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "ggx_sphere_grid",
            "width": W,
            "height": H,
            "spp": spp,
            "max_depth": info["max_depth"],
            "triangles": info["num_triangles"],
            "environment": "2048x1024 lat-long, importance sampled",
            "camera": "thin lens f/2.8, 8 blades",
            "launch": tuned,
            # the timed steps record HIP events around the extension launches,
            # and a render with events launches eagerly (no hipGraph replay)
            "launch_timed": dict(tuned, use_graph=0),
            "parallelism": "single-gpu",
        },
        "rays_per_s": value * rays_per_path,
        "rays_per_path": rays_per_path,
        # traversal work per ray of each kind (the counted render): shadow
        # rays toward the sky find no blocker and visit every node box they
        # cross, extension rays stop shrinking their interval at the hit
        "traversal": {
            "extension_rays": counters["extension_rays"],
            "shadow_rays": counters["shadow_rays"],
            "nodes_per_extension_ray": round((counters["node_visits"] - counters["shadow_node_visits"])
                                             / max(1, counters["extension_rays"]), 2),
            "tris_per_extension_ray": round((counters["triangle_tests"] - counters["shadow_triangle_tests"])
                                            / max(1, counters["extension_rays"]), 2),
            "nodes_per_shadow_ray": round(counters["shadow_node_visits"] / max(1, counters["shadow_rays"]), 2),
            "tris_per_shadow_ray": round(counters["shadow_triangle_tests"] / max(1, counters["shadow_rays"]), 2),
        },
        "load_s": load_s,
        "kernel_ms": {k: round(v[1] / max(1, v[0]), 4) for k, v in kstats.items()},
        "roofline": roofline,
        "cpu_baseline": base,
    }
    for f in films:
        hip.free(f)
    scene.close()
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
