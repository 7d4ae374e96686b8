"""Benchmark: Mrays/s of the wavefront path tracer on one MI355X.

Default workload, the sphere field: 64 level-6 icospheres displaced by 3-octave value
noise plus a ground plane (5,242,882 triangles), Principled materials, an
analytic sky importance-sampled through a 2048^2 envmap, 1920x1080, path
depth 12.  One step = one sample per pixel (2,073,600 paths) traced through
all bounces.  Mrays/s counts closest-hit plus shadow rays traced, the
reference's own accounting (host/Renderer.swift:190-196,249-253).

The same line carries `cache_vs_plain`, the metric's second half: the Cornell box
at 1920x1080 rendered by the plain tracer and by the tracer with the learned
radiance cache, RMSE against a 16384-spp FP64 reference at power-of-two spp,
and the cached tracer's time to the plain tracer's 64-spp RMSE over the plain
time (`time_ratio`).

`roofline` is trace_closest's: `achieved` is its algorithmic bytes (28 B ray +
20 B hit + 80 B per node visit + 48 B per triangle test) over its summed HIP-event
time; `traffic` is its HBM bytes per launch from the FETCH_SIZE / WRITE_SIZE
counter runs recorded in profiles/trace_closest_c2_traffic.json, and
`hbm_measured` the same bytes over this run's time, the rate the HBM actually
delivers.

`--workload graphs` runs the shader-graph workload instead: the Cornell-box geometry
with 32 random node-graph materials, depth 8, shading sorted by material.
`--workload instances` renders a 16 x 16 field of instances of 4 shapes (5.2M
triangles placed, 82k stored) through the two-level acceleration structure.

Prints one JSON line.  `--gpus` above 1 is rejected: the tracer runs on one GPU.
"""
import argparse
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

HBM_PEAK_GBS = 8000.0
NODE_BYTES, TRI_BYTES = 80, 48
RAY_READ, HIT_WRITE = 28, 20  # origin + direction + meta word; t, prim, inst, 2 barycentrics


def copy_probe(mib):
    """One device-to-device copy of `mib` MiB (copy_probe_kernel reads and
    writes mib MiB), the reference a hardware byte counter is scaled by."""
    import raymond_amd.api as api
    dev = api.Device(0)
    ms = dev.copy_probe(mib << 20)
    print("copy probe: %d MiB read and written in %.3f ms (%.0f GB/s)" % (mib, ms, 2 * (mib << 20) / ms / 1e6),
          file=sys.stderr, flush=True)
    dev.close()


def workload_name(graphs, W, H, D, sort, cornell=False, instances=False):
    if instances:
        return ("instanced field: 16x16 instances of 4 noise-displaced level-5 icospheres + ground, "
                "5,242,882 triangles placed, %s, analytic sky, %dx%d, depth %d, plain tracer, 1 spp per step"
                % ("flat tree over all of them" if instances == "flat" else
                   "81,922 stored, two-level acceleration structure", W, H, D))
    if cornell:
        return ("Cornell box: 36 triangles, emitter 15, %dx%d, depth %d, roulette from bounce 3, "
                "plain tracer, 1 spp per step" % (W, H, D))
    if graphs:
        return ("shader graphs: Cornell-box geometry, 32 random 20-60 node Cycles graphs (seed 3) "
                "assigned per triangle, %dx%d, depth %d, plain tracer, 1 spp per step%s" %
                (W, H, D, ", shading sorted by material" if sort else ""))
    return ("sphere field: 64 noise-displaced level-6 icospheres + ground, 5,242,882 "
            "triangles, analytic sky, %dx%d, depth %d, plain tracer, 1 spp per step" % (W, H, D))


def bench_batch(args):
    """Batched uneven jobs: 64 Cornell-box jobs (seed 4) with 1k-2M triangles,
    256^2-1024^2 pixels, 16-256 spp and depth 6-12, rendered as one batch: one
    regenerating bounce loop over all jobs' samples, queue depth and material
    sort picked by the autotuner.  Reports aggregate Mrays/s and each job's
    completion time (its makespan from the batch start)."""
    import numpy as np
    import raymond_amd.api as api
    import raymond_amd.scenes as scenes

    t0 = time.time()
    specs = scenes.batch_jobs(args.jobs, seed=4)
    dev = api.Device(0)
    built = [scenes.build_device_scene(dev, d) for d, *_ in specs]
    batch = api.Batch(dev, [(s, w, h, spp, depth, 0) for s, (_, w, h, spp, depth) in zip(built, specs)])
    setup_s = time.time() - t0
    tris = sum(d.num_triangles for d, *_ in specs)
    paths = sum(w * h * spp for _, w, h, spp, _ in specs)
    print("setup %.1fs: %d jobs, %d triangles, %.2f G samples" % (setup_s, len(specs), tris, paths / 1e9),
          file=sys.stderr, flush=True)
    t1 = time.time()
    cfg, tried, hit = batch.autotune(rr_depth=3)
    print("autotune %.1fs (cache %s): queue %d, sort %d, trace grid %d, sort shift %d; tried %s"
          % (time.time() - t1, hit, cfg.queue_paths, cfg.material_sort, cfg.trace_blocks_per_cu, cfg.sort_key_shift,
             tried), file=sys.stderr, flush=True)
    # the bounce loop replays captured hipGraph chunks (same speed as direct
    # launches at these queue depths: 2999 against 2998 Mrays/s)
    st = batch.render(queue_paths=cfg.queue_paths, material_sort=cfg.material_sort, rr_depth=3, timing=0,
                      use_graph=1, order=api.BATCH_SHORTEST_FIRST, trace_blocks_per_cu=cfg.trace_blocks_per_cu,
                      sort_key_shift=cfg.sort_key_shift)
    secs = st.ms_total / 1e3
    mrays = st.total_rays / secs / 1e6
    times = batch.job_times()
    work = np.array([w * h * spp for _, w, h, spp, _ in specs], np.float64)
    print("batch: %.2fs, %.1f Mrays/s, %d loops, makespan %.1f ms" % (secs, mrays, st.loops, st.makespan_ms),
          file=sys.stderr, flush=True)
    out = {
        "metric": "Mrays/s + RMSE vs high-spp reference at equal time, cache vs plain",
        "value": round(mrays, 2),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "higher_is_better": True,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "batched uneven jobs: %d Cornell-box jobs (seed 4), %d triangles in "
                               "total (1k-2M each), 256^2-1024^2 pixels, 16-256 spp, depth 6-12, roulette from "
                               "bounce 3, one regenerating bounce loop replayed from captured hipGraphs" %
                               (len(specs), tris),
                   "jobs": len(specs), "triangles": tris, "samples": paths,
                   "queue_paths": cfg.queue_paths, "material_sort": bool(cfg.material_sort), "hipgraph": True,
                   "trace_blocks_per_cu": cfg.trace_blocks_per_cu, "sort_key_shift": cfg.sort_key_shift,
                   "order": "shortest job first",
                   "autotune": {"cache_hit": hit, "candidates": [
                       {"queue_paths": q, "material_sort": m, "trace_blocks_per_cu": tb, "sort_key_shift": sh,
                        "mrays_per_s": round(r, 1)} for q, m, tb, sh, r in tried]}},
        "seconds": round(secs, 3),
        "samples_per_s": round(st.samples / secs, 1),
        "bounce_loops": st.loops,
        "makespan_ms": round(st.makespan_ms, 1),
        "job_completion_ms": {"min": round(float(times.min()), 1), "median": round(float(np.median(times)), 1),
                              "max": round(float(times.max()), 1)},
        "jobs": [{"pixels": int(w * h), "spp": spp, "depth": depth, "triangles": d.num_triangles,
                  "done_ms": round(float(t), 1)} for (d, w, h, spp, depth), t in zip(specs, times)],
        "setup_s": round(setup_s, 1),
    }
    print(json.dumps(out), flush=True)


def queue_frames(args, W, H):
    """Samples per pixel in flight: about 16M paths, so that every
    bounce-loop iteration fills the GPU and the drain of the last paths is
    spread over many samples (1080p: 8 frames; 512x512: 64)."""
    return args.spp_per_launch if args.spp_per_launch > 0 else max(1, (1 << 24) // (W * H))


def cache_vs_plain(args, dev=None):
    """Radiance cache against the plain tracer: the Cornell box at full HD,
    depth 12, roulette from bounce 3.  A converged FP64 render of the plain
    tracer is the reference; the plain tracer and the cached tracer (online
    training on 1 path in 2^cache_train_shift, cold start) each render `cache_spp` spp, and the
    RMSE against the reference is recorded at power-of-two sample counts with
    the time spent so far.  Returns the comparison as a dict, with the
    cached tracer's time to the plain tracer's final RMSE over the plain time
    (`time_ratio`)."""
    import numpy as np
    import raymond_amd.api as api
    import raymond_amd.scenes as scenes

    W, H = (args.width, args.height) if args.workload == "cache" else (1920, 1080)
    D = (args.depth or 12) if args.workload == "cache" else 12
    spp = args.steps if args.workload == "cache" else args.cache_spp
    dev = dev or api.Device(0)
    desc = scenes.cornell_box(W, H)
    desc.envmap_resolution = 64  # black world
    scene = scenes.build_device_scene(dev, desc)

    t0 = time.perf_counter()
    ref_r = api.Renderer(scene, W, H, max_depth=D, samples_per_launch=4)
    ref_r.set_uniforms(rr_depth=3, accumulate_f64=1)
    st = ref_r.execute(args.ref_spp)
    ref_s = time.perf_counter() - t0
    ref = (ref_r.image_f64()[..., :3] / args.ref_spp).astype(np.float64)
    del ref_r
    print("reference: %d spp in %.1fs (%.1f Mrays/s)" % (args.ref_spp, ref_s, st.total_rays / ref_s / 1e6),
          file=sys.stderr, flush=True)

    spl = queue_frames(args, W, H)

    def run(cached):
        r = api.Renderer(scene, W, H, max_depth=D, samples_per_launch=spl)
        r.set_uniforms(rr_depth=3)
        cache = None

        def new_cache():
            return api.RadianceCache(dev, log2_table_size=args.cache_log2, seed=1, learning_rate=args.cache_lr,
                                     train_shift=args.cache_train_shift, batch_size=args.cache_batch,
                                     steps_per_launch=args.cache_steps)

        if cached:
            cache = new_cache()
            r.set_cache(cache, mode=1)
        r.execute(1)  # warm-up (first launches, cache training starts)
        r.reset()
        if cache is not None:
            cache.close()
            cache = new_cache()
            r.set_cache(cache, mode=1)
        points, done, spent, rays = [], 0, 0.0, 0
        k = 1
        while done < spp:
            n = min(k, spp) - done
            t = time.perf_counter()
            s = r.execute(n)
            spent += time.perf_counter() - t
            rays += s.total_rays
            done += n
            img = r.image()[..., :3].astype(np.float64) / done
            points.append({"spp": done, "seconds": round(spent, 4),
                           "rmse": float(np.sqrt(np.mean((img - ref) ** 2)))})
            k *= 2
        out = {"points": points, "mrays_per_s": round(rays / spent / 1e6, 2),
               "samples_per_s": round(W * H * done / spent, 1)}
        if cache is not None:
            cs = cache.stats()
            out["cache"] = {"steps": int(cs.steps), "train_samples": int(cs.samples), "queries": int(cs.queries),
                            "last_loss": round(float(cs.last_loss), 5)}
        return out

    plain = run(False)
    print("plain: %s" % plain["points"][-1], file=sys.stderr, flush=True)
    cached = run(True)
    print("cache: %s" % cached["points"][-1], file=sys.stderr, flush=True)

    def time_to(points, target):
        """Seconds until the RMSE first reaches `target` (log-log interpolation)."""
        prev = None
        for p in points:
            if p["rmse"] <= target:
                if prev is None:
                    return p["seconds"]
                a, b = np.log(prev["rmse"]), np.log(p["rmse"])
                f = (np.log(target) - a) / (b - a) if b != a else 1.0
                return float(np.exp(np.log(prev["seconds"]) + f * (np.log(p["seconds"]) - np.log(prev["seconds"]))))
            prev = p
        return None

    target = plain["points"][-1]["rmse"]
    t_plain = plain["points"][-1]["seconds"]
    t_cache = time_to(cached["points"], target)
    return {
        "workload": "radiance cache: Cornell box %dx%d, depth %d, roulette from bounce 3, hash grid 16x2 FP16 x "
                    "2^%d, MLP 48-64-64-3 FP16 on MFMA, Adam %g, 1 path in %d trains, %d training steps of %d "
                    "samples per spp, %d spp each; reference: plain tracer, %d spp, FP64 accumulation"
                    % (W, H, D, args.cache_log2, args.cache_lr, 1 << args.cache_train_shift, args.cache_steps,
                       args.cache_batch, spp, args.ref_spp),
        "spp": spp,
        "plain": plain,
        "cached": cached,
        "plain_rmse": target,
        "cached_rmse_at_equal_spp": cached["points"][-1]["rmse"],
        "plain_seconds": t_plain,
        "cache_seconds_to_plain_rmse": None if t_cache is None else round(t_cache, 4),
        "time_ratio": None if t_cache is None else round(t_cache / t_plain, 4),
        "target_time_ratio": 0.5,
    }


def bench_cache(args):
    """The cache comparison as the headline line (`--workload cache`)."""
    c = cache_vs_plain(args)
    out = {
        "metric": "Mrays/s + RMSE vs high-spp reference at equal time, cache vs plain",
        "value": c["cached"]["mrays_per_s"],
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "higher_is_better": True,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": c["workload"]},
        "plain": c["plain"],
        "cached": c["cached"],
        "plain_rmse_at_steps": c["plain_rmse"],
        "cache_seconds_to_plain_rmse": c["cache_seconds_to_plain_rmse"],
        "time_ratio_cache_vs_plain": c["time_ratio"],
    }
    print(json.dumps(out), flush=True)


def host_cpu():
    """CPU model and core counts of this host."""
    model = None
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
    # the host cores this job may use: the box allots a share of its CPUs to
    # each GPU (OMP_NUM_THREADS names it), which bounds any worker pool here
    share = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or allowed
    return {"model": model, "logical_cpus": os.cpu_count(), "affinity_cpus": allowed,
            "job_share": min(share, allowed)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=64)
    ap.add_argument("--warmup", type=int, default=4)
    ap.add_argument("--spp-per-launch", type=int, default=0,
                    help="path-queue capacity in frames (0: about 16M paths in flight, at least one frame)")
    ap.add_argument("--width", type=int, default=1920)
    ap.add_argument("--height", type=int, default=1080)
    ap.add_argument("--depth", type=int, default=0, help="path depth (0 = the workload's: 12 spheres, 8 graphs)")
    ap.add_argument("--workload", choices=["spheres", "cornell", "graphs", "cache", "batch", "instances"],
                    default="spheres",
                    help="spheres: 5.2M-triangle sphere field; cornell: the 36-triangle Cornell box at 512^2; "
                         "graphs: 32 random node-graph materials; "
                         "cache: learned radiance cache against the plain tracer on the Cornell box; "
                         "batch: 64 uneven Cornell-box jobs in one bounce loop; "
                         "instances: 16x16 instances of 4 shapes through the two-level tree")
    ap.add_argument("--flat", action="store_true",
                    help="instances workload: build the flat tree over the placed triangles instead (for comparison)")
    ap.add_argument("--scene", default="",
                    help="render a scene file of the Blender exporter (JSON + PLY meshes) instead of a workload; "
                         "its render resolution applies unless --width/--height are given")
    ap.add_argument("--jobs", type=int, default=64, help="batch workload: number of jobs")
    ap.add_argument("--ref-spp", type=int, default=16384, help="cache workload: spp of the FP64 reference")
    # cache configuration of the comparison, from a sweep on this workload: table 2^14-2^19 entries per level,
    # one path in 2^3-2^8 training, 1-4 steps of 2048-16384 samples per spp
    ap.add_argument("--cache-batch", type=int, default=4096, help="cache workload: samples per training step")
    ap.add_argument("--cache-steps", type=int, default=2, help="cache workload: training steps per sample")
    ap.add_argument("--cache-log2", type=int, default=15, help="cache workload: log2 of the hash-table entries per level")
    ap.add_argument("--cache-train-shift", type=int, default=7,
                    help="cache workload: one camera path in 2^shift trains")
    ap.add_argument("--cache-lr", type=float, default=1e-2, help="cache workload: Adam learning rate")
    ap.add_argument("--cache-spp", type=int, default=64,
                    help="spp of each side of the cache-vs-plain comparison in the default line")
    ap.add_argument("--no-cache-compare", action="store_true",
                    help="leave the cache-vs-plain comparison out of the default line")
    ap.add_argument("--material-sort", type=int, default=1,
                    help="group hits by material (misses in a bin of their own) before shading")
    ap.add_argument("--inline-materials", type=int, default=-1,
                    help="1: material graphs inlined into the shading kernel, 0: out-of-line calls, -1: by size")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU oracle baseline")
    ap.add_argument("--graph", type=int, default=0, help="replay the bounce loop from a captured hipGraph")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="first copy a buffer of this many MiB on the device (calibrates HBM counters)")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--cpu-pixels", type=int, default=0, help="oracle sample size (0 = full frames)")
    ap.add_argument("--cpu-frames", type=int, default=24, help="oracle sample: frames (spp) of the full image (about 10 s on 16 threads)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on exactly one GPU (the path tracer is single-GPU); "
              "--gpus %d is not supported" % args.gpus, file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import raymond_amd.api as api
    import raymond_amd.scenes as scenes

    if args.copy_probe:
        copy_probe(args.copy_probe)
    if args.workload == "cache":
        return bench_cache(args)
    if args.workload == "batch":
        return bench_batch(args)
    W, H = args.width, args.height
    loaded = None
    if args.scene:
        import raymond_amd.rayjay as rayjay
        t_load = time.time()
        loaded = rayjay.load_scene(args.scene)
        load_s = time.time() - t_load
        if loaded.resolution and "--width" not in sys.argv and "--height" not in sys.argv:
            W, H = loaded.resolution
        print("loaded %s in %.2fs: %d shapes, %d triangles" % (args.scene, load_s, len(loaded.shapes),
                                                             loaded.num_triangles), file=sys.stderr, flush=True)
    graphs = args.workload == "graphs"
    cornell = args.workload == "cornell"
    instances = args.workload == "instances"
    if cornell and (W, H) == (1920, 1080):
        W, H = 512, 512
    D = args.depth or (8 if graphs or cornell or instances else 12)
    # sorting pays for itself on every workload here: on the sphere field it
    # splits misses (sky) from hits and metals from dielectrics (shading
    # 0.53 -> 0.45 ms per spp for 0.05 ms of sorting)
    sort = int(args.material_sort != 0)
    # the Cornell-box workload runs roulette from bounce 3; the sphere field and graph
    # scenes the reference's way, from the first bounce
    rr_depth = 3 if cornell else 0
    t0 = time.time()
    if loaded is not None:
        desc = loaded
    elif graphs:
        desc = scenes.shader_graph_scene(W, H)
    elif cornell:
        desc = scenes.cornell_box(W, H)
        desc.envmap_resolution = 64  # black world
    elif instances:
        desc = scenes.instanced_field(W, H, grid=16, shapes=4, level=5)
    else:
        desc = scenes.sphere_field(W, H)
    dev = api.Device(0)
    # the graph scene inlines its materials (one compile, cached; faster shading)
    inline = args.inline_materials if args.inline_materials >= 0 else (1 if graphs else -1)
    noinline = None if inline < 0 else not inline
    scene = scenes.build_device_scene(dev, desc, noinline=noinline, two_level=instances and not args.flat)
    sst = scene.stats()
    setup_s = time.time() - t0
    print("setup %.1fs: %d triangles, BVH8 %d nodes depth %d (%.1fs), JIT %.1fs, envmap %.2fs" %
          (setup_s, sst.triangles, sst.bvh_nodes, sst.bvh_depth, sst.bvh_build_seconds, sst.jit_seconds,
           sst.envmap_seconds), file=sys.stderr, flush=True)

    spl = queue_frames(args, W, H)
    r = api.Renderer(scene, W, H, max_depth=D, samples_per_launch=spl)
    r.set_uniforms(timing=0 if args.graph else 1, rr_depth=rr_depth, material_sort=sort, use_graph=args.graph)
    if args.warmup:
        r.execute(args.warmup)
    # the timed steps run as one regenerating stream of `steps` samples per pixel
    t_start = time.perf_counter()
    st = r.execute(args.steps)
    wall = time.perf_counter() - t_start
    totals = dict(rays=st.total_rays, closest=st.closest_total, shadow=st.shadow_total,
                  ms_closest=st.ms_trace_closest, ms_shade=st.ms_shade, ms_any=st.ms_trace_any, ms_sort=st.ms_sort,
                  nclosest=st.trace_closest_launches, nany=st.trace_any_launches, iters=st.iterations)
    closest_total = st.closest_total
    mrays = totals["rays"] / wall / 1e6
    samples_per_s = W * H * args.steps / wall

    # traversal statistics pass (not timed): nodes and triangles per ray
    r.set_uniforms(traversal_stats=1, timing=0)
    sst2 = r.execute(1)
    closest_rays = sst2.closest_total
    nodes_per_ray = sst2.nodes_closest / max(closest_rays, 1)
    tris_per_ray = sst2.tris_closest / max(closest_rays, 1)
    bytes_per_ray = RAY_READ + HIT_WRITE + nodes_per_ray * NODE_BYTES + tris_per_ray * TRI_BYTES
    closest_bytes = closest_total * bytes_per_ray
    achieved = closest_bytes / (totals["ms_closest"] * 1e-3) / 1e9 if totals["ms_closest"] > 0 else 0.0
    print("timed %d steps in %.3fs: %.1f Mrays/s, %.1f Msamples/s; closest %.2f ms/step, sort %.2f, shade %.2f, "
          "any %.2f; nodes/ray %.1f tris/ray %.1f" % (args.steps, wall, mrays, samples_per_s / 1e6,
                                                      totals["ms_closest"] / args.steps, totals["ms_sort"] / args.steps,
                                                      totals["ms_shade"] / args.steps, totals["ms_any"] / args.steps,
                                                      nodes_per_ray, tris_per_ray),
          file=sys.stderr, flush=True)

    # HBM traffic of trace_closest: bytes per ray from the FETCH_SIZE /
    # WRITE_SIZE counter runs recorded for this workload, times this run's
    # closest rays per launch
    traffic = None
    traffic_file = os.path.join(HERE, "profiles", "trace_closest_c2_traffic.json")
    if loaded is None and not graphs and not cornell and not instances and os.path.exists(traffic_file):
        with open(traffic_file) as f:
            tr = json.load(f)
        # one closest-hit launch per bounce-loop iteration
        traffic = tr["traffic_bytes_per_ray"] * closest_total / max(totals["iters"], 1)
    # the measured HBM rate of the same launches: counter bytes over their time
    hbm_measured = None if traffic is None or totals["ms_closest"] <= 0 else \
        traffic * max(totals["iters"], 1) / (totals["ms_closest"] * 1e-3) / 1e9

    cpu = None
    if not args.no_cpu:
        from oracle import binding
        cpu_host = host_cpu()
        threads = args.cpu_threads or cpu_host["job_share"]
        orc = binding.OracleScene(desc)
        pdfs, mip = scene.read_envmap()
        orc.set_envmap(pdfs, mip)
        if args.cpu_pixels:
            rng = np.random.default_rng(11)
            pix = rng.choice(W * H, args.cpu_pixels, replace=False).astype(np.uint32)
            frames = 1
            sample = "%d random pixels of the %dx%d frame, 1 spp, depth %d" % (args.cpu_pixels, W, H, D)
        else:
            pix = None
            frames = args.cpu_frames
            sample = "%d full %dx%d frames (%d bench steps), 1 spp each, depth %d" % (frames, W, H, frames, D)
            if instances:
                sample += ", the oracle's flat BVH2 over the placed triangles"
        _, res = orc.render(W, H, D, frames=frames, frame_begin=10_000, pixels=pix, threads=threads, rr_depth=rr_depth)
        cpu = {"value": res.total_rays / res.seconds / 1e6, "unit": "Mrays/s", "cores": int(res.threads_used),
               "kind": "port", "sample": sample, "seconds": round(res.seconds, 2), "cpu_model": cpu_host["model"],
               "host_logical_cpus": cpu_host["logical_cpus"], "host_affinity_cpus": cpu_host["affinity_cpus"]}
        print("cpu oracle: %.3f Mrays/s on %d threads (%.1fs)" % (cpu["value"], cpu["cores"], res.seconds),
              file=sys.stderr, flush=True)

    out = {
        "metric": "Mrays/s + RMSE vs high-spp reference at equal time, cache vs plain",
        "value": round(mrays, 2),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall * 1e3 / args.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": workload_name(graphs, W, H, D, sort, cornell, ("flat" if args.flat else True) if instances
                                      else False) if loaded is None else
            "exporter scene %s: %d triangles, %dx%d, depth %d, plain tracer, 1 spp per step" %
            (os.path.basename(args.scene), int(sst.triangles), W, H, D),
            "width": W, "height": H, "max_depth": D, "queue_frames": spl, "triangles": int(sst.triangles),
            "iterations": int(totals["iters"]), "material_sort": bool(sort),
        },
        "samples_per_s": round(samples_per_s, 1),
        "rays": {"closest": int(totals["closest"]), "shadow": int(totals["shadow"]),
                 "closest_launches": int(totals["iters"])},
        "stage_ms_per_step": {"trace_closest": round(totals["ms_closest"] / args.steps, 3),
                              "material_sort": round(totals["ms_sort"] / args.steps, 3),
                              "shade": round(totals["ms_shade"] / args.steps, 3),
                              "trace_any_shadow": round(totals["ms_any"] / args.steps, 3)},
        "roofline": {"kernel": "trace_closest", "bound": "hbm", "achieved": round(achieved, 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": None if traffic is None else round(traffic),
                     "traffic_source": None if traffic is None else "profiles/trace_closest_c2_traffic.json",
                     "hbm_measured": None if hbm_measured is None else round(hbm_measured, 1),
                     "hbm_measured_frac": None if hbm_measured is None else round(hbm_measured / HBM_PEAK_GBS, 4),
                     "bytes_per_ray": round(bytes_per_ray, 1),
                     "nodes_per_ray": round(nodes_per_ray, 2), "tris_per_ray": round(tris_per_ray, 2)},
        "cpu_baseline": cpu,
    }
    if cpu:
        out["speedup_vs_cpu"] = round(mrays / cpu["value"], 1)
    if loaded is None and not graphs and not cornell and not instances and not args.no_cache_compare:
        # the metric's second half on the same device: the cache against the
        # plain tracer at equal time (the sphere field's buffers are freed first)
        del r, scene
        out["cache_vs_plain"] = cache_vs_plain(args, dev)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
