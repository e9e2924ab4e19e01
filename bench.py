"""Benchmark of the sorted wavefront renderer on one MI355X.

A step is one Pathtracer::process() iteration (every sample of the frame traced through the
sorted wavefront loop, reduced and filtered into the device framebuffer) of the plain method at
scale: a 1024 x 1024-quad height-field terrain (2,097,152 triangles + 2 light
triangles, 64 material groups mixing Lambert and GGX lobes, a sun-sized quad light and a
constant environment) rendered at 1920 x 1080, 64 spp, max depth 8, 2^23-ray waves.

Mrays/s = (closest-hit + shadow rays traced) / wall time of the timed steps. Scene upload and
BVH build happen before the timed region; the rate with the framebuffer copied to host memory
every step is reported beside it (framebuffer_to_host). Prints one JSON line.
"""
import argparse
import json
import os
import statistics
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "msc-project_amd"))

HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="terrain", choices=["terrain", "box_room", "batched", "serving"],
                    help="terrain: the plain render at scale (default); box_room: the CPU-runnable plain case; "
                         "batched: only the batched-vs-plain comparison, batched Mrays/s as the value; "
                         "serving: only the serving run, its Mrays/s as the value")
    ap.add_argument("--spp", type=int, default=None, help="override samples per pixel (must keep the workload name honest)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-samples", type=int, default=0, help="sample ids the CPU baseline traces (0 = auto)")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-kernels", action="store_true",
                    help="skip the sort/traversal kernel bench (16M ray keys, 16M hit keys, 16M rays on the terrain BVH)")
    ap.add_argument("--no-autotune", action="store_true", help="keep the default launch shapes")
    ap.add_argument("--no-stage-alone", action="store_true",
                    help="skip the extra frame with the shadow pass in line (stage_ms_alone; for kernel profiles)")
    ap.add_argument("--tuning", default=None,
                    help="fixed launch shapes instead of the autotuner: keys per thread, ray key bits, hit key bits, "
                         "stack ring entries (e.g. 16,16,16,16)")
    ap.add_argument("--no-batched", action="store_true", help="skip the batched-vs-plain comparison")
    ap.add_argument("--batch-jobs", type=int, default=256, help="jobs of the batched workload")
    ap.add_argument("--no-plain", action="store_true", help="batched workload: skip the job-by-job plain pass")
    ap.add_argument("--batch-paced-only", action="store_true",
                    help="batched workload: only the host-paced loop (no graph replays; for kernel profiles), "
                         "at --batch-wave-rays (default 2^27) with --batch-tuning or the default shapes")
    ap.add_argument("--batch-wave-rays", type=int, default=0,
                    help="rays per wave of the batched loop (0 = the batch autotuner's choice)")
    ap.add_argument("--no-serving", action="store_true", help="skip the serving-loop measurement")
    ap.add_argument("--serve-requests", type=int, default=1000, help="measured requests of the serving run")
    ap.add_argument("--serve-warmup", type=int, default=50, help="requests before the measured ones")
    ap.add_argument("--serve-rate", type=float, default=16.0,
                    help="Poisson arrival rate (requests/s); 16/s loads the device to about 63%%")
    ap.add_argument("--serve-backlog-waves", type=int, default=0,
                    help="serving: admit while fewer paths than this many waves wait (0 = library default)")
    ap.add_argument("--serve-waves-per-check", type=int, default=0,
                    help="serving: graph waves between host checks (0 = library default)")
    ap.add_argument("--serve-saturation", type=int, default=64,
                    help="serving: the first N requests submitted at once (served throughput at saturation), "
                         "the same iterations as one batch, and job by job through the plain loop (0 = skip)")
    ap.add_argument("--serve-state-gib", type=float, default=32.0,
                    help="device memory budget for request state (per-sample arenas + framebuffers)")
    ap.add_argument("--batch-tuning", default=None,
                    help="batched workload: fixed launch shapes (as --tuning, ring 0 = per scene from its BVH depth, at most 16) "
                         "instead of the batch autotuner")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="accepted for older profiling commands: the copy of known size (gather_kernel<4>, "
                         "identity permutation, 2^26 records: 1280 MiB read, 1024 MiB written) that scales "
                         "rocprof FETCH_SIZE/WRITE_SIZE now runs in every terrain or box-room run")
    return ap.parse_args()


def calibrate_copy(ctx):
    import numpy as np
    n = 1 << 26
    src = ctx.alloc(16 * n)
    perm = ctx.upload(np.arange(n, dtype=np.uint32))
    dst = ctx.alloc(16 * n)
    for _ in range(3):
        ctx.gather(src, perm, n, 4, dst)
    ms = ctx.last_device_ms()
    for a in (src, perm, dst):
        a.free()
    return {"records": n, "bytes_read": 20 * n, "bytes_written": 16 * n, "ms": round(ms, 4),
            "GB/s": round(36.0 * n / (ms * 1e-3) / 1e9, 1)}


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU: this build has no multi-GPU path (asked for --gpus {args.gpus}).",
              file=sys.stderr)
        sys.exit(2)
    import numpy as np

    import swr
    from swr import scenes

    if args.workload == "serving":
        ctx = swr.Context(0)
        res = serving(ctx, args)
        ctx.close()
        print(json.dumps({
            "metric": "Mrays/s on 1 MI355X (serving loop, Poisson arrivals); p50/p99 render latency",
            "value": res["mrays_s"], "unit": "Mrays/s", "n_gpus": 1, "steps": 1, "warmup": args.serve_warmup,
            "ms_per_step": res["wall_s"] * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
            "dtype": "f32", "data": "synthetic (seeded procedural scenes, no files)",
            "config": {"workload": "serving_4_scenes_poisson", "method": "served"}, "serving": res}), flush=True)
        return

    if args.workload == "batched":
        ctx = swr.Context(0)
        res = batched_vs_plain(ctx, args, plain=not args.no_plain)
        ctx.close()
        print(json.dumps({
            "metric": "Mrays/s on 1 MI355X (batched persistent loop; closest-hit + shadow rays)",
            "value": res["batched_mrays_s"], "unit": "Mrays/s", "n_gpus": 1, "steps": 1, "warmup": 1,
            "ms_per_step": res["batched_ms"], "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
            "dtype": "f32", "data": "synthetic (seeded procedural scenes, no files)",
            "config": {"workload": "batched_256_jobs_8_scenes_d6", "method": "batched"}, "batched": res}), flush=True)
        return

    t0 = time.time()
    if args.workload == "terrain":
        desc = scenes.terrain(spp=args.spp or 64)
        wname = "terrain_2M_1080p_64spp_d8"
    else:
        desc = scenes.box_room(base=int(round((args.spp or 16) ** 0.5)))
        wname = "box_room_20k_512sq_16spp_d5"
    if args.spp and args.workload == "terrain" and args.spp != 64:
        wname = f"terrain_2M_1080p_{args.spp}spp_d8"
    gen_s = time.time() - t0

    ctx = swr.Context(0)
    copy_cal = calibrate_copy(ctx)
    t0 = time.time()
    scene = swr.Scene(ctx, desc)
    build_s = time.time() - t0
    info = scene.bvh_info()
    rd = swr.render_desc(desc["render"])
    t0 = time.time()
    if args.tuning:
        ctx.set_tuning(*[int(v) for v in args.tuning.split(",")])
        tuning = ctx.tuning()
    else:
        tuning = ctx.tuning() if args.no_autotune else ctx.autotune(scene, rd)
    tune_s = time.time() - t0

    for w in range(args.warmup):
        rd.iteration = 1000 + w
        scene.render(rd)
    # Timed steps keep the image in HBM (the boundary hands the framebuffer to the host only on
    # request); the same steps with the framebuffer copied to host memory are reported beside.
    def steps(to_host):
        times, stats, pix = [], [], None
        ctx.synchronize()
        for k in range(args.steps):
            rd.iteration = k
            t = time.perf_counter()
            pix, _, st = scene.render(rd, pixels=to_host)
            ctx.synchronize()
            times.append(time.perf_counter() - t)
            stats.append(st)
        return times, stats, pix

    times, stats, _ = steps(False)
    total_s = sum(times)
    rays = sum(s["closest_rays"] + s["shadow_rays"] for s in stats)
    value = rays / total_s / 1e6
    ms_step = 1000.0 * total_s / args.steps
    host_times, host_stats, pix = steps(True)
    host_s = sum(host_times)
    # Outside the timed steps: one frame with the shadow pass in line, so that each stage's own
    # time is on record (in the timed steps the shadow pass overlaps the next wave's traversal).
    alone = None
    saved_tuning = ctx.tuning()
    if saved_tuning.get("shadow_overlap", 0) and not args.no_stage_alone:
        ctx.set_tuning(**dict(saved_tuning, shadow_overlap=0))
        scene.render(rd, pixels=False)
        _, _, st_alone = scene.render(rd, pixels=False)
        ctx.set_tuning(**saved_tuning)
        alone = {k[3:]: round(st_alone[k], 3) for k in st_alone if k.startswith("ms_")}
    host_rays = sum(s["closest_rays"] + s["shadow_rays"] for s in host_stats)

    agg = {k: sum(s[k] for s in stats) for k in stats[0]}
    dev_ms = agg["ms_total"]
    # Dominant kernel: closest-hit traversal. Algorithmic bytes per ray: 32 B ray in + 16 B hit out.
    trace_gbs = 48.0 * agg["closest_rays"] / (agg["ms_trace"] * 1e-3) / 1e9
    stage_ms = {k[3:]: round(agg[k] / args.steps, 3) for k in agg if k.startswith("ms_")}
    dominant = max(("trace", "shadow", "sort_rays", "sort_hits", "shade", "raygen", "accumulate", "filter"),
                   key=lambda k: stage_ms.get(k, 0.0))
    # In-loop ray sort: key width as tuned, so 4 B (histogram read) + 16 B per 8-bit pass per element.
    ray_passes = (tuning["ray_key_bits"] + 7) // 8
    sort_gbs = (4.0 + 16.0 * ray_passes) * agg["closest_rays"] / (agg["ms_sort_rays"] * 1e-3) / 1e9

    out = {
        "metric": "Mrays/s on 1 MI355X (plain sorted-wavefront render; closest-hit + shadow rays)",
        "value": round(value, 2),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded procedural scene, no files)",
        "config": {
            "workload": wname,
            "width": rd.width, "height": rd.height, "spp": rd.spp or rd.base * rd.base,
            "max_depth": rd.max_depth, "wave_rays": 1 << rd.bin_exponent, "triangles": scene.num_triangles,
            "method": "plain",
        },
        "roofline": {
            "kernel": "trace_queue_kernel (closest-hit BVH4 traversal)",
            "bound": "latency",
            "yardstick": "hbm (no MFMA work on this path; frac is algorithmic bytes over HBM peak)",
            "achieved": round(trace_gbs, 2),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(trace_gbs / HBM_PEAK_GBS, 5),
            "traffic": None,
            "avg_launch_ms": round(agg["ms_trace"] / agg["waves"], 4),
            "rays_per_launch": agg["closest_rays"] // agg["waves"],
            "note": "algorithmic 48 B/ray (ray in + hit out), node traffic excluded; traversal is bound by "
                    "vector-instruction issue at 8 waves per SIMD over chains of node fetches that miss L2 "
                    "half the time, not by HBM bandwidth; each launch shares the device with the previous "
                    "wave's shadow pass (second stream), which lengthens it (0.72 ms per launch alone)",
            "mrays_s": round(agg["closest_rays"] / (agg["ms_trace"] * 1e-3) / 1e6, 1),
            "l2_hit": None,
        },
        "sort_roofline": {
            "kernel": f"in-loop ray sort ({ray_passes} passes of 8 bits, "
                      f"{['onesweep', 'reduce-then-scan'][tuning['sort_algorithm']]})",
            "bound": "hbm",
            "achieved": round(sort_gbs, 2),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(sort_gbs / HBM_PEAK_GBS, 4),
        },
        "device_mrays_s": round(rays / (dev_ms * 1e-3) / 1e6, 2),
        "framebuffer_to_host": {
            "mrays_s": round(host_rays / host_s / 1e6, 2),
            "ms_per_step": round(1000.0 * host_s / args.steps, 3),
            "note": "the same steps with the filtered framebuffer copied to pageable host memory each step "
                    "(PCIe-inclusive; not the value)",
        },
        "rays_per_step": rays // args.steps,
        "stage_ms_per_step": stage_ms,
        "stage_ms_alone": alone,
        "stage_note": "each wave's shadow pass runs on a second stream beside the next wave's traversal, "
                      "so trace and shadow overlap in stage_ms_per_step; stage_ms_alone is one more frame "
                      "(outside the timed steps) with the pass in line (tuning shadow_overlap 0)",
        "dominant_stage": dominant,
        "waves_per_step": agg["waves"] // args.steps,
        "scene_gen_s": round(gen_s, 2),
        "scene_upload_bvh_s": round(build_s, 3),
        "tuning": tuning,
        "autotune_s": round(tune_s, 3),
        "bvh": info,
        "image_mean": float(np.mean(pix)),
    }

    # measured device-copy ceiling: gather_kernel<4> over 2^26 16-byte records, identity permutation
    out["copy_ceiling"] = copy_cal
    traffic_file = os.path.join(ROOT, "profiles", "trace_queue_kernel_terrain_traffic.md")
    if os.path.exists(traffic_file) and args.workload == "terrain":
        # HBM bytes per traversal ray from the committed rocprofv3 --pmc summary (FETCH_SIZE +
        # WRITE_SIZE per dispatch, scaled by the copy calibration), times this run's rays per launch.
        per_ray = l2_hit = None
        with open(traffic_file) as f:
            for line in f:
                if line.startswith("hbm_bytes_per_ray:"):
                    per_ray = float(line.split(":", 1)[1])
                if line.startswith("l2_hit_pct:"):
                    l2_hit = float(line.split(":", 1)[1])
        if per_ray:
            out["roofline"]["traffic"] = round(per_ray * agg["closest_rays"] / agg["waves"], 1)
            out["roofline"]["traffic_per_ray"] = per_ray
        if l2_hit is not None:
            out["roofline"]["l2_hit"] = round(l2_hit / 100.0, 3)
            out["roofline"]["l2_hit_source"] = "profiles/trace_queue_kernel_terrain_traffic.md (TCC_HIT/(TCC_HIT+TCC_MISS))"

    if not args.no_kernels:
        out["kernels"] = kernels_bench(ctx, scene, desc, args)

    if not args.no_batched:
        scene.close()
        out["batched"] = batched_vs_plain(ctx, args)
    if not args.no_serving:
        out["serving"] = serving(ctx, args)

    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(desc, rd, args)
    ctx.close()
    print(json.dumps(out), flush=True)


def kernels_bench(ctx, scene, desc, args):
    """Sort/traversal kernel bench. 2^24 rays with origins uniform in the unit cube and
    directions uniform on the sphere, keyed by 3-bit octant + 29-bit Morton code and radix
    sorted (key, index); then 2^24 hit records keyed ((mesh + 1) << 21 | face), mesh in [0, 64),
    face in [0, 2^21), sorted by that key (28 bits). Both permutations are checked bit for bit
    against a stable sort of the same keys. Then 2^24 rays placed at terrain scale (origins
    uniform over x, z in [-1, 1], y in [0, 0.5]; directions uniform on the sphere) traced against
    the terrain BVH for closest hits, and 2^24 shadow rays from the same origins to uniform points
    of the sun quad. Device times from HIP events around each call; best of 5."""
    import numpy as np

    from swr import scenes

    n = 1 << 24
    res = {"n": n}
    rays = scenes.sort_bench_rays(n)
    d_rays = ctx.upload(rays)
    d_keys = ctx.alloc(4 * n)
    d_vals = ctx.alloc(4 * n)
    ctx.ray_keys(d_rays, n, (0.0, 0.0, 0.0), (1024.0, 1024.0, 512.0), d_keys)
    keys = d_keys.download(np.uint32, (n,))

    def time_sort(host_keys, bits):
        ms = []
        for _ in range(6):
            d_keys.upload_from(host_keys)
            ctx.sort_pairs(d_keys, d_vals, n, bits, iota=True)
            ms.append(ctx.last_device_ms())
        perm = d_vals.download(np.uint32, (n,))
        m = min(ms[1:])
        gbs = 68.0 * n / (m * 1e-3) / 1e9
        ok = bool(np.array_equal(perm, np.argsort(host_keys, kind="stable").astype(np.uint32)))
        return {"ms": round(m, 4), "GB/s": round(gbs, 1), "frac": round(gbs / HBM_PEAK_GBS, 4),
                "bits": bits, "bit_exact": ok}

    # the sort's launch shapes tuned for this size: the LSD pass structures at every tile size and
    # the bucket sort (whose tiles are fixed), each key type taking the fastest
    names = ["onesweep", "reduce-then-scan", "bucket"]
    saved = ctx.tuning()

    def sweep_sort(host_keys, bits):
        sweep = {}
        for alg in (0, 1):
            for kpt in (8, 12, 16, 20, 24):
                ctx.set_tuning(**dict(saved, sort_algorithm=alg, sort_keys_per_thread=kpt))
                sweep[(alg, kpt)] = time_sort(host_keys, bits)
        ctx.set_tuning(**dict(saved, sort_algorithm=2))
        sweep[(2, 0)] = time_sort(host_keys, bits)
        best = min(sweep, key=lambda c: sweep[c]["ms"])
        return dict(sweep[best], algorithm=names[best[0]], keys_per_thread=best[1] or None,
                    sweep_ms={(f"{['onesweep', 'scan'][a]}/{k}" if a < 2 else "bucket"): v["ms"]
                              for (a, k), v in sweep.items()},
                    all_bit_exact=all(v["bit_exact"] for v in sweep.values()))

    res["ray_sort"] = sweep_sort(keys, 32)
    del keys
    rng = np.random.default_rng(0x5EED0007)
    mesh = rng.integers(0, 64, n, dtype=np.uint32)
    face = rng.integers(0, 1 << 21, n, dtype=np.uint32)
    d_mesh, d_face = ctx.upload(mesh), ctx.upload(face)
    ctx.mesh_face_keys(d_mesh, d_face, n, d_keys)
    hkeys = d_keys.download(np.uint32, (n,))
    res["hit_sort"] = sweep_sort(hkeys, 28)
    res["hit_sort"]["keys_match"] = bool(np.array_equal(hkeys, ((mesh + 1) << 21) | face))
    ctx.set_tuning(**saved)
    for a in (d_mesh, d_face, d_keys, d_vals):
        a.free()

    # traversal against the terrain BVH
    t = np.zeros((n, 8), np.float32)
    t[:, 0] = rays[:, 0] * 2.0 - 1.0
    t[:, 1] = rays[:, 1] * 0.5
    t[:, 2] = rays[:, 2] * 2.0 - 1.0
    t[:, 3] = 0.001
    t[:, 4:7] = rays[:, 4:7]
    t[:, 7] = 1e5
    d_rays.upload_from(t)
    d_hits = ctx.alloc(16 * n)
    ms = []
    for _ in range(5):
        scene.intersect_dev(d_rays, n, d_hits)
        ms.append(ctx.last_device_ms())
    hits = d_hits.download(np.uint32, (n, 4))
    res["closest_hit"] = {"ms": round(min(ms), 4), "mrays_s": round(n / (min(ms) * 1e-3) / 1e6, 1),
                          "hit_fraction": round(float(np.mean(hits[:, 3] != 0xFFFFFFFF)), 4)}
    light = desc["lights"][0]
    sx, sy = light["scale"]
    u = rng.random((n, 2), dtype=np.float32) - 0.5
    # points of the sun quad: translation + R (0, sx u, sy v); R = Rx Ry Rz of the light
    ax, ay, az = [np.radians(v) for v in light["rotation"]]
    rx = np.array([[1, 0, 0], [0, np.cos(ax), -np.sin(ax)], [0, np.sin(ax), np.cos(ax)]])
    ry = np.array([[np.cos(ay), 0, np.sin(ay)], [0, 1, 0], [-np.sin(ay), 0, np.cos(ay)]])
    rz = np.array([[np.cos(az), -np.sin(az), 0], [np.sin(az), np.cos(az), 0], [0, 0, 1]])
    local = np.stack([np.zeros(n, np.float32), sx * u[:, 0], sy * u[:, 1]], 1)
    pts = (local @ (rx @ ry @ rz).T.astype(np.float32)) + np.asarray(light["translation"], np.float32)
    d = pts - t[:, 0:3]
    dist = np.linalg.norm(d, axis=1)
    t[:, 4:7] = d / dist[:, None]
    t[:, 7] = dist
    d_rays.upload_from(t)
    d_flags = ctx.alloc(4 * n)
    ms = []
    for _ in range(5):
        scene.occluded_dev(d_rays, n, d_flags)
        ms.append(ctx.last_device_ms())
    flags = d_flags.download(np.uint32, (n,))
    res["shadow"] = {"ms": round(min(ms), 4), "mrays_s": round(n / (min(ms) * 1e-3) / 1e6, 1),
                     "occluded_fraction": round(float(np.mean(flags != 0)), 4)}
    for a in (d_rays, d_hits, d_flags):
        a.free()
    return res


def batched_vs_plain(ctx, args, plain=True):
    """The batched workload (256 jobs over 8 resident scenes of 10k-2M triangles, pixel counts
    log-uniform in [128^2, 1024^2], spp uniform in {1..64}, depth 6) traced twice: job by job
    through the plain loop, and all together in one persistent graph-replayed loop. Both times
    run from each loop's first raygen to its last resolved framebuffer (device events), each the
    faster of two runs (per job for the plain side); images of every 8th job are compared bit for
    bit."""
    import numpy as np

    import swr
    from swr import scenes

    t0 = time.time()
    descs, jobs = scenes.batch_workload(jobs=args.batch_jobs)
    gen_s = time.time() - t0
    t0 = time.time()
    scs = [swr.Scene(ctx, d) for d in descs]
    build_s = time.time() - t0
    rds = [swr.render_desc(r) for _, r in jobs]
    saved = ctx.tuning()

    # plain: one render per job through the plain loop (run-ahead waves), each scene with the
    # launch shapes the plain autotuner (swr_autotune) picks for it at that scene's largest job;
    # every job is rendered twice and the faster device time counted (the batched side below is
    # likewise the faster of two timed runs)
    t0 = time.time()
    paths_of = lambda rd: rd.width * rd.height * (rd.spp or rd.base * rd.base)
    plain_tunings = {}
    for s in sorted({s for s, _ in jobs}) if plain else ():
        big = max((rd for (js, _), rd in zip(jobs, rds) if js == s), key=paths_of)
        plain_tunings[s] = ctx.autotune(scs[s], big)
    plain_tune_s = time.time() - t0
    plain_ms, plain_rays, plain_pix = 0.0, 0, {}
    wall = time.perf_counter()
    for j, ((s, _), rd) in enumerate(zip(jobs, rds)):
        if not plain:
            break
        ctx.set_tuning(**plain_tunings[s])
        pix, _, st = scs[s].render(rd)
        _, _, st2 = scs[s].render(rd, pixels=False)
        plain_ms += min(st["ms_total"], st2["ms_total"])
        plain_rays += st["closest_rays"] + st["shadow_rays"]
        if j % 8 == 0:
            plain_pix[j] = pix
    plain_wall = (time.perf_counter() - wall) / 2

    b = swr.Batch(ctx)
    for (s, _), rd in zip(jobs, rds):
        b.add(scs[s], rd)
    if args.batch_paced_only:
        if args.batch_tuning:
            ctx.set_tuning(*[int(v) for v in args.batch_tuning.split(",")])
        wave_rays = args.batch_wave_rays or (1 << 27)
        paced = b.run(wave_rays, paced=True)
        used = ctx.tuning()
        b.close()
        ctx.set_tuning(**saved)
        for sc in scs:
            sc.close()
        return {"workload": f"{len(jobs)} jobs over 8 scenes (10k-2M tris), 128^2-1024^2 px, 1-64 spp, depth 6",
                "rays": int(paced["closest_rays"] + paced["shadow_rays"]), "batched_wave_rays": wave_rays,
                "batched_mrays_s": round((paced["closest_rays"] + paced["shadow_rays"]) / (paced["ms_total"] * 1e-3)
                                         / 1e6, 2),
                "batched_ms": round(paced["ms_total"], 2), "paced_only": True,
                "paced_stage_ms": {k[3:]: round(paced[k], 2) for k in paced if k.startswith("ms_")},
                "tuning": used}
    t0 = time.time()
    if args.batch_tuning:
        ctx.set_tuning(*[int(v) for v in args.batch_tuning.split(",")])
        tuning = ctx.tuning()
    else:
        # per-scene stack rings + sort tile and key widths at the default wave size; the batched
        # loop is tuned again at the wave size it runs
        tuning = b.autotune()
    tune_s = time.time() - t0

    wave_rays = args.batch_wave_rays
    if not wave_rays:  # the batched loop's own wave size: the fastest of 2^24 .. 2^28 rays
        trial = {}
        for wr in (1 << 24, 1 << 25, 1 << 26, 1 << 27, 1 << 28):
            b.run(wr)
            trial[wr] = b.run(wr)["ms_total"]
        wave_rays = min(trial, key=trial.get)
    batch_tuning = tuning
    if not args.batch_tuning:
        t0 = time.time()
        batch_tuning = b.autotune(wave_rays)
        tune_s += time.time() - t0
    paced = b.run(wave_rays, paced=True)  # the same batch launched wave by wave from the host: per-stage times
    b.run(wave_rays)  # warm-up of the graph path
    st, batch_wall = None, None
    for _ in range(2):
        wall = time.perf_counter()
        run = b.run(wave_rays)
        w_s = time.perf_counter() - wall
        if st is None or run["ms_total"] < st["ms_total"]:
            st, batch_wall = run, w_s
    batch_rays = st["closest_rays"] + st["shadow_rays"]
    same = all(np.array_equal(b.result(j)[0], p) for j, p in plain_pix.items())
    b.close()
    ctx.set_tuning(**saved)
    rings = [sc.stack_ring() for sc in scs]
    for sc in scs:
        sc.close()
    plain_v = plain_rays / (plain_ms * 1e-3) / 1e6 if plain else None
    batch_v = batch_rays / (st["ms_total"] * 1e-3) / 1e6
    return {
        "workload": f"{len(jobs)} jobs over 8 scenes (10k-2M tris), 128^2-1024^2 px, 1-64 spp, depth 6",
        "paths": int(sum(rd.width * rd.height * (rd.spp or rd.base * rd.base) for rd in rds)),
        "rays": int(batch_rays),
        "plain_mrays_s": round(plain_v, 2) if plain else None,
        "batched_mrays_s": round(batch_v, 2),
        "speedup": round(batch_v / plain_v, 3) if plain else None,
        "plain_ms": round(plain_ms, 2),
        "batched_ms": round(st["ms_total"], 2),
        "plain_wall_s": round(plain_wall, 3),
        "batched_wall_s": round(batch_wall, 3),
        "waves": {"batched": st["waves"]},
        "batched_wave_rays": wave_rays,
        "paced_batched_mrays_s": round((paced["closest_rays"] + paced["shadow_rays"]) / (paced["ms_total"] * 1e-3) / 1e6,
                                       2),
        "paced_stage_ms": {k[3:]: round(paced[k], 2) for k in paced if k.startswith("ms_")},
        "images_equal_sampled": bool(same and plain_rays == batch_rays) if plain else None,
        "scene_rings": rings,
        "plain_tuning": {f"scene{s}": t for s, t in plain_tunings.items()},
        "plain_tuning_note": "swr_autotune per scene at the scene's largest job (sort pass structure and tile, "
                             "ray and hit key widths, stack ring)",
        "batch_tuning": batch_tuning,
        "autotune_s": round(tune_s, 2),
        "plain_autotune_s": round(plain_tune_s, 2),
        "scene_gen_s": round(gen_s, 2),
        "scene_upload_bvh_s": round(build_s, 2),
    }


def serving(ctx, args):
    """The serving loop under Poisson arrivals at a fixed rate: 4 resident scenes (61k-2.1M
    triangles), requests with pixel counts log-uniform in [256^2, 1920x1080] and total spp
    log-uniform in [4, 256] (iterations of up to 16 spp), depth 6. Latency runs from a request's
    submission (its arrival) to its framebuffer in host memory; Mrays/s is every ray traced over
    the wall time from the first arrival to the last delivery."""
    import random
    import threading

    import swr
    from swr import scenes

    n_total = args.serve_warmup + args.serve_requests
    t0 = time.time()
    descs, reqs = scenes.serve_workload(requests=n_total)
    gen_s = time.time() - t0
    scs = [swr.Scene(ctx, d) for d in descs]
    rds = [(s, swr.render_desc(r), it) for s, r, it in reqs]
    saturation = serve_saturation(ctx, scs, rds[:args.serve_saturation], args) if args.serve_saturation else None
    srv = swr.Server(ctx, scs, state_bytes=int(args.serve_state_gib * (1 << 30)), max_requests=256,
                     backlog_waves=args.serve_backlog_waves, waves_per_check=args.serve_waves_per_check)
    arrivals = []
    rng = random.Random(0x5EED0006)
    t = 0.0
    for _ in range(n_total):
        t += rng.expovariate(args.serve_rate)
        arrivals.append(t)
    ids = [None] * n_total
    stats = [None] * n_total
    ready = threading.Semaphore(0)

    def waiter():
        for k in range(n_total):
            ready.acquire()
            _, stats[k] = srv.wait(ids[k])

    wt = threading.Thread(target=waiter)
    wt.start()
    start = time.perf_counter()
    for k, (s, rd, it) in enumerate(rds):
        delay = arrivals[k] - (time.perf_counter() - start)
        if delay > 0:
            time.sleep(delay)
        ids[k] = srv.submit(s, rd, it, keep=False)
        ready.release()
    wt.join()
    wall = time.perf_counter() - start
    tot = srv.totals()
    srv.close()
    for sc in scs:
        sc.close()
    lat = sorted(st["latency_ms"] for st in stats[args.serve_warmup:])
    queue = sorted(st["queue_ms"] for st in stats[args.serve_warmup:])
    pct = lambda v, q: v[min(len(v) - 1, int(q * len(v)))]
    rays = tot["closest_rays"] + tot["shadow_rays"]
    paths = sum(rd.width * rd.height * rd.base * rd.base * it for _, rd, it in rds)
    return {
        "workload": f"{args.serve_requests} requests (+{args.serve_warmup} warm-up) on 4 resident scenes "
                    f"(61k-2.1M tris), 256^2-1920x1080 px, 4-256 spp, depth 6",
        "rate_per_s": args.serve_rate,
        "backlog_waves": args.serve_backlog_waves or 2,
        "waves_per_check": args.serve_waves_per_check or 2,
        "saturation": saturation,
        "mrays_s": round(rays / wall / 1e6, 2),
        "p50_ms": round(pct(lat, 0.50), 2),
        "p99_ms": round(pct(lat, 0.99), 2),
        "mean_ms": round(sum(lat) / len(lat), 2),
        "queue_p50_ms": round(pct(queue, 0.50), 2),
        "queue_p99_ms": round(pct(queue, 0.99), 2),
        "completed": tot["completed"],
        "rays": int(rays),
        "paths": int(paths),
        "wall_s": round(wall, 3),
        "busy_s": round(tot["busy_s"], 3),
        "achieved_rate_per_s": round(n_total / wall, 2),
        "state_gib": round(tot["state_bytes"] / (1 << 30), 2),
        "queue_gib": round(tot["queue_bytes"] / (1 << 30), 2),
        "scene_gen_s": round(gen_s, 2),
    }


def serve_saturation(ctx, scs, reqs, args):
    """Throughput of the served form on the same requests three ways: all submitted at once to a
    serving loop (wall time from the first submission to the last result in host memory), their
    iterations as one batch (device time), and every iteration rendered on its own through the
    plain loop with that scene's plain-autotuned shapes (device time, the faster of two renders
    per iteration). The served result of each request is compared bit for bit with the mean of
    its plain iterations for the first request on every scene."""
    import numpy as np

    import swr

    saved = ctx.tuning()
    paths_of = lambda rd: rd.width * rd.height * (rd.spp or rd.base * rd.base)
    tunings = {}
    for s in sorted({s for s, _, _ in reqs}):
        big = max((rd for js, rd, _ in reqs if js == s), key=paths_of)
        tunings[s] = ctx.autotune(scs[s], big)
    plain_ms, plain_rays, plain_mean = 0.0, 0, {}
    first = {}
    for k, (s, rd, it) in enumerate(reqs):
        first.setdefault(s, k)
    for k, (s, rd, it) in enumerate(reqs):
        ctx.set_tuning(**tunings[s])
        d = swr.RenderDesc.from_buffer_copy(rd)
        acc = None
        for i in range(it):
            d.iteration = rd.iteration + i
            keep = first[s] == k
            pix, _, st = scs[s].render(d, pixels=keep)
            _, _, st2 = scs[s].render(d, pixels=False)
            plain_ms += min(st["ms_total"], st2["ms_total"])
            plain_rays += st["closest_rays"] + st["shadow_rays"]
            if keep:
                acc = pix if acc is None else acc + pix
        if acc is not None:
            plain_mean[k] = acc * np.float32(1.0 / it)
    ctx.set_tuning(**saved)

    b = swr.Batch(ctx)
    for s, rd, it in reqs:
        for i in range(it):
            d = swr.RenderDesc.from_buffer_copy(rd)
            d.iteration = rd.iteration + i
            b.add(scs[s], d)
    b.run()
    bst = b.run()
    b.close()
    batch_rays = bst["closest_rays"] + bst["shadow_rays"]

    best = None
    for _ in range(2):
        srv = swr.Server(ctx, scs, state_bytes=int(args.serve_state_gib * (1 << 30)), max_requests=256,
                         backlog_waves=args.serve_backlog_waves, waves_per_check=args.serve_waves_per_check)
        t = time.perf_counter()
        ids = [srv.submit(s, rd, it, keep=k in plain_mean) for k, (s, rd, it) in enumerate(reqs)]
        res = [srv.wait(i) for i in ids]
        wall = time.perf_counter() - t
        tot = srv.totals()
        srv.close()
        if best is None or wall < best[0]:
            best = (wall, tot, res)
    wall, tot, res = best
    same = all(np.array_equal(res[k][0], m) for k, m in plain_mean.items())
    served_rays = tot["closest_rays"] + tot["shadow_rays"]
    served_v = served_rays / wall / 1e6
    plain_v = plain_rays / (plain_ms * 1e-3) / 1e6
    return {
        "requests": len(reqs),
        "iterations": int(sum(it for _, _, it in reqs)),
        "rays": int(served_rays),
        "served_mrays_s": round(served_v, 2),
        "served_wall_s": round(wall, 3),
        "as_batch_mrays_s": round(batch_rays / (bst["ms_total"] * 1e-3) / 1e6, 2),
        "plain_mrays_s": round(plain_v, 2),
        "served_vs_plain": round(served_v / plain_v, 3),
        "rays_equal": bool(served_rays == plain_rays == batch_rays),
        "results_equal_plain_mean": bool(same),
        "note": "served: wall time, all requests submitted at once (the faster of two runs); as_batch and "
                "plain: device time; plain with swr_autotune's shapes per scene",
    }


def cpu_baseline(desc, rd, args):
    """The CPU restatement (oracle) on this host's cores, over a bounded sample of the same frame."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle_ffi as orc
    import swr

    threads = args.cpu_threads or min(os.cpu_count() or 1, 16)
    t0 = time.time()
    oc = orc.Scene(desc)
    build_s = time.time() - t0
    od = swr.render_desc(desc["render"], orc.RenderDesc)
    total = rd.width * rd.height * (rd.spp or rd.base * rd.base)
    n = args.cpu_samples or total  # the whole frame: a few seconds on 16 cores
    n = min(n, total)
    # a band of whole pixel columns in the middle of the frame (sample ids are column-major)
    start = (total // 2 // n) * n if n < total else 0
    _, _, st = oc.render(od, samples=False, sid_begin=start, sid_end=start + n, threads=threads, pixels=False)
    rays = st.closest_rays + st.shadow_rays
    # the same restatement on one core, over 1/threads of the sample band (same CPU time)
    n1 = max(1, n // threads)
    _, _, st1 = oc.render(od, samples=False, sid_begin=start, sid_end=start + n1, threads=1, pixels=False)
    rays1 = st1.closest_rays + st1.shadow_rays
    return {
        "value": round(rays / st.seconds / 1e6, 3),
        "unit": "Mrays/s",
        "cores": threads,
        "kind": "port",
        "sample": f"{n} of {total} sample ids (ids {start}..{start + n - 1}) of the same frame; "
                  f"{rays} rays in {st.seconds:.2f} s; CPU BVH build {build_s:.1f} s excluded",
        "one_core": {"value": round(rays1 / st1.seconds / 1e6, 3), "cores": 1,
                     "sample": f"ids {start}..{start + n1 - 1}; {rays1} rays in {st1.seconds:.2f} s"},
        "cpu_model": _cpu_model(),
        "host_cpus": os.cpu_count(),
    }


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


if __name__ == "__main__":
    main()
