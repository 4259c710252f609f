"""Throughput of the MI355X wavefront path tracer on the 1080p spectral icosphere-grid workload.

Workload: spectral path tracing (4 hero wavelengths, 360-830 nm, Meng upsampling) of a 32x32
grid of 1,280-triangle icosphere instances over a 2-triangle ground plane (1,310,722
triangles; GGX/UE4 metals with roughness U[0.05,0.8], 10 % smooth dielectrics of IOR 1.5),
lit by a 1024x512 value-noise environment map exp-mapped to [0.01, 50]; 1920x1080, max
path length 16.  A step is one vlrContextRender call: one sample per pixel over the whole
image (generate, then per depth intersect / shade / shadow, then accumulate and convert).
The timed region is K steps starting from firstFrame, i.e. a K-spp render.

The JSON line carries
  value          extension + shadow rays traced per second (device counters / wall clock)
  roofline       the dominant kernel's algorithmic bytes per launch over its mean launch
                 time (HIP events on the stream it runs on, in extra frames after the timed
                 region) against 8 TB/s HBM; the per-ray node,
                 triangle and instance counts come from extra frames rendered after the timed
                 region with the traversal kernels' counting variant; `traffic` is the HBM
                 bytes per launch measured by rocprofv3 counters, read from the committed
                 summary under profiles/ when it exists
  rmse           the learned radiance cache against plain path tracing on the Cornell box,
                 spectral (4 hero wavelengths), 1920x1080, max depth 8: RMSE of both at equal
                 samples per pixel (--rmse-spp, 1024) and over wall time against a plain
                 reference of --rmse-ref-spp (65,536) samples accumulated in FP64 (two
                 independent halves; the reference's own noise variance is removed), and the
                 fraction of the plain render's time the cached one needs to reach its RMSE
  cpu_baseline   the oracle (C++/OpenMP restatement of the same integrator) on the host's
                 cores, whole 1920x1080 frames of the same scene and seeds, about 10 s of work
"""
import argparse
import ctypes as C
import json
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from vlr_amd import api, scenes  # noqa: E402

METRIC = "rays/sec at 1080p + equal-time RMSE vs FP64 64k-spp ref, cached vs plain PT"
HBM_PEAK_GBS = 8000.0
WORKLOAD = ("spectral path tracing, 32x32 grid of 1280-tri icosphere instances + ground "
            "(1,310,722 tris), 1024x512 value-noise env, 1920x1080, max depth 16")

# bytes each kernel moves per unit of work, fixed by the device layout
RAY_IN = 32          # origin.xyz + path index, direction.xyz + tmax
HIT_OUT = 24         # (inst, geomInst, prim, t) + (b1, b2)
SHADOW_IN = 48       # origin, direction, contribution
SHADOW_RMW = 32      # contribution read + write when unoccluded
NODE = 64            # 4-wide BVH node: origin, grid steps, four quantised child boxes, four child codes
TRI = 48             # three vertices + geometry-instance / primitive indices
INST = 48            # world-to-object 3x4 rows read on instance entry
# one shaded hit: ray 32 + hit 24 + path state read and written 2 x 56 + three vertices 144 +
# material 80 + instance 96 + next ray 32 + shadow record 48
SHADE = 568


def hip():
    for name in ("libamdhip64.so", "/opt/rocm/lib/libamdhip64.so"):
        try:
            return C.CDLL(name)
        except OSError:
            continue
    raise RuntimeError("libamdhip64.so not found")


def hip_sync():
    rc = hip().hipDeviceSynchronize()
    if rc != 0:
        raise RuntimeError(f"hipDeviceSynchronize failed: {rc}")


def copy_probe(nbytes):
    """One device-to-device copy of nbytes: the reference dispatch for scaling the HBM
    counters (FETCH_SIZE / WRITE_SIZE of a plain copy kernel of known size)."""
    H = hip()
    src, dst = C.c_void_p(), C.c_void_p()
    for p in (src, dst):
        rc = H.hipMalloc(C.byref(p), C.c_size_t(nbytes))
        if rc != 0:
            raise RuntimeError(f"hipMalloc failed: {rc}")
    H.hipMemset(src, 1, C.c_size_t(nbytes))
    hip_sync()
    rc = H.hipMemcpy(dst, src, C.c_size_t(nbytes), 3)    # hipMemcpyDeviceToDevice
    hip_sync()
    H.hipFree(src)
    H.hipFree(dst)
    if rc != 0:
        raise RuntimeError(f"hipMemcpy failed: {rc}")
    print(f"copy probe: {nbytes} bytes device to device", file=sys.stderr, flush=True)


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor()


def read_traffic(kernel):
    """HBM bytes per launch of `kernel` on this workload, from the committed counter summary
    (rocprofv3 FETCH_SIZE + WRITE_SIZE, scaled by a copy kernel of known size), if present."""
    path = os.path.join(ROOT, "profiles", f"{kernel}_icosphere_grid_1080p_traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return float(json.load(f)["hbm_bytes_per_launch"])


def cpu_threads():
    """Host threads for the CPU baseline: every core in this process's affinity mask, capped by
    OMP_NUM_THREADS when the machine sets it (a GPU box's share of the host's cores)."""
    n = len(os.sched_getaffinity(0))
    cap = os.environ.get("OMP_NUM_THREADS")
    if cap and cap.isdigit() and int(cap) > 0:
        n = min(n, int(cap))
    return n


def cpu_baseline(desc, width, height, depth, budget_s):
    """The oracle path tracer (a C++/OpenMP restatement of the reference's path tracer), built
    for this machine's CPU (-march=native) and run on cpu_threads() threads over whole frames
    of the same scene and per-pixel seeds for about budget_s seconds."""
    import tempfile
    from oracle import binding as ob
    native_dir = tempfile.mkdtemp(prefix="oracle_native_")
    ob.lib(ob.build_native(native_dir))
    threads = cpu_threads()
    scene = ob.Scene(desc)
    rng = ob.seed_states(width, height)
    rays = 0
    frames = 0
    elapsed = 0.0
    while elapsed < budget_s or frames == 0:
        t0 = time.perf_counter()
        _, st, rng, _ = scene.render(width, height, 1, depth, threads=threads, rng_state=rng)
        elapsed += time.perf_counter() - t0
        rays += st["closest_rays"] + st["shadow_rays"]
        frames += 1
    return {"value": rays / elapsed, "unit": "rays/s", "cores": threads, "kind": "port",
            "sample": f"{frames} full {width}x{height} frames (1 spp each, depth {depth}) of the same scene "
                      f"and per-pixel seeds, {elapsed:.1f} s on {threads} threads of {cpu_model()} "
                      f"(oracle built with -O3 -march=native)",
            "samples_per_s": frames * width * height / elapsed,
            "host": {"nproc": os.cpu_count(), "affinity": len(os.sched_getaffinity(0)),
                     "omp_num_threads": os.environ.get("OMP_NUM_THREADS"), "model": cpu_model(),
                     "threads_used": threads}}


def rmse_curve(ctx, camera, frames, checkpoints, ref, ref_var, first_frame=True):
    """Renders `frames` frames; at each checkpoint spp returns (spp, seconds, RMSE vs ref with the
    reference's own noise variance removed).  Read-backs happen outside the timed spans."""
    out = []
    elapsed = 0.0
    hip_sync()
    t0 = time.perf_counter()
    for k in range(frames):
        ctx.render(camera, first_frame=first_frame and k == 0, limit=1 << 30)
        if (k + 1) in checkpoints:
            hip_sync()
            elapsed += time.perf_counter() - t0
            img = ctx.read_output()[..., :3].astype(np.float64)
            mse = float(((img - ref) ** 2).mean()) - ref_var
            out.append({"spp": k + 1, "seconds": elapsed, "rmse": float(np.sqrt(max(mse, 0.0)))})
            hip_sync()
            t0 = time.perf_counter()
    return out


def interp_log(points, x_key, y_key, x):
    """log-log interpolation of y at x over points sorted by x (None outside the range)."""
    xs = [p[x_key] for p in points]
    ys = [p[y_key] for p in points]
    if x < xs[0] or x > xs[-1]:
        return None
    for i in range(1, len(xs)):
        if x <= xs[i]:
            a = (np.log(x) - np.log(xs[i - 1])) / (np.log(xs[i]) - np.log(xs[i - 1]))
            return float(np.exp(np.log(ys[i - 1]) + a * (np.log(ys[i]) - np.log(ys[i - 1]))))
    return float(ys[-1])


def cache_experiment(width, height, ref_spp, spp, warmup_frames, seed=20261202, spectral=True, log=None,
                     full_radiance=True, training_paths=None, threshold=None):
    """Radiance cache vs plain path tracing on the Cornell box (max depth 8; spectral: 4 hero
    wavelengths, the network's 4 outputs): RMSE against an FP64-accumulated plain reference,
    at equal samples per pixel and over wall time.  The reference is two independent halves
    (their own seeds); their difference estimates the reference's noise variance, which is
    removed from every measured MSE.  The cached run's wall time includes its training-only
    warm-up frames."""
    desc = scenes.cornell_box(seed)
    n = width * height
    say = log or (lambda *a: None)

    def context(cache, rng_seed=None):
        ctx = api.Context(spectral=spectral, max_path_length=8)
        built = api.build_scene(ctx, desc)
        ctx.bind_output(width, height)
        if rng_seed is not None:
            ctx.write_rng(np.random.default_rng(rng_seed).integers(1, 2 ** 63, size=n, dtype=np.uint64))
        if cache:
            ctx.set_radiance_cache(True)
            kw = {"fullRadiance": int(full_radiance)}
            if training_paths:
                kw["trainingPaths"] = training_paths
            if threshold is not None:
                kw["terminationThreshold"] = threshold
            ctx.set_cache_settings(**kw)
        return ctx, built

    halves = []
    t_ref = time.perf_counter()
    for h, rs in enumerate((101, 202)):
        ctx, built = context(False, rs)
        acc = np.zeros((height, width, 3))
        chunk = 1024
        done = 0
        while done < ref_spp // 2:
            m = min(chunk, ref_spp // 2 - done)
            for k in range(m):
                ctx.render(built.camera, first_frame=(k == 0), limit=m)
            acc += ctx.read_output()[..., :3].astype(np.float64) * m
            done += m
            say(f"reference half {h}: {done} of {ref_spp // 2} spp, {time.perf_counter() - t_ref:.0f} s")
        halves.append(acc / done)
        ctx.close()
    ref = 0.5 * (halves[0] + halves[1])
    ref_var = float(((halves[0] - halves[1]) ** 2).mean()) / 4.0
    checkpoints = sorted({c for c in (16, 32, 64, 128, 256, 512, 1024, 2048, 4096) if c <= spp} | {spp})

    ctx, built = context(False)
    k = 0
    while k < 8 or not ctx.schedule()[2]:   # schedule tuning and first-use allocations outside the timed region
        ctx.render(built.camera, first_frame=(k == 0), limit=1 << 30)
        k += 1
    plain = rmse_curve(ctx, built.camera, spp, checkpoints, ref, ref_var)
    pst = ctx.stats()
    ctx.close()
    say(f"plain: {spp} spp in {plain[-1]['seconds']:.2f} s, RMSE {plain[-1]['rmse']:.5f}")

    ctx, built = context(True)
    hip_sync()
    t0 = time.perf_counter()
    for k in range(warmup_frames):
        ctx.render(built.camera, first_frame=(k == 0), limit=1 << 30)
    hip_sync()
    warm_s = time.perf_counter() - t0
    ctx.reset_stats()
    ctx.set_kernel_timing(True)
    cached = rmse_curve(ctx, built.camera, spp, checkpoints, ref, ref_var)
    st = ctx.stats()
    loss, steps = ctx.cache_loss()
    ctx.close()
    for p in cached:
        p["seconds"] += warm_s

    # equal time: the plain path tracer's RMSE at each cached checkpoint's wall time
    equal_time = []
    for p in cached:
        r = interp_log(plain, "seconds", "rmse", p["seconds"])
        if r is not None:
            equal_time.append({"seconds": p["seconds"], "cached": p["rmse"], "plain": r})
    # time for the cached form to reach the plain form's RMSE at its largest checkpoint
    target = plain[-1]["rmse"]
    reach = None
    for i, p in enumerate(cached):
        if p["rmse"] <= target:
            if i == 0:
                reach = p["seconds"]
            else:
                q = cached[i - 1]
                a = (np.log(q["rmse"]) - np.log(target)) / (np.log(q["rmse"]) - np.log(p["rmse"]))
                reach = float(np.exp(np.log(q["seconds"]) + a * (np.log(p["seconds"]) - np.log(q["seconds"]))))
            break
    frames = max(1, st["numFrames"])
    # equal samples per pixel: both curves share their checkpoints
    equal_spp = [{"spp": p["spp"], "cached": c["rmse"], "plain": p["rmse"]} for p, c in zip(plain, cached)]
    say(f"cached: {spp} spp in {cached[-1]['seconds']:.2f} s (warm-up included), RMSE {cached[-1]['rmse']:.5f}")
    return {
        "workload": f"Cornell box (C1 scene), {'spectral, 4 hero wavelengths' if spectral else 'RGB'}, "
                    f"max depth 8, {width}x{height}",
        "equal_spp": equal_spp,
        "cache_settings": {"full_radiance": bool(full_radiance), "training_paths": training_paths,
                           "threshold": threshold},
        "rays_per_path": {"plain": (pst["closestRays"] + pst["shadowRays"]) / max(1, pst["paths"]),
                          "cached": (st["closestRays"] + st["shadowRays"]) / max(1, st["paths"])},
        "reference": {"spp": ref_spp, "accumulation": "fp64 on the host over fp32 Kahan chunks of <= 1024 frames",
                      "noise_rmse_removed": float(np.sqrt(ref_var))},
        "plain": plain,
        "cached": cached,
        "cache_warmup": {"frames": warmup_frames, "seconds": warm_s},
        "equal_time": equal_time,
        "plain_rmse_at_spp": {"spp": plain[-1]["spp"], "rmse": target, "seconds": plain[-1]["seconds"]},
        "cached_seconds_to_reach_it": reach,
        "cached_time_fraction": None if reach is None else reach / plain[-1]["seconds"],
        "cache_ms_per_frame": {"query": st["msCacheQuery"] / frames, "train": st["msCacheTrain"] / frames},
        "cache_loss": loss,
        "cache_steps": steps,
    }


def point_light_pixels(desc, width, height):
    """Pixels a point light projects to: the lens connection of LVC-BPT sees the light itself
    there, which path tracing cannot hit, so those pixels are left out of the comparison."""
    cam = desc.camera
    pts = desc.meshes[-2].vertices[:, :3].astype(np.float64)
    th = np.tan(cam.fovy / 2)
    rel = pts - np.array(cam.position)
    lx, ly, lz = -rel[:, 0], rel[:, 1], -rel[:, 2]     # camera looks down -z (qRotateY(pi))
    u = 0.5 - (lx / lz) / (2 * th * cam.aspect)
    v = 0.5 - (ly / lz) / (2 * th)
    mask = np.ones((height, width), bool)
    for x, y in zip((u * width).astype(int), (v * height).astype(int)):
        mask[max(0, y - 1):y + 2, max(0, x - 1):x + 2] = False
    return mask


def bpt_experiment(width, height, spp, ref_spp, light_paths, seed=20261202):
    """LVC-BPT vs path tracing on the Cornell box with 64 point lights and 8 quad lights (RGB,
    max depth 8): throughput of both and RMSE over time against an FP64-accumulated path-traced
    reference (two independent halves; their difference estimates the reference noise, which is
    removed).  Pixels that see a point light directly are masked."""
    desc = scenes.cornell_many_lights(seed)
    n = width * height
    mask = point_light_pixels(desc, width, height)

    def context(renderer, rng_seed=None):
        ctx = api.Context(spectral=False, max_path_length=8)
        built = api.build_scene(ctx, desc)
        ctx.bind_output(width, height)
        ctx.set_renderer(renderer)
        if renderer in (api.RENDERER_BPT, api.RENDERER_LIGHT_TRACING):
            ctx.set_num_light_paths(light_paths)
        if rng_seed is not None:
            ctx.write_rng(np.random.default_rng(rng_seed).integers(1, 2 ** 63, size=n, dtype=np.uint64))
        return ctx, built

    halves = []
    for rs in (303, 404):
        ctx, built = context(api.RENDERER_PATH_TRACING, rs)
        acc = np.zeros((height, width, 3))
        done = 0
        while done < ref_spp // 2:
            m = min(1024, ref_spp // 2 - done)
            for k in range(m):
                ctx.render(built.camera, first_frame=(k == 0), limit=m)
            acc += ctx.read_output()[..., :3].astype(np.float64) * m
            done += m
        halves.append(acc / done)
        ctx.close()
    ref = 0.5 * (halves[0] + halves[1])
    ref_var = float(((halves[0] - halves[1])[mask] ** 2).mean()) / 4.0
    checkpoints = sorted({c for c in (4, 8, 16, 32, 64, 128, 256) if c <= spp} | {spp})

    def curve(renderer):
        ctx, built = context(renderer)
        for k in range(4):
            ctx.render(built.camera, first_frame=(k == 0), limit=1 << 30)
        ctx.reset_stats()
        ctx.set_kernel_timing(True)
        out, elapsed = [], 0.0
        hip_sync()
        t0 = time.perf_counter()
        for k in range(spp):
            ctx.render(built.camera, first_frame=(k == 0), limit=1 << 30)
            if (k + 1) in checkpoints:
                hip_sync()
                elapsed += time.perf_counter() - t0
                img = ctx.read_output()[..., :3].astype(np.float64)
                mse = float(((img - ref)[mask] ** 2).mean()) - ref_var
                out.append({"spp": k + 1, "seconds": elapsed, "rmse": float(np.sqrt(max(mse, 0.0)))})
                hip_sync()
                t0 = time.perf_counter()
        st = ctx.stats()
        ctx.close()
        rays = st["closestRays"] + st["shadowRays"]
        frames = max(1, st["numFrames"])
        return out, {"rays_per_s": rays / elapsed, "samples_per_s": st["paths"] / elapsed,
                     "ms_per_frame": elapsed * 1e3 / spp, "closest_rays": st["closestRays"],
                     "shadow_rays": st["shadowRays"],
                     "ms_light_paths_per_frame": st["msLightPaths"] / frames,
                     "ms_eye_paths_per_frame": st["msEyePaths"] / frames}

    bpt, bst = curve(api.RENDERER_BPT)
    pt, pst = curve(api.RENDERER_PATH_TRACING)
    lt, lst = curve(api.RENDERER_LIGHT_TRACING)
    equal_time = []
    for p in bpt:
        r = interp_log(pt, "seconds", "rmse", p["seconds"])
        rl = interp_log(lt, "seconds", "rmse", p["seconds"])
        if r is not None:
            equal_time.append({"seconds": p["seconds"], "bpt": p["rmse"], "pt": r, "lt": rl})
    return {
        "workload": f"Cornell box + 64 point lights + 8 quad lights, RGB, max depth 8, {width}x{height}, "
                    f"{light_paths} light paths per frame",
        "reference": {"renderer": "path tracing", "spp": ref_spp, "noise_rmse_removed": float(np.sqrt(ref_var)),
                      "masked_pixels": int((~mask).sum())},
        "bpt": bpt, "bpt_stats": bst, "pt": pt, "pt_stats": pst, "lt": lt, "lt_stats": lst,
        "equal_time": equal_time,
    }


def batch_experiment(num_jobs, seed=20261202):
    """C4: render jobs drawn from one seed (tessellated Cornell boxes of 10k-1M unique
    triangles, log-uniform; square images of 128, 256, 512 or 1024 pixels; 1-64 samples;
    max path length 4-16), once as one batch (graph-captured steps, autotuned schedule) and
    once serially, one job after another, each in its own context prepared beforehand.
    Reports aggregate rays/s and the p50 / p99 job completion latency of both."""
    rng = np.random.default_rng(seed)
    specs = []
    for k in range(num_jobs):
        tris = int(round(np.exp(rng.uniform(np.log(1e4), np.log(1e6)))))
        res = int(rng.choice([128, 256, 512, 1024]))
        specs.append((tris, res, int(rng.integers(1, 65)), int(rng.integers(4, 17)), seed + 1000 + k))
    descs = [scenes.tessellated_cornell(sd, tris) for (tris, _, _, _, sd) in specs]

    # batch
    bctx = api.Context(spectral=False)
    built = [api.build_scene(bctx, d) for d in descs]
    jobs = [(built[k].scene, built[k].camera, res, res, spp, depth)
            for k, (_, res, spp, depth, _) in enumerate(specs)]
    t0 = time.perf_counter()
    bctx.render_batch([(b, c, 16, 16, 1, d) for (b, c, _, _, _, d) in jobs])   # scene set-up, first-use allocations
    setup_s = time.perf_counter() - t0
    bctx.reset_stats()
    hip_sync()
    t0 = time.perf_counter()
    done_ms = bctx.render_batch(jobs)
    hip_sync()
    batch_s = time.perf_counter() - t0
    bst = bctx.stats()
    sched = bctx.batch_schedule()
    bctx.close()

    # serial: each job alone, in order, in its own prepared context
    ctxs = []
    for k, d in enumerate(descs):
        c = api.Context(spectral=False, max_path_length=specs[k][3])
        bs = api.build_scene(c, d)
        c.bind_output(specs[k][1], specs[k][1])
        c.set_wavefront_depth(0)           # the fastest fixed schedule for these single jobs
        seeds = c.read_rng()
        c.render(bs.camera, True, 1)          # scene set-up outside the timed region
        c.write_rng(seeds)                    # the timed render starts from the same seeds as the batch
        ctxs.append((c, bs))
    hip_sync()
    serial_done, rays = [], 0
    t0 = time.perf_counter()
    for k, (c, bs) in enumerate(ctxs):
        c.reset_stats()
        for f in range(specs[k][2]):
            c.render(bs.camera, f == 0, specs[k][2])
        hip_sync()
        serial_done.append((time.perf_counter() - t0) * 1e3)
    serial_s = time.perf_counter() - t0
    for c, _ in ctxs:
        st = c.stats()
        rays += st["closestRays"] + st["shadowRays"]
        c.close()
    brays = bst["closestRays"] + bst["shadowRays"]
    q = lambda v, p: float(np.percentile(np.asarray(v, np.float64), p))
    return {
        "workload": f"{num_jobs} jobs: tessellated Cornell boxes of 10k-1M unique triangles (log-uniform), "
                    "128^2-1024^2, 1-64 spp, max depth 4-16, RGB",
        "jobs": [{"triangles": descs[k].num_triangles, "res": r, "spp": p, "depth": d}
                 for k, (_, r, p, d, _) in enumerate(specs)],
        "batch": {"seconds": batch_s, "rays": brays, "rays_per_s": brays / batch_s,
                  "latency_ms_p50": q(done_ms, 50), "latency_ms_p99": q(done_ms, 99),
                  "schedule": {"wavefront_depth": sched[0], "tail_blocks": sched[1]},
                  "setup_seconds": setup_s},
        "serial": {"seconds": serial_s, "rays": rays, "rays_per_s": rays / serial_s,
                   "latency_ms_p50": q(serial_done, 50), "latency_ms_p99": q(serial_done, 99)},
        "rays_equal": brays == rays,
    }


def accel_experiment(seed=20261202):
    """Acceleration-structure builds inside scene setup, device builder against host builder:
    the 32 x 32 icosphere grid (1025 geometry groups, 1,310,722 triangles) and tessellated
    Cornell boxes of about 10k, 100k and 1M unique triangles (the range of the batch workload).
    Per case: a first setup, then a warm rebuild of every BLAS (vertices of one mesh set again),
    then a transform-only update (one instance moved: TLAS rebuilt, BLASes kept).  Times are
    host wall clock of the builds (round trips included) and of the whole scene setup."""
    cases = [("icosphere grid 32x32", lambda: scenes.icosphere_grid(seed))]
    for tris in (10_000, 100_000, 1_000_000):
        cases.append((f"tessellated Cornell ~{tris // 1000}k", lambda t=tris: scenes.tessellated_cornell(seed, t)))
    rows = []
    for name, make in cases:
        desc = make()
        row = {"scene": name}
        for builder, key in ((api.ACCEL_BUILDER_DEVICE, "device"), (api.ACCEL_BUILDER_HOST, "host")):
            ctx = api.Context(spectral=desc.spectral)
            ctx.set_accel_builder(builder)
            bs = api.build_scene(ctx, desc)
            t0 = time.perf_counter()
            ctx.read_accel(bs.scene, api.ACCEL_INSTANCES)
            first = time.perf_counter() - t0
            st_first = ctx.accel_stats()
            mesh = next(i for i, m in enumerate(desc.meshes) if not m.points)
            v = np.ascontiguousarray(desc.meshes[mesh].vertices, np.float32)
            api.call("vlrTriangleMeshSurfaceNodeSetVertices", bs.meshes[mesh], v.ctypes.data, len(v))
            t0 = time.perf_counter()
            ctx.read_accel(bs.scene, api.ACCEL_INSTANCES)
            rebuild = time.perf_counter() - t0
            st = ctx.accel_stats()
            xf = np.array(desc.instances[-1].transform, np.float32).reshape(4, 4, order="F").copy()
            xf[:3, 3] += 0.25
            api.call("vlrParentNodeSetTransform", bs.instances[-1], ctx.transform(xf.reshape(16, order="F")))
            t0 = time.perf_counter()
            ctx.read_accel(bs.scene, api.ACCEL_INSTANCES)
            update = time.perf_counter() - t0
            su = ctx.accel_stats()
            ctx.close()
            row["triangles"] = st["numTriangles"]
            row["blas_count"] = st["numBlas"]
            row[key] = {"first_setup_s": first, "first_ms_blas": st_first["msBlas"],
                        "rebuild_ms_blas": st["msBlas"], "rebuild_ms_tlas": st["msTlas"], "rebuild_setup_s": rebuild,
                        "update_blas_reused": bool(su["blasReused"]), "update_ms_tlas": su["msTlas"],
                        "update_setup_s": update}
        row["blas_speedup"] = row["host"]["rebuild_ms_blas"] / max(1e-6, row["device"]["rebuild_ms_blas"])
        rows.append(row)
        print(json.dumps({"accel_case": row}), flush=True)
    return rows


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=64)
    # warm-up frames; the warm-up also runs on until the frame-schedule autotuner is done
    ap.add_argument("--warmup", type=int, default=24)
    ap.add_argument("--width", type=int, default=1920)
    ap.add_argument("--height", type=int, default=1080)
    ap.add_argument("--depth", type=int, default=16)
    ap.add_argument("--seed", type=int, default=20261202)
    ap.add_argument("--stat-frames", type=int, default=2)
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--wavefront-depth", type=int, default=None,
                    help="depths run as per-depth kernels before the persistent tail kernel (library default)")
    ap.add_argument("--copy-probe", type=int, default=0,
                    help="also run one device-to-device copy of this many bytes (counter calibration)")
    ap.add_argument("--rmse-ref-spp", type=int, default=65536,
                    help="samples per pixel of the FP64-accumulated plain reference of the rmse field")
    ap.add_argument("--rmse-spp", type=int, default=1024,
                    help="samples per pixel of the plain and cached renders of the rmse field")
    ap.add_argument("--no-rmse", action="store_true", help="leave the rmse field out (null)")
    ap.add_argument("--cache-experiment", action="store_true",
                    help="only run the radiance-cache RMSE experiment and print its JSON line")
    ap.add_argument("--cache-ref-spp", type=int, default=2048)
    ap.add_argument("--cache-indirect", action="store_true",
                    help="cache mode: the last vertex keeps its next-event estimate (default: the cache "
                         "supplies all of its reflected radiance)")
    ap.add_argument("--cache-training-paths", type=int, default=None)
    ap.add_argument("--cache-threshold", type=float, default=None)
    ap.add_argument("--cache-rgb", action="store_true")
    ap.add_argument("--cache-spp", type=int, default=256)
    ap.add_argument("--cache-warmup", type=int, default=64)
    ap.add_argument("--batch-experiment", action="store_true",
                    help="only run the batched-vs-serial render-job experiment and print its JSON line")
    ap.add_argument("--batch-jobs", type=int, default=64)
    ap.add_argument("--bpt-experiment", action="store_true",
                    help="only run the LVC-BPT vs path tracing experiment (1280x720) and print its JSON line")
    ap.add_argument("--bpt-spp", type=int, default=32)
    ap.add_argument("--bpt-ref-spp", type=int, default=2048)
    ap.add_argument("--accel-experiment", action="store_true",
                    help="device against host acceleration-structure builds")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (the path shards by pixel "
              f"tiles with one framebuffer gather, which this build does not run)", file=sys.stderr)
        sys.exit(2)
    if not os.path.exists(api.LIB_PATH):
        raise RuntimeError(f"{api.LIB_PATH} is missing: build it with vlr_amd/build.py")
    if args.accel_experiment:
        res = accel_experiment(args.seed)
        print(json.dumps({"accel": res}), flush=True)
        return
    if args.batch_experiment:
        res = batch_experiment(args.batch_jobs)
        print(json.dumps({"batch": res}), flush=True)
        return
    if args.bpt_experiment:
        res = bpt_experiment(1280, 720, args.bpt_spp, args.bpt_ref_spp, 1280 * 720)
        print(json.dumps({"lvc_bpt": res}), flush=True)
        return
    if args.cache_experiment:
        log = lambda m: print(f"cache: {m}", file=sys.stderr, flush=True)
        res = cache_experiment(args.width, args.height, args.cache_ref_spp, args.cache_spp, args.cache_warmup,
                               spectral=not args.cache_rgb, log=log, full_radiance=not args.cache_indirect,
                               training_paths=args.cache_training_paths, threshold=args.cache_threshold)
        print(json.dumps({"radiance_cache": res}), flush=True)
        return

    desc = scenes.icosphere_grid(args.seed)
    ctx = api.Context(spectral=desc.spectral, max_path_length=args.depth)
    built = api.build_scene(ctx, desc)
    ctx.bind_output(args.width, args.height)
    if args.wavefront_depth is not None:
        ctx.set_wavefront_depth(args.wavefront_depth)

    # warm-up: at least --warmup frames, and on until the frame-schedule autotuner has timed
    # every candidate (its frames synchronise), so the timed region runs the tuned schedule only
    t_setup = time.perf_counter()
    warm = 0
    while warm < args.warmup or (args.wavefront_depth is None and not ctx.schedule()[2]):
        ctx.render(built.camera, first_frame=(warm == 0), limit=1 << 30)
        warm += 1
        if warm > 500:
            raise RuntimeError("the schedule autotuner did not finish within 500 frames")
    hip_sync()
    setup_s = time.perf_counter() - t_setup
    accel = ctx.accel_stats()
    sched_tuned_before = ctx.schedule()[2]

    # timed region: HIP events around every launch, on the stream each kernel runs on, read
    # back after the region (they do not synchronise the frames)
    ctx.reset_stats()
    ctx.set_kernel_timing(True)
    hip_sync()
    t0 = time.perf_counter()
    for k in range(args.steps):
        ctx.render(built.camera, first_frame=(k == 0), limit=args.steps)
    hip_sync()
    wall = time.perf_counter() - t0
    st = ctx.stats()
    kt = st
    ctx.set_kernel_timing(False)
    qc, qs = ctx.queue_counts(args.depth)
    sched_wd, sched_tb, sched_tuned = ctx.schedule()
    sched_sort = ctx.shade_sorting()
    img = ctx.read_output()

    rays = st["closestRays"] + st["shadowRays"]
    value = rays / wall
    samples_per_s = st["paths"] / wall

    # traversal work per ray, from extra frames with the counting kernels
    ctx.reset_stats()
    ctx.set_traversal_stats(True)
    for k in range(args.stat_frames):
        ctx.render(built.camera, first_frame=True, limit=1)
    hip_sync()
    tv = ctx.stats()
    ctx.set_traversal_stats(False)
    per_closest = {k: tv[k + "Closest"] / max(1, tv["closestRays"]) for k in ("nodes", "triangles", "instances")}
    per_shadow = {k: tv[k + "Shadow"] / max(1, tv["shadowRays"]) for k in ("nodes", "triangles", "instances")}

    def traversal_bytes(io, per):
        return io + per["nodes"] * NODE + per["triangles"] * TRI + per["instances"] * INST

    b_closest = traversal_bytes(RAY_IN + HIT_OUT, per_closest)
    b_shadow = traversal_bytes(SHADOW_IN + SHADOW_RMW, per_shadow)
    # rays by stage: per-depth kernels for the first depths, the persistent tail kernel for the rest
    wf_c, wf_s = kt["closestRaysWavefront"], kt["shadowRaysWavefront"]
    tail_c, tail_s = kt["closestRays"] - wf_c, kt["shadowRays"] - wf_s
    kernels = {
        "k_intersect": (kt["msIntersect"], kt["launchesIntersect"], wf_c * b_closest),
        "k_shadow": (kt["msShadow"], kt["launchesShadow"], wf_s * b_shadow),
        "k_shade": (kt["msShade"], kt["launchesShade"], wf_c * SHADE),
        "k_tail": (kt["msTail"], kt["launchesTail"], tail_c * (b_closest + SHADE) + tail_s * b_shadow),
    }
    dominant = max(kernels, key=lambda k: kernels[k][0])
    ms, launches, total_bytes = kernels[dominant]
    per_launch_bytes = total_bytes / max(1, launches)
    mean_launch_s = ms / 1e3 / max(1, launches)
    achieved = per_launch_bytes / mean_launch_s / 1e9
    traffic = read_traffic(dominant)

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "rays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": WORKLOAD, "width": args.width, "height": args.height, "spp": args.steps,
                   "max_path_length": args.depth, "scene_seed": args.seed, "spectral": True},
        "samples_per_s": samples_per_s,
        "rays": {"closest": st["closestRays"], "shadow": st["shadowRays"], "paths": st["paths"],
                 "dropped_non_finite": st["droppedNonFinite"]},
        "stage_ms_per_frame": {k: kt["ms" + k.capitalize()] / args.steps
                               for k in ("generate", "intersect", "shade", "shadow", "tail", "accumulate", "convert")},
        "kernel_timing": "HIP events around every launch inside the timed region",
        "schedule": {"wavefront_depth": sched_wd, "tail_blocks": sched_tb, "material_sort": sched_sort,
                     "autotuned": bool(sched_tuned and sched_tuned_before), "warmup_frames_run": warm},
        "rays_by_stage": {"per_depth_closest": wf_c, "per_depth_shadow": wf_s, "tail_closest": tail_c,
                          "tail_shadow": tail_s},
        "traversal_per_ray": {"closest": per_closest, "shadow": per_shadow},
        "last_frame_rays_per_depth": {"closest": qc.tolist(), "shadow": qs.tolist()},
        "roofline": {"bound": "hbm", "kernel": dominant, "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "traffic_source": None if traffic is None else
                     f"profiles/{dominant}_icosphere_grid_1080p_traffic.json (rocprofv3 FETCH_SIZE + WRITE_SIZE "
                     f"of this workload, scaled by a copy of known size; not measured in this run)",
                     "algorithmic_bytes_per_launch": per_launch_bytes, "mean_launch_ms": mean_launch_s * 1e3},
        "image_mean_rgb": [float(x) for x in img[..., :3].reshape(-1, 3).mean(axis=0)],
        "setup_s": setup_s,
        "accel_build": {"builder": "device", "ms_blas": accel["msBlas"], "ms_tlas": accel["msTlas"],
                        "blas_count": accel["numBlas"], "triangles": accel["numTriangles"],
                        "node_program_steps": accel["numNodeProgramSteps"]},
        "rmse": None,
    }
    if not args.no_rmse:
        log = lambda m: print(f"rmse: {m}", file=sys.stderr, flush=True)
        r = cache_experiment(args.width, args.height, args.rmse_ref_spp, args.rmse_spp, args.cache_warmup,
                             seed=20261202, spectral=True, log=log)
        last = r["equal_spp"][-1]
        out["rmse"] = {
            "workload": r["workload"],
            "reference": r["reference"],
            "equal_spp": {"spp": last["spp"], "plain": last["plain"], "cached": last["cached"]},
            "plain_seconds": r["plain_rmse_at_spp"]["seconds"],
            "cached_seconds_to_plain_rmse": r["cached_seconds_to_reach_it"],
            "cached_time_fraction": r["cached_time_fraction"],
            "equal_time": r["equal_time"],
            "curves": {"plain": r["plain"], "cached": r["cached"], "equal_spp": r["equal_spp"]},
            "cache_warmup": r["cache_warmup"],
            "cache_ms_per_frame": r["cache_ms_per_frame"],
            "rays_per_path": r["rays_per_path"],
        }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(desc, args.width, args.height, args.depth, args.cpu_seconds)
        out["gpu_over_cpu"] = value / out["cpu_baseline"]["value"]
    ctx.close()
    if args.copy_probe:
        copy_probe(args.copy_probe)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
