#!/usr/bin/env python3
"""Benchmark of the ray-tracing hot path on one MI355X.

Workload (one "step"): the intersection microbenchmark -- a displaced
icosphere subdivided to 1,310,720 triangles (0.05 * fBm, 6 octaves, seed 1)
plus 10,000 spheres (radius log-uniform 0.001..0.05), a binned-SAH BVH
(12 bins) built on the host, and four launches over rays already resident in
HBM: closest hit and any hit on 2048x2048 coherent pinhole primary rays and
on 4,194,304 incoherent rays (origin uniform in the bounds, direction uniform
on the sphere).  `value` is ray queries per second over the whole step.

Two more legs are reported in the same line (skip them with --legs):
  render  the Cornell box workload (256x256, 4x4 stratified samples per pixel,
          max depth 5, Russian roulette after depth 3) on the plain path:
          samples/s, device-counted Mrays/s, and the image RMSE (luminance)
          against a 4096-spp oracle render committed under tests/golden;
  serve   a Poisson stream of render jobs over 8 resident scenes (the Cornell
          box and 7 procedural box-stack / displaced-grid / icosphere scenes of
          0.1M-8M triangles), each job with its own resolution and spp, run
          once through the served path (continuous batching) and once through
          the plain path (one job at a time, in arrival order): Mrays/s,
          samples/s and p50/p99 job latency per arrival rate, and whether
          every job's image is the same on both paths.
The optional `build` leg (--legs build) times scene creation of the largest
serving scene with the host and the device BVH build.

Prints one JSON line.  Only N=1 is supported: this build targets one GPU.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Mrays/s and p50/p99 job latency, served vs plain GPU path at equal RMSE"
PEAK_HBM_GBS = 8000.0
# algorithmic bytes per unit of work: a ray is 8 floats in, a closest hit
# 4 words out, an any-hit 1 byte out; a node of width w is w child boxes of
# 24 B plus w links of 4 B (28 w B); a triangle test reads 9 floats
B_RAY_IN, B_HIT_OUT, B_OCC_OUT = 32, 16, 1
B_NODE_PER_CHILD, B_TRI = 28, 36
CORNELL_REF = os.path.join(ROOT, "tests", "golden", "cornell256_4096spp.npz")


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def host_threads():
    """CPU threads for the baseline: the host's share (OMP_NUM_THREADS when
    set), never more than the cores visible."""
    n = os.cpu_count() or 1
    return max(1, min(n, int(os.environ.get("OMP_NUM_THREADS", n))))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline legs")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--levels", type=int, default=8)
    ap.add_argument("--res", type=int, default=2048)
    ap.add_argument("--only", default=None,
                    help="time only this launch (e.g. closest/incoherent); for counter runs")
    ap.add_argument("--copy-ref", action="store_true",
                    help="also run a plain 1 GiB device copy, the reference for scaling "
                         "FETCH_SIZE / WRITE_SIZE counters")
    ap.add_argument("--legs", default="intersect,render,divergent,serve",
                    help="comma-separated subset of intersect,render,divergent,serve")
    ap.add_argument("--render-steps", type=int, default=10)
    ap.add_argument("--serve-jobs", type=int, default=32, help="jobs per arrival rate")
    ap.add_argument("--serve-rates", default="5,50,200", help="Poisson arrival rates, jobs/s")
    ap.add_argument("--serve-max-tris", type=float, default=8e6)
    ap.add_argument("--serve-res", default="128,256,512,1024")
    ap.add_argument("--serve-spp", default="4,16,64")
    ap.add_argument("--build-reps", type=int, default=3, help="scene creations per builder in the build leg")
    ap.add_argument("--config", default="",
                    help="tuning knobs, key=value[,key=value] (bgpu_config_set keys)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: only --gpus 1 is supported (this build runs on one GPU; "
              "the multi-GPU sharding is a design note, not a code path)", file=sys.stderr)
        sys.exit(2)
    legs = set(args.legs.split(","))

    # torch's bundled HIP runtime has the same soname as the system one that
    # libbgpu.so links: loading torch first makes the process use one runtime
    import torch  # noqa: F401
    from bling_amd.runtime import Context
    ctx = Context(0)
    knobs = {}
    for kv in filter(None, args.config.split(",")):
        k, v = kv.split("=")
        ctx.config_set(k, int(v))
    for k in ("bvh.width", "bvh.device", "trace.block", "trace.persistent", "pt.block", "pt.pool", "pt.shade_waves",
              "pt.persistent", "pt.sort"):
        knobs[k] = ctx.config_get(k)
    out = {
        "metric": METRIC,
        "value": None,
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": None,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"knobs": knobs},
    }
    if "intersect" in legs:
        intersect_leg(ctx, args, out)
    if "render" in legs:
        out["config"]["render"] = render_leg(ctx, args)
    if "divergent" in legs:
        out["config"]["divergent"] = divergent_leg(ctx, args)
    if "serve" in legs:
        out["config"]["serve"] = serve_leg(ctx, args)
    if "build" in legs:
        out["config"]["build"] = build_leg(ctx, args)
    ctx.close()
    print(json.dumps(out), flush=True)


# ------------------------------------------------------------ intersection
def intersect_leg(ctx, args, out):
    import ctypes as C

    import torch
    from bling_amd import abi, scenes

    t0 = time.time()
    flat = scenes.icosphere_scene(levels=args.levels, num_spheres=10000, res=args.res)
    lo, hi = scenes.scene_bounds(flat)
    nr = args.res * args.res
    batches = {
        "coherent": scenes.primary_rays(flat),
        "incoherent": scenes.incoherent_rays(lo, hi, nr, seed=2),
    }
    log("workload built in %.1fs: %d tris, %d shapes, 2x%d rays" % (
        time.time() - t0, flat.num_triangles, flat.num_shapes, nr))

    t0 = time.time()
    gs = ctx.scene(flat)
    build_s = time.time() - t0
    log("scene upload + BVH build %.2fs, %.1f MB on device" % (build_s, gs.device_bytes() / 1e6))

    dev = torch.device("cuda:0")
    dbat = {}
    for name, rays in batches.items():
        cols = [torch.from_numpy(np.ascontiguousarray(rays[:, k])).to(dev) for k in range(8)]
        outs = [torch.empty(nr, dtype=torch.float32, device=dev) for _ in range(3)]
        outs.append(torch.empty(nr, dtype=torch.int32, device=dev))
        occ = torch.empty(nr, dtype=torch.uint8, device=dev)
        dbat[name] = (cols, outs, occ)
    torch.cuda.synchronize()

    launches = []  # (label, fn)
    for name, (cols, outs, occ) in dbat.items():
        rp = [c.data_ptr() for c in cols]
        hp = [o.data_ptr() for o in outs]
        coh = name == "coherent"  # camera rays: the caller's BGPU_COHERENT hint
        launches.append(("closest/" + name,
                         lambda rp=rp, hp=hp, coh=coh, **kw: gs.intersect_dev(rp, nr, hp, coherent=coh, **kw)))
        launches.append(("any/" + name,
                         lambda rp=rp, op=occ.data_ptr(), coh=coh, **kw: gs.intersects_dev(rp, nr, op, coherent=coh, **kw)))

    # counting pass: node visits and primitive tests per launch (not timed)
    work = {}
    for label, fn in launches:
        ctx.reset_stats()
        kind, batch = label.split("/")
        cols, outs, occ = dbat[batch]
        r = abi.Rays(*[c.data_ptr() for c in cols])
        if kind == "closest":
            h = abi.Hits(*[o.data_ptr() for o in outs])
            abi.check(ctx.lib.bgpu_intersect_closest(gs.h, C.byref(r), nr, C.byref(h),
                                                     abi.DEVICE_PTRS | abi.COUNT))
        else:
            abi.check(ctx.lib.bgpu_intersect_any(gs.h, C.byref(r), nr, occ.data_ptr(),
                                                 abi.DEVICE_PTRS | abi.COUNT))
        st = ctx.stats()
        work[label] = (st["nodes_visited"], st["prims_tested"])
    p_c = dbat["incoherent"][1][3].cpu().numpy()
    shape_frac = float((p_c >= flat.num_triangles).mean())

    b_node = B_NODE_PER_CHILD * gs.bvh_info()["width"]

    def alg_bytes(label):
        nodes, prims = work[label]
        o = B_HIT_OUT if label.startswith("closest") else B_OCC_OUT
        return nr * (B_RAY_IN + o) + nodes * b_node + prims * B_TRI

    if args.copy_ref:
        # an elementwise kernel that reads exactly 1 GiB and writes exactly 1 GiB
        src = torch.ones(256 * 1024 * 1024, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        for _ in range(3):
            torch.add(src, 1.0, out=dst)
        torch.cuda.synchronize()
        del src, dst
    if args.only:
        launches = [(l, f) for l, f in launches if l == args.only]

    # warmup + timed steps
    for _ in range(args.warmup):
        for _, fn in launches:
            fn()
    torch.cuda.synchronize()
    kms = {label: [] for label, _ in launches}
    t0 = time.perf_counter()
    for _ in range(args.steps):
        for label, fn in launches:
            kms[label].append(fn(timed=True))
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    queries = len(launches) * nr * args.steps
    avg = {k: float(np.mean(v)) for k, v in kms.items()}
    dom = max(avg, key=lambda k: avg[k])
    achieved = alg_bytes(dom) / (avg[dom] * 1e-3) / 1e9
    per_launch = {k: {"ms": round(avg[k], 4), "mrays_s": round(nr / avg[k] / 1e3, 1),
                      "nodes_per_ray": round(work[k][0] / nr, 2),
                      "prims_per_ray": round(work[k][1] / nr, 2),
                      "alg_bytes": int(alg_bytes(k)),
                      "alg_GBps": round(alg_bytes(k) / (avg[k] * 1e-3) / 1e9, 1)} for k in avg}
    log("per launch:", json.dumps(per_launch))

    traffic = None
    pmc_path = os.path.join(ROOT, "profiles", "closest_incoherent_c2_pmc.json")
    if os.path.exists(pmc_path):
        with open(pmc_path) as f:
            traffic = json.load(f).get("hbm_bytes_per_launch")

    out["value"] = round(queries / wall / 1e6, 2)
    out["ms_per_step"] = round(wall / args.steps * 1e3, 3)
    out["config"].update({
        "workload": "intersection microbench: displaced icosphere %d tris + %d spheres, "
                    "binned-SAH BVH2 (12 bins); closest+any hit on %d coherent (%dx%d pinhole) "
                    "and %d incoherent rays per step" % (flat.num_triangles, flat.num_shapes,
                                                         nr, args.res, args.res, nr),
        "rays_per_step": len(launches) * nr,
        "bvh_build_s": round(build_s, 2),
        "bvh": gs.bvh_info(),
        "per_launch": per_launch,
        "shape_hit_fraction_incoherent": round(shape_frac, 4),
    })
    out["roofline"] = {
        "bound": "hbm",
        "kernel": dom,
        "achieved": round(achieved, 1),
        "peak": PEAK_HBM_GBS,
        "unit": "GB/s",
        "frac": round(achieved / PEAK_HBM_GBS, 4),
        "traffic": traffic,
    }
    out["cpu_baseline"] = None if args.no_cpu else cpu_baseline(flat, batches, args.cpu_seconds)
    gs.close()


def cpu_baseline(flat, batches, seconds):
    """The oracle (C++ restatement, OpenMP over rays) on a bounded sample of
    the same four query batches, on this host's cores."""
    from oracle.pyoracle import OracleScene
    threads = host_threads()
    o = OracleScene(flat)
    n = 4096
    rng = np.random.default_rng(0)
    samples = {k: v[rng.permutation(len(v))] for k, v in batches.items()}
    total_q, total_t, reps = 0, 0.0, 0
    full = len(next(iter(samples.values())))
    while True:  # grow the sample to the whole batch, then repeat it
        t0 = time.perf_counter()
        for k, v in samples.items():
            o.intersect(v[:n], threads=threads)
            o.intersects(v[:n], threads=threads)
        dt = time.perf_counter() - t0
        if n < full and dt * 4 <= seconds:
            n = min(full, n * 4)
            continue
        total_q += 4 * n
        total_t += dt
        reps += 1
        if total_t >= seconds:
            break
    # the same restatement on one core, over a quarter of the time
    n1 = 4096
    q1, t1 = 0, 0.0
    while t1 < seconds / 4:
        t0 = time.perf_counter()
        for k, v in samples.items():
            o.intersect(v[:n1], threads=1)
            o.intersects(v[:n1], threads=1)
        t1 += time.perf_counter() - t0
        q1 += 4 * n1
    o.close()
    log("cpu baseline: %d queries in %.2fs on %d threads; %d in %.2fs on 1" % (total_q, total_t, threads, q1, t1))
    return {"value": round(total_q / total_t / 1e6, 4), "unit": "Mrays/s", "cores": threads,
            "kind": "port",
            "sample": "%d rays from each of the 4 query batches (closest+any, coherent+incoherent), "
                      "%d time(s) over, %.1f s; oracle C++ restatement with its own median BVH, "
                      "%d OpenMP threads" % (n, reps, total_t, threads),
            "value_1core": round(q1 / t1 / 1e6, 4),
            "sample_1core": "4096 rays from each of the 4 query batches, %d queries in %.1f s, one thread"
                            % (q1, t1)}


# ------------------------------------------------------------------ render
def rays_of(st):
    return st["rays_camera"] + st["rays_extend"] + st["rays_shadow"] + st["rays_mis"]


def luminance(img):
    w = np.maximum(img[..., 0], 1e-30)
    return img[..., 2] / w


def render_leg(ctx, args):
    """Cornell box 256x256, 16 spp (4x4), max depth 5, RR after depth 3."""
    from bling_amd import scenes
    from bling_amd.runtime import GpuRenderer, JobSpec
    flat = scenes.cornell(256)
    gs = ctx.scene(flat)
    mk = lambda seed: JobSpec(spp_nu=4, spp_nv=4, max_depth=5, sample_depth=3, rr_start_depth=3,
                              passes=1, seed=seed)
    for k in range(2):
        GpuRenderer(mk(1000 + k)).render_plain(gs)
    ctx.reset_stats()
    imgs = []
    t0 = time.perf_counter()
    for k in range(args.render_steps):
        img, _ = GpuRenderer(mk(1 + k)).render_plain(gs)
        imgs.append(img)
    wall = time.perf_counter() - t0
    st = ctx.stats()
    gs.close()
    res = {
        "workload": "Cornell box 256x256, 4x4 stratified spp, max depth 5, RR after depth 3, "
                    "plain path, one render per step",
        "steps": args.render_steps,
        "ms_per_render": round(wall / args.render_steps * 1e3, 3),
        "msamples_s": round(st["samples"] / wall / 1e6, 2),
        "mrays_s": round(rays_of(st) / wall / 1e6, 2),
        "rays_per_sample": round(rays_of(st) / max(1, st["samples"]), 3),
        "ray_mix": {k: st[k] for k in ("rays_camera", "rays_extend", "rays_shadow", "rays_mis")},
    }
    if os.path.exists(CORNELL_REF):
        ref = np.load(CORNELL_REF)["y"]
        rm = [float(np.sqrt(np.mean((luminance(i) - ref) ** 2))) for i in imgs]
        res["rmse_y_vs_4096spp"] = round(float(np.mean(rm)), 6)
        res["reference_image"] = "oracle CPU render, 4096 spp (tests/golden/cornell256_4096spp.npz)"
        if not args.no_cpu:
            res["cpu"] = render_cpu(flat, mk, ref, args.cpu_seconds)
    log("render:", json.dumps(res))
    return res


def render_cpu(flat, mk, ref, seconds):
    """The oracle integrator on the same job (16 spp, all 66049 extent
    pixels), repeated with new seeds until `seconds` of work: samples/s on
    this host, and the image RMSE of the first render (equal spp)."""
    from oracle import pyoracle
    threads = host_threads()
    n = (flat.film[0] + 1) * (flat.film[1] + 1) * 16
    dt, rays, renders, rmse = 0.0, 0, 0, None
    while dt < seconds:
        t0 = time.perf_counter()
        img, _, r = pyoracle.render(flat, mk(1 + renders), threads=threads)
        dt += time.perf_counter() - t0
        rays += r
        renders += 1
        if rmse is None:
            rmse = float(np.sqrt(np.mean((luminance(img) - ref) ** 2)))
    return {"msamples_s": round(n * renders / dt / 1e6, 4), "mrays_s": round(rays / dt / 1e6, 3),
            "cores": threads, "kind": "port", "seconds": round(dt, 2), "renders": renders,
            "rmse_y_vs_4096spp": round(rmse, 6)}


# --------------------------------------------------------------- divergent
def divergent_leg(ctx, args):
    """Divergent materials: 1024^2 heightmap + 20 tori of 20k triangles +
    200 spheres (2,493,058 triangles), matte/plastic/metal/glass, a 512x256
    sky map; 1024x1024, 64 spp (8x8), max depth 16, RR after depth 3."""
    from bling_amd import scenes
    from bling_amd.runtime import GpuRenderer, JobSpec
    t0 = time.time()
    flat = scenes.divergent_scene()
    gen_s = time.time() - t0
    t0 = time.time()
    gs = ctx.scene(flat)
    build_s = time.time() - t0
    spec = JobSpec(spp_nu=8, spp_nv=8, max_depth=16, sample_depth=3, rr_start_depth=3, passes=1, seed=5)
    GpuRenderer(JobSpec(spp_nu=1, spp_nv=1, max_depth=16, sample_depth=3, rr_start_depth=3,
                        seed=6, view=flat.view(256, 256))).render_plain(gs)
    ctx.reset_stats()
    t0 = time.perf_counter()
    img, _ = GpuRenderer(spec).render_plain(gs)
    wall = time.perf_counter() - t0
    st = ctx.stats()
    res = {
        "workload": "divergent materials: %d triangles + %d spheres, 1024x1024, 8x8 spp, max depth 16, "
                    "RR after depth 3, sky map 512x256, plain path" % (flat.num_triangles, flat.num_shapes),
        "scene_gen_s": round(gen_s, 2), "bvh_build_s": round(build_s, 2), "bvh": gs.bvh_info(),
        "ms_per_render": round(wall * 1e3, 1),
        "msamples_s": round(st["samples"] / wall / 1e6, 2),
        "mrays_s": round(rays_of(st) / wall / 1e6, 2),
        "rays_per_sample": round(rays_of(st) / max(1, st["samples"]), 3),
        "image_finite": bool(np.isfinite(img).all()),
    }
    gs.close()
    if not args.no_cpu:
        from oracle import pyoracle
        threads = host_threads()
        osc = pyoracle.OracleScene(flat)  # its build is not timed
        dt, rays, renders = 0.0, 0, 0
        while dt < args.cpu_seconds:
            small = JobSpec(spp_nu=2, spp_nv=2, max_depth=16, sample_depth=3, rr_start_depth=3,
                            seed=5 + renders, view=flat.view(256, 256))
            t0 = time.perf_counter()
            _, _, r = pyoracle.render(flat, small, threads=threads, scene=osc)
            dt += time.perf_counter() - t0
            rays += r
            renders += 1
        osc.close()
        n = 257 * 257 * 4 * renders
        res["cpu"] = {"msamples_s": round(n / dt / 1e6, 4), "mrays_s": round(rays / dt / 1e6, 3),
                      "cores": threads, "kind": "port", "seconds": round(dt, 2),
                      "sample": "the same scene at 256x256, 2x2 spp (264,196 camera samples), %d "
                                "render(s) with new seeds" % renders}
    log("divergent:", json.dumps(res))
    return res


# ------------------------------------------------------------------- serve
def build_leg(ctx, args):
    """Scene creation (upload, BVH build and collapse) of the largest serving
    scene, a displaced grid of about --serve-max-tris triangles, with the
    host BVH build and with the device build; best of --build-reps."""
    from bling_amd import scenes
    t0 = time.time()
    flat = scenes.grid_scene(int(args.serve_max_tris), seed=6)
    log("build: grid %d tris generated in %.1fs" % (flat.num_triangles, time.time() - t0))
    res = {}
    for dev in (0, 1):
        ctx.config_set("bvh.device", dev)
        best, info = None, None
        for _ in range(args.build_reps):
            t1 = time.perf_counter()
            g = ctx.scene(flat)
            dt = time.perf_counter() - t1
            info = g.bvh_info()
            g.close()
            best = dt if best is None else min(best, dt)
        res["device" if dev else "host"] = {"create_s": round(best, 4), "bvh2_ms": round(info["build_ms"], 1),
                                            "device_tree": info["device"], "nodes": info["nodes"]}
        log("build: bvh.device=%d: scene created in %.3fs (binary build %.1f ms)" % (dev, best, info["build_ms"]))
    ctx.config_set("bvh.device", 1)
    return {"workload": "scene creation of a displaced grid of %d triangles (upload, BVH build, collapse to "
                        "width %d); best of %d" % (flat.num_triangles, ctx.config_get("bvh.width"), args.build_reps),
            **res, "speedup": round(res["host"]["create_s"] / max(1e-9, res["device"]["create_s"]), 2)}


def serve_scenes(max_tris):
    """The Cornell box plus 7 procedural scenes, 0.1M..max_tris triangles
    (log-spaced), alternating box stacks, displaced grids and an icosphere."""
    from bling_amd import scenes
    sizes = np.exp(np.linspace(np.log(1e5), np.log(max_tris), 7))
    out = [("cornell", scenes.cornell(256))]
    kinds = ["stack", "grid", "icosphere", "stack", "grid", "stack", "grid"]
    for k, (kind, n) in enumerate(zip(kinds, sizes)):
        if kind == "stack":
            out.append(("stack", scenes.stack_scene(int(n), seed=k)))
        elif kind == "grid":
            out.append(("grid", scenes.grid_scene(int(n), seed=k)))
        else:
            lv = int(np.clip(np.round(np.log(n / 20.0) / np.log(4.0)), 3, 9))
            out.append(("icosphere", scenes.icosphere_scene(levels=lv, num_spheres=1000, seed=k, res=256)))
    return out


def job_stream(n, rate, nscenes, res_set, spp_set, seed):
    rng = np.random.default_rng(seed)
    gaps = rng.exponential(1.0 / rate, n)
    arrivals = np.cumsum(gaps) - gaps[0]
    jobs = []
    for k in range(n):
        jobs.append({"arrival": float(arrivals[k]), "scene": int(rng.integers(nscenes)),
                     "res": int(rng.choice(res_set)), "spp": int(rng.choice(spp_set)),
                     "seed": int(seed * 100003 + k)})
    return jobs


def spec_for(job, flat):
    from bling_amd.runtime import JobSpec
    s = int(round(np.sqrt(job["spp"])))
    nu, nv = (s, job["spp"] // s) if s * (job["spp"] // s) == job["spp"] else (job["spp"], 1)
    return JobSpec(spp_nu=nu, spp_nv=nv, max_depth=5, sample_depth=3, rr_start_depth=3, passes=1,
                   seed=job["seed"], view=flat.view(job["res"], job["res"]))


def lat_stats(lat):
    a = np.array(lat) * 1e3
    return {"p50_ms": round(float(np.percentile(a, 50)), 2), "p99_ms": round(float(np.percentile(a, 99)), 2),
            "mean_ms": round(float(a.mean()), 2)}


def run_plain(ctx, gss, flats, jobs):
    from bling_amd.runtime import GpuRenderer
    ctx.reset_stats()
    t_start = time.perf_counter()
    lat, imgs, t_end = [], [], 0.0
    for job in jobs:
        now = time.perf_counter() - t_start
        if now < job["arrival"]:
            time.sleep(job["arrival"] - now)
        img, _ = GpuRenderer(spec_for(job, flats[job["scene"]])).render_plain(gss[job["scene"]])
        t_end = time.perf_counter() - t_start
        lat.append(t_end - job["arrival"])
        imgs.append(img)
    st = ctx.stats()
    return lat, imgs, t_end, st


def run_served(ctx, gss, flats, jobs):
    from bling_amd.runtime import Job
    ctx.reset_stats()
    t_start = time.perf_counter()
    handles, lag = [], []
    for job in jobs:
        now = time.perf_counter() - t_start
        if now < job["arrival"]:
            time.sleep(job["arrival"] - now)
        lag.append(time.perf_counter() - t_start - job["arrival"])
        handles.append(Job(gss[job["scene"]], spec_for(job, flats[job["scene"]])))
    lat, imgs = [], []
    for h, lg in zip(handles, lag):
        h.wait(timeout=600)
    t_end = time.perf_counter() - t_start
    for h, lg in zip(handles, lag):
        lat.append(lg + h.latency_ms() * 1e-3)
        imgs.append(h.poll()[0])
        h.release()
    st = ctx.stats()
    makespan = max(j["arrival"] + l for j, l in zip(jobs, lat))
    return lat, imgs, min(t_end, makespan), st


def serve_leg(ctx, args):
    t0 = time.time()
    named = serve_scenes(args.serve_max_tris)
    flats = [f for _, f in named]
    log("serve: %d scenes generated in %.1fs" % (len(flats), time.time() - t0))
    # scene admission: creation with the host BVH build, then with the device
    # build (the one kept); both include the upload of the shading data
    t0 = time.time()
    gss, admission = [], []
    for (name, f) in named:
        ctx.config_set("bvh.device", 0)
        t1 = time.perf_counter()
        g = ctx.scene(f)
        host_s = time.perf_counter() - t1
        g.close()
        ctx.config_set("bvh.device", 1)
        t1 = time.perf_counter()
        g = ctx.scene(f)
        dev_s = time.perf_counter() - t1
        info = g.bvh_info()
        gss.append(g)
        admission.append({"scene": name, "tris": f.num_triangles, "host_build_s": round(host_s, 3),
                          "device_build_s": round(dev_s, 3), "device_tree": info["device"],
                          "device_bvh2_ms": round(info["build_ms"], 1)})
        log("serve: scene %s %d tris created in %.3fs (host BVH) / %.3fs (device BVH)" % (
            name, f.num_triangles, host_s, dev_s))
    log("serve: uploaded + BVHs in %.1fs, %.1f MB" % (time.time() - t0,
                                                       sum(g.device_bytes() for g in gss) / 1e6))
    res_set = [int(v) for v in args.serve_res.split(",")]
    spp_set = [int(v) for v in args.serve_spp.split(",")]
    # warm both paths (engine creation, graph capture)
    warm = job_stream(2, 1000.0, len(flats), [128], [1], 999)
    run_plain(ctx, gss, flats, warm)
    run_served(ctx, gss, flats, warm)
    rows = []
    for ri, rate in enumerate(float(r) for r in args.serve_rates.split(",")):
        jobs = job_stream(args.serve_jobs, rate, len(flats), res_set, spp_set, seed=ri + 1)
        p_lat, p_img, p_span, p_st = run_plain(ctx, gss, flats, jobs)
        s_lat, s_img, s_span, s_st = run_served(ctx, gss, flats, jobs)
        worst = 0.0
        for a, b in zip(s_img, p_img):
            d = np.abs(a - b) / np.maximum(np.abs(b), 1e-3)
            worst = max(worst, float(d.max()))
        row = {
            "rate_jobs_s": rate,
            "jobs": len(jobs),
            "samples": int(s_st["samples"]),
            "served": dict(mrays_s=round(rays_of(s_st) / s_span / 1e6, 2),
                           msamples_s=round(s_st["samples"] / s_span / 1e6, 2),
                           makespan_s=round(s_span, 3), **lat_stats(s_lat)),
            "plain": dict(mrays_s=round(rays_of(p_st) / p_span / 1e6, 2),
                          msamples_s=round(p_st["samples"] / p_span / 1e6, 2),
                          makespan_s=round(p_span, 3), **lat_stats(p_lat)),
            "rays_served": int(rays_of(s_st)),
            "rays_plain": int(rays_of(p_st)),
            "image_max_rel_diff": worst,
        }
        row["speedup_mrays"] = round(row["served"]["mrays_s"] / max(1e-9, row["plain"]["mrays_s"]), 3)
        log("serve:", json.dumps(row))
        rows.append(row)
    for g in gss:
        g.close()
    return {
        "workload": "Poisson job stream over 8 resident scenes (%s), resolution uniform in %s, "
                    "spp uniform in %s, max depth 5, RR after depth 3; served = continuous batching "
                    "through bgpu_job_submit, plain = bgpu_render one job at a time in arrival order; "
                    "same seeds on both paths (equal RMSE: the images agree to film-atomic rounding)"
                    % (", ".join("%s %d tris" % (k, f.num_triangles) for k, f in named),
                       res_set, spp_set),
        "rates": rows,
        "scene_admission": admission,
    }


if __name__ == "__main__":
    main()
