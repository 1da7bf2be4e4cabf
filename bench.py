#!/usr/bin/env python3
"""Benchmark of the ray-tracing hot path on one MI355X.

Headline (`value`): Mrays/s of the served path on the serving stream at its
saturating arrival rate -- Poisson arrivals (200 jobs/s) of render jobs over
8 resident scenes (the Cornell box and 7 procedural box-stack /
displaced-grid / icosphere scenes of 0.1M-8M triangles), resolution uniform
in {128, 256, 512, 1024}^2, spp uniform in {4, 16, 64}, path integrator with
max depth 5 and Russian roulette after depth 3.  One step is 10 consecutive
jobs of the stream; `--steps K` times 10 K jobs, after `--warmup W` jobs.
The same stream goes through the plain path (bgpu_render, one job at a time
in arrival order) with the same seeds, so both paths render the same images
(equal RMSE); each path reports Mrays/s, samples/s, p50/p99 job latency and
its GPU-busy fraction (device time of the bounce loop / makespan).  Rates 50
and 5 jobs/s are reported beside it.

Further legs, each in `config` (select with --legs):
  batch      64 jobs submitted at once, each on its own scene of log-uniform
             1k-2M triangles at 128^2-1024^2 with 1-32 spp, served vs plain;
  intersect  closest and any hit on 4,194,304 coherent and 4,194,304
             incoherent rays against a displaced icosphere of 1,310,720
             triangles plus 10,000 spheres (rays resident in HBM), checked ray
             for ray against the oracle;
  render     the Cornell box, 256^2, 16 spp, max depth 5, plain path, with the
             image RMSE against a 4096-spp oracle render;
  divergent  matte/plastic/metal/glass over 2.49M triangles and 200 spheres
             under a sky map, 1024^2, 64 spp, max depth 16;
  build      (not default) scene creation with the host and the device BVH build;
  phases     (not default, for counter runs) only the per-phase profiling pass
             of the headline workload, with the items of each phase.

`roofline` is the headline workload's dominant kernel, timed per phase with
HIP events in a profiling pass over its first jobs; `cpu_baseline` is the
oracle (the C++ restatement of the reference) on a bounded sample of the
same jobs, on this host's cores.

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
STEP_JOBS = 10  # jobs of the serving stream per step
# algorithmic bytes per unit of work: a ray is 8 floats in, a closest hit
# 4 words out, an any-hit 1 byte out; a node of width w is w child boxes of
# 24 B plus w links of 4 B (28 w B); a triangle test reads 9 floats; a
# shape test reads its type/parameter index, the 48-byte w2o matrix and 16 B
# of parameters (68 B)
B_RAY_IN, B_HIT_OUT, B_OCC_OUT = 32, 16, 1
B_NODE_PER_CHILD, B_TRI, B_SHAPE = 28, 36, 68
# VALU operations per unit: a child-box slab test 21 (6 sub/mul pairs, 3
# swaps, 6 min/max), a triangle test 52 (51 and a divide), a shape test 40
# (ray transform and quadratic); against the FP32 issue rate of 256 CUs x 128
# lanes x 2.4 GHz (78.6 TFLOP/s counts an FMA as two)
OPS_SLAB, OPS_TRI, OPS_SHAPE = 21, 52, 40
PEAK_VALU_OPS = 78.6e12 / 2
# path-pool bytes the shading kernel reads and writes per active path: reads
# queue entry, job, ray (7 words), hit (3), primitive, depth, pixel, sample,
# pass, throughput (16), radiance (3) = 144 B and the triangle's indices and
# vertices (48 B); writes ray, depth, state, throughput, radiance, the shadow
# ray with its pending XYZ (11 words), the MIS ray with its pending XYZ and
# light (12 words) and two queue entries = 224 B
B_SHADE_PATH = 144 + 48 + 224
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
    ap.add_argument("--steps", type=int, default=20, help="steps of %d serving jobs (saturating rate)" % STEP_JOBS)
    ap.add_argument("--warmup", type=int, default=3, help="warmup jobs per path")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline legs")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--levels", type=int, default=8)
    ap.add_argument("--res", type=int, default=2048)
    ap.add_argument("--isect-steps", type=int, default=20, help="timed repetitions of the 4 intersection launches")
    ap.add_argument("--no-tune", action="store_true",
                    help="time the intersection launches at the current knobs instead of tuning "
                         "trace.block and trace.stack on them first")
    ap.add_argument("--only", default=None,
                    help="time only this launch (e.g. closest/incoherent); for counter runs")
    ap.add_argument("--copy-ref", action="store_true",
                    help="also run a plain 1 GiB device copy, the reference for scaling "
                         "FETCH_SIZE / WRITE_SIZE counters")
    ap.add_argument("--legs", default="serve,batch,intersect,render,divergent",
                    help="comma-separated subset of serve,batch,intersect,render,divergent,build")
    ap.add_argument("--render-steps", type=int, default=10)
    ap.add_argument("--serve-rates", default="200,50,5", help="Poisson arrival rates, jobs/s; the first is the headline")
    ap.add_argument("--serve-slow-jobs", type=int, default=24, help="jobs at rates below 20 jobs/s (arrival-bound)")
    ap.add_argument("--serve-max-tris", type=float, default=8e6)
    ap.add_argument("--serve-res", default="128,256,512,1024")
    ap.add_argument("--serve-spp", default="4,16,64")
    ap.add_argument("--profile-jobs", type=int, default=20, help="jobs of the per-phase profiling pass")
    ap.add_argument("--batch-jobs", type=int, default=64)
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
    for k in ("bvh.width", "bvh.device", "trace.block", "trace.persistent", "trace.stack", "trace.refill",
              "trace.walk", "pt.block", "pt.pool",
              "pt.shade_waves", "pt.diffuse_waves", "pt.stack", "pt.walk", "pt.persistent", "pt.sort"):
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
    if args.copy_ref:
        copy_ref()
    if "serve" in legs:
        serve_leg(ctx, args, out)
    if "batch" in legs:
        out["config"]["batch"] = batch_leg(ctx, args)
    if "intersect" in legs:
        out["config"]["intersect"] = intersect_leg(ctx, args)
    if "render" in legs:
        out["config"]["render"] = render_leg(ctx, args)
    if "divergent" in legs:
        out["config"]["divergent"] = divergent_leg(ctx, args)
    if "build" in legs:
        out["config"]["build"] = build_leg(ctx, args)
    if "phases" in legs:
        out["config"]["phases"] = phases_leg(ctx, args)
    ctx.close()
    print(json.dumps(out), flush=True)


def copy_ref():
    """An elementwise kernel that reads exactly 1 GiB and writes exactly 1 GiB
    (three dispatches), the scale of the HBM counters in counter runs."""
    import torch
    src = torch.ones(256 * 1024 * 1024, dtype=torch.float32, device="cuda:0")
    dst = torch.empty_like(src)
    for _ in range(3):
        torch.add(src, 1.0, out=dst)
    torch.cuda.synchronize()
    del src, dst


# ------------------------------------------------------------ intersection
def intersect_leg(ctx, args):
    """Closest and any hit on 2048^2 coherent and 4,194,304 incoherent rays
    (intersection microbenchmark), resident in HBM; every result checked
    against the oracle's."""
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
    log("intersect: workload built in %.1fs: %d tris, %d shapes, 2x%d rays" % (
        time.time() - t0, flat.num_triangles, flat.num_shapes, nr))

    t0 = time.time()
    gs = ctx.scene(flat)
    build_s = time.time() - t0
    log("intersect: scene upload + BVH build %.2fs, %.1f MB on device" % (build_s, gs.device_bytes() / 1e6))

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

    # counting pass: node visits, triangle and shape tests per launch (not timed)
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
        work[label] = (st["nodes_visited"], st["prims_tested"] - st["shapes_tested"], st["shapes_tested"],
                       st["shape_nodes_visited"])
    p_c = dbat["incoherent"][1][3].cpu().numpy()
    shape_frac = float((p_c >= flat.num_triangles).mean())

    width = gs.bvh_info()["width"]
    b_node = B_NODE_PER_CHILD * width

    def alg_bytes(label):
        nodes, tris, shapes, _ = work[label]
        o = B_HIT_OUT if label.startswith("closest") else B_OCC_OUT
        return nr * (B_RAY_IN + o) + nodes * b_node + tris * B_TRI + shapes * B_SHAPE

    def alg_ops(label):
        nodes, tris, shapes, _ = work[label]
        return nodes * width * OPS_SLAB + tris * OPS_TRI + shapes * OPS_SHAPE

    if args.only:
        launches = [(l, f) for l, f in launches if l == args.only]

    # the ray-batch kernels' block size and LDS stack depth, tuned on the
    # timed launches themselves (results do not depend on them); the timed
    # region runs the winner
    tuning = None
    if not args.no_tune:
        from bling_amd import tune
        space = {"trace.block": (128, 256), "trace.stack": (8, 12, 16), "trace.refill": (16, 32, 48),
                 "trace.walk": (0, 8, 16, 24, 32)}
        best, table = tune.tune_launches(ctx, [fn for _, fn in launches], space, repeats=3)
        tuning = {"space": {k: list(v) for k, v in space.items()}, "best": best, "table": table,
                  "score": "sum over the launches of the median of 3 BGPU_TIMED kernel times (ms)"}
        log("intersect: tuned", json.dumps(best))

    # warmup + timed repetitions
    for _ in range(max(1, args.warmup)):
        for _, fn in launches:
            fn()
    torch.cuda.synchronize()
    kms = {label: [] for label, _ in launches}
    t0 = time.perf_counter()
    for _ in range(args.isect_steps):
        for label, fn in launches:
            kms[label].append(fn(timed=True))
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    queries = len(launches) * nr * args.isect_steps
    avg = {k: float(np.mean(v)) for k, v in kms.items()}
    dom = max(avg, key=lambda k: avg[k])
    per_launch = {k: {"ms": round(avg[k], 4), "mrays_s": round(nr / avg[k] / 1e3, 1),
                      "nodes_per_ray": round(work[k][0] / nr, 2),
                      "shape_nodes_per_ray": round(work[k][3] / nr, 2),
                      "tris_per_ray": round(work[k][1] / nr, 2),
                      "shapes_per_ray": round(work[k][2] / nr, 3),
                      "alg_bytes": int(alg_bytes(k)),
                      "alg_GBps": round(alg_bytes(k) / (avg[k] * 1e-3) / 1e9, 1),
                      "alg_ops": int(alg_ops(k)),
                      "frac_valu": round(alg_ops(k) / (avg[k] * 1e-3) / PEAK_VALU_OPS, 4)} for k in avg}
    log("intersect: per launch:", json.dumps(per_launch))

    # measured HBM traffic of closest/incoherent (committed counter runs)
    traffic = None
    pmc_path = os.path.join(ROOT, "profiles", "closest_incoherent_c2_pmc.json")
    if os.path.exists(pmc_path):
        with open(pmc_path) as f:
            traffic = json.load(f).get("hbm_bytes_per_launch")
    achieved = alg_bytes(dom) / (avg[dom] * 1e-3) / 1e9
    roofline = {
        "bound": "hbm", "kernel": dom, "achieved": round(achieved, 1), "peak": PEAK_HBM_GBS, "unit": "GB/s",
        "frac": round(achieved / PEAK_HBM_GBS, 4), "traffic": traffic,
        "frac_model": round(achieved / PEAK_HBM_GBS, 4),
        "frac_hbm": (round(traffic / (avg[dom] * 1e-3) / 1e9 / PEAK_HBM_GBS, 4)
                     if traffic and dom == "closest/incoherent" else None),
        "frac_valu": per_launch[dom]["frac_valu"],
        "note": "achieved counts every node and leaf fetch of the traversal (the stats build's visit counts, "
                "shape tests at 68 B), cache hits included, so a cache-resident launch can exceed 1; "
                "frac_hbm divides the measured DRAM traffic (FETCH_SIZE + WRITE_SIZE) by the same time; "
                "frac_valu is the model's VALU operations over the FP32 issue rate",
    }
    res = {
        "workload": "intersection microbench: displaced icosphere %d tris + %d spheres, binned-SAH BVH "
                    "(12 bins, width %d, built on the GPU); closest+any hit on %d coherent (%dx%d pinhole) "
                    "and %d incoherent rays, resident in HBM" % (
                        flat.num_triangles, flat.num_shapes, gs.bvh_info()["width"], nr, args.res, args.res, nr),
        "mrays_s": round(queries / wall / 1e6, 2),
        "ms_per_step": round(wall / args.isect_steps * 1e3, 3),
        "steps": args.isect_steps,
        "bvh_build_s": round(build_s, 2),
        "bvh": gs.bvh_info(),
        "per_launch": per_launch,
        "shape_hit_fraction_incoherent": round(shape_frac, 4),
        "roofline": roofline,
        "knobs": {k: ctx.config_get(k) for k in ("trace.block", "trace.stack", "trace.persistent", "trace.refill",
                                                 "trace.walk")},
        "tuning": tuning,
    }
    if not args.no_cpu and not args.only:
        res["parity"], res["cpu_baseline"] = isect_parity_and_cpu(flat, batches, dbat, args.cpu_seconds)
    gs.close()
    return res


def isect_parity_and_cpu(flat, batches, dbat, seconds):
    """The oracle (C++ restatement, OpenMP over rays) on all four full
    batches, timed on this host's cores, and its answers compared with the
    GPU's ray for ray: primitive ids, the bits of t, b1 and b2 of every hit,
    and every any-hit answer."""
    from oracle.pyoracle import OracleScene
    threads = host_threads()
    o = OracleScene(flat)
    parity = {}
    total_q, total_t = 0, 0.0
    for name, rays in batches.items():
        cols, outs, occ = dbat[name]
        t0 = time.perf_counter()
        ot, ob1, ob2, oprim, ties = o.intersect(rays, threads=threads)
        oocc = o.intersects(rays, threads=threads)
        total_t += time.perf_counter() - t0
        total_q += 2 * len(rays)
        gt, gb1, gb2, gprim = (x.cpu().numpy() for x in outs)
        gocc = occ.cpu().numpy().astype(bool)
        hit = (gprim == oprim) & (oprim >= 0)
        parity[name] = {
            "rays": int(len(rays)), "hits": int((oprim >= 0).sum()),
            "prim_mismatch": int((gprim != oprim).sum()),
            "t_bits_mismatch": int((gt[hit].view(np.int32) != ot[hit].view(np.int32)).sum()),
            "b1_bits_mismatch": int((gb1[hit].view(np.int32) != ob1[hit].view(np.int32)).sum()),
            "b2_bits_mismatch": int((gb2[hit].view(np.int32) != ob2[hit].view(np.int32)).sum()),
            "any_mismatch": int((gocc != oocc).sum()),
            "exact_ties": int(ties),
        }
    log("intersect: parity vs oracle:", json.dumps(parity))
    # the same restatement on one core, over a quarter of the time
    rng = np.random.default_rng(0)
    samples = {k: v[rng.permutation(len(v))[:4096]] for k, v in batches.items()}
    q1, t1 = 0, 0.0
    while t1 < seconds / 4:
        t0 = time.perf_counter()
        for v in samples.values():
            o.intersect(v, threads=1)
            o.intersects(v, threads=1)
        t1 += time.perf_counter() - t0
        q1 += 4 * 4096
    o.close()
    log("intersect: cpu %d queries in %.2fs on %d threads; %d in %.2fs on 1" % (total_q, total_t, threads, q1, t1))
    cpu = {"value": round(total_q / total_t / 1e6, 4), "unit": "Mrays/s", "cores": threads, "kind": "port",
           "sample": "the four full batches (closest+any on 4,194,304 coherent and 4,194,304 incoherent rays), "
                     "%.1f s; oracle C++ restatement with its own median BVH, %d OpenMP threads" % (total_t, threads),
           "value_1core": round(q1 / t1 / 1e6, 4),
           "sample_1core": "4096 rays from each batch, %d queries in %.1f s, one thread" % (q1, t1)}
    return parity, cpu


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


def lat_stats(lat):
    a = np.array(lat) * 1e3
    return {"p50_ms": round(float(np.percentile(a, 50)), 2), "p99_ms": round(float(np.percentile(a, 99)), 2),
            "mean_ms": round(float(a.mean()), 2)}


def run_plain(ctx, gss, flats, jobs):
    """The plain path: bgpu_render one job at a time, in arrival order."""
    from bling_amd.runtime import GpuRenderer
    from bling_amd.workloads import spec_for
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
    """The served path: each job submitted when it arrives (bgpu_job_submit),
    continuous batching inside the library."""
    from bling_amd.runtime import Job
    from bling_amd.workloads import spec_for
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
    for h in handles:
        h.wait(timeout=600)
    t_end = time.perf_counter() - t_start
    for h, lg in zip(handles, lag):
        lat.append(lg + h.latency_ms() * 1e-3)
        imgs.append(h.poll()[0])
        h.release()
    st = ctx.stats()
    makespan = max(j["arrival"] + l for j, l in zip(jobs, lat))
    return lat, imgs, min(t_end, makespan), st


def path_row(st, span, lat):
    return dict(mrays_s=round(rays_of(st) / span / 1e6, 2), msamples_s=round(st["samples"] / span / 1e6, 2),
                makespan_s=round(span, 3), gpu_busy=round(st["gpu_ms"] * 1e-3 / span, 3), **lat_stats(lat))


def compare_paths(ctx, gss, flats, jobs):
    """Both paths over the same jobs: per-path figures and the largest
    relative difference between a job's two images."""
    p_lat, p_img, p_span, p_st = run_plain(ctx, gss, flats, jobs)
    s_lat, s_img, s_span, s_st = run_served(ctx, gss, flats, jobs)
    worst, weights_equal = 0.0, True
    for a, b in zip(s_img, p_img):
        weights_equal = weights_equal and bool(np.array_equal(a[..., 0], b[..., 0]))
        d = np.abs(a - b) / np.maximum(np.abs(b), 1e-3)
        worst = max(worst, float(d.max()))
    row = {
        "jobs": len(jobs),
        "samples": int(s_st["samples"]),
        "served": path_row(s_st, s_span, s_lat),
        "plain": path_row(p_st, p_span, p_lat),
        "rays_served": int(rays_of(s_st)),
        "rays_plain": int(rays_of(p_st)),
        "image_max_rel_diff": worst,
        "image_weights_equal": weights_equal,
    }
    row["speedup_mrays"] = round(row["served"]["mrays_s"] / max(1e-9, row["plain"]["mrays_s"]), 3)
    return row


def create_scenes(ctx, named, admission=None):
    """Device scenes of `named` [(name, flat)]; with `admission` a list, also
    times creation with the host BVH build and records both."""
    gss = []
    for (name, f) in named:
        host_s = None
        if admission is not None:
            ctx.config_set("bvh.device", 0)
            t1 = time.perf_counter()
            g = ctx.scene(f)
            host_s = time.perf_counter() - t1
            g.close()
            ctx.config_set("bvh.device", 1)
        t1 = time.perf_counter()
        g = ctx.scene(f)
        dev_s = time.perf_counter() - t1
        gss.append(g)
        if admission is not None:
            info = g.bvh_info()
            admission.append({"scene": name, "tris": f.num_triangles, "host_build_s": round(host_s, 3),
                              "device_build_s": round(dev_s, 3), "device_tree": info["device"],
                              "device_bvh2_ms": round(info["build_ms"], 1)})
    return gss


def profile_pass(ctx, gss, flats, jobs):
    """Per-phase device time of the bounce loop over `jobs` on the plain
    path, the iterations launched one by one with events between the
    phases (pt.profile 1); outside the timed runs."""
    from bling_amd.runtime import GpuRenderer
    from bling_amd.workloads import spec_for
    ctx.config_set("pt.profile", 1)
    ctx.reset_stats()
    for job in jobs:
        GpuRenderer(spec_for(job, flats[job["scene"]])).render_plain(gss[job["scene"]])
    phases, iters = ctx.profile()
    ctx.config_set("pt.profile", 0)
    total = sum(ms for ms, _ in phases.values())
    return {ph: {"ms": round(ms, 3), "share": round(ms / total, 4), "items": int(n),
                 "us_per_iteration": round(1e3 * ms / max(1, iters), 2)} for ph, (ms, n) in phases.items()}, iters


def serve_cpu(flats, jobs, seconds):
    """The oracle integrator on the stream's jobs in arrival order, skipping
    those over 4M samples, until `seconds` of work."""
    from oracle import pyoracle
    from bling_amd.workloads import spec_for
    threads = host_threads()
    osc = {}
    dt, rays, samples, done = 0.0, 0, 0, 0
    for job in jobs:
        n = (job["res"] + 1) ** 2 * job["spp"]
        if n > 4_000_000:
            continue
        k = job["scene"]
        if k not in osc:
            osc[k] = pyoracle.OracleScene(flats[k])  # its build is not timed
        t0 = time.perf_counter()
        _, _, r = pyoracle.render(flats[k], spec_for(job, flats[k]), threads=threads, scene=osc[k])
        dt += time.perf_counter() - t0
        rays += r
        samples += n
        done += 1
        if dt >= seconds:
            break
    for o in osc.values():
        o.close()
    return {"value": round(rays / dt / 1e6, 4), "unit": "Mrays/s", "cores": threads, "kind": "port",
            "msamples_s": round(samples / dt / 1e6, 4),
            "sample": "the first %d jobs of the saturating-rate stream with at most 4M camera samples each "
                      "(%d samples, %.1f s), oracle C++ restatement of the path integrator, %d OpenMP threads"
                      % (done, samples, dt, threads)}


def serve_leg(ctx, args, out):
    """The headline: the serving stream at the saturating rate, served vs plain."""
    from bling_amd.workloads import job_stream, serve_scenes
    t0 = time.time()
    named = serve_scenes(args.serve_max_tris)
    flats = [f for _, f in named]
    log("serve: %d scenes generated in %.1fs" % (len(flats), time.time() - t0))
    t0 = time.time()
    admission = []
    gss = create_scenes(ctx, named, admission)
    log("serve: uploaded + BVHs in %.1fs, %.1f MB" % (time.time() - t0, sum(g.device_bytes() for g in gss) / 1e6))
    res_set = [int(v) for v in args.serve_res.split(",")]
    spp_set = [int(v) for v in args.serve_spp.split(",")]
    # warmup: engine creation and graph capture on both paths
    warm = job_stream(max(1, args.warmup), 1000.0, len(flats), [128], [4], seed=999)
    run_plain(ctx, gss, flats, warm)
    run_served(ctx, gss, flats, warm)
    rows, streams = [], []
    for ri, rate in enumerate(float(r) for r in args.serve_rates.split(",")):
        n = STEP_JOBS * args.steps if rate >= 20 else args.serve_slow_jobs
        jobs = job_stream(n, rate, len(flats), res_set, spp_set, seed=ri + 1)
        row = {"rate_jobs_s": rate, **compare_paths(ctx, gss, flats, jobs)}
        log("serve:", json.dumps(row))
        rows.append(row)
        streams.append(jobs)
    head, jobs = rows[0], streams[0]
    phases, iters = profile_pass(ctx, gss, flats, jobs[:args.profile_jobs])
    log("serve: profile:", json.dumps(phases))
    dom = max(phases, key=lambda k: phases[k]["ms"])
    sh = phases["shade"]
    achieved = sh["items"] * B_SHADE_PATH / (sh["ms"] * 1e-3) / 1e9
    traffic = None
    pmc_path = os.path.join(ROOT, "profiles", "shade_serve_pmc.json")
    if os.path.exists(pmc_path):
        with open(pmc_path) as f:
            traffic = json.load(f).get("hbm_bytes_per_path")
    out["value"] = head["served"]["mrays_s"]
    out["ms_per_step"] = round(head["served"]["makespan_s"] * 1e3 / args.steps, 3)
    out["roofline"] = {
        "bound": "hbm", "kernel": "k_shade (path shading)", "achieved": round(achieved, 1), "peak": PEAK_HBM_GBS,
        "unit": "GB/s", "frac": round(achieved / PEAK_HBM_GBS, 4),
        "traffic": None if traffic is None else int(traffic * sh["items"] / max(1, iters)),
        "share_of_loop": sh["share"], "dominant_phase": dom,
        "note": "achieved = %d path-pool bytes per shaded path x paths shaded / shade time, per-phase HIP "
                "events over the first %d jobs (%d iterations); traffic = measured DRAM bytes per shading "
                "launch (FETCH_SIZE + WRITE_SIZE per path shaded x mean paths per launch)"
                % (B_SHADE_PATH, len(jobs[:args.profile_jobs]), iters),
    }
    out["cpu_baseline"] = None if args.no_cpu else serve_cpu(flats, jobs, args.cpu_seconds)
    for g in gss:
        g.close()
    out["config"].update({
        "workload": "serving stream at %g jobs/s: %d Poisson jobs over 8 resident scenes (%s), resolution "
                    "uniform in %s, spp uniform in %s, path integrator, max depth 5, RR after depth 3; "
                    "value = Mrays/s of the served path (continuous batching through bgpu_job_submit); the "
                    "plain path (bgpu_render, one job at a time in arrival order) renders the same jobs with "
                    "the same seeds (equal RMSE)" % (head["rate_jobs_s"], head["jobs"],
                                                      ", ".join("%s %d tris" % (k, f.num_triangles) for k, f in named),
                                                      res_set, spp_set),
        "step": "%d consecutive jobs of the stream" % STEP_JOBS,
        "served_vs_plain": {"mrays_s": [head["served"]["mrays_s"], head["plain"]["mrays_s"]],
                            "p50_ms": [head["served"]["p50_ms"], head["plain"]["p50_ms"]],
                            "p99_ms": [head["served"]["p99_ms"], head["plain"]["p99_ms"]],
                            "gpu_busy": [head["served"]["gpu_busy"], head["plain"]["gpu_busy"]],
                            "speedup_mrays": head["speedup_mrays"]},
        "serve": {"rates": rows, "scene_admission": admission, "profile": phases, "profile_iterations": iters},
    })


def phases_leg(ctx, args):
    """The headline workload's per-phase profiling pass alone, for counter
    runs: the serving scenes, then the first --profile-jobs jobs of the
    saturating-rate stream on the plain path with pt.profile 1 (every bounce
    kernel a plain launch).  Reports the items of each phase, so that a
    counter total over a phase's dispatches divides into bytes per item."""
    from bling_amd.workloads import job_stream, serve_scenes
    named = serve_scenes(args.serve_max_tris)
    flats = [f for _, f in named]
    gss = create_scenes(ctx, named)
    res_set = [int(v) for v in args.serve_res.split(",")]
    spp_set = [int(v) for v in args.serve_spp.split(",")]
    rate = float(args.serve_rates.split(",")[0])
    jobs = job_stream(STEP_JOBS * args.steps, rate, len(flats), res_set, spp_set, seed=1)[:args.profile_jobs]
    phases, iters = profile_pass(ctx, gss, flats, jobs)
    log("phases:", json.dumps(phases))
    for g in gss:
        g.close()
    return {"workload": "per-phase profiling pass: the first %d jobs of the %g jobs/s serving stream, plain path, "
                        "pt.profile 1" % (len(jobs), rate), "iterations": iters, "phases": phases}


def batch_leg(ctx, args):
    """64 jobs submitted at once, each on its own scene (log-uniform 1k-2M
    triangles, 128^2-1024^2, 1-32 spp): served (one batch, continuous
    batching) against plain (one job at a time)."""
    from bling_amd.workloads import batch_workload
    t0 = time.time()
    flats, jobs = batch_workload(n_jobs=args.batch_jobs)
    log("batch: %d scenes (%d tris) generated in %.1fs" % (len(flats), sum(f.num_triangles for f in flats),
                                                            time.time() - t0))
    t0 = time.time()
    gss = create_scenes(ctx, [("s%d" % k, f) for k, f in enumerate(flats)])
    create_s = time.time() - t0
    warm = [dict(j, res=64, spp=1) for j in jobs[:2]]
    run_plain(ctx, gss, flats, warm)
    run_served(ctx, gss, flats, warm)
    row = compare_paths(ctx, gss, flats, jobs)
    for g in gss:
        g.close()
    res = {"workload": "%d jobs at once, each on its own scene: box stacks / displaced grids of %d..%d "
                       "triangles (log-uniform), resolution %d..%d (log-uniform), spp 1..32; path integrator, "
                       "max depth 5, RR after depth 3" % (len(jobs), min(f.num_triangles for f in flats),
                                                          max(f.num_triangles for f in flats),
                                                          min(j["res"] for j in jobs), max(j["res"] for j in jobs)),
           "scenes_created_s": round(create_s, 2), **row}
    log("batch:", json.dumps(res))
    return res


if __name__ == "__main__":
    main()
