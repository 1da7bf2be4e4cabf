"""The render-job workloads the benchmark and the tests share.

- The serving stream: Poisson arrivals of render jobs over 8 resident scenes
  (the Cornell box and 7 procedural scenes of 0.1M-8M triangles), each job
  with its own resolution and samples per pixel.
- The batched workload: 64 jobs, each on its own procedural scene of
  log-uniform 1k-2M triangles (box stacks and displaced grids), resolution
  log-uniform in 128^2..1024^2 and 1..32 samples per pixel, submitted at once.

A job renders the scene's camera re-made for its own image size (a
RenderJob's camera follows its imageSize, IO/CameraParser.hs:30-38) with the
path integrator, max depth 5 and Russian roulette after depth 3.
"""
import numpy as np

from . import scenes
from .runtime import JobSpec


def serve_scenes(max_tris=8e6):
    """The Cornell box plus 7 procedural scenes of 0.1M..max_tris triangles
    (log-spaced), alternating box stacks, displaced grids and an icosphere
    (under a uniform sky), all Lambertian."""
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
            out.append(("icosphere", scenes.icosphere_scene(levels=lv, num_spheres=1000, seed=k, res=256, lit=True)))
    return out


def job_stream(n, rate, nscenes, res_set=(128, 256, 512, 1024), spp_set=(4, 16, 64), seed=1):
    """n jobs with exponential inter-arrival gaps of mean 1/rate seconds
    (the first at t = 0), each on a uniformly drawn scene, resolution and spp."""
    rng = np.random.default_rng(seed)
    gaps = rng.exponential(1.0 / rate, n)
    arrivals = np.cumsum(gaps) - gaps[0]
    jobs = []
    for k in range(n):
        jobs.append({"arrival": float(arrivals[k]), "scene": int(rng.integers(nscenes)),
                     "res": int(rng.choice(list(res_set))), "spp": int(rng.choice(list(spp_set))),
                     "seed": int(seed * 100003 + k)})
    return jobs


def spec_for(job, flat):
    """The job's JobSpec: spp as an nu x nv stratified grid (square when it
    can be), path integrator, max depth 5, RR after depth 3."""
    spp = job["spp"]
    s = int(np.floor(np.sqrt(spp)))
    while spp % s:
        s -= 1
    return JobSpec(spp_nu=spp // s, spp_nv=s, max_depth=5, sample_depth=3, rr_start_depth=3, passes=1,
                   seed=job["seed"], view=flat.view(job["res"], job["res"]))


def batch_workload(n_jobs=64, min_tris=1e3, max_tris=2e6, min_res=128, max_res=1024, max_spp=32):
    """The batched workload: job k (seed k) on its own scene of log-uniform
    [min_tris, max_tris] triangles (a box stack for even k, a displaced grid
    for odd k), log-uniform resolution in [min_res, max_res] (a multiple of
    8) and uniform spp in 1..max_spp; every job arrives at t = 0.
    Returns (flats, jobs)."""
    flats, jobs = [], []
    for k in range(n_jobs):
        rng = np.random.default_rng(k)
        tris = int(np.exp(rng.uniform(np.log(min_tris), np.log(max_tris))))
        res = max(8, int(np.exp(rng.uniform(np.log(min_res), np.log(max_res)))) // 8 * 8)
        spp = int(rng.integers(1, max_spp + 1))
        make = scenes.stack_scene if k % 2 == 0 else scenes.grid_scene
        flats.append(make(tris, res=256, seed=k))
        jobs.append({"arrival": 0.0, "scene": k, "res": res, "spp": spp, "seed": 7000 + k})
    return flats, jobs
