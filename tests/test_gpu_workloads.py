"""The benchmark's job workloads at their real sizes, on the GPU.

- The serving mix: the 8 resident scenes of 0.1M-8M triangles the serving
  stream uses, jobs at 128^2-1024^2 with 4, 16 and 64 spp, submitted
  together: every job's served image equals its plain image.
- The batched workload: 64 jobs at once, each on its own scene of
  log-uniform 1k-2M triangles at 128^2-1024^2 with 1-32 spp: served equals
  plain for every job, and two of its jobs track the oracle (the CPU
  restatement of Integrator/Path.hs:41-87) at the same seed on reduced views.

The reference renders one job at a time (prender, Rendering.hs:111-140); a
served image equal to the plain one is what makes the served path's
throughput comparable at equal RMSE.
"""
import numpy as np
import pytest

from bling_amd import abi
from bling_amd.runtime import GpuRenderer, Job
from bling_amd.workloads import batch_workload, job_stream, serve_scenes, spec_for
from oracle import pyoracle

pytestmark = pytest.mark.gpu


def _served_equals_plain(ctx, flats, jobs):
    gss = [ctx.scene(f) for f in flats]
    try:
        plain = [GpuRenderer(spec_for(j, flats[j["scene"]])).render_plain(gss[j["scene"]])[0] for j in jobs]
        handles = [Job(gss[j["scene"]], spec_for(j, flats[j["scene"]])) for j in jobs]
        for h, p, j in zip(handles, plain, jobs):
            h.wait(timeout=600)
            img, _, done, state = h.poll()
            assert state == abi.JOB_DONE and done == 1
            assert img.shape == (j["res"], j["res"], 4)
            # same samples on both paths; film sums differ only in atomic order
            assert np.array_equal(img[..., 0], p[..., 0]), j
            np.testing.assert_allclose(img, p, rtol=1e-4, atol=1e-5, err_msg=str(j))
            assert img[..., 2].sum() > 0
            h.release()
    finally:
        for g in gss:
            g.close()
    return plain


def test_serving_mix_at_full_size(gpu_ctx):
    named = serve_scenes(8e6)
    flats = [f for _, f in named]
    assert max(f.num_triangles for f in flats) > 7.5e6 and min(f.num_triangles for f in flats[1:]) >= 9e4
    # every scene once, at every resolution and spp, then four more drawn
    # from the stream's distribution; all arrive together
    res_spp = [(128, 4), (256, 16), (512, 64), (1024, 4), (128, 64), (256, 4), (512, 16), (1024, 16)]
    jobs = [{"arrival": 0.0, "scene": k, "res": r, "spp": s, "seed": 500 + k} for k, (r, s) in enumerate(res_spp)]
    jobs += job_stream(4, 1000.0, len(flats), seed=11)
    _served_equals_plain(gpu_ctx, flats, jobs)


def test_batched_64_jobs_at_full_size(gpu_ctx):
    flats, jobs = batch_workload(64)
    tris = [f.num_triangles for f in flats]
    assert len(jobs) == 64 and min(tris) < 2e3 and max(tris) > 1e6
    assert min(j["res"] for j in jobs) < 192 and max(j["res"] for j in jobs) > 768
    _served_equals_plain(gpu_ctx, flats, jobs)
    # jobs 0 (a 126,657-triangle box stack, 10 spp) and 1 (a 48,925-triangle
    # displaced grid, 2 spp) against the oracle, same seed, at 24 x 24
    for k in (0, 1):
        spec = spec_for(dict(jobs[k], res=24), flats[k])
        gs = gpu_ctx.scene(flats[k])
        try:
            g, _ = GpuRenderer(spec).render_plain(gs)
        finally:
            gs.close()
        o, _, _ = pyoracle.render(flats[k], spec)
        assert np.array_equal(g[..., 0], o[..., 0])
        rel = np.abs(g[..., 1:] - o[..., 1:]) / np.maximum(np.abs(o[..., 1:]), 1e-3)
        assert np.median(rel) < 1e-4, (k, np.median(rel))
