"""Serving on the GPU: HBM-budget admission of served jobs.

The reference renders one job at a time (prender, Rendering.hs:111-140) and
has no notion of a device memory budget; these tests pin the served path's
own contract: a job whose film alone exceeds the context's budget is refused
at submission, and jobs that do not fit beside the ones running wait in
arrival order until a finished job releases its film.
"""
import time

import numpy as np
import pytest

from bling_amd import abi, scenes
from bling_amd.runtime import Context, GpuRenderer, Job, JobSpec

pytestmark = pytest.mark.gpu

FILM_BYTES_PER_PIXEL = 16  # (w, X, Y, Z) float32


def _warm(ctx, flat):
    """Creates the scene and the served engine (one tiny job)."""
    ctx.config_set("pt.pool", 8192)
    gs = ctx.scene(flat)
    j = Job(gs, JobSpec(spp_nu=1, spp_nv=1, max_depth=2, sample_depth=2, seed=1))
    j.wait(timeout=120)
    j.release()
    return gs


def test_hbm_budget_admission(gpu_ctx):
    flat = scenes.cornell(32)
    res = 256
    film = res * res * FILM_BYTES_PER_PIXEL
    specs = [JobSpec(spp_nu=2, spp_nv=2, max_depth=4, sample_depth=3, rr_start_depth=3, passes=2,
                     seed=20 + k, view=flat.view(res, res)) for k in range(3)]
    # device bytes of the scene and the served engine, measured on a context
    # without a budget; a second context gets room for them plus one film
    with Context(0) as a:
        gs = _warm(a, flat)
        used0, _ = a.mem_info()
        gs.close()
    budget = used0 + film + film // 2  # room for one film, not two
    with Context(0, hbm_budget=budget) as b:
        gs = _warm(b, flat)
        try:
            assert b.mem_info() == (used0, budget)
            # a film larger than the whole budget can never be admitted
            big = JobSpec(spp_nu=1, spp_nv=1, max_depth=2, sample_depth=2, view=flat.view(1024, 1024))
            with pytest.raises(abi.BgpuError) as e:
                Job(gs, big)
            assert e.value.code == abi.E_BUDGET
            jobs = [Job(gs, sp) for sp in specs]
            waited = False
            served, lat = {}, {}
            deadline = time.time() + 120
            while len(served) < len(jobs):
                # later jobs first: a later job seen running implies that
                # every earlier one had finished before that observation
                st = [None] * len(jobs)
                for k in reversed(range(len(jobs))):
                    st[k] = abi.JOB_DONE if k in served else jobs[k].poll(want_image=False)[3]
                for k, s in enumerate(st):
                    assert s != abi.JOB_FAILED
                    if s == abi.JOB_RUNNING:
                        assert all(st[m] == abi.JOB_DONE for m in range(k)), st
                        assert all(st[m] == abi.JOB_QUEUED for m in range(k + 1, len(jobs))), st
                        waited = waited or k + 1 < len(jobs)
                used, _ = b.mem_info()
                assert used <= budget
                # a finished job keeps its film until released: take the
                # image and release it, which lets the next job in
                for k, s in enumerate(st):
                    if s == abi.JOB_DONE and k not in served:
                        served[k] = jobs[k].poll()[0]
                        lat[k] = jobs[k].latency_ms()
                        jobs[k].release()
                assert time.time() < deadline, st
                time.sleep(0.0005)
            assert waited, "no job was seen waiting for the budget"
            assert lat[0] < lat[1] < lat[2], lat
            served = [served[k] for k in range(len(jobs))]
            # a job whose film cannot fit even with every other film released fails
            lone = JobSpec(spp_nu=1, spp_nv=1, max_depth=2, sample_depth=2, view=flat.view(512, 512))
            j = Job(gs, lone)
            with pytest.raises(abi.BgpuError) as e:
                j.wait(timeout=60)
            assert e.value.code != abi.E_TIMEOUT
            assert j.poll(want_image=False)[3] == abi.JOB_FAILED
            j.release()
        finally:
            gs.close()
    # each result equals the job rendered alone
    gp = gpu_ctx.scene(flat)
    try:
        for sp, img in zip(specs, served):
            p_img, _ = GpuRenderer(sp).render_plain(gp)
            assert np.array_equal(img[..., 0], p_img[..., 0])
            np.testing.assert_allclose(img, p_img, rtol=1e-4, atol=1e-5)
    finally:
        gp.close()
