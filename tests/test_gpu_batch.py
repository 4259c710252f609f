"""Batches of render jobs (vlrContextRenderBatch) against single renders.

Each job's pixels are seeded like a single render and every path draws and adds in the same
order whatever the schedule, so each job of a batch must give exactly the image of rendering
it alone: scenes, image sizes, sample counts and path lengths differ between the jobs, some
jobs share a scene, and the batch's first steps try different wavefront depths (autotuning).
"""
import numpy as np
import pytest

from vlr_amd import api, scenes

pytestmark = pytest.mark.gpu


def single(desc, W, H, spp, depth, spectral):
    with api.Context(spectral=spectral, max_path_length=depth) as ctx:
        bs = api.build_scene(ctx, desc)
        ctx.bind_output(W, H)
        for k in range(spp):
            ctx.render(bs.camera, k == 0, spp)
        return ctx.read_output()


def run_batch(descs, specs, spectral, graphs=True):
    with api.Context(spectral=spectral) as ctx:
        ctx.set_batch_graphs(graphs)
        built = [api.build_scene(ctx, d) for d in descs]
        jobs = [(built[s].scene, built[s].camera, W, H, spp, depth) for (s, W, H, spp, depth) in specs]
        ms = ctx.render_batch(jobs)
        outs = [ctx.read_batch_output(k) for k in range(len(jobs))]
        st = ctx.stats()
        sched = ctx.batch_schedule()
    return outs, ms, st, sched


@pytest.mark.parametrize("spectral", [False, True])
def test_batch_jobs_equal_single_renders(torch_gpu, spectral):
    descs = [scenes.cornell_box(), scenes.cornell_box(seed=7), scenes.icosphere_grid(grid=3, env_size=(64, 32)),
             scenes.cornell_many_lights()]
    # (scene, width, height, samples, max path length)
    specs = [(0, 64, 48, 5, 8), (1, 33, 17, 3, 4), (2, 80, 45, 9, 12), (3, 40, 40, 1, 16), (0, 20, 30, 8, 6)]
    outs, ms, st, sched = run_batch(descs, specs, spectral)
    print("completion ms", ms, "schedule", sched, "rays", st["closestRays"], st["shadowRays"])
    for k, (s, W, H, spp, depth) in enumerate(specs):
        ref = single(descs[s], W, H, spp, depth, spectral)
        assert np.array_equal(outs[k].view(np.uint32), ref.view(np.uint32)), k
    # a job finishes no later than any job with more samples
    order = np.argsort([spp for (_, _, _, spp, _) in specs], kind="stable")
    assert np.all(np.diff(ms[order]) >= -1e-3)


def test_batch_without_graphs_is_identical(torch_gpu):
    descs = [scenes.cornell_box(), scenes.icosphere_grid(grid=3, env_size=(64, 32))]
    specs = [(0, 48, 32, 10, 8), (1, 40, 24, 6, 10)]
    a, _, _, _ = run_batch(descs, specs, False, graphs=True)
    b, _, _, _ = run_batch(descs, specs, False, graphs=False)
    for x, y in zip(a, b):
        assert np.array_equal(x.view(np.uint32), y.view(np.uint32))


def single_with_rng(desc, W, H, spp, depth, spectral):
    with api.Context(spectral=spectral, max_path_length=depth) as ctx:
        bs = api.build_scene(ctx, desc)
        ctx.bind_output(W, H)
        for k in range(spp):
            ctx.render(bs.camera, k == 0, spp)
        return ctx.read_output(), ctx.read_rng(), ctx.stats()


@pytest.mark.parametrize("spectral", [False, True])
def test_c4_shaped_batch_equals_single_renders(torch_gpu, spectral):
    """Jobs shaped like the 64-job batch workload (tessellated Cornell boxes of unique triangles,
    square images, uneven samples and path lengths) in one batch: a ~1M-triangle box at 1024^2
    with path length 16, a 128^2 job with path length 4 and jobs in between; every job's image
    equals its single render bit for bit, and the batch traces as many rays as the single
    renders together."""
    jobs = ([(1_000_000, 1024, 3, 16), (10_000, 128, 5, 4), (100_000, 512, 4, 10), (50_000, 256, 2, 12)]
            if not spectral else [(200_000, 256, 4, 8), (20_000, 128, 3, 6), (1_000_000, 512, 2, 16)])
    descs = [scenes.tessellated_cornell(7000 + k, tris) for k, (tris, _, _, _) in enumerate(jobs)]
    specs = [(k, res, res, spp, depth) for k, (_, res, spp, depth) in enumerate(jobs)]
    outs, ms, st, sched = run_batch(descs, specs, spectral)
    print("batch rays", st["closestRays"], st["shadowRays"], "schedule", sched, flush=True)
    rays = 0
    for k, (tris, res, spp, depth) in enumerate(jobs):
        print(f"job {k}: {descs[k].num_triangles} triangles, {res}^2, {spp} spp, depth {depth}", flush=True)
        ref, _, sst = single_with_rng(descs[k], res, res, spp, depth, spectral)
        rays += sst["closestRays"] + sst["shadowRays"]
        assert np.array_equal(outs[k].view(np.uint32), ref.view(np.uint32)), k
    assert st["closestRays"] + st["shadowRays"] == rays


def test_batched_job_matches_the_oracle(torch_gpu):
    """A batched job (a tessellated Cornell box) against the oracle path tracer: the job equals
    its single render, whose per-pixel generators agree with the oracle's, and where they agree
    the batched pixels equal the oracle's to float rounding."""
    from oracle import binding as ob
    desc = scenes.tessellated_cornell(7100, 10_000)
    other = scenes.cornell_many_lights()
    W = H = 48
    spp, depth = 6, 6
    outs, _, _, _ = run_batch([desc, other], [(0, W, H, spp, depth), (1, 32, 24, 3, 8)], False)
    solo, rng, _ = single_with_rng(desc, W, H, spp, depth, False)
    assert np.array_equal(outs[0].view(np.uint32), solo.view(np.uint32))
    ref, _, orng, _ = ob.Scene(desc).render(W, H, spp, max_path_length=depth, rng_state=ob.seed_states(W, H))
    eq = (rng == orng).reshape(H, W)
    print("pixels with identical generator states:", eq.mean())
    assert eq.mean() > 0.97
    a, b = outs[0][eq][:, :3].astype(np.float64), ref[eq][:, :3].astype(np.float64)
    rel = np.abs(a - b) / np.maximum(np.abs(b), 1e-3 * max(1e-12, np.abs(b).mean()))
    assert np.quantile(rel, 0.999) < 1e-3
