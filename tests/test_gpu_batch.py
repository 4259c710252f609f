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
