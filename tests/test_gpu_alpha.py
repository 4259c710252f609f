"""Alpha nodes on the GPU: the stochastic alpha test of closest-hit rays (path_tracing.cu:14-25,
light_tracing.cu:19-30) and the fractional visibility of shadow and connection rays
(light_transport_common.h:205-218).  Candidate hits are met in the device BVH's order, which no
other traversal reproduces, so these are statistical checks:

  furnace         a flat albedo-1 sphere with alpha 0.5 in a unit environment vanishes: every
                  pixel equals the camera weight (closed form), for path tracing, LVC-BPT and
                  light tracing
  oracle          the Cornell box with a half-transparent tall box: quadrant means of the GPU
                  render within Monte-Carlo noise of the oracle's
"""
import math

import numpy as np
import pytest

from oracle import binding as ob
from tests.test_oracle_alpha import batch_quadrants
from tests.test_oracle_lights_render import camera_weight
from vlr_amd import api, scenes
from vlr_amd.scenes import NodeDesc

pytestmark = pytest.mark.gpu

PLUG_FLOAT1 = 0


def alpha_node(value):
    return (NodeDesc("Float2", floats={"0": [value], "1": [0.0]}), PLUG_FLOAT1, 0)


def with_alpha(desc, mesh, value):
    desc.meshes[mesh].group_nodes = {0: (None, None, alpha_node(value))}
    desc.meshes[mesh].group_alpha = {0: value}
    return desc


def render(desc, W, H, frames, max_len, renderer=api.RENDERER_PATH_TRACING, light_paths=None, seed_states=None):
    with api.Context(spectral=desc.spectral, max_path_length=max_len) as ctx:
        ctx.set_renderer(renderer)
        if light_paths:
            ctx.set_num_light_paths(light_paths)
        bs = api.build_scene(ctx, desc)
        ctx.bind_output(W, H)
        if seed_states is not None:
            ctx.write_rng(seed_states)
        for f in range(frames):
            ctx.render(bs.camera, f == 0, frames)
        return ctx.read_output()[..., :3], ctx.stats()


@pytest.mark.parametrize("renderer", [api.RENDERER_PATH_TRACING, api.RENDERER_BPT, api.RENDERER_LIGHT_TRACING])
def test_alpha_furnace(torch_gpu, renderer):
    desc = with_alpha(scenes.white_furnace(spectral=False, flat=True), 0, 0.5)
    W = H = 32
    img, st = render(desc, W, H, 64, 8, renderer, light_paths=16 * W * H)
    assert st["droppedNonFinite"] == 0
    r = img / camera_weight(desc.camera, W, H)[..., None]
    sigma = r.std() / math.sqrt(W * H)
    print("furnace mean", r.mean(), "sigma", sigma)
    assert abs(r.mean() - 1.0) < 4 * sigma + 2e-3, (r.mean(), sigma)


def test_alpha_cornell_matches_oracle(torch_gpu):
    W = H = 24
    spp, batches = 16, 6
    desc = with_alpha(scenes.cornell_box(), 3, 0.5)
    ref, sref = batch_quadrants(desc, W, H, spp, batches, 700)
    out = []
    for k in range(batches):
        img, _ = render(desc, W, H, spp, 8, seed_states=ob.mt19937_64(900 + k, W * H))
        q = []
        for ys in (slice(0, H // 2), slice(H // 2, H)):
            for xs in (slice(0, W // 2), slice(W // 2, W)):
                q.append(img[ys, xs].astype(np.float64).mean(axis=(0, 1)))
        out.append(np.concatenate(q))
    out = np.array(out)
    gpu, sgpu = out.mean(axis=0), out.std(axis=0, ddof=1) / math.sqrt(batches)
    sig = np.sqrt(sref ** 2 + sgpu ** 2)
    print("gpu", gpu, "oracle", ref, "sigma", sig)
    assert np.all(np.abs(gpu - ref) < 4 * sig + 1e-5)


def test_batches_refuse_alpha_scenes(torch_gpu):
    desc = with_alpha(scenes.cornell_box(), 3, 0.5)
    with api.Context(spectral=False, max_path_length=4) as ctx:
        bs = api.build_scene(ctx, desc)
        ctx.bind_output(16, 16)
        ctx.render(bs.camera, True, 1)   # the single-job renderer takes it
        with pytest.raises(api.VLRError):
            ctx.render_batch([(bs.scene, bs.camera, 16, 16, 1, 4)])
