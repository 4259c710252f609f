"""Light tracing on the GPU (VLRRenderer_LightTracing) against the oracle's restatement of
light_tracing.cu:49-263.

Both sides seed the light-path PCG32 states from mt19937_64(459182033132123413) and draw each
path's wavelengths, light sample, lens samples and BSDF samples in the reference's order.

  RNG states      light-path states after the frames: the same draws through the same branches
  per pixel       when every light path matched, the splatted images match (float atomics in
                  either order)
  image mean      within Monte-Carlo noise of the oracle's
  white furnace   every pixel equals the camera weight (closed form, independent of the oracle)
"""
import math

import numpy as np
import pytest

from oracle import binding as ob
from tests.test_oracle_lights_render import camera_weight
from vlr_amd import api, scenes

pytestmark = pytest.mark.gpu

LIGHT_SEED = 459182033132123413


def render_lt_gpu(desc, W, H, frames, max_len, spectral, num_light_paths):
    with api.Context(spectral=spectral, max_path_length=max_len) as ctx:
        ctx.set_renderer(api.RENDERER_LIGHT_TRACING)
        ctx.set_num_light_paths(num_light_paths)
        bs = api.build_scene(ctx, desc)
        ctx.bind_output(W, H)
        imgs = []
        for f in range(frames):
            ctx.render(bs.camera, f == 0, frames)
            imgs.append(ctx.read_output())
        return imgs, ctx.stats(), ctx.read_light_rng()


def compare_with_oracle(desc, W, H, frames, max_len, spectral, L):
    imgs, st, lrng = render_lt_gpu(desc, W, H, frames, max_len, spectral, L)
    assert st["numFrames"] == frames and st["closestRays"] > 0 and st["shadowRays"] > 0
    ref, ost, lrng_ref, _ = ob.Scene(desc, spectral=spectral).render_lt(
        W, H, frames, max_len, num_light_paths=L, light_rng_state=ob.mt19937_64(LIGHT_SEED, L))
    same_light = np.mean(lrng == lrng_ref)
    print("identical light RNG states:", same_light)
    print("rays gpu", st["closestRays"], st["shadowRays"], "oracle", ost["closest_rays"], ost["shadow_rays"])
    assert same_light > 0.97
    img = imgs[-1][..., :3].astype(np.float64)
    r = ref[..., :3].astype(np.float64)
    # image mean: per-frame splat images are independent samples
    frames_img = np.stack([imgs[0][..., :3]] + [imgs[k][..., :3] * (k + 1) - imgs[k - 1][..., :3] * k
                                                for k in range(1, frames)])
    sigma = frames_img.reshape(frames, -1, 3).mean(axis=1).std(axis=0, ddof=1) / math.sqrt(frames)
    diff = img.mean(axis=(0, 1)) - r.mean(axis=(0, 1))
    print("gpu mean", img.mean(axis=(0, 1)), "oracle mean", r.mean(axis=(0, 1)), "sigma", sigma)
    assert np.all(np.abs(diff) < 4 * math.sqrt(2) * sigma + 1e-6)
    if same_light == 1.0:
        # every light path matched, so every splat did: equal up to the order of float atomics
        assert st["closestRays"] == ost["closest_rays"] and st["shadowRays"] == ost["shadow_rays"]
        rel = np.abs(img - r) / np.maximum(np.abs(r), 1e-3 * max(1e-12, np.abs(r).mean()))
        q = np.quantile(rel, 0.999)
        print("per-pixel relative difference, 99.9th percentile:", q)
        assert q < 1e-3


def test_lt_cornell_rgb(torch_gpu):
    compare_with_oracle(scenes.cornell_box(), 32, 32, 4, 8, False, 8192)


def test_lt_many_lights(torch_gpu):
    # 64 point lights (delta positions) and 8 quad lights
    compare_with_oracle(scenes.cornell_many_lights(), 32, 32, 4, 8, False, 8192)


def test_lt_icosphere_grid_spectral(torch_gpu):
    # environment light: light paths start on the scene's bounding disc; 16-strata splats;
    # dielectric spheres exercise the dispersive single-wavelength branch
    compare_with_oracle(scenes.icosphere_grid(grid=4, env_size=(64, 32)), 48, 27, 3, 8, True, 8192)


def test_lt_white_furnace(torch_gpu):
    desc = scenes.white_furnace(spectral=False, flat=True)
    W = H = 32
    imgs, st, _ = render_lt_gpu(desc, W, H, 64, 8, False, 16 * W * H)
    r = imgs[-1][..., :3] / camera_weight(desc.camera, W, H)[..., None]
    sigma = r.std() / math.sqrt(W * H)
    assert abs(r.mean() - 1.0) < 4 * sigma + 1e-3, (r.mean(), sigma)
    assert st["droppedNonFinite"] == 0
