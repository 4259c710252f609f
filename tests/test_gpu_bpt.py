"""LVC-BPT on the GPU (VLRRenderer_BPT) against the oracle's restatement of lvc_bpt.cu:144-971.

Both sides draw each frame's common wavelengths from a default std::mt19937, seed the light-path
PCG32 states from mt19937_64(459182033132123413) and keep the light-vertex cache in light-path
order, so an eye path connects to the same cache vertex as in the oracle wherever every light
path took the same branches.  Per scene:

  RNG states      eye-path and light-path PCG32 states after the frames: the same draws
                  through the same branches
  per pixel       where the eye states agree and every light path agreed, the pixels match
                  (lens-connection splats are summed with float atomics in either order)
  image mean      within Monte-Carlo noise of the oracle's
  white furnace   every pixel equals the camera weight (closed form, independent of the oracle)
"""
import math

import numpy as np
import pytest

from oracle import binding as ob
from tests.test_gpu_render import assert_pixels_match_where_paths_match, per_frame_images
from tests.test_oracle_lights_render import camera_weight
from vlr_amd import api, scenes

pytestmark = pytest.mark.gpu

LIGHT_SEED = 459182033132123413


def render_bpt_gpu(desc, W, H, frames, max_len, spectral, num_light_paths):
    with api.Context(spectral=spectral, max_path_length=max_len) as ctx:
        ctx.set_renderer(api.RENDERER_BPT)
        ctx.set_num_light_paths(num_light_paths)
        bs = api.build_scene(ctx, desc)
        ctx.bind_output(W, H)
        imgs = []
        for f in range(frames):
            ctx.render(bs.camera, f == 0, frames)
            imgs.append(ctx.read_output())
        return imgs, ctx.stats(), ctx.read_rng(), ctx.read_light_rng()


def compare_with_oracle(desc, W, H, frames, max_len, spectral, sigma_factor=4.0):
    L = W * H
    imgs, st, rng, lrng = render_bpt_gpu(desc, W, H, frames, max_len, spectral, L)
    assert st["numFrames"] == frames and st["closestRays"] > 0 and st["shadowRays"] > 0
    ref, ost, rng_ref, lrng_ref, _ = ob.Scene(desc, spectral=spectral).render_bpt(
        W, H, frames, max_len, num_light_paths=L, rng_state=ob.seed_states(W, H),
        light_rng_state=ob.mt19937_64(LIGHT_SEED, L))
    same_eye, same_light = np.mean(rng == rng_ref), np.mean(lrng == lrng_ref)
    print("identical eye / light RNG states:", same_eye, same_light)
    print("rays gpu", st["closestRays"], st["shadowRays"], "oracle", ost["closest_rays"], ost["shadow_rays"])
    assert same_eye > 0.97 and same_light > 0.97
    img = imgs[-1][..., :3]
    samples = per_frame_images(imgs)[..., :3]
    sigma = np.sqrt((samples.var(axis=0, ddof=1) / frames).sum(axis=(0, 1))) / (W * H)
    diff = img.mean(axis=(0, 1)) - ref[..., :3].mean(axis=(0, 1))
    print("gpu mean", img.mean(axis=(0, 1)), "oracle mean", ref[..., :3].mean(axis=(0, 1)), "sigma", sigma)
    assert np.all(np.abs(diff) < sigma_factor * math.sqrt(2) * sigma + 1e-6)
    if same_light == 1.0:
        # every light path matched, so every cache vertex and every lens splat did
        assert_pixels_match_where_paths_match(img, ref, rng, rng_ref, W, H, 1e-3)
    assert st["droppedNonFinite"] == ost["dropped_non_finite"]


def test_bpt_cornell_rgb(torch_gpu):
    compare_with_oracle(scenes.cornell_box(), 32, 32, 4, 8, False)


def test_bpt_many_lights(torch_gpu):
    # 64 point lights (delta positions: lens connections only) and 8 quad lights
    compare_with_oracle(scenes.cornell_many_lights(), 32, 32, 4, 8, False)


def test_bpt_icosphere_grid_spectral(torch_gpu):
    # environment light: light paths start on the scene's bounding disc; 16-strata splats
    compare_with_oracle(scenes.icosphere_grid(grid=4, env_size=(64, 32)), 48, 27, 3, 8, True)


def test_bpt_white_furnace(torch_gpu):
    desc = scenes.white_furnace(spectral=False, flat=True)
    W = H = 32
    frames = 64
    imgs, st, _, _ = render_bpt_gpu(desc, W, H, frames, 8, False, W * H)
    assert st["droppedNonFinite"] == 0
    r = imgs[-1][..., :3] / camera_weight(desc.camera, W, H)[..., None]
    sigma = r.std() / math.sqrt(W * H)
    print("furnace mean", r.mean(), "sigma", sigma)
    assert abs(r.mean() - 1.0) < 4 * sigma + 1e-4
    assert abs(r[12:20, 12:20].mean() - 1.0) < 4 * r[12:20, 12:20].std() / 8 + 1e-4


def test_bpt_light_rng_state_round_trip(torch_gpu):
    """The light-path states are seeded once per context and can be read back and restored."""
    with api.Context() as ctx:
        ctx.set_num_light_paths(4096)
        s0 = ctx.read_light_rng()
        assert np.array_equal(s0, ob.mt19937_64(LIGHT_SEED, 4096))
        s1 = s0[::-1].copy()
        ctx.write_light_rng(s1)
        assert np.array_equal(ctx.read_light_rng(), s1)
