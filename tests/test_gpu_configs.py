"""The GPU path against the oracle at the benchmark configurations' own sizes, not scaled-down
scenes: C1 (Cornell box, RGB, 256x256, 16 spp, max depth 8), C2 (the 1,310,722-triangle
spectral icosphere grid with its 1024x512 environment at 1920x1080, max depth 16; 2 spp, the
full 64 being a rate, not a different computation) and C3 (the many-lights box with LVC-BPT at
1280x720, one light path per pixel, max depth 8; 1 spp).

Both sides start every pixel (and light path) from the reference's seeds and consume the
generators in the reference's order, so wherever the per-pixel generator states agree after the
render the two took the same branches and their pixels must agree to float rounding (1e-3
relative at the 99.9th percentile; the device contracts some products into fused
multiply-adds); the share of such pixels is asserted, and the image means must agree within
Monte-Carlo noise."""
import math

import numpy as np
import pytest

from oracle import binding as ob
from tests.test_gpu_bpt import LIGHT_SEED, render_bpt_gpu
from tests.test_gpu_render import assert_pixels_match_where_paths_match, per_frame_images, render_gpu
from vlr_amd import api, scenes

pytestmark = pytest.mark.gpu


def mean_within(imgs, ref, W, H, factor):
    frames = len(imgs)
    img = imgs[-1][..., :3].astype(np.float64)
    if frames > 1:
        samples = per_frame_images(imgs)[..., :3]
        sigma = np.sqrt((samples.var(axis=0, ddof=1) / frames).sum(axis=(0, 1))) / (W * H)
    else:
        # one frame: the per-pixel spread over the image bounds the mean's noise
        sigma = img.reshape(-1, 3).std(axis=0) / math.sqrt(W * H)
    diff = img.mean(axis=(0, 1)) - ref[..., :3].astype(np.float64).mean(axis=(0, 1))
    print("gpu mean", img.mean(axis=(0, 1)), "oracle mean", ref[..., :3].mean(axis=(0, 1)), "sigma", sigma)
    assert np.all(np.abs(diff) < factor * math.sqrt(2) * sigma + 1e-7)


def test_c1_cornell_256_16spp(torch_gpu):
    desc = scenes.cornell_box()
    W = H = 256
    imgs, st, rng = render_gpu(desc, W, H, 16, 8, False)
    ref, ost, orng, _ = ob.Scene(desc).render(W, H, 16, max_path_length=8, rng_state=ob.seed_states(W, H))
    same = float(np.mean(rng == orng))
    print("C1: identical generator states", same, "rays", st["closestRays"], ost["closest_rays"])
    assert same > 0.97
    assert_pixels_match_where_paths_match(imgs[-1][..., :3], ref, rng, orng, W, H, 1e-3)
    mean_within(imgs, ref, W, H, 3.0)


def test_c2_icosphere_grid_1080p(torch_gpu):
    desc = scenes.icosphere_grid()
    assert desc.num_triangles == 1_310_722
    W, H, frames = 1920, 1080, 2
    imgs, st, rng = render_gpu(desc, W, H, frames, 16, True)
    ref, ost, orng, _ = ob.Scene(desc, spectral=True).render(W, H, frames, max_path_length=16,
                                                             rng_state=ob.seed_states(W, H))
    same = float(np.mean(rng == orng))
    print("C2: identical generator states", same, "rays gpu", st["closestRays"], st["shadowRays"], "oracle",
          ost["closest_rays"], ost["shadow_rays"])
    assert same > 0.9
    assert_pixels_match_where_paths_match(imgs[-1][..., :3], ref, rng, orng, W, H, 1e-3)
    mean_within(imgs, ref, W, H, 3.0)
    assert st["droppedNonFinite"] == 0


def test_c3_many_lights_bpt_720p(torch_gpu):
    desc = scenes.cornell_many_lights()
    W, H = 1280, 720
    L = W * H
    imgs, st, rng, lrng = render_bpt_gpu(desc, W, H, 1, 8, False, L)
    ref, ost, orng, olrng, _ = ob.Scene(desc).render_bpt(W, H, 1, 8, num_light_paths=L, rng_state=ob.seed_states(W, H),
                                                         light_rng_state=ob.mt19937_64(LIGHT_SEED, L))
    same_eye, same_light = float(np.mean(rng == orng)), float(np.mean(lrng == olrng))
    print("C3: identical eye / light generator states", same_eye, same_light, "rays gpu", st["closestRays"],
          st["shadowRays"], "oracle", ost["closest_rays"], ost["shadow_rays"])
    assert same_eye > 0.97 and same_light > 0.97
    mean_within(imgs, ref, W, H, 4.0)
    if same_light == 1.0:
        assert_pixels_match_where_paths_match(imgs[-1][..., :3], ref, rng, orng, W, H, 1e-3)
