"""Pins of the oracle's light tracer (light_tracing.cu:49-263) against a closed form and
against the independently pinned path tracer (no GPU).

  white furnace   a flat albedo-1 sphere in a unit environment vanishes: every pixel equals
                  the camera weight, reached only through light paths that start on the scene's
                  bounding disc and splat their lens connections
  LT = PT         both estimators converge to the same image on the Cornell box (compared over
                  independent seed batches, per image quadrant)
  RNG states      light tracing advances only the light-path states, the same for any thread count
"""
import math

import numpy as np

from oracle import binding as ob
from tests.test_oracle_lights_render import camera_weight
from vlr_amd import scenes


def test_lt_white_furnace():
    desc = scenes.white_furnace(spectral=False, flat=True)
    W = H = 32
    img, st, _, _ = ob.Scene(desc).render_lt(W, H, 64, 8, num_light_paths=16 * W * H)
    assert st["paths"] == 64 * 16 * W * H
    r = img[..., :3] / camera_weight(desc.camera, W, H)[..., None]
    # per-pixel noise of splatted estimates: compare the image mean and the sphere's block
    sigma = r.std() / math.sqrt(W * H)
    assert abs(r.mean() - 1.0) < 4 * sigma + 1e-3, (r.mean(), sigma)
    blk = r[12:20, 12:20]
    assert abs(blk.mean() - 1.0) < 4 * blk.std() / 8 + 1e-3, blk.mean()


def test_lt_white_furnace_spectral():
    desc = scenes.white_furnace(spectral=True, flat=True)
    W = H = 32
    img, _, _, _ = ob.Scene(desc).render_lt(W, H, 64, 8, num_light_paths=16 * W * H)
    r = img[..., :3] / camera_weight(desc.camera, W, H)[..., None]
    sigma = r.reshape(-1, 3).std(axis=0) / math.sqrt(W * H)
    # the white light source through 16 strata of piecewise-constant spectra: within 1 %
    assert np.all(np.abs(r.reshape(-1, 3).mean(axis=0) - 1.0) < 4 * sigma + 1e-2), r.reshape(-1, 3).mean(axis=0)


def quadrant_means(img, W, H):
    q = []
    for ys in (slice(0, H // 2), slice(H // 2, H)):
        for xs in (slice(0, W // 2), slice(W // 2, W)):
            q.append(img[ys, xs, :3].astype(np.float64).mean(axis=(0, 1)))
    return np.concatenate(q)


def test_lt_equals_pt_on_the_cornell_box():
    desc = scenes.cornell_box()
    scene = ob.Scene(desc)
    W = H = 32
    lt, pt = [], []
    for k in range(8):
        img, _, _, _ = scene.render_lt(W, H, 8, 25, num_light_paths=8 * W * H,
                                       light_rng_state=ob.mt19937_64(5000 + k, 8 * W * H))
        lt.append(quadrant_means(img, W, H))
        img, _, _, _ = scene.render(W, H, 64, 25, rng_state=ob.mt19937_64(6000 + k, W * H))
        pt.append(quadrant_means(img, W, H))
    lt, pt = np.array(lt), np.array(pt)
    diff = lt.mean(axis=0) - pt.mean(axis=0)
    sig = np.sqrt(lt.var(axis=0, ddof=1) / len(lt) + pt.var(axis=0, ddof=1) / len(pt))
    print("LT", lt.mean(axis=0), "PT", pt.mean(axis=0), "sigma", sig)
    assert np.all(np.abs(diff) < 4 * sig + 1e-4), (diff, sig)


def test_lt_light_states_and_thread_count():
    desc = scenes.cornell_box()
    scene = ob.Scene(desc)
    W = H = 16
    n = 2048
    s0 = ob.mt19937_64(77, n)
    a, _, sa, _ = scene.render_lt(W, H, 2, 8, num_light_paths=n, threads=1, light_rng_state=s0.copy())
    b, _, sb, _ = scene.render_lt(W, H, 2, 8, num_light_paths=n, threads=4, light_rng_state=s0.copy())
    assert np.array_equal(sa, sb) and not np.array_equal(sa, s0)
    # splats are added atomically in any order: equal to float rounding
    np.testing.assert_allclose(a, b, rtol=1e-5, atol=1e-7)
