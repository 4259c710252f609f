"""Pins of the oracle's alpha test (path_tracing.cu:14-25 for closest-hit rays,
light_transport_common.h:205-218 for shadow rays) on the CPU.

  furnace          a flat albedo-1 sphere with alpha 0.5 in a unit environment still vanishes:
                   light either passes it or is reflected without loss, so every pixel equals the
                   camera weight (closed form)
  alpha 0          a fully transparent box renders as the scene without it (statistically: the
                   test still draws random numbers)
  alpha 1          a fully opaque alpha box renders as the plain box (statistically)
"""
import math

import numpy as np

from oracle import binding as ob
from tests.test_oracle_lights_render import camera_weight
from vlr_amd import scenes


def furnace_with_alpha(alpha, spectral=False):
    desc = scenes.white_furnace(spectral=spectral, flat=True)
    desc.meshes[0].group_alpha = {0: alpha}
    return desc


def test_alpha_furnace():
    desc = furnace_with_alpha(0.5)
    W = H = 32
    img, st, _, _ = ob.Scene(desc).render(W, H, 256, 8)
    r = img[..., :3] / camera_weight(desc.camera, W, H)[..., None]
    sigma = r.std() / math.sqrt(W * H)
    assert abs(r.mean() - 1.0) < 4 * sigma + 1e-4, (r.mean(), sigma)
    blk = r[12:20, 12:20]
    assert abs(blk.mean() - 1.0) < 4 * blk.std() / 8 + 1e-4


def batch_quadrants(desc, W, H, spp, batches, seed):
    scene = ob.Scene(desc)
    out = []
    for k in range(batches):
        img, _, _, _ = scene.render(W, H, spp, 8, rng_state=ob.mt19937_64(seed + k, W * H))
        q = []
        for ys in (slice(0, H // 2), slice(H // 2, H)):
            for xs in (slice(0, W // 2), slice(W // 2, W)):
                q.append(img[ys, xs, :3].astype(np.float64).mean(axis=(0, 1)))
        out.append(np.concatenate(q))
    out = np.array(out)
    return out.mean(axis=0), out.std(axis=0, ddof=1) / math.sqrt(batches)


def cornell_tall_box_alpha(alpha):
    desc = scenes.cornell_box()
    if alpha is not None:
        desc.meshes[3].group_alpha = {0: alpha}
    return desc


def test_transparent_box_renders_as_no_box():
    W = H = 24
    a, sa = batch_quadrants(cornell_tall_box_alpha(0.0), W, H, 16, 6, 100)
    removed = scenes.cornell_box()
    removed.instances = [i for i in removed.instances if i.name != "TallBoxInst"]
    b, sb = batch_quadrants(removed, W, H, 16, 6, 200)
    sig = np.sqrt(sa ** 2 + sb ** 2)
    assert np.all(np.abs(a - b) < 4 * sig + 1e-5), (a - b, sig)
    # and differs from the scene with the box
    c, sc = batch_quadrants(cornell_tall_box_alpha(None), W, H, 16, 6, 300)
    assert np.any(np.abs(a - c) > 6 * np.sqrt(sa ** 2 + sc ** 2))


def test_opaque_alpha_box_renders_as_plain_box():
    W = H = 24
    a, sa = batch_quadrants(cornell_tall_box_alpha(1.0), W, H, 16, 6, 400)
    b, sb = batch_quadrants(cornell_tall_box_alpha(None), W, H, 16, 6, 500)
    sig = np.sqrt(sa ** 2 + sb ** 2)
    assert np.all(np.abs(a - b) < 4 * sig + 1e-5), (a - b, sig)
