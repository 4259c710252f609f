"""Pins of the oracle's LVC-BPT restatement (lvc_bpt.cu:144-971) against closed forms and
against the independently pinned path tracer (no GPU).

  white furnace     a flat albedo-1 sphere in a unit environment vanishes: every pixel equals
                    the camera weight, through light paths that start on the scene's bounding
                    disc (EnvironmentEDF), lens splats and eye connections
  BPT = PT          both estimators converge to the same image; compared over independent
                    seed batches on the Cornell box and on the box with 64 point lights and 8
                    quad lights, where the pixels that see a point light directly are masked
                    (path tracing cannot hit a point light; the lens connection of LVC-BPT can)
"""
import math

import numpy as np

from oracle import binding as ob
from tests.test_oracle_lights_render import camera_weight
from vlr_amd import scenes


def test_bpt_white_furnace():
    desc = scenes.white_furnace(spectral=False, flat=True)
    W = H = 32
    img, st, _, _, _ = ob.Scene(desc).render_bpt(W, H, 256, 8, num_light_paths=W * H)
    assert st["dropped_non_finite"] == 0
    r = img[..., :3] / camera_weight(desc.camera, W, H)[..., None]
    sigma = r.std() / math.sqrt(W * H)
    assert abs(r.mean() - 1.0) < 4 * sigma + 1e-4, (r.mean(), sigma)
    assert abs(r[12:20, 12:20].mean() - 1.0) < 4 * r[12:20, 12:20].std() / 8 + 1e-4


def batch_means(scene, W, H, spp, batches, mask, bpt):
    out = []
    for k in range(batches):
        if bpt:
            img, _, _, _, _ = scene.render_bpt(W, H, spp, 25, num_light_paths=W * H,
                                               rng_state=ob.mt19937_64(2000 + k, W * H),
                                               light_rng_state=ob.mt19937_64(3000 + k, W * H))
        else:
            img, _, _, _ = scene.render(W, H, spp, 25, rng_state=ob.mt19937_64(1000 + k, W * H))
        out.append(img[mask][:, :3].astype(np.float64).mean(axis=0))
    out = np.array(out)
    return out.mean(axis=0), out.std(axis=0, ddof=1) / math.sqrt(batches)


def point_light_mask(desc, W, H):
    cam = desc.camera
    pts = desc.meshes[-2].vertices[:, :3].astype(np.float64)
    th = math.tan(cam.fovy / 2)
    rel = pts - np.array(cam.position)
    # the camera looks down -z (qRotateY(pi)): camera-local x = -world x
    lx, ly, lz = -rel[:, 0], rel[:, 1], -rel[:, 2]
    u = 0.5 - (lx / lz) / (2 * th * cam.aspect)
    v = 0.5 - (ly / lz) / (2 * th)
    mask = np.ones((H, W), bool)
    for x, y in zip((u * W).astype(int), (v * H).astype(int)):
        if 0 <= x < W and 0 <= y < H:
            mask[y, x] = False
    return mask


def test_bpt_matches_path_tracing_cornell():
    s = ob.Scene(scenes.cornell_box())
    W = H = 24
    mask = np.ones((H, W), bool)
    mp, sp = batch_means(s, W, H, 256, 6, mask, bpt=False)
    mb, sb = batch_means(s, W, H, 128, 6, mask, bpt=True)
    print("PT", mp, sp, "BPT", mb, sb)
    assert np.all(np.abs(mp - mb) < 4 * np.sqrt(sp ** 2 + sb ** 2) + 2e-3 * mp)


def test_bpt_matches_path_tracing_many_lights():
    desc = scenes.cornell_many_lights()
    s = ob.Scene(desc)
    W = H = 32
    mask = point_light_mask(desc, W, H)
    mp, sp = batch_means(s, W, H, 256, 6, mask, bpt=False)
    mb, sb = batch_means(s, W, H, 256, 6, mask, bpt=True)
    print("PT", mp, sp, "BPT", mb, sb)
    # point lights near walls make both estimators heavy-tailed: 4 sigma plus 2 %
    assert np.all(np.abs(mp - mb) < 4 * np.sqrt(sp ** 2 + sb ** 2) + 0.02 * mp)
