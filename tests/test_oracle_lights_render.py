"""Pins of the oracle's light sampling and of its whole integrator (no GPU).

  triangle-light sampling (triangle.cu:113-177, light_transport_common.h:179-192)
                        one quad light: selection probability 1, area pdf = 1 / area,
                        positions on the quad and uniform over it (chi-square)
  environment sampling (infinite_sphere.cu:60-114, shared.h:276-298, context.cpp:1890-1921)
                        sampled pdf = pdf looked up along the sampled direction; the
                        solid-angle pdf integrates to 1; the pdf is the 4x-shrunk luminance
                        map weighted by sin(theta), recomputed here with numpy
  path tracer (path_tracing.cu:30-337)
                        white furnace: a flat-shaded albedo-1 sphere in a unit environment
                        vanishes, every pixel equals the camera weight (pixel-averaged cos^4
                        times 4 tan^2(fovy/2) aspect, integrated here with numpy); RGB exactly,
                        spectral within the 16-strata discretisation of the white spectrum
                        unbiasedness: RMSE against a 4096-spp render with independent seeds
                        falls as spp^-1/2
"""
import math

import numpy as np
import pytest

from oracle import binding as ob
from vlr_amd import scenes


def test_quad_light_sampling_is_uniform():
    desc = scenes.cornell_box()
    s = ob.Scene(desc)
    rng = np.random.default_rng(2)
    n = 64000
    ls = s.sample_lights(rng.uniform(size=(n, 3)))
    assert np.all(ls["valid"] == 1) and np.all(ls["atInfinity"] == 0)
    np.testing.assert_allclose(ls["lightProb"], 1.0, rtol=1e-6)
    np.testing.assert_allclose(ls["areaPDF"], 1.0 / 0.25 ** 2, rtol=1e-5)
    p = ls["position"].astype(np.float64)
    np.testing.assert_allclose(p[:, 1], 2.99, atol=1e-5)
    np.testing.assert_allclose(ls["normal"], np.tile([0.0, -1.0, 0.0], (n, 1)), atol=1e-6)
    x, z = p[:, 0], p[:, 2]
    assert np.all(np.abs(x) <= 0.125 + 1e-6) and np.all(np.abs(z) <= 0.125 + 1e-6)
    counts, _, _ = np.histogram2d(x, z, bins=8, range=[[-0.125, 0.125], [-0.125, 0.125]])
    expected = n / 64
    chi2 = ((counts - expected) ** 2 / expected).sum()
    assert chi2 < 63 + 5 * math.sqrt(2 * 63)


@pytest.fixture(scope="module")
def env_scene():
    desc = scenes.icosphere_grid(grid=2, env_size=(64, 32))
    return desc, ob.Scene(desc, spectral=False)


def test_env_sample_pdf_matches_lookup(env_scene):
    desc, s = env_scene
    rng = np.random.default_rng(3)
    ls = s.sample_lights(rng.uniform(size=(20000, 3)))
    assert np.all(ls["atInfinity"] == 1)          # the environment is the only emitter
    np.testing.assert_allclose(ls["lightProb"], 1.0, rtol=1e-6)
    look = s.env_lookup(ls["position"])
    # within 2 degrees of a pole acos(y) of the float direction no longer resolves theta
    away = np.abs(ls["position"][:, 1]) < math.cos(math.radians(2.0))
    assert away.mean() > 0.99
    np.testing.assert_allclose(look[away, 2], ls["areaPDF"][away], rtol=2e-4)


def test_env_pdf_integrates_to_one(env_scene):
    _, s = env_scene
    rng = np.random.default_rng(4)
    d = rng.normal(size=(400000, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    est = s.env_lookup(d)[:, 2].astype(np.float64) * 4 * math.pi
    mean, sigma = est.mean(), est.std() / math.sqrt(len(est))
    assert abs(mean - 1.0) < 4 * sigma + 1e-3, (mean, sigma)


def test_env_pdf_is_the_luminance_map(env_scene):
    desc, s = env_scene
    tex = desc.env.texture_half.view(np.float16).astype(np.float64)   # (H, W, 4)
    H, W = tex.shape[:2]
    mh, mw = H // 4, W // 4
    shrunk = tex[..., :3].reshape(mh, 4, mw, 4, 3).mean(axis=(1, 3))
    lum = shrunk @ np.array([0.2126729, 0.7151522, 0.0721750])
    theta_rows = math.pi * (np.arange(mh) + 0.5) / mh
    weight = lum * np.sin(theta_rows)[:, None]
    pdf_uv = weight / weight.mean()
    # query at the cell centres: u = (x + 0.5) / mw, v = (y + 0.5) / mh
    uu, vv = np.meshgrid((np.arange(mw) + 0.5) / mw, (np.arange(mh) + 0.5) / mh)
    phi, theta = 2 * math.pi * uu - desc.env.rotation, math.pi * vv
    dirs = np.stack([-np.sin(phi) * np.sin(theta), np.cos(theta), np.cos(phi) * np.sin(theta)], axis=-1).reshape(-1, 3)
    look = s.env_lookup(dirs)
    np.testing.assert_allclose(look[:, 0], uu.ravel(), atol=1e-5)
    np.testing.assert_allclose(look[:, 1], vv.ravel(), atol=1e-5)
    expected = pdf_uv.ravel() / (2 * math.pi ** 2 * np.sin(theta.ravel()))
    np.testing.assert_allclose(look[:, 2], expected, rtol=2e-4)


# ---------------------------------------------------------------------------------------------
def camera_weight(cam, W, H, sub=16):
    th = math.tan(cam.fovy / 2)
    o = (np.arange(sub) + 0.5) / sub
    ys, xs = np.mgrid[0:H, 0:W]
    acc = np.zeros((H, W))
    for a in o:
        for b in o:
            px = (0.5 - (xs + a) / W) * 2 * th * cam.aspect
            py = (0.5 - (ys + b) / H) * 2 * th
            acc += (1.0 / (px * px + py * py + 1.0)) ** 2
    return acc / sub ** 2 * 4 * th * th * cam.aspect


def test_white_furnace_rgb():
    desc = scenes.white_furnace(spectral=False, flat=True)
    W = H = 64
    img, st, _, _ = ob.Scene(desc).render(W, H, 256, 8)
    assert st["dropped_non_finite"] == 0
    r = img[..., :3] / camera_weight(desc.camera, W, H)[..., None]
    sigma = r.std() / math.sqrt(W * H)
    assert abs(r.mean() - 1.0) < 4 * sigma + 1e-5, (r.mean(), sigma)
    sphere = r[24:40, 24:40]
    assert abs(sphere.mean() - 1.0) < 4 * sphere.std() / 16 + 1e-5
    assert np.abs(r - 1.0).max() < 0.1


def test_white_furnace_spectral():
    desc = scenes.white_furnace(spectral=True, flat=True)
    W = H = 64
    img, st, _, _ = ob.Scene(desc).render(W, H, 256, 8)
    r = img[..., :3] / camera_weight(desc.camera, W, H)[..., None]
    sphere = r[24:40, 24:40].reshape(-1, 3)
    env = np.concatenate([r[:10].reshape(-1, 3), r[-10:].reshape(-1, 3)])
    # the sphere is invisible: its pixels match the directly seen environment
    diff = sphere.mean(axis=0) - env.mean(axis=0)
    sig = np.sqrt(sphere.var(axis=0) / len(sphere) + env.var(axis=0) / len(env))
    assert np.all(np.abs(diff) < 4 * sig), (diff, sig)
    # the white light source through 16 strata of piecewise-constant spectra: within 1 %
    np.testing.assert_allclose(env.mean(axis=0), 1.0, atol=1e-2)


def test_rmse_falls_as_inverse_sqrt_spp():
    desc = scenes.cornell_box()
    s = ob.Scene(desc)
    W = H = 24
    # independent seeds for the reference
    ref_seeds = ob.mt19937_64(12345, W * H)
    ref, _, _, _ = s.render(W, H, 4096, 8, rng_state=ref_seeds)
    rmse = []
    spps = (16, 64, 256)
    for spp in spps:
        img, _, _, _ = s.render(W, H, spp, 8, rng_state=ob.seed_states(W, H))
        rmse.append(math.sqrt(np.mean((img[..., :3].astype(np.float64) - ref[..., :3]) ** 2)))
    slope = np.polyfit(np.log(spps), np.log(rmse), 1)[0]
    assert -0.6 < slope < -0.4, (rmse, slope)
