"""Pins of the oracle's denoiser guides (first-hit albedo and normal, path_tracing.cu:95-108,
converted as copyBuffers does, post_process.cu:70-90) by closed forms (no GPU).

  a camera that sees only one Lambertian quad of albedo rho:
      albedo guide = rho (RGB: exactly, up to the RGB -> XYZ -> RGB matrices in float;
      spectral: the white point and the wavelength pdf cancel in the CMF normalisation, so the
      average over samples converges to rho)
      normal guide = the quad's normal in the camera's frame with x negated:
      (-n.x_c, n.y_c, n.z_c) for the camera's axes x_c, y_c, z_c (z_c the viewing direction)
  a camera that sees nothing: both guides are zero
  an emitter seen directly: its BSDF is the null BSDF, albedo zero
"""
import math

import numpy as np
import pytest

from oracle import binding as ob
from vlr_amd import scenes


def quad_scene(material, normal, spectral, cam_pos=(1.2, 0.7, 5.0), target=(0.0, 0.0, 0.0)):
    n = np.asarray(normal, np.float64)
    n /= np.linalg.norm(n)
    # a 40 x 40 quad through the origin with normal n
    t = np.cross(n, [0.0, 1.0, 0.0] if abs(n[1]) < 0.9 else [1.0, 0.0, 0.0])
    t /= np.linalg.norm(t)
    b = np.cross(n, t)
    corners = [20 * (-t - b), 20 * (t - b), 20 * (t + b), 20 * (-t + b)]
    v = np.zeros((4, 11), np.float32)
    for i, c in enumerate(corners):
        v[i, 0:3] = c
        v[i, 3:6] = n
        v[i, 6:9] = t
    # winding so that the geometric normal agrees with n
    idx = np.array([0, 1, 2, 0, 2, 3], np.uint32)
    if np.dot(np.cross(corners[1] - corners[0], corners[2] - corners[0]), n) < 0:
        idx = np.array([0, 2, 1, 0, 3, 2], np.uint32)
    cam = scenes.CameraDesc(tuple(cam_pos), scenes.q_look_at(cam_pos, target), 1.0, math.radians(30.0))
    return scenes.SceneDesc([material], [scenes.MeshDesc("Quad", v, [(idx, 0)])],
                            [scenes.InstanceDesc("QuadInst", scenes.identity4(), [0])], cam, None, spectral, "quad")


def camera_axes(cam_pos, target, up=(0.0, 1.0, 0.0)):
    z = np.asarray(target, np.float64) - np.asarray(cam_pos, np.float64)
    z /= np.linalg.norm(z)
    x = np.cross(up, z)
    x /= np.linalg.norm(x)
    return x, np.cross(z, x), z


NORMAL = np.array([0.9, 0.3, 1.0]) / np.linalg.norm([0.9, 0.3, 1.0])


def test_rgb_albedo_and_normal_of_a_single_quad():
    rho = (0.6, 0.3, 0.2)
    m = scenes.MaterialDesc(scenes.MATTE, {"albedo": (scenes.REC709,) + rho})
    desc = quad_scene(m, NORMAL, False)
    _, alb, nrm, _, _ = ob.Scene(desc).render_guides(24, 24, 3, 4)
    np.testing.assert_allclose(alb[..., :3], np.broadcast_to(rho, (24, 24, 3)), rtol=2e-4)
    assert np.all(alb[..., 3] == 1.0)
    x, y, z = camera_axes((1.2, 0.7, 5.0), (0.0, 0.0, 0.0))
    expect = np.array([-NORMAL @ x, NORMAL @ y, NORMAL @ z])
    np.testing.assert_allclose(nrm[..., :3], np.broadcast_to(expect, (24, 24, 3)), atol=2e-6)
    # a wrong handedness (x kept) or the rotation applied instead of its inverse would differ
    assert abs(expect[0]) > 0.1


def test_spectral_albedo_converges_to_the_reflectance():
    rho = 0.45
    m = scenes.MaterialDesc(scenes.MATTE, {"albedo": (scenes.REC709, rho, rho, rho)})
    desc = quad_scene(m, NORMAL, True)
    _, alb, nrm, _, _ = ob.Scene(desc).render_guides(8, 8, 512, 2)
    mean = alb[..., :3].reshape(-1, 3).mean(axis=0)
    np.testing.assert_allclose(mean, rho, rtol=1e-2)
    # without the white point (or without dividing by the wavelength pdf) the scale is far off
    assert np.all(np.abs(mean / rho - 1) < 0.01)


def test_misses_and_emitters_give_zero_guides():
    m = scenes.MaterialDesc(scenes.MATTE, {"albedo": (scenes.REC709, 0.5, 0.5, 0.5)})
    # the camera looks away from the quad
    desc = quad_scene(m, NORMAL, False, cam_pos=(0.0, 0.0, 5.0), target=(0.0, 0.0, 10.0))
    _, alb, nrm, _, _ = ob.Scene(desc).render_guides(8, 8, 2, 2)
    assert np.all(alb[..., :3] == 0.0) and np.all(nrm[..., :3] == 0.0)
    e = scenes.MaterialDesc(scenes.DIFFUSE_EMITTER, {"emittance": (scenes.REC709, 3.0, 3.0, 3.0)}, {"scale": 1.0})
    _, alb, nrm, _, _ = ob.Scene(quad_scene(e, NORMAL, False)).render_guides(8, 8, 2, 2)
    assert np.all(alb[..., :3] == 0.0)
    assert np.all(np.abs(np.linalg.norm(nrm[..., :3], axis=-1) - 1.0) < 1e-5)
