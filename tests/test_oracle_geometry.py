"""Pins of the oracle's ray casting and scene bounds (no GPU).

  BVH traversal (stands in for optixTrace, path_tracing.cu:89-93)
                            closest hits through the oracle's BVH equal its O(N) brute force
                            over every triangle of every instance, bit for bit
  watertight triangle test (Woop, Benthin, Wald, JCGT 2(1) 2013)
                            rays from the centre of a closed icosphere, including rays aimed
                            exactly at its vertices and edge midpoints, never escape
  hit distance              rays dropped vertically onto the ground plane hit at t = height
                            with barycentrics inside the triangle
  any-hit                   occluded in [tmin, tmax) iff the closest hit lies before tmax
  scene bounds (setup_scene.cu:7-92)
                            centre / radius / disc area of the union of the transformed
                            object-space boxes of the instances, recomputed here with numpy
"""
import math

import numpy as np
import pytest

from oracle import binding as ob
from tests.helpers import random_rays_in_box, rays_towards
from vlr_amd import scenes


@pytest.fixture(scope="module")
def cornell():
    return ob.Scene(scenes.cornell_box())


@pytest.fixture(scope="module")
def grid_small():
    return ob.Scene(scenes.icosphere_grid(grid=5, env_size=(64, 32)))


def assert_hits_equal(a, b):
    assert np.array_equal(a["inst"], b["inst"])
    assert np.array_equal(a["prim"], b["prim"])
    assert np.array_equal(a["geomInst"], b["geomInst"])
    assert np.array_equal(a["t"].view(np.uint32), b["t"].view(np.uint32))
    assert np.array_equal(a["b1"].view(np.uint32), b["b1"].view(np.uint32))
    assert np.array_equal(a["b2"].view(np.uint32), b["b2"].view(np.uint32))


def test_bvh_equals_brute_force_cornell(cornell):
    rays = random_rays_in_box(20000, (-1.4, 0.05, -1.4), (1.4, 2.95, 1.4), seed=3)
    assert_hits_equal(cornell.trace_closest(rays, use_bvh=True), cornell.trace_closest(rays, use_bvh=False))


def test_bvh_equals_brute_force_instanced_grid(grid_small):
    rays = rays_towards(6000, (-8, 1, 8), (8, 6, 14), (-6, 0, -6), (6, 1.5, 6), seed=4)
    a = grid_small.trace_closest(rays, use_bvh=True)
    assert (a["inst"] >= 0).mean() > 0.5
    assert_hits_equal(a, grid_small.trace_closest(rays, use_bvh=False))


def test_watertight_closed_mesh():
    v, idx = scenes.icosphere(3)
    desc = scenes.white_furnace()
    s = ob.Scene(desc)
    p = v[:, :3].astype(np.float64)
    tri = idx.reshape(-1, 3)
    targets = [p, (p[tri[:, 0]] + p[tri[:, 1]]) / 2, (p[tri[:, 1]] + p[tri[:, 2]]) / 2]
    rng = np.random.default_rng(8)
    d = rng.normal(size=(20000, 3))
    targets.append(d / np.linalg.norm(d, axis=1, keepdims=True))
    dirs = np.concatenate(targets)
    dirs /= np.linalg.norm(dirs, axis=1, keepdims=True)
    rays = np.zeros(len(dirs), ob.RAY_DTYPE)
    rays["dir"] = dirs.astype(np.float32)
    rays["tmax"] = np.float32(3.4e38)
    hits = s.trace_closest(rays)
    assert np.all(hits["inst"] == 0), int((hits["inst"] < 0).sum())
    assert np.all((hits["t"] > 0.9) & (hits["t"] <= 1.0 + 1e-6))


def test_vertical_rays_hit_ground_at_height(grid_small):
    rng = np.random.default_rng(12)
    n = 4000
    rays = np.zeros(n, ob.RAY_DTYPE)
    # between the spheres' footprints: spheres sit on grid points spaced 2.2 apart with radius <= 1
    gx = rng.integers(0, 4, n) * 2.2 - 4.4 + 1.1
    gz = rng.integers(0, 4, n) * 2.2 - 4.4 + 1.1
    h = rng.uniform(0.5, 3.0, n).astype(np.float32)
    rays["org"] = np.stack([gx + rng.uniform(-0.05, 0.05, n), h, gz + rng.uniform(-0.05, 0.05, n)], axis=1)
    rays["dir"] = (0.0, -1.0, 0.0)
    rays["tmax"] = np.float32(3.4e38)
    hits = grid_small.trace_closest(rays)
    assert np.all(hits["inst"] == 0)
    np.testing.assert_allclose(hits["t"], h, rtol=2e-7)
    assert np.all((hits["b1"] >= 0) & (hits["b2"] >= 0) & (hits["b1"] + hits["b2"] <= 1.0 + 1e-6))


def test_any_hit_consistent_with_closest(cornell, grid_small):
    for s, lo, hi in ((cornell, (-1.4, 0.05, -1.4), (1.4, 2.95, 1.4)), (grid_small, (-6, 0.1, -6), (6, 3, 6))):
        rays = random_rays_in_box(8000, lo, hi, seed=21)
        rng = np.random.default_rng(22)
        rays["tmax"] = rng.uniform(0.05, 4.0, len(rays)).astype(np.float32)
        closest = s.trace_closest(rays)
        occ = s.trace_any(rays)
        expected = (closest["inst"] >= 0) & (closest["t"] < rays["tmax"])
        assert np.array_equal(occ.astype(bool), expected)
        assert np.array_equal(s.trace_any(rays, use_bvh=False), occ)


@pytest.mark.parametrize("factory", [scenes.cornell_box, lambda: scenes.icosphere_grid(grid=4, env_size=(64, 32))])
def test_scene_bounds(factory):
    desc = factory()
    lo, hi = np.full(3, np.inf), np.full(3, -np.inf)
    for inst in desc.instances:
        pts = np.concatenate([desc.meshes[m].vertices[:, :3] for m in inst.meshes]).astype(np.float64)
        blo, bhi = pts.min(axis=0), pts.max(axis=0)
        corners = np.array([[x, y, z] for x in (blo[0], bhi[0]) for y in (blo[1], bhi[1]) for z in (blo[2], bhi[2])])
        m = np.asarray(inst.transform, np.float64).reshape(4, 4).T    # column-major
        w = corners @ m[:3, :3].T + m[:3, 3]
        lo, hi = np.minimum(lo, w.min(axis=0)), np.maximum(hi, w.max(axis=0))
    center = (lo + hi) / 2
    radius = np.linalg.norm(hi - center)
    b = ob.Scene(desc).bounds()
    np.testing.assert_allclose(b[:3], center, rtol=1e-5, atol=1e-5)
    np.testing.assert_allclose(b[3], radius, rtol=1e-5)
    np.testing.assert_allclose(b[4], math.pi * radius ** 2, rtol=2e-5)
