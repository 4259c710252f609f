"""Closest-hit / any-hit parity of the HIP traversal against the oracle's brute force.

The oracle tests every triangle of every instance with the watertight test on the
object-space ray (oracle/oracle.cpp traceInstanceBrute); the device walks its own two-level
BVH.  Primitive ids, t and barycentrics must agree bit for bit; rays whose closest t is
shared by two primitives (exact ties) may pick either and are counted.
"""
import numpy as np
import pytest

from oracle import binding as ob
from tests.helpers import gpu_trace, random_rays_in_box, rays_towards
from vlr_amd import api, scenes

pytestmark = pytest.mark.gpu


def _compare(gpu, ref):
    same_prim = (gpu["inst"] == ref["inst"]) & (gpu["geomInst"] == ref["geomInst"]) & (gpu["prim"] == ref["prim"])
    both_miss = (gpu["inst"] < 0) & (ref["inst"] < 0)
    tie = (~same_prim) & (gpu["inst"] >= 0) & (ref["inst"] >= 0) & (gpu["t"] == ref["t"])
    bad = ~(same_prim | tie)
    exact = same_prim & ~both_miss
    t_ok = np.all(gpu["t"][exact].view(np.uint32) == ref["t"][exact].view(np.uint32))
    b_ok = np.all(gpu["b1"][exact].view(np.uint32) == ref["b1"][exact].view(np.uint32)) and \
        np.all(gpu["b2"][exact].view(np.uint32) == ref["b2"][exact].view(np.uint32))
    return int(bad.sum()), int(tie.sum()), bool(t_ok), bool(b_ok)


def test_cornell_closest_hit_bitexact(torch_gpu):
    desc = scenes.cornell_box()
    oracle = ob.Scene(desc)
    rays = random_rays_in_box(20000, (-1.49, 0.01, -1.49), (1.49, 2.98, 1.49), seed=1)
    # plus camera-like rays from outside the box
    rays = np.concatenate([rays, rays_towards(5000, (-0.5, 1.0, 5.5), (0.5, 2.0, 6.5), (-1.5, 0, -1.5), (1.5, 3, 1.5), 2)])
    ref = oracle.trace_closest(rays, use_bvh=False)
    with api.Context() as ctx:
        bs = api.build_scene(ctx, desc)
        gpu = gpu_trace(torch_gpu, ctx, bs.scene, rays)
    bad, ties, t_ok, b_ok = _compare(gpu, ref)
    print(f"cornell: {len(rays)} rays, mismatches {bad}, exact ties {ties}")
    assert bad == 0
    assert t_ok and b_ok
    # the box is open towards the camera (+z), so some rays escape
    assert (ref["inst"] >= 0).mean() > 0.8


def test_icosphere_grid_closest_hit_bitexact(torch_gpu):
    desc = scenes.icosphere_grid(grid=6, env_size=(64, 32))
    oracle = ob.Scene(desc)
    half = 5 * 2.2 / 2
    rays = rays_towards(20000, (-half, 3.0, half + 5), (half, 8.0, half + 10), (-half, 0, -half), (half, 1.5, half), 3)
    rays = np.concatenate([rays, random_rays_in_box(10000, (-half, 0.05, -half), (half, 2.0, half), seed=4)])
    ref = oracle.trace_closest(rays, use_bvh=False)
    with api.Context() as ctx:
        bs = api.build_scene(ctx, desc)
        gpu = gpu_trace(torch_gpu, ctx, bs.scene, rays)
    bad, ties, t_ok, b_ok = _compare(gpu, ref)
    print(f"icosphere grid: {len(rays)} rays, mismatches {bad}, exact ties {ties}, hit rate {(ref['inst'] >= 0).mean():.3f}")
    assert bad == 0
    assert t_ok and b_ok


def test_any_hit_matches_brute_force(torch_gpu):
    desc = scenes.icosphere_grid(grid=4, env_size=(64, 32))
    oracle = ob.Scene(desc)
    half = 3 * 2.2 / 2
    rays = random_rays_in_box(20000, (-half, 0.05, -half), (half, 2.0, half), seed=5)
    rays["tmax"] = np.random.default_rng(6).uniform(0.1, 4.0, len(rays)).astype(np.float32)
    ref = oracle.trace_any(rays, use_bvh=False)
    with api.Context() as ctx:
        bs = api.build_scene(ctx, desc)
        gpu = gpu_trace(torch_gpu, ctx, bs.scene, rays, any_hit=True)
    occ = (gpu["inst"] >= 0).astype(np.uint8)
    print(f"any-hit: occluded {occ.mean():.3f}")
    assert np.array_equal(occ, ref)


def degenerate_scene(kind, n):
    """One mesh whose BVH is degenerate: `identical` n copies of one triangle (no SAH split
    separates their centroids), `geometric` triangles in the planes x = 2^-(k mod 120) (the SAH
    peels one plane per level, past the builder's depth limit)."""
    k = np.arange(n)
    if kind == "identical":
        x = np.full(n, 0.5, np.float32)
    else:
        x = np.ldexp(np.float32(1.0), -(k % 120).astype(np.int32)).astype(np.float32)
    v = np.zeros((3 * n, 11), np.float32)
    v[0::3, 0:3] = np.stack([x, np.full(n, -0.5), np.full(n, -0.5)], 1)
    v[1::3, 0:3] = np.stack([x, np.full(n, 0.5), np.full(n, -0.5)], 1)
    v[2::3, 0:3] = np.stack([x, np.zeros(n), np.full(n, 0.5)], 1)
    v[:, 3] = 1.0
    v[:, 7] = 1.0
    m = scenes.MaterialDesc(scenes.MATTE, {"albedo": (scenes.REC709, 0.5, 0.5, 0.5)})
    cam = scenes.CameraDesc((3.0, 0.0, 0.0), scenes.q_rotate_y(-np.pi / 2), 1.0, 0.7)
    return scenes.SceneDesc([m], [scenes.MeshDesc("Degenerate", v, [(np.arange(3 * n, dtype=np.uint32), 0)])],
                            [scenes.InstanceDesc("Inst", scenes.identity4(), [0])], cam, None, False, kind)


@pytest.mark.parametrize("kind,n", [("identical", 2500), ("geometric", 5000)])
def test_closest_hit_on_degenerate_trees(torch_gpu, kind, n):
    """The deepest trees the builder makes within the traversal stack's bound: every hit equals
    the brute force (exact ties aside), so no stack entry is lost."""
    desc = degenerate_scene(kind, n)
    oracle = ob.Scene(desc)
    rays = rays_towards(20000, (1.5, -0.4, -0.4), (3.0, 0.4, 0.4), (-0.1, -0.45, -0.45), (1.1, 0.45, 0.45), 12)
    rays = np.concatenate([rays, random_rays_in_box(5000, (-0.2, -0.6, -0.6), (1.2, 0.6, 0.6), seed=13)])
    ref = oracle.trace_closest(rays, use_bvh=False)
    with api.Context() as ctx:
        bs = api.build_scene(ctx, desc)
        gpu = gpu_trace(torch_gpu, ctx, bs.scene, rays)
        st = ctx.accel_stats()
    bad, ties, t_ok, b_ok = _compare(gpu, ref)
    print(f"{kind}: depths TLAS {st['tlasDepth']} BLAS {st['blasDepth']}, mismatches {bad}, ties {ties}, "
          f"hit rate {(ref['inst'] >= 0).mean():.3f}")
    assert 3 * (st["tlasDepth"] + st["blasDepth"]) + 1 <= 80
    if kind == "geometric":
        assert st["blasDepth"] >= 20
    assert bad == 0 and t_ok and b_ok
    assert (ref["inst"] >= 0).mean() > 0.3


def test_tree_deeper_than_the_stack_is_refused(torch_gpu):
    """A mesh whose tree would need more stack entries than a traversal holds fails at scene
    set-up with InvalidArgument instead of rendering with lost entries."""
    desc = degenerate_scene("geometric", 100000)
    with api.Context() as ctx:
        bs = api.build_scene(ctx, desc)
        ctx.bind_output(8, 8)
        with pytest.raises(api.VLRError) as e:
            ctx.render(bs.camera, True, 1)
        assert e.value.code == 3
