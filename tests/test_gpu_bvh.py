"""The device BVH builder (bvh_build.hip) against the host builder (bvh_host.cpp).

Both implement one algorithm with the same float arithmetic (binned SAH, stable partition,
greedy 4-wide collapse, quantisation), so the trees must be equal byte for byte once node
numbering is made canonical, and the primitive orders must be equal.  Covered: random,
clustered, identical (no SAH split: middle of the range), geometric (past the depth limit:
median split, and middle split for nodes of more than 1024 primitives), signed zeros, tiny
inputs, leaf sizes 1 and 4, leaf remapping, large inputs that take several wide rounds, and the
acceleration structures of whole scenes (TLAS, every BLAS, triangle records, instance bases).
A render with device-built structures is bit-identical to one with host-built structures, and
the traversal parity tests (tests/test_gpu_trace.py) run on device-built trees.
"""
import numpy as np
import pytest

from tests import bvh_check as bc
from vlr_amd import api, scenes

pytestmark = pytest.mark.gpu


def device_build(torch, ctx, boxes, max_leaf, remap=None):
    n = len(boxes)
    cap = api.bvh4_node_capacity(n)
    b = torch.from_numpy(np.ascontiguousarray(boxes, np.float32).reshape(-1)).cuda()
    nodes = torch.zeros(cap * 64, dtype=torch.uint8, device="cuda")
    order = torch.zeros(max(n, 1), dtype=torch.int32, device="cuda")
    r = None if remap is None else torch.from_numpy(np.asarray(remap, np.int32)).cuda()
    torch.cuda.synchronize()
    cnt = ctx.build_bvh4(b.data_ptr(), n, max_leaf, 0 if r is None else r.data_ptr(), nodes.data_ptr(), cap,
                         order.data_ptr())
    torch.cuda.synchronize()
    dn = nodes.cpu().numpy().view(api.BVH4_NODE)[:cnt].copy()
    do = order.cpu().numpy().view(np.uint32)[:n].copy()
    return dn, do


def assert_same_tree(dn, do, hn, ho):
    assert np.array_equal(do, ho)
    assert len(dn) == len(hn)
    if len(hn):
        assert bc.canonical(dn).tobytes() == bc.canonical(hn).tobytes()


@pytest.mark.parametrize("name", list(bc.test_inputs().keys()))
@pytest.mark.parametrize("max_leaf", [1, 4])
def test_device_build_equals_host_build(torch_gpu, name, max_leaf):
    boxes = bc.test_inputs()[name]
    print(f"case {name} max_leaf {max_leaf} n {len(boxes)}", flush=True)
    with api.Context() as ctx:
        dn, do = device_build(torch_gpu, ctx, boxes, max_leaf)
    hn, ho = api.build_bvh4_host(boxes, max_leaf)
    bc.check_bvh4(dn, do, boxes, max_leaf)
    assert_same_tree(dn, do, hn, ho)


def test_leaf_remap_and_repeated_builds(torch_gpu):
    boxes = bc.random_boxes(3000, 21)
    remap = (np.random.default_rng(22).permutation(3000) * 5 + 1).astype(np.uint32)
    with api.Context() as ctx:
        first = device_build(torch_gpu, ctx, boxes, 1, remap)
        small = device_build(torch_gpu, ctx, boxes[:700], 4)      # workspace reuse, smaller input
        again = device_build(torch_gpu, ctx, boxes, 1, remap)
    hn, ho = api.build_bvh4_host(boxes, 1, remap=remap)
    assert_same_tree(*first, hn, ho)
    assert_same_tree(*again, hn, ho)
    assert_same_tree(*small, *api.build_bvh4_host(boxes[:700], 4))
    bc.check_bvh4(first[0], first[1], boxes, 1, remap=remap)


@pytest.mark.parametrize("n", [300_000, 1_000_000])
def test_large_builds(torch_gpu, n):
    rng = np.random.default_rng(n)
    # clustered primitives of very different sizes: several wide rounds, many subtrees
    centres = rng.uniform(-100, 100, (64, 3)).astype(np.float32)
    c = (centres[rng.integers(0, 64, n)] + rng.normal(0, 3, (n, 3))).astype(np.float32)
    e = rng.exponential(0.05, (n, 3)).astype(np.float32)
    boxes = np.concatenate([c - e, c + e], axis=1)
    print(f"large build n {n}", flush=True)
    with api.Context() as ctx:
        dn, do = device_build(torch_gpu, ctx, boxes, 4)
    hn, ho = api.build_bvh4_host(boxes, 4)
    assert_same_tree(dn, do, hn, ho)


def read_all(torch, desc, builder):
    with api.Context() as ctx:
        ctx.set_accel_builder(builder)
        bs = api.build_scene(ctx, desc)
        parts = [ctx.read_accel(bs.scene, w) for w in range(4)]
        st = ctx.accel_stats()
    return parts, st


def blas_trees(nodes, inst):
    """Canonical form of every distinct BLAS (root node and record base from the instances)."""
    bases = sorted(set((int(a), int(b)) for a, b in inst if a != 0xFFFFFFFF), key=lambda p: p[1])
    return [(tri_base, bc.canonical(nodes, node_base).tobytes()) for node_base, tri_base in bases]


@pytest.mark.parametrize("scene", ["cornell", "icosphere_grid", "icosphere_grid_32", "tessellated"])
def test_scene_acceleration_structures_match(torch_gpu, scene):
    desc = {"cornell": lambda: scenes.cornell_box(),
            "icosphere_grid": lambda: scenes.icosphere_grid(grid=8, env_size=(64, 32)),
            "icosphere_grid_32": lambda: scenes.icosphere_grid(grid=32, env_size=(64, 32)),
            "tessellated": lambda: scenes.tessellated_cornell(5, 200_000)}[scene]()
    dev, dst = read_all(torch_gpu, desc, api.ACCEL_BUILDER_DEVICE)
    host, hst = read_all(torch_gpu, desc, api.ACCEL_BUILDER_HOST)
    print(scene, "device", dst, "host", hst)
    tlas_d, blas_d, tris_d, inst_d = dev
    tlas_h, blas_h, tris_h, inst_h = host
    assert np.array_equal(inst_d[:, 1], inst_h[:, 1])          # record bases
    assert tris_d.tobytes() == tris_h.tobytes()
    assert bc.canonical(tlas_d).tobytes() == bc.canonical(tlas_h).tobytes()
    td, th = blas_trees(blas_d, inst_d), blas_trees(blas_h, inst_h)
    assert len(td) == len(th) == dst["numBlas"]
    assert td == th
    assert len(blas_d) == len(blas_h)


def test_render_identical_with_either_builder(torch_gpu):
    desc = scenes.icosphere_grid(grid=6, env_size=(64, 32))
    out = []
    for builder in (api.ACCEL_BUILDER_DEVICE, api.ACCEL_BUILDER_HOST):
        with api.Context(spectral=True, max_path_length=8) as ctx:
            ctx.set_accel_builder(builder)
            bs = api.build_scene(ctx, desc)
            ctx.bind_output(160, 90)
            for f in range(4):
                ctx.render(bs.camera, f == 0, 4)
            out.append((ctx.read_output(), ctx.read_rng(), ctx.stats()["closestRays"]))
    assert np.array_equal(out[0][0].view(np.uint32), out[1][0].view(np.uint32))
    assert np.array_equal(out[0][1], out[1][1])
    assert out[0][2] == out[1][2]


def _render(ctx, bs, W, H, frames):
    for f in range(frames):
        ctx.render(bs.camera, f == 0, frames)
    return ctx.read_output()


@pytest.mark.parametrize("builder", [api.ACCEL_BUILDER_DEVICE, api.ACCEL_BUILDER_HOST])
def test_transform_change_keeps_blases(torch_gpu, builder):
    """Moving an instance rebuilds only the TLAS (the BLASes are kept) and renders exactly what a
    fresh setup of the moved scene renders; changing vertices rebuilds the BLASes."""
    import copy
    desc = scenes.icosphere_grid(grid=6, env_size=(64, 32))
    moved = copy.deepcopy(desc)
    xf = np.array(moved.instances[7].transform, np.float32).reshape(4, 4, order="F").copy()
    xf[:3, 3] += np.array([0.4, 0.9, -0.3], np.float32)
    moved.instances[7].transform = xf.reshape(16, order="F")
    W, H = 96, 54
    with api.Context(max_path_length=6) as ctx:
        ctx.set_accel_builder(builder)
        bs = api.build_scene(ctx, desc)
        ctx.bind_output(W, H)
        ctx.read_accel(bs.scene, api.ACCEL_TRIS)                  # first setup
        assert ctx.accel_stats()["blasReused"] == 0
        api.call("vlrParentNodeSetTransform", bs.instances[7], ctx.transform(moved.instances[7].transform))
        img = _render(ctx, bs, W, H, 3)
        st = ctx.accel_stats()
        assert st["blasReused"] == 1, st
        assert st["tlasRefit"] == 1, st
        # new vertices of one mesh: BLASes are built again
        v = np.ascontiguousarray(desc.meshes[0].vertices, np.float32)
        api.call("vlrTriangleMeshSurfaceNodeSetVertices", bs.meshes[0], v.ctypes.data, len(v))
        ctx.read_accel(bs.scene, api.ACCEL_TRIS)
        assert ctx.accel_stats()["blasReused"] == 0
    with api.Context(max_path_length=6) as ctx:
        ctx.set_accel_builder(builder)
        bs = api.build_scene(ctx, moved)
        ctx.bind_output(W, H)
        ref = _render(ctx, bs, W, H, 3)
    assert np.array_equal(img.view(np.uint32), ref.view(np.uint32))


def test_refitted_tlas_traces_exactly(torch_gpu):
    """After a transform-only update the TLAS is refitted (same tree, new boxes): closest hits
    on the moved scene equal the oracle's brute force bit for bit; with refit off it is rebuilt."""
    import copy
    from oracle import binding as ob
    from tests.helpers import gpu_trace, random_rays_in_box
    from tests.test_gpu_trace import _compare
    desc = scenes.icosphere_grid(grid=5, env_size=(64, 32))
    moved = copy.deepcopy(desc)
    rng = np.random.default_rng(3)
    for k in (2, 9, 17):
        xf = np.array(moved.instances[k].transform, np.float32).reshape(4, 4, order="F").copy()
        xf[:3, 3] += rng.uniform(-0.6, 0.6, 3).astype(np.float32)
        moved.instances[k].transform = xf.reshape(16, order="F")
    half = 4 * 2.2 / 2
    rays = random_rays_in_box(20000, (-half, 0.05, -half), (half, 2.5, half), seed=8)
    ref = ob.Scene(moved).trace_closest(rays, use_bvh=False)
    for refit in (True, False):
        with api.Context() as ctx:
            ctx.set_tlas_refit(refit)
            bs = api.build_scene(ctx, desc)
            ctx.read_accel(bs.scene, api.ACCEL_TRIS)        # first setup
            for k in (2, 9, 17):
                api.call("vlrParentNodeSetTransform", bs.instances[k], ctx.transform(moved.instances[k].transform))
            gpu = gpu_trace(torch_gpu, ctx, bs.scene, rays)
            st = ctx.accel_stats()
        assert st["tlasRefit"] == (1 if refit else 0), st
        bad, ties, t_ok, b_ok = _compare(gpu, ref)
        print(f"refit {refit}: mismatches {bad}, ties {ties}")
        assert bad == 0 and t_ok and b_ok


def test_shuffled_instances_rebuild_the_tlas(torch_gpu):
    """Shuffling every instance's transform of an 8 x 8 grid leaves a refitted tree whose node
    boxes overlap across the whole scene: the sum of inner-node areas over the root's area grows
    from 1.84 to 5.36 (2.9 times the built cost; on a 5 x 5 grid only 1.4 times, where a refit is
    kept), past the limit of twice the built cost, so the TLAS is rebuilt; the render equals a
    fresh setup of the shuffled scene."""
    import copy
    desc = scenes.icosphere_grid(grid=8, env_size=(64, 32))
    shuffled = copy.deepcopy(desc)
    perm = np.random.default_rng(5).permutation(len(desc.instances) - 1) + 1   # instance 0: the ground
    for dst, src in zip(range(1, len(desc.instances)), perm):
        shuffled.instances[dst].transform = np.array(desc.instances[src].transform, np.float32)
    W, H = 64, 36
    with api.Context(max_path_length=6) as ctx:
        bs = api.build_scene(ctx, desc)
        ctx.bind_output(W, H)
        ctx.read_accel(bs.scene, api.ACCEL_TRIS)
        for k in range(1, len(desc.instances)):
            api.call("vlrParentNodeSetTransform", bs.instances[k], ctx.transform(shuffled.instances[k].transform))
        img = _render(ctx, bs, W, H, 2)
        st = ctx.accel_stats()
    assert st["blasReused"] == 1 and st["tlasRefit"] == 0, st
    with api.Context(max_path_length=6) as ctx:
        bs = api.build_scene(ctx, shuffled)
        ctx.bind_output(W, H)
        ref = _render(ctx, bs, W, H, 2)
    assert np.array_equal(img.view(np.uint32), ref.view(np.uint32))
