"""vlrBuildBvh4NoAlloc, the BVH build that allocates no device memory, against
vlrContextBuildBvh4 (the same build with the context's own scratch): identical node bytes,
node counts and primitive orders, on two calls in a row (the second smaller, so it runs over
scratch the first call left behind), for leaf sizes 4 and 1 (with a leaf remap); free device
memory is the same before and after a call; a build above the stated maximum is refused."""
import ctypes as C

import numpy as np
import pytest

from tests import bvh_check as bc
from vlr_amd import api

pytestmark = pytest.mark.gpu

MAX_PRIMS = 131072


def _bind():
    fn = api.lib().vlrBuildBvh4NoAlloc
    fn.argtypes = [C.c_void_p, C.c_void_p, C.c_uint32, C.c_uint32, C.c_void_p, C.c_void_p, C.c_uint32, C.c_void_p,
                   C.POINTER(C.c_uint32)]
    fn.restype = C.c_int
    return fn


def no_alloc_build(torch, boxes_dev, n, max_leaf, remap_dev, cap):
    nodes = torch.zeros(max(cap, 1) * 64, dtype=torch.uint8, device="cuda")
    order = torch.zeros(max(n, 1), dtype=torch.int32, device="cuda")
    torch.cuda.synchronize()
    cnt = C.c_uint32()
    rc = _bind()(None, C.c_void_p(boxes_dev.data_ptr()), n, max_leaf,
                 C.c_void_p(remap_dev.data_ptr()) if remap_dev is not None else None,
                 C.c_void_p(nodes.data_ptr()), cap, C.c_void_p(order.data_ptr()), C.byref(cnt))
    assert rc == 0, rc
    torch.cuda.synchronize()
    return nodes.cpu().numpy().view(api.BVH4_NODE)[:cnt.value].copy(), order.cpu().numpy().view(np.uint32)[:n].copy()


def context_build(torch, ctx, boxes_dev, n, max_leaf, remap_dev, cap):
    nodes = torch.zeros(max(cap, 1) * 64, dtype=torch.uint8, device="cuda")
    order = torch.zeros(max(n, 1), dtype=torch.int32, device="cuda")
    torch.cuda.synchronize()
    cnt = ctx.build_bvh4(boxes_dev.data_ptr(), n, max_leaf, 0 if remap_dev is None else remap_dev.data_ptr(),
                         nodes.data_ptr(), cap, order.data_ptr())
    torch.cuda.synchronize()
    return nodes.cpu().numpy().view(api.BVH4_NODE)[:cnt].copy(), order.cpu().numpy().view(np.uint32)[:n].copy()


@pytest.mark.parametrize("max_leaf", [4, 1])
def test_no_alloc_build_equals_context_build(torch_gpu, max_leaf):
    torch = torch_gpu
    rng = np.random.default_rng(41 + max_leaf)
    inputs = bc.test_inputs()
    cases = [("random 120k", (lambda n: np.concatenate([c := rng.uniform(-50, 50, (n, 3)),
                                                         c + rng.uniform(0.01, 2, (n, 3))], axis=1))(120_000)),
             ("clustered 3000", inputs["clustered_3000"]), ("identical 2500", inputs["identical_2500"])]
    with api.Context() as ctx:
        prev_n = None
        for name, boxes in cases:
            boxes = np.ascontiguousarray(boxes, np.float32).reshape(-1, 6)
            n = len(boxes)
            assert n <= MAX_PRIMS
            if prev_n is not None:
                assert n < prev_n                    # each call smaller than the one before
            prev_n = n
            print(f"case {name} n {n} max_leaf {max_leaf}", flush=True)
            cap = api.bvh4_node_capacity(n)
            b = torch.from_numpy(boxes.reshape(-1)).cuda()
            remap = None
            if max_leaf == 1:
                remap = torch.from_numpy(rng.permutation(n).astype(np.int32)).cuda()
            want_nodes, want_order = context_build(torch, ctx, b, n, max_leaf, remap, cap)
            got_nodes, got_order = no_alloc_build(torch, b, n, max_leaf, remap, cap)
            assert len(got_nodes) == len(want_nodes)
            assert got_nodes.tobytes() == want_nodes.tobytes()
            assert np.array_equal(got_order, want_order)


def test_no_alloc_build_allocates_nothing_and_refuses_oversize(torch_gpu):
    torch = torch_gpu
    rng = np.random.default_rng(5)
    n = 50_000
    c = rng.uniform(-10, 10, (n, 3))
    boxes = np.concatenate([c, c + rng.uniform(0.01, 1, (n, 3))], axis=1).astype(np.float32)
    b = torch.from_numpy(boxes.reshape(-1)).cuda()
    cap = api.bvh4_node_capacity(n)
    nodes = torch.zeros(cap * 64, dtype=torch.uint8, device="cuda")
    order = torch.zeros(n, dtype=torch.int32, device="cuda")
    cnt = C.c_uint32()
    args = (None, C.c_void_p(b.data_ptr()), n, 4, None, C.c_void_p(nodes.data_ptr()), cap,
            C.c_void_p(order.data_ptr()), C.byref(cnt))
    torch.cuda.synchronize()
    assert _bind()(*args) == 0                       # first call on this GPU: looks up the region
    torch.cuda.synchronize()
    free0, _ = torch.cuda.mem_get_info()
    assert _bind()(*args) == 0
    torch.cuda.synchronize()
    free1, _ = torch.cuda.mem_get_info()
    print("free device memory before / after:", free0, free1)
    assert free1 == free0
    # one primitive past the maximum: refused before any work
    big = MAX_PRIMS + 1
    rc = _bind()(None, C.c_void_p(b.data_ptr()), big, 4, None, C.c_void_p(nodes.data_ptr()),
                 api.bvh4_node_capacity(big), C.c_void_p(order.data_ptr()), C.byref(cnt))
    assert rc == 3                                  # VLRResult_InvalidArgument
