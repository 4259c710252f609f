"""The host BVH builder (vlrBuildBvh4Host) on the CPU.

The device builder is tested against this one byte for byte (tests/test_gpu_bvh.py); here the
host build itself is checked against what any valid tree must satisfy (tests/bvh_check.py):
a permutation of the primitives, each in exactly one leaf of at most max_leaf, and decoded
quantised boxes that contain everything below them.  Also: the degenerate inputs end in a
bounded depth, a leaf remap puts each remapped index in exactly one leaf, and -0 / +0 in the
input give the same tree.
"""
import numpy as np
import pytest

from tests import bvh_check as bc
from vlr_amd import api


@pytest.mark.parametrize("name", list(bc.test_inputs().keys()))
@pytest.mark.parametrize("max_leaf", [1, 4, 8])
def test_host_build_is_a_valid_tree(name, max_leaf):
    boxes = bc.test_inputs()[name]
    nodes, order = api.build_bvh4_host(boxes, max_leaf)
    assert len(nodes) <= api.bvh4_node_capacity(len(boxes))
    depth = bc.check_bvh4(nodes, order, boxes, max_leaf)
    # the degenerate guards keep the 4-wide depth within the traversal stack's reach
    assert depth <= 40, depth


def test_empty_and_tiny():
    nodes, order = api.build_bvh4_host(np.zeros((0, 6), np.float32), 4)
    assert len(nodes) == 0 and len(order) == 0
    b = bc.random_boxes(4, 9)
    nodes, order = api.build_bvh4_host(b, 4)
    assert len(nodes) == 1
    assert list(nodes["child"][0][1:]) == [0, 0, 0]
    assert bc.leaf_range(nodes["child"][0][0]) == (0, 4)


def test_leaf_remap():
    boxes = bc.random_boxes(777, 10)
    remap = (np.random.default_rng(11).permutation(777) * 3 + 5).astype(np.uint32)
    nodes, order = api.build_bvh4_host(boxes, 1, remap=remap)
    bc.check_bvh4(nodes, order, boxes, 1, remap=remap)
    plain, plain_order = api.build_bvh4_host(boxes, 1)
    assert np.array_equal(order, plain_order)
    assert np.array_equal(nodes["child"] > 0, plain["child"] > 0)


def test_signed_zero_does_not_change_the_tree():
    boxes = bc.random_boxes(3000, 12)
    boxes[::3, 1] = 0.0
    boxes[::3, 4] = 0.0
    flipped = boxes.copy()
    flipped[::6, 1] = -0.0
    flipped[::6, 4] = -0.0
    a, oa = api.build_bvh4_host(boxes, 4)
    b, ob = api.build_bvh4_host(flipped, 4)
    assert np.array_equal(oa, ob)
    assert a.tobytes() == b.tobytes()
