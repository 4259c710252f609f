"""Checks of 4-wide BVHs read back from the library (host and device builders).

  check_bvh4   walks the tree from node 0: every node is reached once, the leaves cover every
               primitive exactly once with at most max_leaf per leaf, and each decoded
               (quantised) child box contains the exact boxes of every primitive below it
  canonical    renumbers a tree in depth-first preorder so that trees built with different
               node numberings compare byte for byte
"""
import numpy as np

from vlr_amd import api

LEAF_SHIFT = 3


def leaf_range(code):
    u = (~int(code)) & 0xFFFFFFFF
    return u >> LEAF_SHIFT, (u & 7) + 1


def check_bvh4(nodes, order, boxes, max_leaf, remap=None):
    """Returns the depth of the tree; raises AssertionError on any violation."""
    boxes = np.asarray(boxes, np.float32).reshape(-1, 6)
    n = len(boxes)
    if n == 0:
        assert len(nodes) == 0
        return 0
    assert np.array_equal(np.sort(order), np.arange(n)), "order is not a permutation"
    lo, hi = api.decode_bvh4(nodes)
    blo = boxes[:, :3] + np.float32(0.0)
    bhi = boxes[:, 3:] + np.float32(0.0)
    inv = np.empty(n, np.int64)
    inv[order] = np.arange(n)
    covered = np.zeros(n, np.int32)
    visited = np.zeros(len(nodes), np.int32)
    # preorder walk; post-order pass for the exact subtree boxes
    pre, depth_of = [], {}
    stack = [(0, 1)]
    while stack:
        i, d = stack.pop()
        visited[i] += 1
        pre.append(i)
        depth_of[i] = d
        for k in range(4):
            c = int(nodes["child"][i][k])
            if c > 0:
                assert c < len(nodes)
                stack.append((c, d + 1))
    assert np.all(visited == 1), "node reached more than once or never"
    sub_lo = np.zeros((len(nodes), 3), np.float32)
    sub_hi = np.zeros((len(nodes), 3), np.float32)
    for i in reversed(pre):
        nl = np.full(3, np.inf, np.float32)
        nh = np.full(3, -np.inf, np.float32)
        used = 0
        for k in range(4):
            c = int(nodes["child"][i][k])
            if c == 0:
                continue
            used += 1
            if c > 0:
                cl, ch = sub_lo[c], sub_hi[c]
            else:
                if remap is not None:
                    u = (~c) & 0xFFFFFFFF
                    assert u & 7 == 0
                    prims = np.nonzero(remap == (u >> LEAF_SHIFT))[0]
                    assert len(prims) == 1
                else:
                    first, cnt = leaf_range(c)
                    assert 1 <= cnt <= max_leaf
                    assert first + cnt <= n
                    prims = order[first:first + cnt]
                covered[prims] += 1
                cl, ch = blo[prims].min(axis=0), bhi[prims].max(axis=0)
            assert np.all(lo[i, k] <= cl) and np.all(hi[i, k] >= ch), (i, k, lo[i, k], cl, hi[i, k], ch)
            nl, nh = np.minimum(nl, cl), np.maximum(nh, ch)
        assert used >= 1
        sub_lo[i], sub_hi[i] = nl, nh
    assert np.all(covered == 1), "primitive in no leaf or in several"
    return max(depth_of.values())


def canonical(nodes, base=0):
    """The tree whose root is node `base` and whose inner child codes count from `base` (as in a
    BLAS), renumbered in depth-first preorder (children in slot order) from 0."""
    pre = []
    stack = [base]
    while stack:
        i = stack.pop()
        pre.append(i)
        for k in reversed(range(4)):
            c = int(nodes["child"][i][k])
            if c > 0:
                stack.append(base + c)
    index = {old: new for new, old in enumerate(pre)}
    out = nodes[pre].copy()
    ch = out["child"]
    for r in range(len(out)):
        for k in range(4):
            if ch[r][k] > 0:
                ch[r][k] = index[base + int(ch[r][k])]
    out["child"] = ch
    return out


def random_boxes(n, seed, spread=10.0, size=0.3):
    rng = np.random.default_rng(seed)
    c = rng.uniform(-spread, spread, (n, 3)).astype(np.float32)
    e = rng.uniform(0, size, (n, 3)).astype(np.float32)
    return np.concatenate([c - e, c + e], axis=1)


def test_inputs():
    """Named box sets covering the builder's branches."""
    rng = np.random.default_rng(7)
    sets = {}
    sets["random_5000"] = random_boxes(5000, 1)
    sets["clustered_3000"] = np.concatenate([random_boxes(1500, 2, spread=0.5, size=0.01),
                                             random_boxes(1500, 3, spread=50.0, size=2.0)])
    sets["identical_2500"] = np.tile(np.array([[0, 0, 0, 1, 1, 1]], np.float32), (2500, 1))
    # centroids on 2^-k: the SAH peels one box per level, past the depth limit
    k = np.arange(300, dtype=np.float32)
    x = np.ldexp(np.float32(1.0), -(k % 120).astype(np.int32)).astype(np.float32)
    sets["geometric_300"] = np.stack([x, np.zeros_like(x), np.zeros_like(x), x, np.zeros_like(x), np.zeros_like(x)], 1)
    k = np.arange(5000)
    x = np.ldexp(np.float32(1.0), -(k % 120).astype(np.int32)).astype(np.float32)
    sets["geometric_5000"] = np.stack([x, np.zeros_like(x), np.zeros_like(x), x, np.zeros_like(x), np.zeros_like(x)], 1)
    flat = random_boxes(2000, 4)
    flat[:, 1] = 0.0
    flat[:, 4] = -0.0
    sets["flat_signed_zero_2000"] = flat
    sets["tiny_3"] = random_boxes(3, 5)
    sets["single"] = random_boxes(1, 6)
    return sets
