"""The oracle's treelet restructuring, pinned on the CPU before it checks the
device pass (oracle/treelet.cpp; lucida_bvh_build_treelet).

The reference has no counterpart: its GPU path traces the LBVH as built
(crates/core/src/lbvh.rs:422). So the oracle is pinned by what the method
itself fixes:
  * the subset search returns the least cost over every binary tree on its
    leaves, found here by enumerating all of them (up to 6 leaves), and the
    split it records rebuilds a tree of exactly that cost;
  * every tree it emits passes validate() (bvh.rs:544-606), indexes each
    primitive once, keeps leaves of at most 4 primitives and the 32-entry
    traversal stack's depth bound;
  * closest hits along random rays over the restructured tree equal those over
    the LBVH (the same distance for every ray);
  * on real meshes its surface-area cost is below the LBVH's.
"""
from __future__ import annotations

import itertools

import numpy as np
import pytest

import lucida_amd as L

CI, CT, MAX_LEAF = 1.2, 1.0, 4


def adversarial(kind: str):
    if kind == "duplicates":
        pos = np.array([[0, 0, 0, 0], [1, 0, 0, 0], [0, 1, 0, 0]], np.float32)
        tris = np.zeros(64, L.TRIANGLE)
        tris["i1"], tris["i2"] = 1, 2
        return tris, pos
    if kind == "flat":
        rng = np.random.default_rng(1)
        pos = rng.random((300, 4), dtype=np.float32)
        pos[:, 1] = 5.0
    else:
        n = int(kind[2:])
        rng = np.random.default_rng(n)
        pos = rng.random((3 * n, 4), dtype=np.float32) * 10 - 5
    n = len(pos) // 3
    tris = np.zeros(n, L.TRIANGLE)
    tris["i0"], tris["i1"], tris["i2"] = np.arange(n) * 3, np.arange(n) * 3 + 1, np.arange(n) * 3 + 2
    return tris, pos


def geometry(case: str):
    if case in ("flat", "duplicates") or case.startswith("n="):
        return adversarial(case)
    a = L.Blob(case).arrays()
    return a["triangles"], a["positions"]


def area(lo, hi) -> float:
    d = np.asarray(hi, np.float64) - np.asarray(lo, np.float64)
    return 2.0 * (d[0] * d[1] + d[1] * d[2] + d[2] * d[0])


def all_trees(leaves):
    """Every rooted binary tree with these labelled leaves ((2k-3)!! of them)."""
    if len(leaves) == 1:
        yield leaves[0]
        return
    last = leaves[-1]
    for t in all_trees(leaves[:-1]):
        yield from _insert(t, last)


def _insert(t, leaf):
    yield (t, leaf)
    if isinstance(t, tuple):
        for l in _insert(t[0], leaf):
            yield (l, t[1])
        for r in _insert(t[1], leaf):
            yield (t[0], r)


def tree_cost(t, lo, hi, cost, size):
    """(cost, leaves, lo, hi, primitives) of a topology under the cost model,
    each node taking the cheaper of splitting and staying one leaf."""
    if not isinstance(t, tuple):
        return float(cost[t]), lo[t], hi[t], int(size[t])
    cl, llo, lhi, nl = tree_cost(t[0], lo, hi, cost, size)
    cr, rlo, rhi, nr = tree_cost(t[1], lo, hi, cost, size)
    blo, bhi = np.minimum(llo, rlo), np.maximum(lhi, rhi)
    a = area(blo, bhi)
    c = CI * a + cl + cr
    if nl + nr <= MAX_LEAF:
        c = min(c, CT * a * (nl + nr))
    return c, blo, bhi, nl + nr


@pytest.mark.parametrize("k", [2, 3, 4, 5, 6])
def test_subset_search_finds_the_cheapest_tree(oracle, k):
    rng = np.random.default_rng(100 + k)
    for trial in range(6):
        lo = rng.random((k, 3)).astype(np.float32) * 4
        hi = lo + rng.random((k, 3)).astype(np.float32) * (3 if trial % 2 else 0.5)
        size = rng.integers(1, 3 if trial < 3 else 6, k).astype(np.uint32)
        cost = np.array([CT * area(lo[i], hi[i]) * size[i] * rng.uniform(0.6, 1.0) for i in range(k)], np.float32)
        boxes = np.zeros(k, L.AABB)
        boxes["lo"], boxes["hi"] = lo, hi
        c, part, leafy = oracle.treelet_dp(boxes, cost, size)
        full = (1 << k) - 1
        brute = min(tree_cost(t, lo, hi, cost, size)[0] for t in all_trees(list(range(k))))
        assert c[full] == pytest.approx(brute, rel=1e-5), (k, trial)

        # the recorded splits rebuild a tree of that cost
        def rebuild(s):
            if s & (s - 1) == 0:
                return int(s).bit_length() - 1
            p = int(part[s])
            assert p & s == p and 0 < p < s and p & (s & -s)
            return (rebuild(p), rebuild(s ^ p))

        assert tree_cost(rebuild(full), lo, hi, cost, size)[0] == pytest.approx(float(c[full]), rel=1e-5)


def test_subset_search_prefers_a_leaf_on_a_tie(oracle):
    # Two leaves sharing one unit box: splitting costs CI*6 + 6 + 6 = 19.2,
    # one leaf of two primitives CT*6*2 = 12.
    boxes = np.zeros(2, L.AABB)
    boxes["hi"] = 1.0
    c, part, leafy = oracle.treelet_dp(boxes, np.array([6, 6], np.float32), np.array([1, 1], np.uint32))
    assert c[3] == 12.0 and leafy[3] == 1
    # five primitives can not be one leaf
    c, part, leafy = oracle.treelet_dp(boxes, np.array([6, 6], np.float32), np.array([2, 3], np.uint32))
    assert leafy[3] == 0 and c[3] == pytest.approx(19.2, rel=1e-6)


def test_subset_search_breaks_ties_towards_even_splits(oracle):
    # Four coinciding boxes of two primitives each (too many for one leaf):
    # every topology costs the same, and the even split {0,1} | {2,3} is kept
    # (of the even splits, the one of smallest mask), not a chain.
    boxes = np.zeros(4, L.AABB)
    boxes["hi"] = 1.0
    c, part, leafy = oracle.treelet_dp(boxes, np.full(4, 12.0, np.float32), np.full(4, 2, np.uint32))
    assert part[15] == 0b0011 and leafy[15] == 0


def test_coinciding_triangles_keep_the_tree_shallow(oracle):
    # Thousands of identical triangles: all boxes coincide, so every treelet
    # topology ties; the restructured tree stays about as deep as the LBVH.
    for n in (1000, 20000):
        pos = np.array([[0, 0, 0, 0], [1, 0, 0, 0], [0, 1, 0, 0]], np.float32)
        tris = np.zeros(n, L.TRIANGLE)
        tris["i1"], tris["i2"] = 1, 2
        t = oracle.treelet_build(tris, pos)
        assert oracle.validate(t["nodes"], t["prim_indices"], n) == 0
        lb = depth_of_leaves(oracle.lbvh_build(tris, pos)["nodes"]).max()
        assert depth_of_leaves(t["nodes"]).max() <= lb + 2, (n, lb)


def depth_of_leaves(nodes):
    d = np.zeros(len(nodes), np.int64)
    for i in range(len(nodes)):
        if nodes["count"][i] == 0:
            c = nodes["left_first"][i]
            d[c] = d[c + 1] = d[i] + 1
    return d[nodes["count"] > 0]


CASES = ["icosphere-0", "icosphere-3", "cornell-mesh", "box-c1", "flat", "duplicates", "n=1", "n=2", "n=3", "n=5",
         "n=33", "n=4097"]


@pytest.mark.parametrize("case", CASES)
def test_restructured_tree_is_valid(oracle, case):
    tris, pos = geometry(case)
    n = len(tris)
    for rounds in (1, 3):
        t = oracle.treelet_build(tris, pos, rounds)
        nodes, prims = t["nodes"], t["prim_indices"]
        assert oracle.validate(nodes, prims, n) == 0
        assert np.array_equal(np.sort(prims), np.arange(n))
        assert nodes["count"].max() <= MAX_LEAF or n <= MAX_LEAF
        assert depth_of_leaves(nodes).max() <= 32
        if n >= 64 and case not in ("duplicates",):
            assert t["sah_cost"] < oracle.lbvh_build(tris, pos)["sah_cost"], case


def test_cost_gain_on_meshes(oracle):
    for case, bound in (("cornell-mesh", 0.8), ("icosphere-3", 0.95)):
        tris, pos = geometry(case)
        ratio = oracle.treelet_build(tris, pos)["sah_cost"] / oracle.lbvh_build(tris, pos)["sah_cost"]
        assert ratio < bound, (case, ratio)


def test_closest_hits_equal_the_lbvh_hits(oracle):
    blob = L.Blob("cornell-mesh")
    a = blob.arrays()
    tris, pos = a["triangles"], a["positions"]
    lb = oracle.lbvh_build(tris, pos)
    tl = oracle.treelet_build(tris, pos)
    rng = np.random.default_rng(5)
    n = 20_000
    rays = np.zeros((n, 8), np.float32)
    rays[:, :3] = rng.uniform(0.05, 0.95, (n, 3))
    d = rng.normal(size=(n, 3))
    rays[:, 4:7] = d / np.linalg.norm(d, axis=1, keepdims=True)
    rays[:, 7] = 1e30
    hits = []
    for t in (lb, tl):
        b = L.Blob("cornell-mesh")
        b.set_bvh(t["nodes"], tris[t["prim_indices"]])
        h, _ = oracle.scene_cast(b.desc(), rays)
        tri = h[:, 3].view(np.uint32)
        hit = h[:, 0] < 1e29
        orig = np.where(hit & (tri < len(tris)), t["prim_indices"][np.minimum(tri, len(tris) - 1)], tri)
        hits.append((h[:, 0].copy(), orig))
    assert (hits[0][0] < 1e29).mean() > 0.9
    np.testing.assert_array_equal(hits[0][0], hits[1][0])
    assert (hits[0][1] == hits[1][1]).mean() > 0.999  # ties at one distance may pick either triangle
