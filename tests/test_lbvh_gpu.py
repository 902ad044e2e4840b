"""GPU LBVH parity: every stage bit-exact against the CPU oracle.

Mirrors the reference's stage-by-stage checks (crates/gpu/tests/lbvh_gpu.rs,
radix_sort.rs) and strengthens them: Morton codes, the sorted order, the
Karras topology, the fitted boxes, subtree sizes and the relayout into
GpuBvhNode must equal the CPU twin's bit for bit, not merely agree to a cell.
"""
from __future__ import annotations

import sys

import numpy as np
import pytest

import lucida_amd as L

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def xorshift(seed: int, n: int) -> np.ndarray:
    out = np.empty(n, np.uint32)
    s = np.uint32(seed)
    with np.errstate(over="ignore"):
        for i in range(n):
            s ^= np.uint32(s << np.uint32(13))
            s ^= np.uint32(s >> np.uint32(17))
            s ^= np.uint32(s << np.uint32(5))
            out[i] = s
    return out


def check_sort(gpu, oracle, keys, label):
    say(f"sort case: {label} n={len(keys)}")
    vals = np.arange(len(keys), dtype=np.uint32)
    ko, vo = L.LbvhBuilder(gpu).sort(keys, vals)
    ek, ev = oracle.sort_pairs(keys, vals)
    np.testing.assert_array_equal(ko, ek, err_msg=f"{label}: keys")
    np.testing.assert_array_equal(vo, ev, err_msg=f"{label}: payload (stability)")


@pytest.mark.parametrize("n", [1, 2, 255, 256, 257, 511, 512, 1000, 2047, 2048, 2049, 4096, 100_000, 3_000_001])
def test_sorts_random_keys(gpu, oracle, n):
    check_sort(gpu, oracle, xorshift(0x12345678, n), f"random n={n}")


def test_sorts_duplicates_and_adversarial_orders(gpu, oracle):
    check_sort(gpu, oracle, xorshift(99, 20_000) % 64, "few distinct values")
    check_sort(gpu, oracle, np.full(20_000, 7, np.uint32), "all identical")
    n = 10_000
    i = np.arange(n, dtype=np.uint32)
    check_sort(gpu, oracle, i.copy(), "already sorted")
    check_sort(gpu, oracle, i[::-1].copy(), "reversed")
    check_sort(gpu, oracle, (i % 16) << 28, "high digit only")
    check_sort(gpu, oracle, i % 16, "low digit only")
    check_sort(gpu, oracle, np.array([0xFFFFFFFF, 0, 0x7FFFFFFF, 1, 0xFFFFFFFE, 0x80000000], np.uint32), "extremes")


def test_empty_sort_is_not_an_error(gpu):
    k, v = L.LbvhBuilder(gpu).sort(np.zeros(0, np.uint32), np.zeros(0, np.uint32))
    assert len(k) == 0 and len(v) == 0


def geometry(name: str, **kw):
    a = L.Blob(name, **kw).arrays()
    return a["triangles"], a["positions"]


def adversarial(kind: str):
    if kind == "flat":
        rng = np.random.default_rng(1)
        pos = rng.random((300, 4), dtype=np.float32)
        pos[:, 1] = 5.0  # every centroid on one plane: a zero-extent axis
    elif kind == "duplicates":
        pos = np.array([[0, 0, 0, 0], [1, 0, 0, 0], [0, 1, 0, 0]], np.float32)
        tris = np.zeros(64, L.TRIANGLE)
        tris["i1"], tris["i2"] = 1, 2
        return tris, pos
    elif kind == "nan":
        rng = np.random.default_rng(2)
        pos = rng.random((300, 4), dtype=np.float32)
        pos[7, 0] = np.nan
        pos[40, :3] = np.nan
    elif kind.startswith("n="):
        n = int(kind[2:])
        rng = np.random.default_rng(n)
        pos = rng.random((3 * n, 4), dtype=np.float32) * 10 - 5
    else:
        raise ValueError(kind)
    n = len(pos) // 3
    tris = np.zeros(n, L.TRIANGLE)
    tris["i0"], tris["i1"], tris["i2"] = np.arange(n) * 3, np.arange(n) * 3 + 1, np.arange(n) * 3 + 2
    return tris, pos


CASES = ["icosphere-0", "icosphere-3", "icosphere-5", "cornell-mesh", "box-c1", "flat", "duplicates", "nan",
         "n=1", "n=2", "n=3", "n=5", "n=4097"]


def case_geometry(case):
    if case in ("flat", "duplicates", "nan") or case.startswith("n="):
        return adversarial(case)
    return geometry(case)


def bits(a):
    return np.ascontiguousarray(a).view(np.uint32)


@pytest.mark.parametrize("case", CASES)
def test_build_parts_bit_exact(gpu, oracle, case):
    tris, pos = case_geometry(case)
    say(f"lbvh build_parts case: {case} n={len(tris)}")
    g = L.LbvhBuilder(gpu).build_parts(tris, pos)
    o = oracle.lbvh_build(tris, pos)
    n = len(tris)
    np.testing.assert_array_equal(g["codes"], o["codes"], err_msg="Morton codes")
    np.testing.assert_array_equal(g["order"], o["order"], err_msg="sorted order")
    sb_g, sb_o = g["scene_bounds"][0], o["scene_bounds"][0]
    # min/max of centroids: equal as floats (a zero's sign may differ)
    assert np.array_equal(sb_g["lo"], sb_o["lo"]) and np.array_equal(sb_g["hi"], sb_o["hi"])
    if n >= 2:
        np.testing.assert_array_equal(g["karras"], o["karras"], err_msg="Karras topology")
        np.testing.assert_array_equal(g["parent"], o["parent"], err_msg="parents")
        np.testing.assert_array_equal(g["subtree_size"], o["subtree_size"], err_msg="subtree sizes")
        assert g["subtree_size"][0] == n
    for f in ("lo", "hi"):
        np.testing.assert_array_equal(bits(g["node_bounds"][f]), bits(o["node_bounds"][f]), err_msg=f"box {f}")


@pytest.mark.parametrize("case", CASES)
def test_build_nodes_bit_exact(gpu, oracle, case):
    tris, pos = case_geometry(case)
    say(f"lbvh build case: {case} n={len(tris)}")
    nodes, prims = L.LbvhBuilder(gpu).build(tris, pos)
    o = oracle.lbvh_build(tris, pos)
    assert len(nodes) == len(o["nodes"])
    np.testing.assert_array_equal(bits(nodes), bits(o["nodes"]), err_msg="GpuBvhNode relayout")
    np.testing.assert_array_equal(prims, o["prim_indices"])
    assert oracle.validate(nodes, prims, len(tris)) == 0
    assert oracle.sah_cost(nodes) == o["sah_cost"]


def test_full_size_c2_build_is_bit_exact(gpu, oracle):
    """The benchmark geometry at its real size: 1,048,820 + 1,024 triangles."""
    tris, pos = geometry("sphere-c2")
    say(f"lbvh full-size case: sphere-c2 n={len(tris)}")
    assert len(tris) == 1_049_844
    nodes, prims = L.LbvhBuilder(gpu).build(tris, pos)
    o = oracle.lbvh_build(tris, pos)
    np.testing.assert_array_equal(prims, o["prim_indices"])
    np.testing.assert_array_equal(bits(nodes), bits(o["nodes"]))


@pytest.mark.parametrize("case", ["icosphere-3", "cornell-mesh", "duplicates", "n=2", "n=5", "n=1023", "n=1025", "n=5000", "sphere-c2"])
def test_fit_orders_agree(gpu, oracle, case):
    """The two-phase fit, the atomic bottom-up fit and the level-synchronous
    fit produce the same tree, the oracle's."""
    tris, pos = geometry(case) if case == "sphere-c2" else case_geometry(case)
    say(f"lbvh fit modes case: {case} n={len(tris)}")
    o = oracle.lbvh_build(tris, pos)
    try:
        for fit in ("atomic", "levels", "blocks"):
            nodes, prims = L.LbvhBuilder(gpu, fit=fit).build(tris, pos)
            np.testing.assert_array_equal(bits(nodes), bits(o["nodes"]), err_msg=fit)
            np.testing.assert_array_equal(prims, o["prim_indices"], err_msg=fit)
    finally:
        L.LbvhBuilder(gpu, fit="blocks")


def test_scene_lbvh_is_device_resident_and_compacted(gpu, oracle):
    b = L.Blob("cornell-mesh")
    a = b.arrays()
    say("scene-create LBVH case: cornell-mesh")
    scene = L.Scene(gpu, b, build_lbvh=True)
    nodes, tris = scene.download_bvh()
    o = oracle.lbvh_build(a["triangles"], a["positions"])
    np.testing.assert_array_equal(bits(nodes), bits(o["nodes"]))
    # triangles permuted into leaf order (Bvh::compact_primitives, bvh.rs:517)
    np.testing.assert_array_equal(bits(tris), bits(a["triangles"][o["prim_indices"]]))


def test_build_rejects_out_of_range_indices(gpu):
    tris = np.zeros(2, L.TRIANGLE)
    tris["i2"] = 99
    pos = np.zeros((3, 4), np.float32)
    with pytest.raises(L.GpuError) as e:
        L.LbvhBuilder(gpu).build(tris, pos)
    assert e.value.kind == "InvalidArg"
