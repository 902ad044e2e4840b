"""Treelet restructuring on the GPU (lucida_bvh_build_treelet) against the
oracle (oracle/treelet.cpp), and renders over the restructured tree against
renders over the LBVH.

No reference counterpart: the reference traces its LBVH as built. The device
pass must reproduce the oracle's tree exactly: the same nodes (boxes equal as
floats, a zero's sign aside; links and counts equal) and the same leaf order.
"""
from __future__ import annotations

import sys

import numpy as np
import pytest

import lucida_amd as L
from test_lbvh_gpu import adversarial, case_geometry, geometry

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def assert_same_tree(nodes, prims, ref, label):
    assert len(nodes) == len(ref["nodes"]), label
    for f in ("bounds_min", "bounds_max"):
        np.testing.assert_array_equal(nodes[f], ref["nodes"][f], err_msg=f"{label}: {f}")
    np.testing.assert_array_equal(nodes["left_first"], ref["nodes"]["left_first"], err_msg=f"{label}: links")
    np.testing.assert_array_equal(nodes["count"], ref["nodes"]["count"], err_msg=f"{label}: counts")
    np.testing.assert_array_equal(prims, ref["prim_indices"], err_msg=f"{label}: leaf order")


CASES = ["icosphere-0", "icosphere-3", "icosphere-5", "cornell-mesh", "box-c1", "flat", "duplicates", "nan",
         "n=1", "n=2", "n=3", "n=5", "n=33", "n=1023", "n=1025", "n=4097"]


def coinciding(n: int):
    pos = np.array([[0, 0, 0, 0], [1, 0, 0, 0], [0, 1, 0, 0]], np.float32)
    tris = np.zeros(n, L.TRIANGLE)
    tris["i1"], tris["i2"] = 1, 2
    return tris, pos


@pytest.mark.parametrize("case", CASES + ["coinciding=20000"])
def test_treelet_tree_matches_oracle(gpu, oracle, case):
    if case.startswith("coinciding="):
        tris, pos = coinciding(int(case.split("=")[1]))
    elif case in ("flat", "duplicates", "nan") or case.startswith("n="):
        tris, pos = case_geometry(case)
    else:
        tris, pos = geometry(case)
    for rounds in (1, 3):
        say(f"treelet case: {case} n={len(tris)} rounds={rounds}")
        nodes, prims = L.LbvhBuilder(gpu).build_treelet(tris, pos, rounds)
        ref = oracle.treelet_build(tris, pos, rounds)
        assert_same_tree(nodes, prims, ref, f"{case} rounds={rounds}")
        assert oracle.validate(nodes, prims, len(tris)) == 0 or case == "nan"


def test_treelet_full_size_c2_matches_oracle(gpu, oracle):
    """The benchmark geometry: 1,049,844 triangles."""
    tris, pos = geometry("sphere-c2")
    say(f"treelet full-size case: sphere-c2 n={len(tris)}")
    nodes, prims = L.LbvhBuilder(gpu).build_treelet(tris, pos)
    ref = oracle.treelet_build(tris, pos)
    assert_same_tree(nodes, prims, ref, "sphere-c2")
    assert ref["sah_cost"] < 0.8 * oracle.lbvh_build(tris, pos)["sah_cost"]


def test_treelet_after_other_builds_reuses_buffers(gpu, oracle):
    """A larger build, then a smaller one, then the LBVH again: the pass keeps
    no state between builds."""
    b = L.LbvhBuilder(gpu)
    for case in ("n=4097", "n=33"):
        tris, pos = adversarial(case)
        say(f"treelet reuse case: {case}")
        nodes, prims = b.build_treelet(tris, pos)
        assert_same_tree(nodes, prims, oracle.treelet_build(tris, pos), case)
    tris, pos = adversarial("n=1025")
    nodes, prims = b.build(tris, pos)
    ref = oracle.lbvh_build(tris, pos)
    np.testing.assert_array_equal(nodes.view(np.uint32), ref["nodes"].view(np.uint32))


def test_treelet_rounds_are_bounded(gpu, oracle):
    """More than 64 rounds is rejected before any work; 64 and 0 (taken as
    one) still build the oracle's tree."""
    tris, pos = adversarial("n=33")
    b = L.LbvhBuilder(gpu)
    say("treelet rounds case: n=33 rounds=65")
    with pytest.raises(L.GpuError, match="rounds"):
        b.build_treelet(tris, pos, 65)
    for rounds in (64, 0):
        say(f"treelet rounds case: n=33 rounds={rounds}")
        nodes, prims = b.build_treelet(tris, pos, rounds)
        assert_same_tree(nodes, prims, oracle.treelet_build(tris, pos, max(rounds, 1)), f"rounds={rounds}")


def test_scene_treelet_is_device_resident_and_compacted(gpu, oracle):
    b = L.Blob("cornell-mesh")
    a = b.arrays()
    say("scene-create treelet case: cornell-mesh")
    scene = L.Scene(gpu, b, build_lbvh=True, treelet=True)
    nodes, tris = scene.download_bvh()
    ref = oracle.treelet_build(a["triangles"], a["positions"])
    assert_same_tree(nodes, ref["prim_indices"], ref, "scene")
    np.testing.assert_array_equal(tris.view(np.uint32), a["triangles"][ref["prim_indices"]].view(np.uint32))


@pytest.mark.parametrize("name,kw", [("cornell-mesh", {}), ("sphere-c2", {"frequency": 40})])
def test_render_over_treelet_tree_equals_lbvh_render(gpu, name, kw):
    """The same paths over both trees: closest hits agree, so the films agree
    (up to triangles tied at one distance)."""
    blob = L.Blob(name, **kw)
    films = []
    for treelet in (False, True):
        say(f"treelet render case: {name} treelet={treelet}")
        scene = L.Scene(gpu, blob, build_lbvh=True, treelet=treelet)
        film = L.Film(gpu, 128, 128)
        u = blob.uniforms(128, 128, samples=8, max_depth=8, sampling=L.SAMPLING_MIS, sampler=L.SAMPLER_SOBOL)
        L.render_wavefront(gpu, scene, u, 8, film)
        films.append(L.film_mean(film.download(), 128, 128).astype(np.float64))
    gpu.check_errors()
    a, b = films
    assert np.isfinite(b).all()
    diff = np.abs(a - b)
    assert diff.mean() <= 1e-5, diff.mean()
    assert (diff > 1e-4).mean() < 1e-3


def test_device_entry_point_matches_oracle(gpu, oracle):
    """lucida_bvh_build_treelet_device on device arrays: nodes, leaf order and
    the triangles gathered into leaf order."""
    import ctypes as C
    lib = L.lib()
    tris, pos = geometry("cornell-mesh")
    tris = np.ascontiguousarray(tris)
    pos = np.ascontiguousarray(pos, np.float32)
    n = len(tris)
    say(f"treelet device entry case: cornell-mesh n={n}")
    sizes = {"tris": tris.nbytes, "pos": pos.nbytes, "nodes": (2 * n - 1) * 32, "prims": n * 4, "sorted": tris.nbytes}
    ptrs = {}
    try:
        for k, b in sizes.items():
            p = C.c_void_p()
            L._check(lib.lucida_device_alloc(gpu._h, b, C.byref(p)))
            ptrs[k] = p
        L._check(lib.lucida_copy_to_device(gpu._h, ptrs["tris"], tris.ctypes.data, tris.nbytes))
        L._check(lib.lucida_copy_to_device(gpu._h, ptrs["pos"], pos.ctypes.data, pos.nbytes))
        count = C.c_size_t()
        stage = np.zeros(3, np.float32)
        L._check(lib.lucida_bvh_build_treelet_device(gpu._h, ptrs["tris"], n, ptrs["pos"], L.TREELET_ROUNDS,
                                                     ptrs["nodes"], C.byref(count), ptrs["prims"], ptrs["sorted"],
                                                     stage.ctypes.data))
        nodes = np.empty(count.value, L.BVH_NODE)
        prims = np.empty(n, np.uint32)
        sorted_tris = np.empty(n, L.TRIANGLE)
        L._check(lib.lucida_copy_to_host(gpu._h, nodes.ctypes.data, ptrs["nodes"], nodes.nbytes))
        L._check(lib.lucida_copy_to_host(gpu._h, prims.ctypes.data, ptrs["prims"], prims.nbytes))
        L._check(lib.lucida_copy_to_host(gpu._h, sorted_tris.ctypes.data, ptrs["sorted"], sorted_tris.nbytes))
    finally:
        for p in ptrs.values():
            lib.lucida_device_free(gpu._h, p)
    ref = oracle.treelet_build(tris, pos)
    assert_same_tree(nodes, prims, ref, "device entry")
    np.testing.assert_array_equal(sorted_tris.view(np.uint32), tris[ref["prim_indices"]].view(np.uint32))
    assert (stage > 0).all()
