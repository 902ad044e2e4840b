"""Host restatement of the device's two-level build, for the tests.

Mirrors lucida_scene_create_instanced step by step with the oracle's LBVH
(which is bit-exact with the device's): one LBVH per mesh range with the
triangles compacted into leaf order and the nodes rebased, the placements'
inverses and padded world bounds in the same floating-point order, an LBVH
over two-point "triangles" spanning those bounds, and the instances appended
in TLAS leaf order. Scene::build_instances (crates/core/src/scene.rs:252) is
the reference recipe; it builds both levels with binned SAH instead.
"""
from __future__ import annotations


import numpy as np

import lucida_amd as L


def invert_affine(f):
    """[R | t]^-1 in double precision, cofactor order as the device builder."""
    a = [[float(f[4 * r + c]) for c in range(3)] for r in range(3)]
    t = [float(f[4 * r + 3]) for r in range(3)]
    c00 = a[1][1] * a[2][2] - a[1][2] * a[2][1]
    c01 = a[1][2] * a[2][0] - a[1][0] * a[2][2]
    c02 = a[1][0] * a[2][1] - a[1][1] * a[2][0]
    det = a[0][0] * c00 + a[0][1] * c01 + a[0][2] * c02
    inv_det = 1.0 / det
    r = [[0.0] * 3 for _ in range(3)]
    r[0][0] = c00 * inv_det
    r[1][0] = c01 * inv_det
    r[2][0] = c02 * inv_det
    r[0][1] = (a[0][2] * a[2][1] - a[0][1] * a[2][2]) * inv_det
    r[1][1] = (a[0][0] * a[2][2] - a[0][2] * a[2][0]) * inv_det
    r[2][1] = (a[0][1] * a[2][0] - a[0][0] * a[2][1]) * inv_det
    r[0][2] = (a[0][1] * a[1][2] - a[0][2] * a[1][1]) * inv_det
    r[1][2] = (a[0][2] * a[1][0] - a[0][0] * a[1][2]) * inv_det
    r[2][2] = (a[0][0] * a[1][1] - a[0][1] * a[1][0]) * inv_det
    out = np.zeros((3, 4))
    for i in range(3):
        out[i, :3] = r[i]
        out[i, 3] = -(r[i][0] * t[0] + r[i][1] * t[1] + r[i][2] * t[2])
    return out


def world_bounds(f, root):
    """Eight corners through the forward transform, then Aabb::pad."""
    lo = [np.float32(np.inf)] * 3
    hi = [np.float32(-np.inf)] * 3
    for c in range(8):
        p = [float(root["bounds_max"][k] if (c >> k) & 1 else root["bounds_min"][k]) for k in range(3)]
        for k in range(3):
            w = np.float32(float(f[4 * k]) * p[0] + float(f[4 * k + 1]) * p[1] + float(f[4 * k + 2]) * p[2]
                           + float(f[4 * k + 3]))
            lo[k] = min(lo[k], w)
            hi[k] = max(hi[k], w)
    for k in range(3):
        pad = np.float32(1.0e-6) * (np.float32(1.0) + max(abs(lo[k]), abs(hi[k])))
        lo[k] = np.float32(lo[k] - pad)
        hi[k] = np.float32(hi[k] + pad)
    return lo, hi


def build_two_level(blob: L.Blob, oracle) -> dict:
    a = blob.arrays()
    meshes, placements = blob.instancing()
    tris = a["triangles"].copy()
    pos = a["positions"]
    parts, offsets, roots = [], [], []
    off = 0
    for m in meshes:
        f, c = int(m["first_triangle"]), int(m["triangle_count"])
        o = oracle.lbvh_build(tris[f:f + c].copy(), pos)
        nodes = o["nodes"].copy()
        roots.append(o["nodes"][0].copy())
        tris[f:f + c] = tris[f:f + c][o["prim_indices"]]
        internal = nodes["count"] == 0
        nodes["left_first"][internal] += off
        nodes["left_first"][~internal] += f
        parts.append(nodes)
        offsets.append(off)
        off += len(nodes)
    blas_total = off
    n = len(placements)
    box_pos = np.zeros((2 * n, 4), np.float32)
    box_tris = np.zeros(n, L.TRIANGLE)
    inst = np.zeros(n, L.INSTANCE)
    for i, pl in enumerate(placements):
        f = pl["object_to_world"]
        w2o = invert_affine(f)
        lo, hi = world_bounds(f, roots[pl["mesh"]])
        box_pos[2 * i, :3] = lo
        box_pos[2 * i + 1, :3] = hi
        box_tris[i] = (2 * i, 2 * i + 1, 2 * i, 0)
        inst[i]["inv_row0"] = w2o[0, :3]
        inst[i]["inv_row1"] = w2o[1, :3]
        inst[i]["inv_row2"] = w2o[2, :3]
        inst[i]["inv_translation"] = w2o[:, 3]
        inst[i]["blas_root"] = offsets[pl["mesh"]]
        inst[i]["kind"] = 2
        inst[i]["material"] = pl["material"]
    o = oracle.lbvh_build(box_tris, box_pos)
    tlas = o["nodes"].copy()
    internal = tlas["count"] == 0
    tlas["left_first"][internal] += blas_total
    tlas["left_first"][~internal] += len(a["primitives"])
    return {"nodes": np.concatenate(parts + [tlas]), "triangles": tris, "instances": inst[o["prim_indices"]],
            "tlas_root": blas_total}


def host_desc(blob: L.Blob, built: dict):
    """The blob's desc pointing at a host two-level structure (keep `built` alive)."""
    d = blob.desc()
    d.bvh_nodes, d.num_bvh_nodes = built["nodes"].ctypes.data, len(built["nodes"])
    d.triangles = built["triangles"].ctypes.data
    d.instances, d.num_instances = built["instances"].ctypes.data, len(built["instances"])
    d.tlas_root = built["tlas_root"]
    return d


def random_rays(n: int, lo, hi, seed: int, t_max: float = 1e30) -> np.ndarray:
    """Origins uniform in the box [lo, hi]^3 widened by 20%, unit directions."""
    rng = np.random.default_rng(seed)
    span = hi - lo
    o = lo - 0.2 * span + 1.4 * span * rng.random((n, 3))
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays = np.zeros((n, 8), np.float32)
    rays[:, 0:3] = o
    rays[:, 4:7] = d
    rays[:, 7] = t_max
    return rays


def world_triangles(blob: L.Blob):
    """Every placed triangle in world space (float64): (n, 3, 3), instance of each."""
    a = blob.arrays()
    meshes, placements = blob.instancing()
    pos = a["positions"][:, :3].astype(np.float64)
    tris = a["triangles"]
    out, owner = [], []
    for i, pl in enumerate(placements):
        m = meshes[pl["mesh"]]
        t = tris[int(m["first_triangle"]):int(m["first_triangle"]) + int(m["triangle_count"])]
        M = pl["object_to_world"].astype(np.float64).reshape(3, 4)
        v = np.stack([pos[t["i0"]], pos[t["i1"]], pos[t["i2"]]], axis=1)  # (k, 3, 3)
        out.append(v @ M[:, :3].T + M[:, 3])
        owner.append(np.full(len(t), i))
    return np.concatenate(out), np.concatenate(owner)


def brute_force_closest(v, rays, t_min: float = 1.0e-4):
    """Double-sided Moller-Trumbore in float64 over all world triangles."""
    p0, e1, e2 = v[:, 0], v[:, 1] - v[:, 0], v[:, 2] - v[:, 0]
    ts = np.full(len(rays), np.inf)
    idx = np.full(len(rays), -1)
    for r in range(len(rays)):
        o = rays[r, 0:3].astype(np.float64)
        d = rays[r, 4:7].astype(np.float64)
        pv = np.cross(d, e2)
        det = np.einsum("ij,ij->i", e1, pv)
        ok = np.abs(det) > 1e-12
        inv = np.where(ok, 1.0 / np.where(ok, det, 1.0), 0.0)
        tv = o - p0
        u = np.einsum("ij,ij->i", tv, pv) * inv
        qv = np.cross(tv, e1)
        vv = (qv @ d) * inv
        t = np.einsum("ij,ij->i", e2, qv) * inv
        hit = ok & (u >= 0) & (u <= 1) & (vv >= 0) & (u + vv <= 1) & (t >= t_min) & (t <= rays[r, 7])
        if hit.any():
            k = np.argmin(np.where(hit, t, np.inf))
            ts[r], idx[r] = t[k], k
    return ts, idx
