"""The reference's built-in scenes (crates/core/src/scenes.rs:56-815), as the
library generates them, against the packed assets the reference ships
(web/public/scenes/*.bin via the fixtures tests/golden/make_golden.py writes).

Materials, analytic primitives and lights compare bit for bit; so do the
environment maps procedural_sky writes (envmap.rs:343) and their CDFs
(EnvMap::new, envmap.rs:112, packed as codegen.rs:448-460 packs them); the
cameras equal the records in web/src/generated/scenes.ts. The instanced
scene's 49 placements are checked against the reference's own world-to-object
records: each is the inverse of exactly one placement. bvh-stress ships no
asset (the reference fetches it remotely), so its lattice is checked against
its definition. No GPU is used.
"""
from __future__ import annotations

import hashlib
import json

import numpy as np
import pytest

import lucida_amd as L
from conftest import GOLDEN

PACKED = ["furnace-test", "metal-sweep", "mis-scene", "glass-box", "sunset", "instance-forest"]


def fixture(name: str) -> dict:
    return json.loads((GOLDEN / f"{name.replace('-', '_')}_sections.json").read_text())


@pytest.mark.parametrize("name", PACKED)
def test_scene_matches_reference_asset(name):
    ref = fixture(name)
    b = L.Blob(name)
    a = b.arrays()
    u32 = lambda arr: arr.view(np.uint32).ravel().tolist()
    assert u32(a["materials"]) == ref["materials"]
    n_analytic = ref["instancing"]["analytic_primitives"]
    assert len(a["primitives"]) == n_analytic
    assert u32(a["primitives"]) == ref["primitives"][:16 * n_analytic]
    assert u32(a["lights"]) == ref["lights"]
    assert a["env_width"] == ref["env"]["width"] and a["env_height"] == ref["env"]["height"]
    u = b.uniforms(64, 64)
    assert list(u.background) == list(np.float32(ref["background"]))
    cam = b.camera()
    rc = ref["camera"]
    for key in ("eye", "look_at", "up"):
        assert list(getattr(cam, key)) == list(np.float32(rc[key])), key
    assert cam.vfov_deg == rc["vfov_deg"] and cam.aperture == rc["aperture"]
    # focus distance = |look_at - eye| in f32; scenes.ts prints its shortest
    # round-trip decimal, so the f32 of that decimal is the same number
    assert cam.focus_distance == np.float32(rc["focus_distance"])


@pytest.mark.parametrize("name", ["sunset", "instance-forest"])
def test_procedural_sky_and_cdf_match_reference_sections(name):
    e = fixture(name)["env_sections"]
    a = L.Blob(name).arrays()
    w, h = a["env_width"], a["env_height"]
    px = a["env_pixels"].reshape(h, w, 4)
    assert hashlib.sha256(px.tobytes()).hexdigest() == e["sha256_envRadiance"]
    cdf = np.zeros((h + 1, w + 1), np.float32)
    cdf[:h] = a["env_conditional"].reshape(h, w + 1)
    cdf[h, :h + 1] = a["env_marginal"]
    assert hashlib.sha256(cdf.tobytes()).hexdigest() == e["sha256_envCdf"]
    # the reductions the fixture also keeps, as a readable second statement
    lum = px[..., :3].astype(np.float64) @ np.array([0.2126, 0.7152, 0.0722])
    assert np.flatnonzero(lum.ravel() > 50.0).tolist() == e["sun_texels"]
    assert lum.sum(axis=1).tolist() == e["row_luminance_sums"]
    assert a["env_marginal"].tolist() == np.float32(e["marginal"]).tolist()


def test_instance_forest_mesh_matches_reference_asset():
    ref = fixture("instance-forest")
    a = L.Blob("instance-forest").arrays()
    sha = lambda arr: hashlib.sha256(np.ascontiguousarray(arr).tobytes()).hexdigest()
    assert len(a["triangles"]) == ref["counts"]["triangles"] == 320
    assert sha(a["positions"]) == ref["sha256_positions"]
    va = a["vertex_attrs"].copy()
    va["uv"] = 0
    assert sha(va) == ref["sha256_vertexAttrs_no_uv"]
    tris = a["triangles"].view(np.uint32).reshape(-1, 4)
    # the reference stores the BLAS's triangles in its leaf order
    assert sha(tris[np.lexsort(tris.T[::-1])]) == ref["sha256_triangles_sorted"]


def test_instance_forest_placements_invert_the_reference_instances():
    """Every GpuInstance the reference packed (world-to-object rows + inverse
    translation, in its TLAS leaf order) is the inverse of exactly one of the
    library's placements (object-to-world from the same PCG stream), with the
    same material override."""
    ref = fixture("instance-forest")
    recs = np.array(ref["primitives"][16:], np.uint32).reshape(-1, 16)
    assert len(recs) == ref["instancing"]["count"] == 49
    f = recs.view(np.float32)
    inv_r = np.stack([f[:, 0:3], f[:, 4:7], f[:, 8:11]], axis=1).astype(np.float64)  # (49, 3, 3)
    inv_t = f[:, 12:15].astype(np.float64)
    kinds, mats = recs[:, 7], recs[:, 11]
    assert np.all(kinds == 2)
    _, pl = L.Blob("instance-forest").instancing()
    assert len(pl) == 49
    matched = set()
    for p in pl:
        m = p["object_to_world"].astype(np.float64).reshape(3, 4)
        # world-to-object applied to object-to-world must give the identity
        err = np.abs(np.einsum("kij,jl->kil", inv_r, m[:, :3]) - np.eye(3)).max(axis=(1, 2))
        err_t = np.abs(np.einsum("kij,j->ki", inv_r, m[:, 3]) + inv_t).max(axis=1)
        k = int(np.argmin(err + err_t))
        assert err[k] < 2e-6 and err_t[k] < 2e-5, (err[k], err_t[k])
        assert k not in matched
        matched.add(k)
        assert int(mats[k]) == int(p["material"])
    assert len(matched) == 49


def test_bvh_stress_lattice():
    """72 level-4 icospheres of radius 30 on the 6 x 2 x 6 lattice of
    scenes.rs:232-255 inside the Cornell box's six quads (368,640 triangles,
    the reference's published size)."""
    b = L.Blob("bvh-stress")
    a = b.arrays()
    assert len(a["triangles"]) == 72 * 5120 == 368_640
    assert len(a["primitives"]) == 6 and np.all(a["primitives"]["kind"] == 1)
    assert len(a["lights"]) == 1
    box = L.Blob("cornell-box").arrays()
    assert a["materials"].tobytes() == box["materials"].tobytes()
    pos = a["positions"][:, :3].reshape(72, -1, 3).astype(np.float64)
    S = 555.0
    k = 0
    for ix in range(6):
        for iy in range(2):
            for iz in range(6):
                f = lambda i, n: np.float32(np.float32(i + 1.0) / np.float32(n + 1.0)) * np.float32(S)
                c = np.array([f(ix, 6), np.float32(f(iy, 2) * np.float32(0.62)) + np.float32(95.0), f(iz, 6)])
                r = np.linalg.norm(pos[k] - c, axis=1)
                assert np.abs(r - 30.0).max() < 1e-3
                mats = a["triangles"]["material"][k * 5120:(k + 1) * 5120]
                assert np.all(mats == (4 if (ix + iy + iz) % 2 == 0 else 5))
                k += 1
