"""The oracle's tracer, pinned on the CPU before it is trusted as a checker.

Known answers and invariants come from the reference's own tests: sobol.rs
:363-600 (bit reversal, (0,2)-sequence, bijective and stratification-preserving
scramble), rng.rs:88-140 (uniformity, decorrelation), bsdf.rs (white furnace,
sample/pdf consistency), the shipped GGX energy table
(shaders/common/ggx_energy.wgsl), denoise.rs:230-432, and integrator
invariants (furnace, progressive == one-shot, sampling modes agree).
"""
from __future__ import annotations

import ctypes as C
import json

import numpy as np
import pytest

import lucida_amd as L
from conftest import GOLDEN


# --- Sobol and RNG ---------------------------------------------------------------

def test_sobol_direction_numbers_match_shipped_table(oracle):
    v = json.loads((GOLDEN / "sobol_v.json").read_text())
    for d in range(2):
        for k in range(32):
            assert oracle.sobol(1 << k, d) == v[d][k]


def test_sobol_first_dimension_is_bit_reversal(oracle):
    for i in range(1000):
        assert oracle.sobol(i, 0) == int(f"{i:032b}"[::-1], 2)
    assert oracle.sobol(0, 0) == 0 and oracle.sobol(0, 1) == 0


def t_value(oracle, max_k):
    pts = np.array([[oracle.sobol(i, 0), oracle.sobol(i, 1)] for i in range(1 << max_k)], np.uint64)
    for k in range(1, max_k + 1):
        n = 1 << k
        for split in range(k + 1):
            a = pts[:n, 0] >> np.uint64(32 - split) if split else np.zeros(n, np.uint64)
            b = pts[:n, 1] >> np.uint64(32 - (k - split)) if k - split else np.zeros(n, np.uint64)
            cells = (a << np.uint64(k - split)) | b
            if len(np.unique(cells)) != n:
                return False
    return True


def test_first_two_dimensions_are_a_02_sequence(oracle):
    assert t_value(oracle, 10)


def test_owen_scramble_is_a_bijection_and_keeps_strata(oracle):
    for seed in (0, 1, 0xDEADBEEF, 12345):
        ys = {oracle.owen_scramble(x << 16, seed) >> 16 for x in range(1 << 12)}
        assert len(ys) == 1 << 12
    for seed in (1, 7, 0xABCD):
        for k in range(1, 9):
            n = 1 << k
            cells = set()
            for i in range(n):
                sh = oracle.owen_scramble(i, oracle.hash_u32(seed))
                v = oracle.owen_scramble(oracle.sobol(sh, 0), oracle.hash_u32(seed + 1))
                cells.add(v >> (32 - k))
            assert len(cells) == n


def test_samples_in_half_open_interval(oracle):
    for seed in range(0, 64, 7):
        for i in range(0, 256, 5):
            for d in range(12):
                v = oracle.sobol_sample(i, d, seed)
                assert 0.0 <= v < 1.0


def test_sobol_converges_faster_than_independent(oracle):
    exact = (1 - np.exp(-1.0)) ** 2
    ratios = []
    for k in (6, 10):
        n = 1 << k
        es, er = 0.0, 0.0
        for t in range(16):
            s = sum(np.exp(-oracle.sobol_sample(i, 0, t) - oracle.sobol_sample(i, 1, t)) for i in range(n))
            es += abs(s / n - exact)
            r = oracle.sampler_draws(0, t, 0, 0, [2] * n).reshape(-1, 2)
            er += abs(np.exp(-r[:, 0] - r[:, 1]).mean() - exact)
        ratios.append(er / es)
    assert ratios[1] > ratios[0] and ratios[1] > 8


def test_pcg_stream_is_uniform_and_decorrelated(oracle):
    u = oracle.sampler_draws(0, 12345, 7, 0, [1] * 200_000)
    assert np.all((u >= 0) & (u < 1))
    assert abs(u.mean() - 0.5) < 3e-3 and abs(u.var() - 1 / 12) < 2e-3
    a = np.array([oracle.sampler_draws(0, 0, i, 0, [1])[0] for i in range(20_000)]) - 0.5
    b = np.array([oracle.sampler_draws(0, 0, i + 1, 0, [1])[0] for i in range(20_000)]) - 0.5
    assert abs((a * b).mean()) < 5 * (1 / 12) / np.sqrt(20_000)


def test_sobol_vec2_aligns_to_even_dimensions(oracle):
    # next_f32 uses dim 0; next_vec2 then starts at dim 2 (sobol.rs:349-360).
    seed = oracle.hash_u32(5 + 9 * 0x9E3779B9 & 0xFFFFFFFF)
    d = oracle.sampler_draws(1, 9, 5, 3, [1, 2])
    assert d[0] == np.float32(oracle.sobol_sample(3, 0, seed))
    assert d[1] == np.float32(oracle.sobol_sample(3, 2, seed))
    assert d[2] == np.float32(oracle.sobol_sample(3, 3, seed))


# --- BSDF -------------------------------------------------------------------------------

def test_ggx_energy_table_matches_shipped_table(oracle):
    ref = np.array(json.loads((GOLDEN / "ggx_energy_table.json").read_text()), np.float32)
    # the shipped table is printed to 6 decimals
    assert np.abs(oracle.ggx_table() - ref).max() <= 5.5e-7


def material(base=(0.8, 0.8, 0.8), roughness=0.5, metallic=0.0, ior=1.5, transmission=0.0):
    m = np.zeros(1, L.MATERIAL)
    m["base_color"] = base
    m["roughness"] = roughness
    m["metallic"] = metallic
    m["ior"] = ior
    m["transmission"] = transmission
    return m


@pytest.mark.parametrize("roughness", [0.2, 0.5, 1.0])
@pytest.mark.parametrize("cos_o", [0.9, 0.5, 0.2])
def test_white_furnace_metal(oracle, roughness, cos_o):
    """A white metal with energy compensation reflects everything (bsdf.rs white_furnace_test)."""
    e = oracle.bsdf_albedo(material((1, 1, 1), roughness, metallic=1.0), cos_o, 400_000)
    assert abs(e - 1.0) < 0.02, e


@pytest.mark.parametrize("roughness", [0.1, 0.6])
def test_white_layered_dielectric_conserves_energy(oracle, roughness):
    e = oracle.bsdf_albedo(material((1, 1, 1), roughness), 0.7, 400_000)
    assert abs(e - 1.0) < 0.03, e


def test_sample_reports_its_own_pdf_and_weight(oracle):
    rng = np.random.default_rng(3)
    for m in (material((0.7, 0.4, 0.2), 0.3), material((0.9, 0.9, 0.9), 0.05, metallic=1.0),
              material((1, 1, 1), 0.2, ior=1.5, transmission=1.0)):
        checked = 0
        for _ in range(400):
            wo = rng.normal(size=3)
            wo[2] = abs(wo[2]) + 0.05
            wo = (wo / np.linalg.norm(wo)).astype(np.float32)
            pdf, wi, w = oracle.bsdf_sample(m, True, wo, *rng.random(3))
            if pdf <= 0:
                continue
            pdf2, f = oracle.bsdf_eval_pdf(m, True, wo, wi)
            assert abs(pdf2 - pdf) <= 1e-5 * pdf
            np.testing.assert_allclose(w, f * abs(wi[2]) / pdf, rtol=1e-5, atol=1e-7)
            checked += 1
        assert checked > 300


def test_offset_ray_origin_moves_off_the_surface(oracle):
    rng = np.random.default_rng(1)
    for _ in range(2000):
        p = (rng.normal(size=3) * rng.choice([1e-3, 1.0, 100.0])).astype(np.float32)
        n = rng.normal(size=3)
        n = (n / np.linalg.norm(n)).astype(np.float32)
        o = oracle.offset_ray_origin(p, n)
        d = o.astype(np.float64) - p
        assert np.dot(d, n) >= 0
        assert np.linalg.norm(d) < 1e-4 * max(1.0, np.abs(p).max())


# --- integrator ------------------------------------------------------------------------------


class Desc:
    """A lucida_scene_desc over numpy arrays (kept alive with it)."""

    def __init__(self, materials, primitives, lights=None):
        self.materials = np.ascontiguousarray(materials, L.MATERIAL)
        self.primitives = np.ascontiguousarray(primitives, L.PRIMITIVE)
        self.lights = np.zeros(0, L.LIGHT) if lights is None else np.ascontiguousarray(lights, L.LIGHT)
        self.dummy = np.zeros(4, np.float32)
        d = L.SceneDesc()
        d.materials = self.materials.ctypes.data
        d.num_materials = len(self.materials)
        d.primitives = self.primitives.ctypes.data
        d.num_primitives = len(self.primitives)
        d.positions = self.dummy.ctypes.data
        d.lights = self.lights.ctypes.data if len(self.lights) else None
        d.num_lights = len(self.lights)
        self.desc = d


def camera_uniforms(width, height, spp, depth, eye, look, vfov, background, sampler=0):
    """Camera::write_uniforms for a plain look-at camera (reused from the
    library's cornell-box uniforms with the camera replaced)."""
    u = L.Blob("cornell-box").uniforms(width, height, samples=spp, max_depth=depth, sampler=sampler)
    eye, look = np.array(eye, np.float32), np.array(look, np.float32)
    f = (look - eye) / np.linalg.norm(look - eye)
    r = np.cross(f, [0, 1, 0]).astype(np.float32)
    r /= np.linalg.norm(r)
    up = np.cross(r, f)
    dist = np.float32(np.linalg.norm(look - eye))
    hh = np.float32(np.tan(np.radians(vfov) / 2)) * dist
    hw = hh * np.float32(width / height)
    u.cam_origin[:] = eye
    u.cam_upper_left[:] = eye + f * dist - r * hw + up * hh
    u.cam_horizontal[:] = r * (2 * hw)
    u.cam_vertical[:] = -up * (2 * hh)
    u.focus_distance = dist
    u.num_primitives = 1
    u.num_triangles = u.num_bvh_nodes = u.num_lights = 0
    u.background[:] = background
    return u


def furnace_desc():
    """The reference's furnace_test (scenes.rs:275-310): five white metal
    spheres, roughness 0 to 1, in a uniform environment of radiance 1."""
    mats = np.concatenate([material((1, 1, 1), i / 4, metallic=1.0) for i in range(5)])
    prim = np.zeros(5, L.PRIMITIVE)
    prim["position"][:, 0] = (np.arange(5) - 2.0) * 2.4
    prim["radius"] = 1.0
    prim["material"] = np.arange(5)
    return Desc(mats, prim)


def test_furnace_spheres_are_invisible(oracle):
    """Energy-compensated GGX renders the furnace as a flat field (the
    reference measures mean radiance 0.99994 on the CPU)."""
    d = furnace_desc()
    u = camera_uniforms(64, 32, 32, 16, (0, 0.8, -12), (0, 0, 0), 45, (1, 1, 1))
    u.num_primitives = 5
    rgb, _, _ = oracle.render(d.desc, u, background=(1, 1, 1), guides=False)
    assert abs(rgb.mean() - 1.0) < 2e-3, rgb.mean()


def test_render_is_deterministic_and_progressive(oracle):
    b = L.Blob("box-c1")
    a = b.arrays()
    r = oracle.lbvh_build(a["triangles"], a["positions"])
    b.set_bvh(r["nodes"], a["triangles"][r["prim_indices"]])
    u16 = b.uniforms(32, 32, samples=16, max_depth=6)
    one, _, _ = oracle.render(b.desc(), u16, guides=False)
    again, _, _ = oracle.render(b.desc(), u16, guides=False)
    assert one.tobytes() == again.tobytes()
    a8, _, _ = oracle.render(b.desc(), b.uniforms(32, 32, samples=8, max_depth=6), guides=False)
    b8, _, _ = oracle.render(b.desc(), b.uniforms(32, 32, samples=8, max_depth=6, sample_offset=8), guides=False)
    assert np.abs((a8 + b8) / 2 - one).max() < 1e-5


def test_sampling_modes_converge_together(oracle):
    b = L.Blob("box-c1")
    a = b.arrays()
    r = oracle.lbvh_build(a["triangles"], a["positions"])
    b.set_bvh(r["nodes"], a["triangles"][r["prim_indices"]])
    means = []
    for mode in (L.SAMPLING_BSDF, L.SAMPLING_NEE, L.SAMPLING_MIS):
        u = b.uniforms(32, 32, samples=256, max_depth=8, sampling=mode)
        rgb, _, _ = oracle.render(b.desc(), u, guides=False)
        means.append(rgb.mean())
    assert abs(means[0] / means[2] - 1) < 0.03 and abs(means[1] / means[2] - 1) < 0.03, means


# --- denoiser ---------------------------------------------------------------------------------

def run_denoise(oracle, film, albedo, normal, depth, iterations=4):
    p = L.DenoiseParams.default()
    p.iterations = iterations
    return oracle.denoise(film, albedo, normal, depth, p)


def test_denoise_constant_image_unchanged(oracle):
    h = w = 32
    film = np.tile(np.array([0.4, 0.5, 0.6], np.float32), (h, w, 1))
    out = run_denoise(oracle, film, np.full((h, w, 3), 0.7, np.float32), np.tile([0, 0, 1.0], (h, w, 1)),
                      np.full((h, w), 5.0))
    assert np.abs(out - film).max() < 1e-4


def test_denoise_reduces_noise(oracle):
    h = w = 64
    rng = np.random.default_rng(7)
    film = (0.5 * (0.4 + 1.2 * rng.random((h, w, 1)))).repeat(3, axis=2).astype(np.float32)
    out = run_denoise(oracle, film, np.ones((h, w, 3)), np.tile([0, 0, 1.0], (h, w, 1)), np.full((h, w), 5.0))
    before = np.sqrt(((film - 0.5) ** 2).sum(axis=2).mean())
    after = np.sqrt(((out - 0.5) ** 2).sum(axis=2).mean())
    assert after * 4 < before


@pytest.mark.parametrize("edge", ["normal", "depth"])
def test_denoise_holds_edges(oracle, edge):
    h = w = 64
    left = np.zeros((h, w), bool)
    left[:, : w // 2] = True
    film = np.where(left[..., None], 0.9, 0.1).repeat(3, axis=2).astype(np.float32)
    normal = np.tile([0, 0, 1.0], (h, w, 1))
    depth = np.full((h, w), 5.0)
    if edge == "normal":
        normal = np.where(left[..., None], [1.0, 0, 0], [0, 1.0, 0])
    else:
        depth = np.where(left, 2.0, 40.0)
    out = run_denoise(oracle, film, np.ones((h, w, 3)), normal, depth)
    assert out[h // 2, w // 2 - 1, 0] > 0.85 and out[h // 2, w // 2, 0] < 0.15


def test_denoise_zero_iterations_is_identity(oracle):
    h = w = 16
    rng = np.random.default_rng(2)
    film = rng.random((h, w, 3)).astype(np.float32)
    alb = (0.2 + rng.random((h, w, 3))).astype(np.float32)
    out = run_denoise(oracle, film, alb, np.tile([0, 0, 1.0], (h, w, 1)), np.full((h, w), 1.0), iterations=0)
    np.testing.assert_allclose(out, film, rtol=1e-6)


def test_render_rows_equals_the_banded_render(oracle):
    """The row-list render (the CPU baseline's sampler) traces the same
    pixels as the banded render: rows in any order, one per work item."""
    b = L.Blob("box-c1")
    u = b.uniforms(48, 40, samples=4, max_depth=6)
    full, _, _ = oracle.render(b.desc(), u, guides=False)
    rows = np.array([39, 0, 17, 18, 5], np.uint32)
    part, _ = oracle.render_rows(b.desc(), u, rows, threads=3)
    assert np.array_equal(part, full[rows])
