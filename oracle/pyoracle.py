"""ctypes wrapper of the CPU oracle (liblucida_oracle.so).

TEST INFRASTRUCTURE. The oracle restates the reference's Rust CPU twin
(crates/core) in C++; it is the checker the HIP path is compared with. Only the
test suite, the smoke check and the bench's CPU-baseline leg import this
module; the product never does.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "liblucida_oracle.so"

AABB = np.dtype([("lo", "<f4", 3), ("_p0", "<f4"), ("hi", "<f4", 3), ("_p1", "<f4")])
TRIANGLE = np.dtype([("i0", "<u4"), ("i1", "<u4"), ("i2", "<u4"), ("material", "<u4")])
BVH_NODE = np.dtype([("bounds_min", "<f4", 3), ("left_first", "<u4"), ("bounds_max", "<f4", 3), ("count", "<u4")])

_lib = None


def build() -> Path:
    """Builds the oracle with its Makefile (g++, contraction off)."""
    subprocess.run(["make", "-s", "-C", str(HERE), "-j8"], check=True)
    return LIB_PATH


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            build()
        L = C.CDLL(str(LIB_PATH))
        P, S, U32, F, I = C.c_void_p, C.c_size_t, C.c_uint32, C.c_float, C.c_int
        sigs = {
            "oracle_expand_bits": ([U32], U32),
            "oracle_morton3d": ([F, F, F], U32),
            "oracle_normalise": ([P, P, P, P], None),
            "oracle_lcp": ([P, S, I, I], I),
            "oracle_triangle_bounds": ([P, S, P, P], None),
            "oracle_lbvh_build": ([P, S, P, S, P, P, P, P, P, P, P, P, P, P], S),
            "oracle_sah_cost": ([P, S], F),
            "oracle_validate": ([P, S, P, S, S], S),
            "oracle_sort_pairs": ([P, P, S, P, P], None),
            "oracle_treelet_build": ([P, S, P, U32, P, P, P], S),
            "oracle_treelet_build_with": ([P, S, P, U32, F, F, U32, U32, P, P, P], S),
            "oracle_treelet_dp": ([I, P, P, P, P, P, P], None),
            "oracle_sobol": ([U32, U32], U32),
            "oracle_owen_scramble": ([U32, U32], U32),
            "oracle_hash_u32": ([U32], U32),
            "oracle_sobol_sample": ([U32, U32, U32], F),
            "oracle_sampler_draws": ([U32, U32, U32, U32, P, S, P], None),
            "oracle_ggx_table": ([P], None),
            "oracle_directional_albedo_grid": ([F, F, U32], F),
            "oracle_offset_ray_origin": ([P, P, P], None),
            "oracle_render": ([P, P, P, U32, U32, U32, P, P, P], None),
            "oracle_denoise": ([P, P, P, P, U32, U32, P, P], None),
            "oracle_compare": ([P, P, S, P], None),
            "oracle_scene_cast": ([P, P, S, I, P, P], None),
            "oracle_bsdf_sample": ([P, U32, P, F, F, F, P], F),
            "oracle_bsdf_eval_pdf": ([P, U32, P, P, P], F),
            "oracle_bsdf_albedo": ([P, F, U32, U32], C.c_double),
            "oracle_fresnel_dielectric": ([F, F], F),
            "oracle_fresnel_conductor": ([F, P, P, P], None),
            "oracle_refract": ([P, P, F, P], I),
            "oracle_bsdf_split_albedo": ([P, U32, P, U32, U32, P], None),
            "oracle_env_build": ([U32, U32, P, P, P], F),
            "oracle_env_pdf": ([U32, U32, P, P, P, F, P, P], F),
            "oracle_env_sample": ([U32, U32, P, P, P, F, F, F, P], F),
            "oracle_env_direction_from_uv": ([F, F, P], None),
            "oracle_env_uv_from_direction": ([P, P], None),
            "oracle_convergence": ([P, S], C.c_double),
            "oracle_render_rows": ([P, P, P, U32, P, U32, P, P], None),
        }
        for name, (args, res) in sigs.items():
            try:
                fn = getattr(L, name)
            except AttributeError:
                continue
            fn.argtypes = args
            fn.restype = res
        _lib = L
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data


def expand_bits(v: int) -> int:
    return lib().oracle_expand_bits(v)


def morton3d(x: float, y: float, z: float) -> int:
    return lib().oracle_morton3d(x, y, z)


def normalise(p, lo, hi):
    p, lo, hi = (np.asarray(v, np.float32) for v in (p, lo, hi))
    out = np.empty(3, np.float32)
    lib().oracle_normalise(_p(p), _p(lo), _p(hi), _p(out))
    return out


def lcp(codes, i: int, j: int) -> int:
    codes = np.ascontiguousarray(codes, np.uint32)
    return lib().oracle_lcp(_p(codes), len(codes), i, j)


def triangle_bounds(tris, positions) -> np.ndarray:
    tris = np.ascontiguousarray(tris, TRIANGLE)
    pos = np.ascontiguousarray(positions, np.float32).reshape(-1, 4)
    out = np.empty(len(tris), AABB)
    lib().oracle_triangle_bounds(_p(tris), len(tris), _p(pos), _p(out))
    return out


def lbvh_build(tris, positions, max_leaf: int = 4) -> dict:
    """Lbvh::build_with_leaf_size (crates/core/src/lbvh.rs:422) with every stage."""
    tris = np.ascontiguousarray(tris, TRIANGLE)
    pos = np.ascontiguousarray(positions, np.float32).reshape(-1, 4)
    n = len(tris)
    r = {
        "codes": np.empty(n, np.uint32), "order": np.empty(n, np.uint32),
        "karras": np.empty((max(n - 1, 1), 2), np.uint32), "parent": np.empty(max(2 * n - 1, 1), np.uint32),
        "node_bounds": np.empty(max(2 * n - 1, 1), AABB), "subtree_size": np.empty(max(n - 1, 1), np.uint32),
        "scene_bounds": np.empty(1, AABB), "nodes": np.empty(max(2 * n - 1, 1), BVH_NODE),
        "prim_indices": np.empty(max(n, 1), np.uint32),
    }
    sah = C.c_float()
    count = lib().oracle_lbvh_build(_p(tris), n, _p(pos), max_leaf, _p(r["codes"]), _p(r["order"]),
                                    _p(r["karras"]), _p(r["parent"]), _p(r["node_bounds"]),
                                    _p(r["subtree_size"]), _p(r["scene_bounds"]), _p(r["nodes"]),
                                    _p(r["prim_indices"]), C.byref(sah))
    r["nodes"] = r["nodes"][:count]
    r["prim_indices"] = r["prim_indices"][:n]
    if n < 2:
        r["karras"] = r["karras"][:0]
        r["subtree_size"] = r["subtree_size"][:0]
    r["sah_cost"] = sah.value
    return r


def treelet_build(tris, positions, rounds: int = 3) -> dict:
    """The LBVH restructured by treelets (oracle/treelet.cpp; no reference
    counterpart): nodes, prim_indices and the SAH cost."""
    tris = np.ascontiguousarray(tris, TRIANGLE)
    pos = np.ascontiguousarray(positions, np.float32).reshape(-1, 4)
    n = len(tris)
    nodes = np.empty(max(2 * n - 1, 1), BVH_NODE)
    prims = np.empty(max(n, 1), np.uint32)
    sah = C.c_float()
    count = lib().oracle_treelet_build(_p(tris), n, _p(pos), rounds, _p(nodes), _p(prims), C.byref(sah))
    return {"nodes": nodes[:count].copy(), "prim_indices": prims[:n].copy(), "sah_cost": sah.value}


def treelet_dp(boxes, costs, sizes) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
    """The treelet subset search over k <= 7 leaves: (least cost, best split's
    left side, leaf choice) per subset mask."""
    boxes = np.ascontiguousarray(boxes, AABB)
    costs = np.ascontiguousarray(costs, np.float32)
    sizes = np.ascontiguousarray(sizes, np.uint32)
    k = len(boxes)
    assert 1 <= k <= 7 and len(costs) == k and len(sizes) == k
    c = np.empty(1 << k, np.float32)
    part = np.empty(1 << k, np.uint8)
    leafy = np.empty(1 << k, np.uint8)
    lib().oracle_treelet_dp(k, _p(boxes), _p(costs), _p(sizes), _p(c), _p(part), _p(leafy))
    return c, part, leafy


def sah_cost(nodes) -> float:
    nodes = np.ascontiguousarray(nodes, BVH_NODE)
    return lib().oracle_sah_cost(_p(nodes), len(nodes))


def validate(nodes, prim_indices, triangle_count: int) -> int:
    nodes = np.ascontiguousarray(nodes, BVH_NODE)
    prims = np.ascontiguousarray(prim_indices, np.uint32)
    return lib().oracle_validate(_p(nodes), len(nodes), _p(prims), len(prims), triangle_count)


def sort_pairs(keys, vals):
    keys = np.ascontiguousarray(keys, np.uint32)
    vals = np.ascontiguousarray(vals, np.uint32)
    ko, vo = np.empty_like(keys), np.empty_like(vals)
    lib().oracle_sort_pairs(_p(keys), _p(vals), len(keys), _p(ko), _p(vo))
    return ko, vo


# --- tracer ------------------------------------------------------------------


class RenderCounts(C.Structure):
    _fields_ = [("extension_rays", C.c_uint64), ("shadow_rays", C.c_uint64), ("node_visits", C.c_uint64),
                ("triangle_tests", C.c_uint64)]


def sobol(index: int, dim: int) -> int:
    return lib().oracle_sobol(index, dim)


def owen_scramble(x: int, seed: int) -> int:
    return lib().oracle_owen_scramble(x, seed)


def hash_u32(x: int) -> int:
    return lib().oracle_hash_u32(x)


def sobol_sample(index: int, dim: int, seed: int) -> float:
    return lib().oracle_sobol_sample(index, dim, seed)


def sampler_draws(kind: int, frame: int, pixel: int, sample: int, pattern) -> np.ndarray:
    """Draws from a Sampler: pattern entries 1 (next_f32) or 2 (next_vec2)."""
    pat = np.ascontiguousarray(pattern, np.uint32)
    out = np.empty(int(pat.sum()), np.float32)
    lib().oracle_sampler_draws(kind, frame, pixel, sample, _p(pat), len(pat), _p(out))
    return out


def ggx_table() -> np.ndarray:
    out = np.empty(1024, np.float32)
    lib().oracle_ggx_table(_p(out))
    return out


def directional_albedo_grid(roughness: float, cos_o: float, grid: int) -> float:
    return lib().oracle_directional_albedo_grid(roughness, cos_o, grid)


def offset_ray_origin(p, n) -> np.ndarray:
    p = np.asarray(p, np.float32)
    n = np.asarray(n, np.float32)
    out = np.empty(3, np.float32)
    lib().oracle_offset_ray_origin(_p(p), _p(n), _p(out))
    return out


def render(desc, uniforms, background=(0.0, 0.0, 0.0), threads: int = 0, rows=None, guides: bool = True):
    """integrator::render_with_guides over rows [r0, r1): (rgb (h,w,3), guides (h,w,8), counts)."""
    r0, r1 = rows if rows is not None else (0, uniforms.height)
    h = r1 - r0
    rgb = np.zeros((h, uniforms.width, 3), np.float32)
    g = np.zeros((h, uniforms.width, 8), np.float32) if guides else None
    bg = np.asarray(background, np.float32)
    counts = RenderCounts()
    lib().oracle_render(C.byref(desc), C.byref(uniforms), _p(bg), threads, r0, r1, _p(rgb), _p(g), C.byref(counts))
    return rgb, g, counts


def render_rows(desc, uniforms, rows, background=(0.0, 0.0, 0.0), threads: int = 0):
    """The render restricted to the listed image rows, one row per work item
    of the thread pool: (rgb (len(rows), w, 3), counts)."""
    rows = np.ascontiguousarray(rows, np.uint32)
    rgb = np.zeros((len(rows), uniforms.width, 3), np.float32)
    bg = np.asarray(background, np.float32)
    counts = RenderCounts()
    lib().oracle_render_rows(C.byref(desc), C.byref(uniforms), _p(bg), threads, _p(rows), len(rows), _p(rgb),
                             C.byref(counts))
    return rgb, counts


def denoise(film, albedo, normal, depth, params) -> np.ndarray:
    """denoise::denoise (crates/core/src/denoise.rs:134)."""
    h, w, _ = film.shape
    f, a, n = (np.ascontiguousarray(x, np.float32) for x in (film, albedo, normal))
    d = np.ascontiguousarray(depth, np.float32)
    out = np.empty_like(f)
    lib().oracle_denoise(_p(f), _p(a), _p(n), _p(d), w, h, C.byref(params), _p(out))
    return out


def scene_cast(desc, rays, any_hit: bool = False):
    """Scene::intersect / occluded per ray without shading: rays (n, 8) float32
    (origin, unused, direction, t_max) -> hits (n, 4) float32 (t, u, v, id
    bits) and the instance of each triangle hit (instanced scenes)."""
    rays = np.ascontiguousarray(rays, np.float32)
    n = len(rays)
    hits = np.zeros((n, 4), np.float32)
    inst = np.zeros(n, np.uint32)
    lib().oracle_scene_cast(C.byref(desc), _p(rays), n, 1 if any_hit else 0, _p(hits), _p(inst))
    return hits, inst


def compare(a, b) -> dict:
    """image::compare (crates/core/src/image.rs:178) plus the mean |a-b| per channel."""
    a = np.ascontiguousarray(a, np.float32)
    b = np.ascontiguousarray(b, np.float32)
    assert a.shape == b.shape
    out = np.empty(6, np.float64)
    lib().oracle_compare(_p(a), _p(b), a.size // 3, _p(out))
    return dict(zip(("rmse", "max_abs", "mean_rel", "mean_a", "mean_b", "mean_abs"), out.tolist()))


def bsdf_sample(material: np.ndarray, front: bool, wo, u_lobe: float, ux: float, uy: float):
    """surface::sample (bsdf.rs:1571): (pdf, wi, weight)."""
    m = np.ascontiguousarray(material)
    wo = np.asarray(wo, np.float32)
    out = np.empty(6, np.float32)
    pdf = lib().oracle_bsdf_sample(_p(m), int(front), _p(wo), u_lobe, ux, uy, _p(out))
    return pdf, out[:3], out[3:]


def bsdf_eval_pdf(material: np.ndarray, front: bool, wo, wi):
    """surface::eval and surface::pdf (bsdf.rs:1493/1543): (pdf, f)."""
    m = np.ascontiguousarray(material)
    wo = np.asarray(wo, np.float32)
    wi = np.asarray(wi, np.float32)
    f = np.empty(3, np.float32)
    pdf = lib().oracle_bsdf_eval_pdf(_p(m), int(front), _p(wo), _p(wi), _p(f))
    return pdf, f


def bsdf_albedo(material: np.ndarray, cos_o: float, n: int = 200_000, seed: int = 1) -> float:
    m = np.ascontiguousarray(material)
    return lib().oracle_bsdf_albedo(_p(m), cos_o, n, seed)


# --- closed-form pins: Fresnel, refraction, split albedo ------------------------


def fresnel_dielectric(cos_i: float, eta: float) -> float:
    """fresnel_dielectric (crates/core/src/bsdf.rs:314)."""
    return lib().oracle_fresnel_dielectric(cos_i, eta)


def fresnel_conductor(cos_i: float, eta, k) -> np.ndarray:
    """fresnel_conductor (crates/core/src/bsdf.rs:358), per RGB channel."""
    eta = np.asarray(eta, np.float32)
    k = np.asarray(k, np.float32)
    out = np.empty(3, np.float32)
    lib().oracle_fresnel_conductor(cos_i, _p(eta), _p(k), _p(out))
    return out


def refract(wo, n, eta: float):
    """refract (crates/core/src/bsdf.rs:555): the transmitted direction or None (TIR)."""
    wo = np.asarray(wo, np.float32)
    n = np.asarray(n, np.float32)
    out = np.empty(3, np.float32)
    return out if lib().oracle_refract(_p(wo), _p(n), eta, _p(out)) else None


def bsdf_split_albedo(material: np.ndarray, front: bool, wo, n: int, seed: int) -> tuple[float, float]:
    """(reflected, transmitted) directional albedo: the mean luminance weight of
    surface::sample split by the side of the surface wi lands on."""
    m = np.ascontiguousarray(material)
    wo = np.asarray(wo, np.float32)
    out = np.empty(2, np.float64)
    lib().oracle_bsdf_split_albedo(_p(m), int(front), _p(wo), n, seed, _p(out))
    return float(out[0]), float(out[1])


# --- environment map (envmap.rs) ----------------------------------------------------


class EnvMap:
    """EnvMap (crates/core/src/envmap.rs:61) over RGBA32F pixels (h, w, 4), with
    the oracle's own CDF build (EnvMap::new, envmap.rs:112)."""

    def __init__(self, pixels: np.ndarray):
        self.pixels = np.ascontiguousarray(pixels, np.float32)
        self.height, self.width, _ = self.pixels.shape
        self.conditional = np.zeros((self.width + 1) * self.height, np.float32)
        self.marginal = np.zeros(self.height + 1, np.float32)
        self.total_weight = lib().oracle_env_build(self.width, self.height, _p(self.pixels), _p(self.conditional),
                                                   _p(self.marginal))

    def _args(self):
        return (self.width, self.height, _p(self.pixels), _p(self.conditional), _p(self.marginal), self.total_weight)

    def pdf(self, direction) -> float:
        d = np.asarray(direction, np.float32)
        return lib().oracle_env_pdf(*self._args(), _p(d), None)

    def radiance(self, direction) -> np.ndarray:
        d = np.asarray(direction, np.float32)
        out = np.empty(3, np.float32)
        lib().oracle_env_pdf(*self._args(), _p(d), _p(out))
        return out

    def sample(self, u: float, v: float):
        """EnvMap::sample (envmap.rs:235): (pdf, direction, radiance)."""
        out = np.empty(6, np.float32)
        pdf = lib().oracle_env_sample(*self._args(), u, v, _p(out))
        return pdf, out[:3], out[3:]

    @staticmethod
    def direction_from_uv(u: float, v: float) -> np.ndarray:
        out = np.empty(3, np.float32)
        lib().oracle_env_direction_from_uv(u, v, _p(out))
        return out

    @staticmethod
    def uv_from_direction(d) -> np.ndarray:
        d = np.asarray(d, np.float32)
        out = np.empty(2, np.float32)
        lib().oracle_env_uv_from_direction(_p(d), _p(out))
        return out


def convergence(accum: np.ndarray) -> float:
    """The convergence meter (shaders/stats/convergence.wgsl) over GpuAccum
    records (a structured array or a float array of 16 per pixel)."""
    a = np.ascontiguousarray(accum).view(np.float32).reshape(-1, 16)
    a = np.ascontiguousarray(a)
    return lib().oracle_convergence(_p(a), len(a))
