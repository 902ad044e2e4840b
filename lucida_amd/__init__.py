"""lucida_amd — MI355X-native hot path of Lucida (GPU LBVH, wavefront path
tracer, a-trous denoiser) behind a C ABI, with a thin Python mirror of the
reference's ``pt-gpu`` harness API (crates/gpu/src/lib.rs, wavefront.rs,
lbvh.rs, denoise.rs).

The compute runs in hand-written HIP kernels inside ``liblucida.so``; this
module only marshals numpy arrays across the C ABI. There is no CPU fallback:
if the library is missing, or no GPU is visible, calls raise.
"""
from __future__ import annotations

import ctypes as C
import os
import weakref
from pathlib import Path

import numpy as np

_HERE = Path(__file__).resolve().parent
LIB_PATH = _HERE / "liblucida.so"

# ---------------------------------------------------------------------------
# Layout (include/lucida.h, mirroring crates/core/src/gpu_layout.rs)
# ---------------------------------------------------------------------------

MATERIAL = np.dtype([("base_color", "<f4", 3), ("metallic", "<f4"), ("emissive", "<f4", 3),
                     ("roughness", "<f4"), ("ior", "<f4"), ("transmission", "<f4"), ("_pad0", "<f4"),
                     ("_pad1", "<f4"), ("eta", "<f4", 3), ("_pad2", "<f4"), ("k", "<f4", 3), ("_pad3", "<f4")])
PRIMITIVE = np.dtype([("position", "<f4", 3), ("radius", "<f4"), ("edge_u", "<f4", 3), ("kind", "<u4"),
                      ("edge_v", "<f4", 3), ("material", "<u4"), ("normal", "<f4", 3), ("_pad0", "<f4")])
VERTEX_ATTR = np.dtype([("normal", "<f4", 3), ("_pad0", "<f4"), ("uv", "<f4", 2), ("_pad1", "<f4", 2)])
TRIANGLE = np.dtype([("i0", "<u4"), ("i1", "<u4"), ("i2", "<u4"), ("material", "<u4")])
BVH_NODE = np.dtype([("bounds_min", "<f4", 3), ("left_first", "<u4"), ("bounds_max", "<f4", 3), ("count", "<u4")])
ACCUM = np.dtype([("radiance", "<f4", 3), ("samples", "<f4"), ("albedo", "<f4", 3), ("depth", "<f4"),
                  ("normal", "<f4", 3), ("traversal", "<f4"), ("radiance_sq", "<f4", 3), ("_pad0", "<f4")])
LIGHT = np.dtype([("origin", "<f4", 3), ("area", "<f4"), ("edge_u", "<f4", 3), ("kind", "<u4"),
                  ("edge_v", "<f4", 3), ("material", "<u4"), ("normal", "<f4", 3), ("_pad0", "<f4")])
AABB = np.dtype([("lo", "<f4", 3), ("_p0", "<f4"), ("hi", "<f4", 3), ("_p1", "<f4")])
INSTANCE = np.dtype([("inv_row0", "<f4", 3), ("blas_root", "<u4"), ("inv_row1", "<f4", 3), ("kind", "<u4"),
                     ("inv_row2", "<f4", 3), ("material", "<u4"), ("inv_translation", "<f4", 3), ("_pad0", "<f4")])
MESH_RANGE = np.dtype([("first_triangle", "<u4"), ("triangle_count", "<u4")])
PLACEMENT = np.dtype([("object_to_world", "<f4", 12), ("mesh", "<u4"), ("material", "<u4")])
for _dt, _size in ((MATERIAL, 80), (PRIMITIVE, 64), (VERTEX_ATTR, 32), (TRIANGLE, 16), (BVH_NODE, 32),
                   (ACCUM, 64), (LIGHT, 64), (AABB, 32), (INSTANCE, 64), (MESH_RANGE, 8), (PLACEMENT, 56)):
    assert _dt.itemsize == _size


class Uniforms(C.Structure):
    """GpuUniforms (gpu_layout.rs:395), 160 bytes."""
    _fields_ = [("cam_origin", C.c_float * 3), ("lens_radius", C.c_float), ("cam_upper_left", C.c_float * 3),
                ("focus_distance", C.c_float), ("cam_horizontal", C.c_float * 3), ("_pad1", C.c_float),
                ("cam_vertical", C.c_float * 3), ("_pad2", C.c_float), ("width", C.c_uint32),
                ("height", C.c_uint32), ("sample_offset", C.c_uint32), ("samples_per_launch", C.c_uint32),
                ("max_depth", C.c_uint32), ("frame_seed", C.c_uint32), ("num_primitives", C.c_uint32),
                ("num_lights", C.c_uint32), ("background", C.c_float * 3), ("num_triangles", C.c_uint32),
                ("num_bvh_nodes", C.c_uint32), ("render_mode", C.c_uint32), ("sampling_mode", C.c_uint32),
                ("env_width", C.c_uint32), ("env_height", C.c_uint32), ("env_total_weight", C.c_float),
                ("sampler_kind", C.c_uint32), ("num_instances", C.c_uint32), ("tlas_root", C.c_uint32),
                ("_pad3", C.c_uint32), ("_pad4", C.c_uint32), ("_pad5", C.c_uint32)]


assert C.sizeof(Uniforms) == 160


class SceneDesc(C.Structure):
    """lucida_scene_desc (include/lucida.h)."""
    _fields_ = [("materials", C.c_void_p), ("num_materials", C.c_size_t), ("primitives", C.c_void_p),
                ("num_primitives", C.c_size_t), ("instances", C.c_void_p), ("num_instances", C.c_size_t),
                ("positions", C.c_void_p), ("num_positions", C.c_size_t), ("vertex_attrs", C.c_void_p),
                ("triangles", C.c_void_p), ("num_triangles", C.c_size_t), ("bvh_nodes", C.c_void_p),
                ("num_bvh_nodes", C.c_size_t), ("lights", C.c_void_p), ("num_lights", C.c_size_t),
                ("tlas_root", C.c_uint32), ("env_width", C.c_uint32), ("env_height", C.c_uint32),
                ("env_pixels", C.c_void_p), ("env_conditional", C.c_void_p), ("env_marginal", C.c_void_p),
                ("env_total_weight", C.c_float), ("build_lbvh", C.c_uint32)]


class LbvhParts(C.Structure):
    _fields_ = [("codes", C.c_void_p), ("order", C.c_void_p), ("karras", C.c_void_p), ("parent", C.c_void_p),
                ("node_bounds", C.c_void_p), ("subtree_size", C.c_void_p), ("scene_bounds", C.c_void_p)]


class RenderStats(C.Structure):
    """lucida_render_stats; kernel order GENERATE, EXTEND, SHADE, CONNECT, RESOLVE."""
    _fields_ = [("extension_rays", C.c_uint64), ("shadow_rays", C.c_uint64), ("node_visits", C.c_uint64),
                ("triangle_tests", C.c_uint64), ("extend_node_visits", C.c_uint64),
                ("extend_triangle_tests", C.c_uint64), ("stack_overflows", C.c_uint64), ("elapsed_ms", C.c_float),
                ("kernel_ms", C.c_float * 5), ("kernel_launches", C.c_uint32 * 5),
                ("extend_lane_slots", C.c_uint64), ("connect_lane_slots", C.c_uint64),
                ("queue_sort_ms", C.c_float), ("queue_sorts", C.c_uint32)]

    KERNELS = ("generate", "extend", "shade", "connect", "resolve")


class WavefrontConfig(C.Structure):
    _fields_ = [("samples_per_batch", C.c_uint32), ("max_paths", C.c_uint32), ("bvh_width", C.c_uint32),
                ("use_graph", C.c_uint32), ("count_traversal", C.c_uint32), ("profile_kernels", C.c_uint32),
                ("ray_order", C.c_uint32), ("serialize", C.c_uint32)]

    RAY_ORDER_AUTO, RAY_ORDER_NEVER, RAY_ORDER_ALWAYS = 0, 1, 2


class Job(C.Structure):
    """lucida_job: one render request of a batch."""
    _fields_ = [("scene", C.c_void_p), ("uniforms", Uniforms), ("samples", C.c_uint32), ("film", C.c_void_p)]


class DenoiseParams(C.Structure):
    """DenoiseParams (crates/core/src/denoise.rs:62-94), defaults as there."""
    _fields_ = [("iterations", C.c_uint32), ("sigma_normal", C.c_float), ("sigma_depth", C.c_float),
                ("sigma_luminance", C.c_float), ("sigma_albedo", C.c_float)]

    @classmethod
    def default(cls) -> "DenoiseParams":
        return cls(4, 64.0, 0.08, 4.0, 0.1)


class Camera(C.Structure):
    """lucida_camera: Camera (crates/core/src/camera.rs:14-26)."""
    _fields_ = [("eye", C.c_float * 3), ("look_at", C.c_float * 3), ("up", C.c_float * 3), ("vfov_deg", C.c_float),
                ("aperture", C.c_float), ("focus_distance", C.c_float)]

    def write_uniforms(self, u: "Uniforms", aspect: float) -> "Uniforms":
        """Camera::write_uniforms (camera.rs:55) into the camera fields of u."""
        _check(lib().lucida_camera_write_uniforms(C.byref(self), aspect, C.byref(u)))
        return u


class BlobParams(C.Structure):
    _fields_ = [("seed", C.c_uint32), ("frequency", C.c_uint32), ("displacement", C.c_float)]


SAMPLING_BSDF, SAMPLING_NEE, SAMPLING_MIS = 0, 1, 2
SAMPLER_INDEPENDENT, SAMPLER_SOBOL = 0, 1
# lucida_bvh_build_treelet's default number of restructuring rounds
TREELET_ROUNDS = 3
SORT_NOALLOC_MAX_KEYS = 1 << 24  # LUCIDA_SORT_NOALLOC_MAX_KEYS

# ---------------------------------------------------------------------------
# Library loading
# ---------------------------------------------------------------------------

_lib = None


class GpuError(RuntimeError):
    """Mirrors the reference's GpuError (crates/gpu/src/lib.rs:100-112)."""

    KINDS = {1: "NoAdapter", 2: "NoDevice", 3: "Shader", 4: "Readback", 5: "Validation", 6: "InvalidArg", 7: "Oom"}

    def __init__(self, code: int, msg: str):
        self.code = code
        self.kind = self.KINDS.get(code, f"Status{code}")
        super().__init__(f"{self.kind}: {msg}")


def lib() -> C.CDLL:
    """The loaded liblucida.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(f"{LIB_PATH} is missing: run `python -m lucida_amd.build` (there is no CPU fallback)")
        L = C.CDLL(str(LIB_PATH))
        P, S, U32, I = C.c_void_p, C.c_size_t, C.c_uint32, C.c_int
        sigs = {
            "lucida_last_error": ([], C.c_char_p),
            "lucida_ctx_create": ([I, C.POINTER(P)], I),
            "lucida_ctx_destroy": ([P], None),
            "lucida_ctx_synchronize": ([P], I),
            "lucida_ctx_info": ([P, C.c_char_p, S, C.POINTER(I)], I),
            "lucida_device_alloc": ([P, S, C.POINTER(P)], I),
            "lucida_device_free": ([P, P], I),
            "lucida_copy_to_device": ([P, P, P, S], I),
            "lucida_copy_to_host": ([P, P, P, S], I),
            "lucida_copy_device": ([P, P, P, S], I),
            "lucida_ctx_set_lbvh_fit": ([P, I], I),
            "lucida_sort_u32": ([P, P, P, S, P, P], I),
            "lucida_sort_u32_device": ([P, P, P, P, P, S], I),
            "lucida_sort_u32_device_noalloc": ([P, P, P, P, P, S], I),
            "lucida_lbvh_build_parts": ([P, P, S, P, S, C.POINTER(LbvhParts)], I),
            "lucida_lbvh_build": ([P, P, S, P, S, P, C.POINTER(S), P], I),
            "lucida_bvh_build_treelet": ([P, P, S, P, S, U32, P, C.POINTER(S), P], I),
            "lucida_bvh_build_treelet_device": ([P, P, S, P, U32, P, C.POINTER(S), P, P, P], I),
            "lucida_lbvh_build_device": ([P, P, S, P, P, C.POINTER(S), P, P, P], I),
            "lucida_scene_create": ([P, C.POINTER(SceneDesc), C.POINTER(P)], I),
            "lucida_scene_destroy": ([P], None),
            "lucida_scene_num_bvh_nodes": ([P, C.POINTER(S)], I),
            "lucida_scene_download_bvh": ([P, P, P], I),
            "lucida_scene_create_instanced": ([P, C.POINTER(SceneDesc), P, S, P, S, C.POINTER(P)], I),
            "lucida_scene_counts": ([P, C.POINTER(S), C.POINTER(S), C.POINTER(S), C.POINTER(U32)], I),
            "lucida_scene_download_instances": ([P, P], I),
            "lucida_scene_cast": ([P, P, P, S, I, P, P], I),
            "lucida_render_batch": ([P, P, S, C.POINTER(WavefrontConfig), P, C.POINTER(RenderStats)], I),
            "lucida_blob_instancing": ([P, C.POINTER(P), C.POINTER(S), C.POINTER(P), C.POINTER(S)], I),
            "lucida_film_create": ([P, U32, U32, C.POINTER(P)], I),
            "lucida_film_destroy": ([P], None),
            "lucida_film_clear": ([P], I),
            "lucida_film_download": ([P, P], I),
            "lucida_film_upload": ([P, P], I),
            "lucida_film_device_ptr": ([P], P),
            "lucida_render_wavefront": ([P, P, C.POINTER(Uniforms), U32, P, C.POINTER(WavefrontConfig),
                                         C.POINTER(RenderStats)], I),
            "lucida_denoise_atrous": ([P, P, C.POINTER(DenoiseParams), P], I),
            "lucida_energy_table": ([P, P], I),
            "lucida_net_create": ([P, U32, C.POINTER(P)], I),
            "lucida_net_destroy": ([P], None),
            "lucida_net_param_count": ([], S),
            "lucida_net_get_params": ([P, P, I], I),
            "lucida_net_set_params": ([P, P], I),
            "lucida_net_forward_backward": ([P, P, P, U32, P, P, U32, U32, C.POINTER(C.c_float)], I),
            "lucida_net_adam": ([P, C.c_float], I),
            "lucida_net_train_step": ([P, P, P, U32, P, P, U32, U32, C.c_float, C.POINTER(C.c_float)], I),
            "lucida_net_denoise": ([P, P, P], I),
            "lucida_net_read_tensor": ([P, I, I, P, S], I),
            "lucida_blob_create": ([C.c_char_p, C.POINTER(BlobParams), C.POINTER(P)], I),
            "lucida_blob_destroy": ([P], None),
            "lucida_blob_desc": ([P, C.POINTER(SceneDesc)], I),
            "lucida_blob_uniforms": ([P, U32, U32, C.POINTER(Uniforms)], I),
            "lucida_blob_set_bvh": ([P, P, S, P], I),
            "lucida_blob_set_env": ([P, U32, U32, P], I),
            "lucida_blob_camera": ([P, C.POINTER(Camera)], I),
            "lucida_film_convergence": ([P, C.POINTER(C.c_float)], I),
            "lucida_render_megakernel": ([P, P, C.POINTER(Uniforms), U32, P, C.POINTER(WavefrontConfig),
                                          C.POINTER(RenderStats)], I),
            "lucida_camera_write_uniforms": ([C.POINTER(Camera), C.c_float, C.POINTER(Uniforms)], I),
        }
        for name, (args, res) in sigs.items():
            fn = getattr(L, name, None)
            if fn is None:
                raise RuntimeError(f"{LIB_PATH} does not export {name}: rebuild with `python -m lucida_amd.build`")
            fn.argtypes = args
            fn.restype = res
        _lib = L
    return _lib


def exported_symbols() -> list[str]:
    """Every entry point include/lucida.h and include/lucida_scenes.h declare."""
    import re
    names = []
    for h in ("lucida.h", "lucida_scenes.h"):
        text = (_HERE.parent / "include" / h).read_text()
        names += re.findall(r"^\s*(?:const\s+)?[\w\*\s]+?\b(lucida_\w+)\s*\(", text, flags=re.M)
    return sorted(set(names))


def _check(status: int) -> None:
    if status != 0:
        msg = lib().lucida_last_error().decode(errors="replace")
        raise GpuError(status, msg)


def _ptr(a: np.ndarray | None) -> int | None:
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"], "arrays crossing the C ABI must be contiguous"
    return a.ctypes.data


# ---------------------------------------------------------------------------
# Scene generation (host)
# ---------------------------------------------------------------------------


class Blob:
    """A generated scene (SceneBlob + camera + environment) held by the library."""

    def __init__(self, name: str, seed: int = 0, frequency: int = 0, displacement: float = 0.0):
        h = C.c_void_p()
        p = BlobParams(seed, frequency, displacement)
        _check(lib().lucida_blob_create(name.encode(), C.byref(p), C.byref(h)))
        self._h = h
        self.name = name

    def __del__(self):
        if getattr(self, "_h", None):
            lib().lucida_blob_destroy(self._h)
            self._h = None

    def desc(self) -> SceneDesc:
        d = SceneDesc()
        _check(lib().lucida_blob_desc(self._h, C.byref(d)))
        return d

    def _arr(self, ptr, count, dtype, cols=None):
        if count == 0 or not ptr:
            shape = (0,) if cols is None else (0, cols)
            return np.zeros(shape, dtype)
        n = count * (cols or 1)
        buf = (C.c_byte * (n * np.dtype(dtype).itemsize)).from_address(ptr)
        a = np.frombuffer(buf, dtype=dtype, count=n).copy()
        return a if cols is None else a.reshape(count, cols)

    def arrays(self) -> dict:
        """Copies of the blob's arrays as numpy structured arrays."""
        d = self.desc()
        return {
            "materials": self._arr(d.materials, d.num_materials, MATERIAL),
            "primitives": self._arr(d.primitives, d.num_primitives, PRIMITIVE),
            "positions": self._arr(d.positions, d.num_positions, np.float32, 4),
            "vertex_attrs": self._arr(d.vertex_attrs, d.num_positions, VERTEX_ATTR),
            "triangles": self._arr(d.triangles, d.num_triangles, TRIANGLE),
            "bvh_nodes": self._arr(d.bvh_nodes, d.num_bvh_nodes, BVH_NODE),
            "lights": self._arr(d.lights, d.num_lights, LIGHT),
            "env_pixels": self._arr(d.env_pixels, d.env_width * d.env_height, np.float32, 4),
            "env_conditional": self._arr(d.env_conditional, (d.env_width + 1) * d.env_height if d.env_width else 0,
                                         np.float32),
            "env_marginal": self._arr(d.env_marginal, d.env_height + 1 if d.env_width else 0, np.float32),
            "env_width": d.env_width, "env_height": d.env_height, "env_total_weight": d.env_total_weight,
            "tlas_root": d.tlas_root,
        }

    def instancing(self) -> tuple[np.ndarray, np.ndarray]:
        """Copies of an instanced scene's mesh ranges and placements (empty otherwise)."""
        mp, mn, pp, pn = C.c_void_p(), C.c_size_t(), C.c_void_p(), C.c_size_t()
        _check(lib().lucida_blob_instancing(self._h, C.byref(mp), C.byref(mn), C.byref(pp), C.byref(pn)))
        return self._arr(mp.value, mn.value, MESH_RANGE), self._arr(pp.value, pn.value, PLACEMENT)

    def uniforms(self, width: int, height: int, *, samples: int = 1, max_depth: int = 8,
                 frame_seed: int = 0x5EED1234, sampling: int = SAMPLING_MIS, sampler: int = SAMPLER_SOBOL,
                 sample_offset: int = 0) -> Uniforms:
        """build_uniforms (crates/core/src/integrator.rs:153)."""
        u = Uniforms()
        _check(lib().lucida_blob_uniforms(self._h, width, height, C.byref(u)))
        u.samples_per_launch = samples
        u.max_depth = max_depth
        u.frame_seed = frame_seed & 0xFFFFFFFF
        u.sampling_mode = sampling
        u.sampler_kind = sampler
        u.sample_offset = sample_offset
        return u

    def camera(self) -> Camera:
        """The scene's camera (SceneDef::camera)."""
        c = Camera()
        _check(lib().lucida_blob_camera(self._h, C.byref(c)))
        return c

    def set_bvh(self, nodes: np.ndarray, triangles_in_leaf_order: np.ndarray | None) -> None:
        nodes = np.ascontiguousarray(nodes, dtype=BVH_NODE)
        tris = None if triangles_in_leaf_order is None else np.ascontiguousarray(triangles_in_leaf_order, TRIANGLE)
        _check(lib().lucida_blob_set_bvh(self._h, _ptr(nodes), len(nodes), _ptr(tris)))

    def set_env(self, rgba: np.ndarray) -> None:
        rgba = np.ascontiguousarray(rgba, dtype=np.float32)
        h, w, _ = rgba.shape
        _check(lib().lucida_blob_set_env(self._h, w, h, _ptr(rgba)))


# ---------------------------------------------------------------------------
# Device context and the harness API
# ---------------------------------------------------------------------------


class Gpu:
    """Gpu::new (crates/gpu/src/lib.rs:144): one HIP device and its stream."""

    def __init__(self, device: int = 0):
        h = C.c_void_p()
        _check(lib().lucida_ctx_create(device, C.byref(h)))
        self._h = h
        # Scenes, films and networks of this context: the C ABI requires them
        # to be destroyed before their context, so close() releases them first.
        self._children = weakref.WeakSet()

    def _adopt(self, obj) -> None:
        self._children.add(obj)

    def close(self) -> None:
        if getattr(self, "_h", None):
            for child in list(self._children):
                child._release()
            lib().lucida_ctx_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def check_errors(self) -> None:
        """Gpu::check_errors (crates/gpu/src/lib.rs:204)."""
        _check(lib().lucida_ctx_synchronize(self._h))

    def info(self) -> tuple[str, int]:
        name = C.create_string_buffer(128)
        cu = C.c_int()
        _check(lib().lucida_ctx_info(self._h, name, 128, C.byref(cu)))
        return name.value.decode(), cu.value


class LbvhBuilder:
    """LbvhBuilder (crates/gpu/src/lbvh.rs:153-758). `fit` picks the fit's
    ordering: "blocks" (block-local subtrees, then the nodes spanning blocks),
    "atomic" (bottom-up walk over the whole tree) or "levels" (one launch per
    level)."""

    FITS = {"blocks": 0, "atomic": 1, "levels": 2}

    def __init__(self, gpu: Gpu, fit: str = "blocks"):
        self.gpu = gpu
        _check(lib().lucida_ctx_set_lbvh_fit(gpu._h, self.FITS[fit]))

    def sort(self, keys: np.ndarray, values: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
        """LbvhBuilder::sort (lbvh.rs:669): stable ascending sort carrying values."""
        keys = np.ascontiguousarray(keys, dtype=np.uint32)
        values = np.ascontiguousarray(values, dtype=np.uint32)
        assert keys.shape == values.shape, "keys and values must be parallel"
        ko = np.empty_like(keys)
        vo = np.empty_like(values)
        _check(lib().lucida_sort_u32(self.gpu._h, _ptr(keys), _ptr(values), len(keys), _ptr(ko), _ptr(vo)))
        return ko, vo

    def build_parts(self, triangles: np.ndarray, positions: np.ndarray) -> dict:
        """LbvhBuilder::build_parts (lbvh.rs:369) -> GpuBuildOutput fields."""
        tris = np.ascontiguousarray(triangles, dtype=TRIANGLE)
        pos = np.ascontiguousarray(positions, dtype=np.float32).reshape(-1, 4)
        n = len(tris)
        out = {
            "codes": np.empty(n, np.uint32), "order": np.empty(n, np.uint32),
            "karras": np.empty((max(n - 1, 1), 2), np.uint32), "parent": np.empty(2 * n - 1, np.uint32),
            "node_bounds": np.empty(2 * n - 1, AABB), "subtree_size": np.empty(max(n - 1, 1), np.uint32),
            "scene_bounds": np.empty(1, AABB),
        }
        parts = LbvhParts(*[_ptr(out[k]) for k in ("codes", "order", "karras", "parent", "node_bounds",
                                                   "subtree_size", "scene_bounds")])
        _check(lib().lucida_lbvh_build_parts(self.gpu._h, _ptr(tris), n, _ptr(pos), len(pos), C.byref(parts)))
        if n < 2:
            out["karras"] = out["karras"][:0]
            out["subtree_size"] = out["subtree_size"][:0]
        return out

    def build(self, triangles: np.ndarray, positions: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
        """LbvhBuilder::build (lbvh.rs:651) + Lbvh::from_parts: (nodes, prim_indices)."""
        tris = np.ascontiguousarray(triangles, dtype=TRIANGLE)
        pos = np.ascontiguousarray(positions, dtype=np.float32).reshape(-1, 4)
        n = len(tris)
        nodes = np.empty(max(2 * n - 1, 1), BVH_NODE)
        prims = np.empty(max(n, 1), np.uint32)
        count = C.c_size_t()
        _check(lib().lucida_lbvh_build(self.gpu._h, _ptr(tris), n, _ptr(pos), len(pos), _ptr(nodes),
                                       C.byref(count), _ptr(prims)))
        return nodes[:count.value].copy(), prims[:n].copy()

    def build_treelet(self, triangles: np.ndarray, positions: np.ndarray,
                      rounds: int = TREELET_ROUNDS) -> tuple[np.ndarray, np.ndarray]:
        """The LBVH restructured by treelets (lucida_bvh_build_treelet; no
        reference counterpart): (nodes, prim_indices) in build()'s layout."""
        tris = np.ascontiguousarray(triangles, dtype=TRIANGLE)
        pos = np.ascontiguousarray(positions, dtype=np.float32).reshape(-1, 4)
        n = len(tris)
        nodes = np.empty(max(2 * n - 1, 1), BVH_NODE)
        prims = np.empty(max(n, 1), np.uint32)
        count = C.c_size_t()
        _check(lib().lucida_bvh_build_treelet(self.gpu._h, _ptr(tris), n, _ptr(pos), len(pos), rounds, _ptr(nodes),
                                              C.byref(count), _ptr(prims)))
        return nodes[:count.value].copy(), prims[:n].copy()


class Scene:
    """A device-resident scene (the buffers render_native_on uploads)."""

    def __init__(self, gpu: Gpu, blob: Blob, build_lbvh: bool = True, treelet: bool = False):
        """build_lbvh: build the tree on the device (else the blob's own nodes
        are uploaded); treelet: restructure that LBVH by treelets."""
        self.gpu = gpu
        d = blob.desc()
        d.build_lbvh = (2 if treelet else 1) if build_lbvh else 0
        h = C.c_void_p()
        meshes, placements = blob.instancing()
        if len(placements):
            # Scene::build_instances on the device: BLAS per mesh + TLAS
            _check(lib().lucida_scene_create_instanced(gpu._h, C.byref(d), _ptr(meshes), len(meshes),
                                                       _ptr(placements), len(placements), C.byref(h)))
        else:
            _check(lib().lucida_scene_create(gpu._h, C.byref(d), C.byref(h)))
        self._h = h
        gpu._adopt(self)
        self.num_triangles = d.num_triangles

    @classmethod
    def from_desc(cls, gpu: Gpu, desc: SceneDesc) -> "Scene":
        """Uploads a caller-built lucida_scene_desc (its arrays are copied)."""
        self = cls.__new__(cls)
        self.gpu = gpu
        h = C.c_void_p()
        _check(lib().lucida_scene_create(gpu._h, C.byref(desc), C.byref(h)))
        self._h = h
        gpu._adopt(self)
        self.num_triangles = desc.num_triangles
        return self

    def counts(self) -> dict:
        nt, nn, ni, root = C.c_size_t(), C.c_size_t(), C.c_size_t(), C.c_uint32()
        _check(lib().lucida_scene_counts(self._h, C.byref(nt), C.byref(nn), C.byref(ni), C.byref(root)))
        return {"triangles": nt.value, "bvh_nodes": nn.value, "instances": ni.value, "tlas_root": root.value}

    def download_instances(self) -> np.ndarray:
        out = np.empty(self.counts()["instances"], INSTANCE)
        _check(lib().lucida_scene_download_instances(self._h, _ptr(out)))
        return out

    def host_desc(self, blob: Blob) -> tuple[SceneDesc, list]:
        """The blob's desc with the device's BVH nodes, leaf-ordered triangles,
        instances and TLAS root, for a host tracer to walk the same structure.
        Returns the desc and the arrays it points into (keep them alive)."""
        d = blob.desc()
        c = self.counts()
        nodes, tris = self.download_bvh()
        inst = self.download_instances()
        keep = [nodes, tris, inst]
        d.bvh_nodes, d.num_bvh_nodes = nodes.ctypes.data, len(nodes)
        d.triangles = tris.ctypes.data
        d.instances, d.num_instances = (inst.ctypes.data if len(inst) else None), len(inst)
        d.tlas_root = c["tlas_root"]
        return d, keep

    def _release(self) -> None:
        if getattr(self, "_h", None):
            lib().lucida_scene_destroy(self._h)
            self._h = None

    def __del__(self):
        self._release()

    def num_bvh_nodes(self) -> int:
        n = C.c_size_t()
        _check(lib().lucida_scene_num_bvh_nodes(self._h, C.byref(n)))
        return n.value

    def download_bvh(self) -> tuple[np.ndarray, np.ndarray]:
        nodes = np.empty(self.num_bvh_nodes(), BVH_NODE)
        tris = np.empty(self.num_triangles, TRIANGLE)
        _check(lib().lucida_scene_download_bvh(self._h, _ptr(nodes), _ptr(tris)))
        return nodes, tris


class Film:
    """A device-resident GpuAccum buffer."""

    def __init__(self, gpu: Gpu, width: int, height: int):
        self.gpu = gpu
        self.width, self.height = width, height
        h = C.c_void_p()
        _check(lib().lucida_film_create(gpu._h, width, height, C.byref(h)))
        self._h = h
        gpu._adopt(self)

    def _release(self) -> None:
        if getattr(self, "_h", None):
            lib().lucida_film_destroy(self._h)
            self._h = None

    def __del__(self):
        self._release()

    def clear(self) -> None:
        _check(lib().lucida_film_clear(self._h))

    def download(self) -> np.ndarray:
        a = np.empty(self.width * self.height, ACCUM)
        _check(lib().lucida_film_download(self._h, _ptr(a)))
        return a

    def convergence(self) -> float:
        """ConvergenceMeter (crates/gpu/src/stats.rs): the mean relative standard
        error of the film's pixels, -1 when no pixel is measurable."""
        v = C.c_float()
        _check(lib().lucida_film_convergence(self._h, C.byref(v)))
        return float(v.value)

    def upload(self, accum: np.ndarray) -> None:
        a = np.ascontiguousarray(accum, dtype=ACCUM)
        assert len(a) == self.width * self.height
        _check(lib().lucida_film_upload(self._h, _ptr(a)))


def render_wavefront(gpu: Gpu, scene: Scene, uniforms: Uniforms, samples: int, film: Film,
                     config: WavefrontConfig | None = None) -> RenderStats:
    """wavefront::render_native_on (crates/gpu/src/wavefront.rs:134), accumulating into `film`."""
    st = RenderStats()
    cfg = C.byref(config) if config is not None else None
    _check(lib().lucida_render_wavefront(gpu._h, scene._h, C.byref(uniforms), samples, film._h, cfg,
                                         C.byref(st)))
    return st


def render_megakernel(gpu: Gpu, scene: Scene, uniforms: Uniforms, samples: int, film: Film,
                      config: WavefrontConfig | None = None) -> RenderStats:
    """The megakernel render (pt_gpu::render_native_on, crates/gpu/src/lib.rs:391): one
    thread per pixel traces its paths start to finish; the film equals the wavefront's."""
    st = RenderStats()
    cfg = C.byref(config) if config is not None else None
    _check(lib().lucida_render_megakernel(gpu._h, scene._h, C.byref(uniforms), samples, film._h, cfg,
                                          C.byref(st)))
    return st


ARCHITECTURES = ("megakernel", "wavefront")


def render_with(gpu: Gpu, scene: Scene, uniforms: Uniforms, samples: int, film: Film, architecture: str = "wavefront",
                config: WavefrontConfig | None = None) -> RenderStats:
    """render_with (crates/gpu/src/lib.rs:69): either architecture; both give the same film."""
    if architecture == "megakernel":
        return render_megakernel(gpu, scene, uniforms, samples, film, config)
    if architecture == "wavefront":
        return render_wavefront(gpu, scene, uniforms, samples, film, config)
    raise ValueError(f"unknown architecture {architecture!r}")


def render_batch(gpu: Gpu, jobs, config: WavefrontConfig | None = None):
    """Many render requests in one wavefront (lucida_render_batch). `jobs` is a
    list of (scene, uniforms, samples, film); returns (stats, per-job completion
    times in ms from the start of the batch)."""
    arr = (Job * len(jobs))()
    for k, (scene, u, samples, film) in enumerate(jobs):
        arr[k].scene = scene._h.value
        arr[k].uniforms = u
        arr[k].samples = samples
        arr[k].film = film._h.value
    done = np.zeros(len(jobs), np.float32)
    st = RenderStats()
    cfg = C.byref(config) if config is not None else None
    _check(lib().lucida_render_batch(gpu._h, arr, len(jobs), cfg, _ptr(done), C.byref(st)))
    return st, done


def scene_cast(gpu: Gpu, scene: Scene, rays: np.ndarray, any_hit: bool = False):
    """Batch closest-hit (or occlusion) queries through the renderer's own
    traversal: rays (n, 8) float32 = origin, unused, direction, t_max. Returns
    hits (n, 4) float32 = t, u, v, id bits (triangle, 0x80000000 | primitive or
    0xFFFFFFFF) and the instance of each instanced hit."""
    rays = np.ascontiguousarray(rays, np.float32)
    n = len(rays)
    hits = np.zeros((n, 4), np.float32)
    inst = np.zeros(n, np.uint32)
    flags = 1 if any_hit else 0
    _check(lib().lucida_scene_cast(gpu._h, scene._h, _ptr(rays), n, flags, _ptr(hits), _ptr(inst)))
    return hits, inst


def energy_table(gpu: Gpu) -> np.ndarray:
    """The device's GGX directional-albedo table (bsdf.rs:1200-1227), 1024 floats."""
    out = np.empty(1024, np.float32)
    _check(lib().lucida_energy_table(gpu._h, _ptr(out)))
    return out


def film_mean(accum: np.ndarray, width: int, height: int) -> np.ndarray:
    """Film pixels = radiance / samples (crates/gpu/src/wavefront.rs:702-706), shape (h, w, 3)."""
    s = np.maximum(accum["samples"], 1.0)[:, None]
    return (accum["radiance"] / s).reshape(height, width, 3)


def render_native_on(gpu: Gpu, blob: Blob, width: int, height: int, samples: int, **kw) -> np.ndarray:
    """Convenience: upload `blob`, render, return the Film (h, w, 3) like the reference's call."""
    scene = Scene(gpu, blob)
    film = Film(gpu, width, height)
    u = blob.uniforms(width, height, samples=samples, **kw)
    render_wavefront(gpu, scene, u, samples, film)
    return film_mean(film.download(), width, height)


class Denoiser:
    """Denoiser (crates/gpu/src/denoise.rs:238-447): a-trous on device films."""

    def __init__(self, gpu: Gpu):
        self.gpu = gpu

    def run(self, film_in: Film, film_out: Film, params: DenoiseParams | None = None) -> None:
        p = params or DenoiseParams.default()
        _check(lib().lucida_denoise_atrous(self.gpu._h, film_in._h, C.byref(p), film_out._h))


# Layer table of the learned denoiser: (cin, cout, relu) and resolution level.
NET_LAYERS = [(16, 32, 1), (32, 32, 1), (32, 48, 1), (48, 48, 1), (48, 64, 1), (64, 64, 1), (64, 96, 1),
              (96, 96, 1), (160, 64, 1), (64, 64, 1), (112, 48, 1), (48, 48, 1), (80, 32, 1), (32, 32, 1),
              (32, 16, 0)]
NET_LEVEL = [0, 0, 1, 1, 2, 2, 3, 3, 2, 2, 1, 1, 0, 0, 0]


class LearnedDenoiser:
    """The BF16 encoder-decoder denoiser trained and run on the device
    (no reference counterpart; inputs are Accum films)."""

    def __init__(self, gpu: Gpu, seed: int = 1):
        self.gpu = gpu
        h = C.c_void_p()
        _check(lib().lucida_net_create(gpu._h, seed, C.byref(h)))
        self._h = h
        gpu._adopt(self)
        self._crop = None

    def _release(self) -> None:
        if getattr(self, "_h", None):
            lib().lucida_net_destroy(self._h)
            self._h = None

    def __del__(self):
        self._release()

    @staticmethod
    def param_count() -> int:
        return int(lib().lucida_net_param_count())

    def params(self, gradients: bool = False) -> np.ndarray:
        out = np.empty(self.param_count(), np.float32)
        _check(lib().lucida_net_get_params(self._h, _ptr(out), 1 if gradients else 0))
        return out

    def set_params(self, p: np.ndarray) -> None:
        a = np.ascontiguousarray(p, np.float32)
        assert a.size == self.param_count()
        _check(lib().lucida_net_set_params(self._h, _ptr(a)))

    @staticmethod
    def _batch(inputs, refs, crops):
        k = len(inputs)
        ins = (C.c_void_p * k)(*[f._h.value for f in inputs])
        rfs = (C.c_void_p * k)(*[f._h.value for f in refs]) if refs is not None else None
        if crops is None:
            return k, ins, rfs, None, None
        x0 = np.ascontiguousarray([c[0] for c in crops], np.uint32)
        y0 = np.ascontiguousarray([c[1] for c in crops], np.uint32)
        return k, ins, rfs, x0, y0

    def forward_backward(self, inputs, refs, width: int, height: int, crops=None) -> float:
        k, ins, rfs, x0, y0 = self._batch(inputs, refs, crops)
        loss = C.c_float()
        _check(lib().lucida_net_forward_backward(self._h, ins, rfs, k, None if x0 is None else _ptr(x0),
                                                 None if y0 is None else _ptr(y0), width, height,
                                                 C.byref(loss)))
        self._crop = (k, height, width)
        return float(loss.value)

    def adam(self, lr: float = 1e-3) -> None:
        _check(lib().lucida_net_adam(self._h, lr))

    def train_step(self, inputs, refs, width: int, height: int, crops=None, lr: float = 1e-3) -> float:
        k, ins, rfs, x0, y0 = self._batch(inputs, refs, crops)
        loss = C.c_float()
        _check(lib().lucida_net_train_step(self._h, ins, rfs, k, None if x0 is None else _ptr(x0),
                                           None if y0 is None else _ptr(y0), width, height, lr,
                                           C.byref(loss)))
        self._crop = (k, height, width)
        return float(loss.value)

    def denoise(self, film_in: Film, film_out: Film) -> None:
        _check(lib().lucida_net_denoise(self._h, film_in._h, film_out._h))
        self._crop = (1, film_in.height, film_in.width)

    def tensor(self, index: int, gradient: bool = False) -> np.ndarray:
        """Activation `index` (0 = features, 1..15 = layer outputs) of the last
        forward pass, or its gradient, shape (n, h, w, c)."""
        n, h, w = self._crop
        lvl = 0 if index == 0 else NET_LEVEL[index - 1]
        c = 16 if index == 0 else NET_LAYERS[index - 1][1]
        shape = (n, h >> lvl, w >> lvl, c)
        out = np.empty(shape, np.float32)
        _check(lib().lucida_net_read_tensor(self._h, index, 1 if gradient else 0, _ptr(out), out.size))
        return out
