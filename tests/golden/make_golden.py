"""Writes the golden fixtures under tests/golden/ from data files the reference
holds. Run in a container where the reference checkout is mounted (it is read
only as data, never executed); the fixtures it writes are committed, so the
test suite never needs the reference itself.

Sources (paths relative to the reference root):
  web/public/scenes/cornell-box.bin   packed SceneBlob of cornell_box()
  web/public/scenes/cornell-mesh.bin  packed SceneBlob of cornell_mesh()
  web/public/scenes/{furnace-test,metal-sweep,mis-scene,glass-box,sunset,
      instance-forest}.bin            the other packed built-in scenes
  web/src/generated/scenes.ts         section offsets and CAMERA_FIXTURES
  shaders/common/ggx_energy.wgsl      the 32x32 GGX directional-albedo table
  shaders/common/generated.wgsl       the Sobol direction numbers (SOBOL_V)

The packed format is codegen.rs:434-491: 16-byte aligned sections in the
order materials, primitives, lights, positions, vertexAttrs, triangles,
bvhNodes, envRadiance, envCdf.

Usage: python tests/golden/make_golden.py [REFERENCE_ROOT]
"""
from __future__ import annotations

import hashlib
import json
import re
import sys
from pathlib import Path

import numpy as np

OUT = Path(__file__).resolve().parent
SIZES = {"materials": 80, "primitives": 64, "lights": 64, "positions": 16, "vertexAttrs": 32, "triangles": 16,
         "bvhNodes": 32, "envRadiance": 16, "envCdf": 4}
LUM = np.array([0.2126, 0.7152, 0.0722])


def sections(ts: str, name: str) -> dict:
    block = ts[ts.index(f'name: "{name}"'):]
    block = block[:block.index("camera:")]
    out = {}
    for m in re.finditer(r"(\w+): \{ byteOffset: (\d+), count: (\d+) \}", block):
        out[m.group(1)] = (int(m.group(2)), int(m.group(3)))
    return out


def section_bytes(data: bytes, secs: dict, key: str) -> bytes:
    off, cnt = secs[key]
    return data[off:off + cnt * SIZES[key]]


def scene_meta(ts: str, name: str) -> dict:
    """camera, env, background and instancing records of one scene in scenes.ts."""
    block = ts[ts.index(f'name: "{name}"'):]
    block = block[:block.index("\n  },")]
    f = lambda s: [float(x) for x in s.split(",")]
    cam = re.search(r"camera: \{ eye: \[([^\]]+)\], lookAt: \[([^\]]+)\], up: \[([^\]]+)\], vfovDeg: ([\d.]+), "
                    r"aperture: ([\d.]+), focusDistance: ([\d.]+) \}", block)
    env = re.search(r"env: \{ width: (\d+), height: (\d+), totalWeight: ([\d.]+) \}", block)
    inst = re.search(r"instancing: \{ count: (\d+), analyticPrimitives: (\d+), tlasRoot: (\d+) \}", block)
    bg = re.search(r"background: \[([^\]]+)\]", block)
    return {"camera": {"eye": f(cam.group(1)), "look_at": f(cam.group(2)), "up": f(cam.group(3)),
                       "vfov_deg": float(cam.group(4)), "aperture": float(cam.group(5)),
                       "focus_distance": float(cam.group(6))},
            "env": {"width": int(env.group(1)), "height": int(env.group(2)), "total_weight": float(env.group(3))},
            "instancing": {"count": int(inst.group(1)), "analytic_primitives": int(inst.group(2)),
                           "tlas_root": int(inst.group(3))},
            "background": f(bg.group(1))}


def env_record(data: bytes, secs: dict, width: int, height: int) -> dict:
    """The environment sections of a packed scene, reduced to small checks:
    exact hashes, every row's luminance sum (float64), the texels brighter
    than 50 (the sun disc), sampled texels, the whole marginal CDF and a
    sample of the conditional CDFs (codegen.rs packs rows 0..h-1 of w+1
    conditional entries, then the h+1 marginal entries, in a (w+1) x (h+1)
    rectangle)."""
    rad = np.frombuffer(section_bytes(data, secs, "envRadiance"), "<f4").reshape(height, width, 4)
    cdf = np.frombuffer(section_bytes(data, secs, "envCdf"), "<f4").reshape(height + 1, width + 1)
    lum = rad[..., :3].astype(np.float64) @ LUM
    sun = np.flatnonzero(lum.ravel() > 50.0)
    idx = np.unique(np.concatenate([np.arange(0, width * height, 61), sun]))
    cidx = np.arange(0, height * (width + 1), 37)
    return {"sha256_envRadiance": hashlib.sha256(rad.tobytes()).hexdigest(),
            "sha256_envCdf": hashlib.sha256(cdf.tobytes()).hexdigest(),
            "row_luminance_sums": lum.sum(axis=1).tolist(),
            "sun_texels": sun.tolist(),
            "texel_samples": {"index": idx.tolist(), "rgba_bits": rad.reshape(-1, 4)[idx].view("<u4").tolist()},
            "marginal": cdf[height, :height + 1].tolist(),
            "conditional_samples": {"index": cidx.tolist(),
                                    "values": cdf[:height].reshape(-1)[cidx].tolist()}}


def main(ref: Path) -> None:
    ts = (ref / "web/src/generated/scenes.ts").read_text()

    # --- cornell-box: small, decoded in full ---------------------------------
    data = (ref / "web/public/scenes/cornell-box.bin").read_bytes()
    secs = sections(ts, "cornell-box")
    box = {k: np.frombuffer(section_bytes(data, secs, k), "<u4").tolist()
           for k in ("materials", "primitives", "lights")}
    (OUT / "cornell_box_sections.json").write_text(json.dumps(box))

    # --- cornell-mesh: hashes plus samples -----------------------------------
    data = (ref / "web/public/scenes/cornell-mesh.bin").read_bytes()
    secs = sections(ts, "cornell-mesh")
    rec = {"counts": {k: secs[k][1] for k in SIZES}}
    for k in ("materials", "lights", "positions", "vertexAttrs"):
        rec[f"sha256_{k}"] = hashlib.sha256(section_bytes(data, secs, k)).hexdigest()
    va = np.frombuffer(section_bytes(data, secs, "vertexAttrs"), "<u4").reshape(-1, 8).copy()
    va[:, 4:6] = 0  # uv: libm-dependent last bits
    rec["sha256_vertexAttrs_no_uv"] = hashlib.sha256(va.tobytes()).hexdigest()
    uv = np.frombuffer(section_bytes(data, secs, "vertexAttrs"), "<f4").reshape(-1, 8)[:, 4:6]
    rec["uv_samples"] = {"index": list(range(0, len(uv), 97)), "values": uv[::97].tolist()}
    tris = np.frombuffer(section_bytes(data, secs, "triangles"), "<u4").reshape(-1, 4)
    order = np.lexsort(tris.T[::-1])
    rec["sha256_triangles_sorted"] = hashlib.sha256(np.ascontiguousarray(tris[order]).tobytes()).hexdigest()
    pos = np.frombuffer(section_bytes(data, secs, "positions"), "<u4").reshape(-1, 4)
    idx = np.linspace(0, len(pos) - 1, 200).astype(int)
    rec["position_samples"] = {"index": idx.tolist(), "bits": pos[idx].tolist()}
    (OUT / "cornell_mesh_sections.json").write_text(json.dumps(rec))

    # --- the other packed scenes ------------------------------------------------
    for name in ("furnace-test", "metal-sweep", "mis-scene", "glass-box", "sunset", "instance-forest"):
        data = (ref / f"web/public/scenes/{name}.bin").read_bytes()
        secs = sections(ts, name)
        meta = scene_meta(ts, name)
        rec = {"counts": {k: secs[k][1] for k in SIZES}, **meta}
        for k in ("materials", "primitives", "lights"):
            rec[k] = np.frombuffer(section_bytes(data, secs, k), "<u4").tolist()
        if secs["triangles"][1]:
            rec["sha256_positions"] = hashlib.sha256(section_bytes(data, secs, "positions")).hexdigest()
            va = np.frombuffer(section_bytes(data, secs, "vertexAttrs"), "<u4").reshape(-1, 8).copy()
            va[:, 4:6] = 0  # uv: libm-dependent last bits
            rec["sha256_vertexAttrs_no_uv"] = hashlib.sha256(va.tobytes()).hexdigest()
            tris = np.frombuffer(section_bytes(data, secs, "triangles"), "<u4").reshape(-1, 4)
            order = np.lexsort(tris.T[::-1])
            rec["sha256_triangles_sorted"] = hashlib.sha256(np.ascontiguousarray(tris[order]).tobytes()).hexdigest()
        if meta["env"]["width"]:
            rec["env_sections"] = env_record(data, secs, meta["env"]["width"], meta["env"]["height"])
        (OUT / f"{name.replace('-', '_')}_sections.json").write_text(json.dumps(rec))

    # --- camera fixtures -------------------------------------------------------
    cams = []
    body = ts[ts.index("export const CAMERA_FIXTURES"):]
    for m in re.finditer(r"\{ label: \"([\w-]+)\", width: (\d+), height: (\d+), camera: \{ eye: \[([^\]]+)\], "
                         r"lookAt: \[([^\]]+)\], up: \[([^\]]+)\], vfovDeg: ([\d.]+), aperture: ([\d.]+), "
                         r"focusDistance: ([\d.]+) \}, camOrigin: \[([^\]]+)\], camUpperLeft: \[([^\]]+)\], "
                         r"camHorizontal: \[([^\]]+)\], camVertical: \[([^\]]+)\] \}", body):
        f = lambda s: [float(x) for x in s.split(",")]
        cams.append({"label": m.group(1), "width": int(m.group(2)), "height": int(m.group(3)),
                     "eye": f(m.group(4)), "look_at": f(m.group(5)), "up": f(m.group(6)),
                     "vfov_deg": float(m.group(7)), "aperture": float(m.group(8)),
                     "focus_distance": float(m.group(9)), "cam_origin": f(m.group(10)),
                     "cam_upper_left": f(m.group(11)), "cam_horizontal": f(m.group(12)),
                     "cam_vertical": f(m.group(13))})
    assert len(cams) == 7, len(cams)
    (OUT / "camera_fixtures.json").write_text(json.dumps(cams, indent=1))

    # --- GGX energy table --------------------------------------------------------
    wgsl = (ref / "shaders/common/ggx_energy.wgsl").read_text()
    body = wgsl[wgsl.index("array<f32, 1024>("):]
    body = re.sub(r"//[^\n]*", "", body)  # row comments carry numbers too
    vals = [float(v) for v in re.findall(r"\d+\.\d+", body)][:1024]
    assert len(vals) == 1024
    (OUT / "ggx_energy_table.json").write_text(json.dumps(vals))

    # --- Sobol direction numbers -----------------------------------------------------
    gen = (ref / "shaders/common/generated.wgsl").read_text()
    body = gen[gen.index("const SOBOL_V"):]
    v = [int(x, 16) for x in re.findall(r"0x([0-9a-f]{8})u", body)][:64]
    assert len(v) == 64
    (OUT / "sobol_v.json").write_text(json.dumps([v[:32], v[32:]]))
    print("golden fixtures written to", OUT)


if __name__ == "__main__":
    main(Path(sys.argv[1] if len(sys.argv) > 1 else "/root/reference"))
