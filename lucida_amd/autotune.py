"""Autotuner for batched renders (lucida_render_batch).

Tunes the launch shape of a batch on the device it runs on: the path budget
of a round (how many jobs' chunks share one set of launches) and whether the
rounds are replayed as one hipGraph. Every candidate renders the same probe
batch into scratch films (each job's first chunk of samples: the images of
the real jobs are unaffected by the choice, which only changes how chunks are
grouped into launches), the fastest wins, and the result is cached as JSON
keyed by the device and a signature of the batch's shape.

    cfg = autotune.tune(gpu, jobs)          # cached after the first call
    stats, done = L.render_batch(gpu, jobs, cfg)
"""
from __future__ import annotations

import hashlib
import json
import os
import time
from pathlib import Path

import lucida_amd as L

MAX_PATHS = (1 << 21, 1 << 22, 1 << 23, 1 << 24)
GRAPH = (0, 1)


def cache_path() -> Path:
    p = os.environ.get("LUCIDA_AUTOTUNE_CACHE")
    return Path(p) if p else Path.home() / ".cache" / "lucida_amd" / "autotune.json"


def signature(gpu: L.Gpu, jobs) -> str:
    """Device name + the batch's shape (resolutions, samples, depths, scene sizes)."""
    name, cus = gpu.info()
    shape = sorted((u.width, u.height, spp, u.max_depth, scene.num_triangles) for scene, u, spp, _ in jobs)
    h = hashlib.sha256(json.dumps([name, cus, shape]).encode()).hexdigest()[:16]
    return f"{name}/{cus}/{h}"


def _load() -> dict:
    try:
        return json.loads(cache_path().read_text())
    except (OSError, ValueError):
        return {}


def _store(data: dict) -> None:
    p = cache_path()
    try:
        p.parent.mkdir(parents=True, exist_ok=True)
        tmp = p.with_suffix(".tmp")
        tmp.write_text(json.dumps(data, indent=1, sort_keys=True))
        tmp.replace(p)
    except OSError:
        pass  # a read-only home only costs the cache


def probe_jobs(gpu: L.Gpu, jobs):
    """Each job's first chunk of samples (the single render's batch size),
    rendered into scratch films."""
    out = []
    for scene, u, spp, film in jobs:
        pixels = u.width * u.height
        chunk = max(1, min(spp, (1 << 23) // pixels))
        out.append((scene, u, chunk, L.Film(gpu, u.width, u.height)))
    return out


def tune(gpu: L.Gpu, jobs, use_cache: bool = True, repeats: int = 2, log=None) -> L.WavefrontConfig:
    key = signature(gpu, jobs)
    data = _load() if use_cache else {}
    if key in data:
        c = data[key]["config"]
        return L.WavefrontConfig(0, c["max_paths"], 0, c["use_graph"], 0, 0)
    probe = probe_jobs(gpu, jobs)
    results = []
    L.render_batch(gpu, probe, L.WavefrontConfig(0, 0, 0, 0, 0, 0))  # warm-up
    for mp in MAX_PATHS:
        for g in GRAPH:
            cfg = L.WavefrontConfig(0, mp, 0, g, 0, 0)
            best = float("inf")
            for _ in range(repeats):
                for _, _, _, f in probe:
                    f.clear()
                t0 = time.perf_counter()
                L.render_batch(gpu, probe, cfg)
                best = min(best, (time.perf_counter() - t0) * 1e3)
            results.append({"max_paths": mp, "use_graph": g, "ms": round(best, 3)})
            if log:
                log(f"autotune: max_paths {mp:>9} graph {g}: {best:.2f} ms")
    win = min(results, key=lambda r: r["ms"])
    data[key] = {"config": {"max_paths": win["max_paths"], "use_graph": win["use_graph"]}, "trials": results}
    if use_cache:
        _store(data)
    return L.WavefrontConfig(0, win["max_paths"], 0, win["use_graph"], 0, 0)


def last_trials(gpu: L.Gpu, jobs) -> list:
    """The cached trial timings of a batch shape (empty if never tuned)."""
    return _load().get(signature(gpu, jobs), {}).get("trials", [])
