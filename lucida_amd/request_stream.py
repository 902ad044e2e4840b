"""The batched request stream: 64 render jobs drawn by a PCG64 stream (seed 6)
over resolution {128, 256, 512, 1024}^2 x samples {1, 4, 16, 64} x scene
{the 1x1x1 Cornell box, four displaced-sphere rooms (seeds 1-4, geodesic
frequency 64: 81,920 + 1,024 triangles)}, 8 bounces, frame seeds 0x1000 + k.
The benchmark times it; the tests check every job's film against its own
single render."""
from __future__ import annotations

import numpy as np

import lucida_amd as L

RESOLUTIONS = (128, 256, 512, 1024)
SAMPLES = (1, 4, 16, 64)
MAX_DEPTH = 8


def stream_blobs() -> list:
    return [L.Blob("box-c1")] + [L.Blob("sphere-c2", seed=k, frequency=64) for k in range(1, 5)]


def stream_requests(jobs_n: int = 64, seed: int = 6, num_scenes: int = 5) -> list[tuple[int, int, int, int]]:
    """(scene index, resolution, samples, frame seed) per job, in arrival order."""
    rng = np.random.default_rng(seed)
    reqs = []
    for k in range(jobs_n):
        res = int(rng.choice(RESOLUTIONS))
        spp = int(rng.choice(SAMPLES))
        si = int(rng.integers(0, num_scenes))
        reqs.append((si, res, spp, 0x1000 + k))
    return reqs


def make_jobs(gpu: L.Gpu, blobs: list, scenes: list, reqs: list) -> list:
    """(scene, uniforms, samples, film) per request, each with a fresh film."""
    out = []
    for si, res, spp, fs in reqs:
        u = blobs[si].uniforms(res, res, samples=spp, max_depth=MAX_DEPTH, frame_seed=fs)
        out.append((scenes[si], u, spp, L.Film(gpu, res, res)))
    return out
