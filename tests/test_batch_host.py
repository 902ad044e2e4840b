"""Host logic of batched renders, on the CPU: the benchmark's seeded request
stream (lucida_amd/request_stream.py) and the autotuner's choice and cache
(lucida_amd/autotune.py), the latter with a stand-in device, batch renderer
and clock, so that each candidate launch shape takes a known time.

No reference counterpart: the reference has no batched renders."""
from __future__ import annotations

import json
from types import SimpleNamespace

import numpy as np
import pytest

import lucida_amd as L
from lucida_amd import autotune as A
from lucida_amd import request_stream as RS


def test_request_stream_is_seeded_and_in_range():
    a, b = RS.stream_requests(), RS.stream_requests()
    assert a == b and len(a) == 64
    assert RS.stream_requests(seed=7) != a
    for k, (si, res, spp, fs) in enumerate(a):
        assert 0 <= si < 5 and res in RS.RESOLUTIONS and spp in RS.SAMPLES and fs == 0x1000 + k
    # every resolution, sample count and scene is drawn
    assert {r[1] for r in a} == set(RS.RESOLUTIONS)
    assert {r[2] for r in a} == set(RS.SAMPLES)
    assert {r[0] for r in a} == set(range(5))
    # the draws are numpy's PCG64 stream in the order resolution, samples, scene
    rng = np.random.default_rng(6)
    first = (int(rng.choice(RS.RESOLUTIONS)), int(rng.choice(RS.SAMPLES)), int(rng.integers(0, 5)))
    assert (a[0][1], a[0][2], a[0][0]) == first


class FakeFilm:
    def __init__(self, gpu, w, h):
        self.width, self.height, self.cleared = w, h, 0

    def clear(self):
        self.cleared += 1


class Clock:
    def __init__(self):
        self.t = 0.0

    def __call__(self):
        return self.t


def fake_setup(monkeypatch, tmp_path, cost_ms):
    """A stand-in device and batch renderer: a render under config c takes
    cost_ms(max_paths, use_graph) on the stand-in clock."""
    clock = Clock()
    calls = []

    def render_batch(gpu, jobs, cfg):
        calls.append((cfg.max_paths, cfg.use_graph, [j[2] for j in jobs]))
        clock.t += cost_ms(cfg.max_paths, cfg.use_graph) / 1e3
        return None, None

    monkeypatch.setattr(A.L, "render_batch", render_batch)
    monkeypatch.setattr(A.L, "Film", FakeFilm)
    monkeypatch.setattr(A, "time", SimpleNamespace(perf_counter=clock))  # the autotuner's clock only
    monkeypatch.setenv("LUCIDA_AUTOTUNE_CACHE", str(tmp_path / "tune.json"))
    gpu = SimpleNamespace(info=lambda: ("gfx950", 256))
    scene = SimpleNamespace(num_triangles=1000)
    jobs = [(scene, SimpleNamespace(width=w, height=w, max_depth=8), spp, None)
            for w, spp in ((128, 64), (1024, 16), (1024, 1), (256, 4))]
    return gpu, jobs, calls


def test_autotuner_picks_the_fastest_shape_and_caches_it(monkeypatch, tmp_path):
    costs = {(1 << 21, 0): 9.0, (1 << 21, 1): 8.0, (1 << 22, 0): 7.0, (1 << 22, 1): 3.0,
             (1 << 23, 0): 5.0, (1 << 23, 1): 6.0, (1 << 24, 0): 4.0, (1 << 24, 1): 10.0}
    gpu, jobs, calls = fake_setup(monkeypatch, tmp_path, lambda mp, g: costs.get((mp, g), 1.0))  # warm-up: max_paths 0
    cfg = A.tune(gpu, jobs, repeats=2)
    assert isinstance(cfg, L.WavefrontConfig) and (cfg.max_paths, cfg.use_graph) == (1 << 22, 1)
    # one warm-up, then every shape `repeats` times, on each job's first chunk only
    assert len(calls) == 1 + 2 * len(costs)
    assert {(mp, g) for mp, g, _ in calls[1:]} == set(costs)
    assert all(spp == [64, 8, 1, 4] for _, _, spp in calls)  # 1024^2: 2^23 / 2^20 = 8 spp per chunk
    trials = A.last_trials(gpu, jobs)
    assert {(t["max_paths"], t["use_graph"]): t["ms"] for t in trials} == pytest.approx(costs)
    # the cache answers the same shape without rendering
    n = len(calls)
    again = A.tune(gpu, jobs)
    assert (again.max_paths, again.use_graph) == (1 << 22, 1) and len(calls) == n
    stored = json.loads((tmp_path / "tune.json").read_text())
    assert list(stored.values())[0]["config"] == {"max_paths": 1 << 22, "use_graph": 1}


def test_autotuner_keys_the_cache_by_batch_shape(monkeypatch, tmp_path):
    gpu, jobs, calls = fake_setup(monkeypatch, tmp_path, lambda mp, g: 1.0 + g + mp / (1 << 24))
    first = A.tune(gpu, jobs, repeats=1)
    assert (first.max_paths, first.use_graph) == (1 << 21, 0)
    n = len(calls)
    # another shape (one more job) is tuned afresh; the order of jobs does not matter
    more = jobs + [jobs[0]]
    A.tune(gpu, more, repeats=1)
    assert len(calls) > n
    m = len(calls)
    A.tune(gpu, list(reversed(jobs)), repeats=1)
    assert len(calls) == m
    assert A.signature(gpu, jobs) != A.signature(gpu, more)
    assert A.signature(gpu, jobs) == A.signature(gpu, list(reversed(jobs)))


def test_autotuner_without_cache_keeps_no_file(monkeypatch, tmp_path):
    gpu, jobs, calls = fake_setup(monkeypatch, tmp_path, lambda mp, g: 2.0 - g)
    cfg = A.tune(gpu, jobs, use_cache=False, repeats=1)
    assert cfg.use_graph == 1 and not (tmp_path / "tune.json").exists()
