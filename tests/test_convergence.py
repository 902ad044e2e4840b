"""The convergence meter (crates/gpu/src/stats.rs, shaders/stats/
convergence.wgsl): the oracle's reduction against closed forms on the CPU,
then the device meter against the oracle and against the reference's own
statistical checks (crates/gpu/tests/convergence.rs) on the GPU."""
from __future__ import annotations

import math
import sys

import numpy as np
import pytest

import lucida_amd as L


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def synthetic_film(n_samples, means, deltas):
    """Accum records of pixels whose N samples alternate mean - d and mean + d
    (grey): mean m, per-sample variance d^2, so the relative standard error of
    the mean is d / (m sqrt(N)) exactly."""
    acc = np.zeros(len(means), L.ACCUM)
    for i, (n, m, d) in enumerate(zip(n_samples, means, deltas)):
        acc["radiance"][i] = n * m
        acc["samples"][i] = n
        acc["radiance_sq"][i] = n * (m * m + d * d)
    return acc


def test_oracle_meter_closed_form(oracle):
    n = np.array([4, 16, 64, 256], np.float64)
    m = np.array([0.5, 2.0, 0.25, 8.0])
    d = np.array([0.25, 1.0, 0.05, 3.0])
    got = oracle.convergence(synthetic_film(n, m, d))
    want = float(np.mean(d / (m * np.sqrt(n))))
    assert got == pytest.approx(want, rel=1e-6)


def test_oracle_meter_leaves_out_dark_and_single_sample_pixels(oracle):
    base = synthetic_film([16], [1.0], [0.5])
    want = 0.5 / 4.0
    extra = synthetic_film([1, 16, 16], [1.0, 1e-4, 0.0], [0.5, 1e-5, 0.0])
    got = oracle.convergence(np.concatenate([base, extra]))
    assert got == pytest.approx(want, rel=1e-6)
    assert oracle.convergence(extra) == -1.0
    # a constant pixel has no variance
    assert oracle.convergence(synthetic_film([8], [3.0], [0.0])) == 0.0


@pytest.mark.gpu
def test_device_meter_matches_oracle(gpu, oracle):
    rng = np.random.default_rng(3)
    n = rng.integers(1, 300, 200_003).astype(np.float64)
    m = np.exp(rng.normal(size=n.size)) * (rng.random(n.size) > 0.05)
    d = m * rng.random(n.size)
    acc = synthetic_film(n, m, d)
    film = L.Film(gpu, 200_003, 1)
    film.upload(acc)
    got = film.convergence()
    want = oracle.convergence(acc)
    say(f"meter: device {got:.7f} oracle {want:.7f}")
    assert got == pytest.approx(want, rel=1e-5)
    empty = L.Film(gpu, 64, 64)
    assert empty.convergence() == -1.0


def render_meter(gpu, scene, blob, spp, sampling=L.SAMPLING_MIS, sampler=L.SAMPLER_SOBOL, seed=0x5EED1234):
    u = blob.uniforms(96, 96, samples=spp, max_depth=5, sampling=sampling, sampler=sampler, frame_seed=seed)
    film = L.Film(gpu, 96, 96)
    L.render_wavefront(gpu, scene, u, spp, film)
    return film


@pytest.mark.gpu
def test_readout_falls_with_sample_count(gpu):
    """convergence.rs:43: falls as samples accumulate, within 2x of 1/sqrt(N)
    at 1024 spp."""
    blob = L.Blob("cornell-box")
    scene = L.Scene(gpu, blob)
    es = {spp: render_meter(gpu, scene, blob, spp).convergence() for spp in (16, 64, 256, 1024)}
    say("meter vs spp: " + ", ".join(f"{k}: {v:.5f}" for k, v in es.items()))
    vals = list(es.values())
    assert all(v > 0 for v in vals) and all(b < a for a, b in zip(vals, vals[1:]))
    expected = es[16] * math.sqrt(16 / 1024)
    assert 0.5 * expected < es[1024] < 2.0 * expected


@pytest.mark.gpu
def test_worse_sampling_reports_more_error(gpu):
    """convergence.rs:101: BSDF-only sampling of the Cornell light reports at
    least 1.5x MIS's error."""
    blob = L.Blob("cornell-box")
    scene = L.Scene(gpu, blob)
    good = render_meter(gpu, scene, blob, 64, sampling=L.SAMPLING_MIS).convergence()
    bad = render_meter(gpu, scene, blob, 64, sampling=L.SAMPLING_BSDF).convergence()
    say(f"mis {good:.5f}, bsdf-only {bad:.5f}")
    assert bad > 1.5 * good


def spread_of_independent_renders(gpu, scene, blob, sampler, runs=16):
    """convergence.rs:151: the c4-corrected spread of `runs` renders with
    independent frame seeds, luminance-weighted like the meter."""
    imgs = []
    for k in range(runs):
        f = render_meter(gpu, scene, blob, 64, sampler=sampler, seed=(0x5EED1234 + k * 0x9E3779B9) & 0xFFFFFFFF)
        imgs.append(L.film_mean(f.download(), 96, 96).astype(np.float64))
    x = np.stack(imgs)
    mean = x.mean(0)
    lw = np.array([0.2126, 0.7152, 0.0722])
    mean_l = mean @ lw
    var = np.maximum(x.var(0, ddof=1), 0.0)
    c4 = math.sqrt(2.0 / (runs - 1)) * math.exp(math.lgamma(runs / 2) - math.lgamma((runs - 1) / 2))
    keep = mean_l > 1e-3
    return float(np.mean(np.sqrt(np.maximum(var @ lw, 0.0))[keep] / c4 / mean_l[keep]))


@pytest.mark.gpu
def test_readout_matches_the_spread_of_independent_renders(gpu):
    """convergence.rs:215: with independent samples the meter equals the
    measured spread of 16 renders within [0.8, 1.25]; with Sobol it is an
    upper bound (convergence.rs:246)."""
    blob = L.Blob("cornell-box")
    scene = L.Scene(gpu, blob)
    for sampler, label in ((L.SAMPLER_INDEPENDENT, "independent"), (L.SAMPLER_SOBOL, "sobol")):
        reported = render_meter(gpu, scene, blob, 64, sampler=sampler).convergence()
        actual = spread_of_independent_renders(gpu, scene, blob, sampler)
        say(f"{label}: meter {reported:.5f}, spread of 16 renders {actual:.5f} ({reported / actual:.3f}x)")
        if sampler == L.SAMPLER_INDEPENDENT:
            assert 0.8 < reported / actual < 1.25
        else:
            assert reported >= 0.95 * actual
