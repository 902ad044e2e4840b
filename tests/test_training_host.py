"""The denoiser training loop's host logic (lucida_amd/denoiser_training.py
train), on the CPU with a stand-in network: the learning-rate schedule it
hands each step, the crops it draws, and the moving average of the weights
it ends with. No GPU: the stand-in records what train() asks of it.

No reference counterpart: the reference has no learned denoiser."""
from __future__ import annotations

import math

import numpy as np
import pytest

from lucida_amd import denoiser_training as T


class FakeFilm:
    def __init__(self, w: int, h: int):
        self.width, self.height = w, h


class FakeNet:
    """Parameters are the step count so far (plus a per-entry offset), so the
    weights after step s are known in closed form."""

    def __init__(self):
        self.p = np.arange(4, dtype=np.float32)
        self.steps = 0
        self.rates, self.crops = [], []
        self.set_calls = 0

    def train_step(self, ins, refs, cw, ch, crops, lr):
        self.steps += 1
        self.p = np.arange(4, dtype=np.float32) + self.steps
        self.rates.append(lr)
        self.crops.append((cw, ch, list(crops), [f.width for f in ins]))
        return 1.0 / self.steps

    def params(self):
        return self.p.copy()

    def set_params(self, v):
        self.set_calls += 1
        self.p = np.asarray(v, np.float32).copy()


class FakeGpu:
    def check_errors(self):
        pass


def pairs(n=5, w=64, h=48):
    return [(FakeFilm(w, h), FakeFilm(w, h)) for _ in range(n)]


def test_without_average_the_last_weights_stay():
    net = FakeNet()
    losses = T.train(FakeGpu(), net, pairs(), 30, crop_w=32, crop_h=16)
    assert net.set_calls == 0
    assert np.array_equal(net.p, np.arange(4) + 30)
    assert losses == [1.0 / s for s in range(1, 31)]


@pytest.mark.parametrize("steps,ema,every", [(100, 0.9, 10), (101, 0.99, 7), (20, 0.5, 1)])
def test_moving_average_of_the_second_half(steps, ema, every):
    net = FakeNet()
    T.train(FakeGpu(), net, pairs(), steps, crop_w=32, crop_h=16, ema=ema, ema_every=every)
    # sampled after steps s (1-based) with s - 1 >= steps // 2 and s % every == 0
    samples = [s for s in range(1, steps + 1) if s - 1 >= steps // 2 and s % every == 0]
    beta = ema ** every
    k = len(samples)
    # closed form: the first sample weighs beta^(k-1), sample i >= 1 weighs (1 - beta) beta^(k-1-i)
    w = [beta ** (k - 1)] + [(1 - beta) * beta ** (k - 1 - i) for i in range(1, k)]
    assert math.isclose(sum(w), 1.0, rel_tol=1e-12)
    expect = sum(wi * (np.arange(4) + s) for wi, s in zip(w, samples))
    assert net.set_calls == 1
    np.testing.assert_allclose(net.p, expect.astype(np.float32), rtol=1e-6)
    # the average lies between the first and the last sampled weights
    assert (np.arange(4) + samples[0] <= net.p + 1e-4).all() and (net.p <= np.arange(4) + samples[-1] + 1e-4).all()


def test_learning_rate_decays_by_a_cosine_to_a_tenth():
    net = FakeNet()
    T.train(FakeGpu(), net, pairs(), 50, crop_w=32, crop_h=16, lr=2e-3)
    r = np.array(net.rates)
    assert r[0] == pytest.approx(2e-3)
    assert np.all(np.diff(r) <= 1e-15)
    # a cosine from lr to lr / 10: halfway at half the steps, symmetric about it, above lr / 10 at the end
    assert r[25] == pytest.approx(0.55 * 2e-3)
    np.testing.assert_allclose(r[25 - np.arange(1, 25)] + r[25 + np.arange(1, 25)], 2 * 0.55 * 2e-3, rtol=1e-9)
    assert 2e-4 < r[-1] < 1.02 * 2e-4


def test_crops_lie_inside_their_films():
    net = FakeNet()
    T.train(FakeGpu(), net, pairs(w=64, h=48), 40, batch=8, crop_w=32, crop_h=16)
    for cw, ch, crops, widths in net.crops:
        assert (cw, ch) == (32, 16) and len(crops) == 8 and widths == [64] * 8
        for x, y in crops:
            assert 0 <= x <= 64 - 32 and 0 <= y <= 48 - 16
