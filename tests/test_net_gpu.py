"""Learned denoiser on the device against the float64 oracle.

The oracle (oracle/net.py) rounds to BF16 wherever the device stores BF16,
so the forward activations, the loss and the parameter gradients are compared
with tolerances written against the BF16 step (2^-8 relative): what remains
is the order of FP32 accumulation and the rounding flips it causes.
"""
from __future__ import annotations

import sys

import numpy as np
import pytest

import lucida_amd as L
from oracle import net as N

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def synthetic_accum(w: int, h: int, seed: int, spp: float = 4.0, noise: float = 0.6,
                    noise_seed: int | None = None) -> np.ndarray:
    """A plausible Accum film: radiance sums, albedo, normals, depth, squares.
    The smooth image is fixed by `seed`; `noise` scales a log-normal per-pixel
    perturbation drawn from `noise_seed` (default: seed)."""
    rng = np.random.default_rng(seed)
    a = np.zeros(w * h, L.ACCUM)
    n = w * h
    yy, xx = np.divmod(np.arange(n), w)
    al = 0.1 + 0.8 * rng.random((n, 3))
    nrm = rng.normal(size=(n, 3))
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    nz = np.random.default_rng(seed if noise_seed is None else noise_seed).normal(size=(n, 3))
    shade = np.exp(0.5 * np.sin(xx / 7.0)[:, None] + 0.3 * np.cos(yy / 5.0)[:, None] + noise * nz)
    rad = al * shade
    a["samples"] = spp
    a["radiance"] = rad * spp
    a["radiance_sq"] = rad * rad * spp * (1.0 + rng.random((n, 3)))
    a["albedo"] = al * spp
    a["normal"] = nrm * spp
    a["depth"] = (1.0 + 4.0 * rng.random(n)) * spp
    return a


def as_float16(a: np.ndarray) -> np.ndarray:
    """GpuAccum records as [..., 16] float32 channels."""
    return np.ascontiguousarray(a).view(np.float32).reshape(-1, 16)


def films(gpu, w, h, seeds):
    out, raw = [], []
    for s in seeds:
        acc = synthetic_accum(w, h, s)
        f = L.Film(gpu, w, h)
        f.upload(acc)
        out.append(f)
        raw.append(as_float16(acc).reshape(h, w, 16))
    return out, raw


def atrous_of(gpu, film):
    """The device a-trous output (default parameters) of a film, [h, w, 16]."""
    tmp = L.Film(gpu, film.width, film.height)
    L.Denoiser(gpu).run(film, tmp)
    return as_float16(tmp.download()).reshape(film.height, film.width, 16)


def random_params(seed: int) -> np.ndarray:
    rng = np.random.default_rng(seed)
    layers = N.unpack(N.init_params(seed))
    for _, b in layers:
        b[:] = rng.normal(size=b.shape) * 0.05
    return N.pack(layers).astype(np.float32)


def rel_l2(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))


FW, FH, CW, CH = 256, 40, 128, 16  # film size and crop size (width multiple of 128, height of 8)
CROPS = [(0, 0), (128, 24)]


@pytest.fixture(scope="module")
def setup(gpu):
    ins, raw_in = films(gpu, FW, FH, [1, 2])
    refs, raw_ref = films(gpu, FW, FH, [3, 4])
    net = L.LearnedDenoiser(gpu, seed=5)
    p = random_params(6)
    net.set_params(p)
    say(f"net forward_backward case: batch 2 crops {CW}x{CH}")
    loss = net.forward_backward(ins, refs, CW, CH, crops=CROPS)
    crop = lambda r, c: r[c[1]:c[1] + CH, c[0]:c[0] + CW]
    base = [atrous_of(gpu, f) for f in ins]
    x_o = np.stack([N.features(crop(r, c), crop(b, c)) for r, b, c in zip(raw_in, base, CROPS)])
    t_o = np.stack([N.target(crop(ri, c), crop(rr, c)) for ri, rr, c in zip(raw_in, raw_ref, CROPS)])
    return dict(ins=ins, refs=refs, raw_in=raw_in, net=net, p=p, loss=loss, x_o=x_o, t_o=t_o,
                base_o=x_o[..., 12:15])


def test_parameter_count(setup):
    assert L.LearnedDenoiser.param_count() == N.param_count()


def test_features_match_oracle(setup):
    x_g = setup["net"].tensor(0)
    # the device stores features in BF16: one BF16 step of slack for libm differences
    np.testing.assert_allclose(x_g, N.to_bf16(setup["x_o"]), rtol=2 ** -7, atol=1e-6)


def test_forward_matches_oracle_layer_by_layer(setup):
    net = setup["net"]
    x_g = net.tensor(0).astype(np.float64)
    A = N.forward(setup["p"].astype(np.float64), x_g, bf16=True)
    for i in range(15):
        g = net.tensor(i + 1)
        o = A[f"l{i}"]
        err = rel_l2(g, o)
        say(f"layer {i}: rel_l2 {err:.2e}")
        assert err < 1.5e-2, (i, err)


def test_loss_and_gradients_match_oracle(setup):
    net = setup["net"]
    x_g = net.tensor(0).astype(np.float64)
    loss_o, g_o, _, D = N.loss_and_grad(setup["p"].astype(np.float64), x_g, setup["t_o"], bf16=True,
                                        base=setup["base_o"])
    assert abs(setup["loss"] - loss_o) <= 2e-3 * loss_o, (setup["loss"], loss_o)
    g_g = net.params(gradients=True).astype(np.float64)
    off = 0
    for i, (ci, co, _) in enumerate(N.LAYERS):
        nw = 9 * ci * co
        gw, ow = g_g[off:off + nw], g_o[off:off + nw]
        gb, ob = g_g[off + nw:off + nw + co], g_o[off + nw:off + nw + co]
        off += nw + co
        ew, eb = rel_l2(gw, ow), rel_l2(gb, ob)
        say(f"layer {i}: dW rel_l2 {ew:.2e}, db rel_l2 {eb:.2e}")
        assert ew < 3e-2 and eb < 3e-2, (i, ew, eb)


def test_backward_on_device_activations_matches_oracle(setup):
    """The oracle's backward pass run on the device's own forward activations:
    no ReLU mask can differ, so what remains is the backward kernels' own
    rounding (one BF16 step per stored activation gradient)."""
    net = setup["net"]
    x_g = net.tensor(0).astype(np.float64)
    A = N.activations_from_layers(x_g, [net.tensor(i + 1) for i in range(15)], bf16=True)
    _, g_o, _, D = N.loss_and_grad(setup["p"].astype(np.float64), x_g, setup["t_o"], bf16=True, acts=A,
                                   base=setup["base_o"])
    g_g = net.params(gradients=True).astype(np.float64)
    off = 0
    for i, (ci, co, _) in enumerate(N.LAYERS):
        n = 9 * ci * co + co
        e = rel_l2(g_g[off:off + n], g_o[off:off + n])
        off += n
        say(f"layer {i}: dW,db rel_l2 {e:.2e} (device activations)")
        assert e < 1e-2, (i, e)
    # every stored activation gradient: d(layer output), after the layer's ReLU mask
    for i in range(15):
        dg = net.tensor(i + 1, gradient=True).astype(np.float64)
        do = D[f"l{i}"] * ((A[f"l{i}"] > 0) if N.LAYERS[i][2] else 1.0)
        e = rel_l2(dg, do)
        say(f"d l{i}: rel_l2 {e:.2e}")
        assert e < 2e-2, (i, e)


def test_adam_step_matches_closed_form(gpu):
    ins, _ = films(gpu, CW, CH, [7])
    refs, _ = films(gpu, CW, CH, [8])
    net = L.LearnedDenoiser(gpu, seed=9)
    net.set_params(random_params(10))
    net.forward_backward(ins, refs, CW, CH)
    p0 = net.params().astype(np.float64)
    g = net.params(gradients=True).astype(np.float64)
    lr = 1e-3
    net.adam(lr)
    p1 = net.params().astype(np.float64)
    exp, _, _ = N.adam(p0, g, np.zeros_like(g), np.zeros_like(g), 1, lr)
    np.testing.assert_allclose(p1, exp, rtol=0, atol=1e-6 * lr + 1.2e-7 * np.abs(p0).max())
    assert np.all(p1[g == 0] == p0[g == 0])


def test_partial_weight_gradient_chunk(gpu):
    """One 128x8 crop: the 1/8 level has 16 pixels, half a 32-pixel chunk."""
    ins, raw_in = films(gpu, CW, 8, [11])
    refs, raw_ref = films(gpu, CW, 8, [12])
    net = L.LearnedDenoiser(gpu, seed=13)
    p = random_params(14)
    net.set_params(p)
    say("net forward_backward case: 1 crop 128x8")
    loss = net.forward_backward(ins, refs, CW, 8)
    x_g = net.tensor(0).astype(np.float64)
    base_o = N.features(raw_in[0], atrous_of(gpu, ins[0]))[None, ..., 12:15]
    loss_o, g_o, _, _ = N.loss_and_grad(p.astype(np.float64), x_g, N.target(raw_in[0], raw_ref[0])[None], bf16=True,
                                        base=base_o)
    assert abs(loss - loss_o) <= 2e-3 * loss_o
    assert rel_l2(net.params(gradients=True).astype(np.float64), g_o) < 3e-2


def test_training_reduces_the_loss(gpu):
    """Noisy inputs against their own clean images: the loss must fall well
    below where the a-trous estimate alone leaves it."""
    ins, refs = [], []
    for s in (21, 22):
        f = L.Film(gpu, FW, FH)
        f.upload(synthetic_accum(FW, FH, s, noise=0.6))
        r = L.Film(gpu, FW, FH)
        r.upload(synthetic_accum(FW, FH, s, noise=0.0))
        ins.append(f)
        refs.append(r)
    net = L.LearnedDenoiser(gpu, seed=25)
    say("net training case: 100 Adam steps, batch 2 crops 128x16")
    losses = [net.train_step(ins, refs, CW, CH, crops=CROPS, lr=2e-3) for _ in range(100)]
    assert all(np.isfinite(losses))
    first, last = np.mean(losses[:5]), np.mean(losses[-5:])
    say(f"loss {first:.4f} -> {last:.4f}")
    assert last < 0.5 * first, (first, last)


def test_training_is_deterministic(gpu):
    """The weight gradients are summed in a fixed order (per-slice partials,
    then one ordered sum), so two training runs from the same start on the
    same crops end with bit-identical weights."""
    ins, refs = [], []
    for s in (41, 42):
        f = L.Film(gpu, FW, FH)
        f.upload(synthetic_accum(FW, FH, s, noise=0.6))
        r = L.Film(gpu, FW, FH)
        r.upload(synthetic_accum(FW, FH, s, noise=0.0))
        ins.append(f)
        refs.append(r)
    say("net determinism case: two runs of 20 Adam steps, batch 2 crops 128x16")
    runs = []
    for _ in range(2):
        net = L.LearnedDenoiser(gpu, seed=26)
        losses = [net.train_step(ins, refs, CW, CH, crops=CROPS, lr=2e-3) for _ in range(20)]
        runs.append((net.params().copy(), net.params(gradients=True).copy(), losses))
    np.testing.assert_array_equal(runs[0][1].view(np.uint32), runs[1][1].view(np.uint32))
    np.testing.assert_array_equal(runs[0][0].view(np.uint32), runs[1][0].view(np.uint32))
    assert not np.array_equal(runs[0][0], L.LearnedDenoiser(gpu, seed=26).params())  # the steps moved the weights


def test_denoise_matches_oracle(gpu):
    ins, raw = films(gpu, 128, 32, [31])
    out = L.Film(gpu, 128, 32)
    net = L.LearnedDenoiser(gpu, seed=32)
    p = random_params(33)
    net.set_params(p)
    say("net denoise case: 128x32")
    net.denoise(ins[0], out)
    got = as_float16(out.download()).reshape(32, 128, 16)
    x_g = net.tensor(0).astype(np.float64)
    A = N.forward(p.astype(np.float64), x_g, bf16=True)
    base_o = N.features(raw[0], atrous_of(gpu, ins[0]))[..., 12:15]
    exp = N.compose(A["l14"][0], raw[0], x_g[0], base=base_o)
    assert rel_l2(got[..., 0:3], exp[..., 0:3]) < 1.5e-2
    np.testing.assert_array_equal(got[..., 3], 1.0)
    np.testing.assert_allclose(got[..., 4:11], exp[..., 4:11], rtol=1e-6, atol=1e-7)


def test_rejects_bad_shapes(gpu):
    ins, _ = films(gpu, 100, 16, [41])
    out = L.Film(gpu, 100, 16)
    net = L.LearnedDenoiser(gpu, seed=1)
    with pytest.raises(L.GpuError) as e:
        net.denoise(ins[0], out)
    assert e.value.kind == "InvalidArg"
    big, _ = films(gpu, 128, 16, [42])
    with pytest.raises(L.GpuError) as e:
        net.forward_backward(big, big, 128, 16, crops=[(8, 0)])  # crop leaves the film
    assert e.value.kind == "InvalidArg"
    with pytest.raises(L.GpuError) as e:
        net.denoise(big[0], big[0])
    assert e.value.kind == "InvalidArg"
