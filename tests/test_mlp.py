"""The warm-start network: CPU checker pinned by finite differences and closed forms; the GPU
BF16 MFMA forward, backward and training checked against it.

Tolerances: the GPU rounds inputs, weights, activations and back-propagated errors to BF16
(8 significant bits, relative rounding 2^-9) and accumulates in FP32; the checker emulates the
same roundings and sums in double, so results agree to the BF16 level: outputs within 2% of their
scale, gradients within 2% in norm.
"""
import numpy as np
import pytest

NT = 20099


def random_params(rng, scale=1.0):
    from oracle import oracle

    p = np.zeros(oracle.MLP_NUM_PARAMS, np.float64)
    p[0:24 * 128] = rng.normal(0, np.sqrt(2 / 24), 24 * 128) * scale
    p[3072:3200] = rng.normal(0, 0.1, 128)
    p[3200:3200 + 16384] = rng.normal(0, np.sqrt(2 / 128), 16384) * scale
    p[19584:19712] = rng.normal(0, 0.1, 128)
    p[19712:19712 + 384] = rng.normal(0, np.sqrt(1 / 128), 384) * scale
    p[20096:20099] = rng.normal(0, 0.1, 3)
    p[20099:20123] = rng.normal(0, 1, 24)          # x_mean
    p[20123:20147] = rng.uniform(0.5, 2.0, 24)     # x_scale
    p[20147:20150] = rng.uniform(0.5, 2.0, 3)      # y_scale
    return p


def random_rows(rng, p, n):
    X = (p[20099:20123] + p[20123:20147] * rng.normal(0, 1, (n, 24))).astype(np.float32)
    T = (p[20147:20150] * rng.normal(0, 1, (n, 3))).astype(np.float32)
    return X, T


def test_gelu_closed_form(oracle_mod):
    """With W1 = e_k (one input to one unit) and everything downstream the identity path, the
    output is the known GELU(tanh) values: gelu(1) = 0.8411919906, gelu(-1) = -0.1588080094."""
    p = np.zeros(oracle_mod.MLP_NUM_PARAMS)
    p[20123:20147] = 1.0
    p[20147:20150] = 1.0
    p[0 * 128 + 0] = 1.0            # W1[0][0]
    p[3200 + 0 * 128 + 0] = 1.0     # W2[0][0]
    p[19712 + 0 * 3 + 0] = 1.0      # W3[0][0]
    X = np.zeros((2, 24), np.float32)
    X[0, 0], X[1, 0] = 1.0, -1.0
    y = oracle_mod.mlp_forward(p, X, emulate_bf16=False)
    g1, gm1 = 0.8411919906082768, -0.15880800939172324
    gelu = lambda x: 0.5 * x * (1 + np.tanh(0.7978845608028654 * (x + 0.044715 * x ** 3)))
    assert y[0, 0] == pytest.approx(gelu(g1), abs=1e-12)
    assert y[1, 0] == pytest.approx(gelu(gm1), abs=1e-12)
    assert gelu(1.0) == pytest.approx(g1, abs=1e-10)
    assert np.all(y[:, 1:] == 0)


def test_zero_last_layer_closed_form(oracle_mod):
    """W3 = 0: the output is b3, dL/db3 = 2 mean(b3 - t / y_scale), every other gradient is 0."""
    rng = np.random.default_rng(1)
    p = random_params(rng)
    p[19712:19712 + 384] = 0.0
    X, T = random_rows(rng, p, 50)
    g, loss = oracle_mod.mlp_gradients(p, X, T, emulate_bf16=False)
    t = T.astype(np.float64) / p[20147:20150].astype(np.float32)
    e = p[20096:20099] - t
    assert loss == pytest.approx(np.mean(np.sum(e * e, axis=1)), rel=1e-6)
    np.testing.assert_allclose(g[20096:20099], 2 * e.mean(axis=0), rtol=1e-6)
    assert np.all(g[:19712] == 0)


def test_oracle_gradients_finite_differences(oracle_mod):
    rng = np.random.default_rng(2)
    p = random_params(rng)
    X, T = random_rows(rng, p, 40)
    g, _ = oracle_mod.mlp_gradients(p, X, T, emulate_bf16=False)
    idx = np.concatenate([rng.choice(np.arange(a, b), 8, replace=False) for a, b in
                          ((0, 3072), (3072, 3200), (3200, 19584), (19584, 19712), (19712, 20096), (20096, 20099))
                          if b - a >= 8] + [np.arange(20096, 20099)])
    h = 1e-6
    for i in idx:
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        lp = oracle_mod.mlp_gradients(pp, X, T, emulate_bf16=False)[1]
        lm = oracle_mod.mlp_gradients(pm, X, T, emulate_bf16=False)[1]
        fd = (lp - lm) / (2 * h)
        assert fd == pytest.approx(g[i], rel=1e-5, abs=1e-9), i


def test_bf16_emulation_close_to_exact(oracle_mod):
    rng = np.random.default_rng(3)
    p = random_params(rng)
    X, T = random_rows(rng, p, 200)
    ye = oracle_mod.mlp_forward(p, X, emulate_bf16=True)
    yx = oracle_mod.mlp_forward(p, X, emulate_bf16=False)
    assert np.abs(ye - yx).max() < 0.05 * np.abs(yx).max()
    ge, _ = oracle_mod.mlp_gradients(p, X, T, emulate_bf16=True)
    gx, _ = oracle_mod.mlp_gradients(p, X, T, emulate_bf16=False)
    assert np.linalg.norm(ge - gx) < 0.05 * np.linalg.norm(gx)


@pytest.fixture(scope="module")
def net(nphys):
    from nphysics_amd import learned

    return learned.WarmStartNet(seed=5)


@pytest.mark.gpu
def test_gpu_forward_vs_checker(net, oracle_mod):
    rng = np.random.default_rng(4)
    p = random_params(rng)
    net.params = p.astype(np.float32)
    for n in (1, 63, 1000):
        X, _ = random_rows(rng, p, n)
        y = net.forward(X).astype(np.float64)
        ref = oracle_mod.mlp_forward(net.params.astype(np.float64), X, emulate_bf16=True)
        scale = np.abs(ref).max()
        err = np.abs(y - ref).max()
        print(f"forward n={n}: max |gpu - checker| {err:.3e} (output scale {scale:.3e})")
        assert err < 0.02 * scale


@pytest.mark.gpu
def test_gpu_gradients_vs_checker(net, oracle_mod):
    rng = np.random.default_rng(6)
    p = random_params(rng)
    net.params = p.astype(np.float32)
    for n in (64, 777):
        X, T = random_rows(rng, p, n)
        g, loss = net.gradients(X, T)
        gr, lr = oracle_mod.mlp_gradients(net.params.astype(np.float64), X, T, emulate_bf16=True)
        rel = np.linalg.norm(g - gr) / np.linalg.norm(gr)
        print(f"gradients n={n}: relative error {rel:.3e}, loss {loss:.6f} vs {lr:.6f}")
        assert rel < 0.02
        assert loss == pytest.approx(lr, rel=2e-3)
        # every parameter group agrees on its own, not only in the total
        for a, b in ((0, 3072), (3072, 3200), (3200, 19584), (19584, 19712), (19712, 20096), (20096, 20099)):
            assert np.linalg.norm(g[a:b] - gr[a:b]) <= 0.03 * np.linalg.norm(gr[a:b]) + 1e-7, (a, b)


@pytest.mark.gpu
def test_gpu_training_fits_a_smooth_target(net, oracle_mod):
    """Trained on the GPU, the network fits a smooth function of the inputs; the trained weights
    evaluate the same on the checker."""
    from nphysics_amd import learned

    rng = np.random.default_rng(7)
    n = 32768
    X = rng.normal(0, 1, (n, 24)).astype(np.float32)
    T = np.stack([np.tanh(X[:, 0] + 0.5 * X[:, 1]), X[:, 2] * X[:, 3] * 0.5, np.sin(X[:, 4])], 1).astype(np.float32)
    m = learned.WarmStartNet(seed=11)
    m.set_data(X, T)
    losses = m.train(epochs=12, batch=1024, lr=2e-3)
    print("epoch losses", losses)
    assert np.all(np.isfinite(losses))
    assert losses[-1] < 0.25 * losses[0]
    Xv = rng.normal(0, 1, (2000, 24)).astype(np.float32)
    y = m.forward(Xv).astype(np.float64)
    ref = oracle_mod.mlp_forward(m.params.astype(np.float64), Xv, emulate_bf16=True)
    assert np.abs(y - ref).max() < 0.02 * np.abs(ref).max()
    m.close()
