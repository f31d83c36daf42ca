"""Pins of the training restatement (oracle/train.py): gradients against central finite
differences of the loss, the composite loss's special cases, and the Adam step's closed form."""
import numpy as np
import pytest


@pytest.fixture(scope="module")
def small(nb, orc):
    from oracle import train
    cfg = nb.network_config(n_levels=3, n_features_per_level=2, log2_hashmap_size=7, base_resolution=3,
                            per_level_scale=2.0)
    tables, mlp = nb.network_init_params(cfg, seed=5, table_range=0.5)
    net = orc.Net(cfg, tables, mlp, (-1, -1, -1), (1, 1, 1))
    lv = net.levels()
    assert lv["dense"][0] == 1 and lv["dense"][2] == 0  # one dense and one hashed level at least
    rng = np.random.default_rng(3)
    n = 40
    x = rng.uniform(0.02, 0.98, (n, 9))
    t = rng.uniform(0, 1, (n, 8))
    t[:, 0] = rng.integers(0, 2, n)
    return train, cfg, lv, tables.astype(np.float64), mlp.astype(np.float64), x, t


def test_forward_matches_the_query_oracle(nb, orc, small):
    train, cfg, lv, tables, mlp, x, t = small
    net = orc.Net(cfg, tables.astype(np.float32), mlp.astype(np.float32), (-1, -1, -1), (1, 1, 1))
    out, _ = net.network(x.astype(np.float32), mode=orc.FP64)
    _, _, _, y = train.train_gradients(lv, cfg, tables, mlp, x.astype(np.float32).astype(np.float64), t)
    # The restatement forms grid positions in FP32 (as the kernels do), the query oracle's FP64
    # mode in FP64: the outputs agree to the FP32 resolution of the finest grid positions.
    assert np.allclose(y, out, rtol=0, atol=1e-5)


def test_gradients_match_finite_differences(small):
    # The relative-L2 (albedo) gradient holds its denominator constant, so the finite differences
    # are of the loss with the denominators frozen at the base point.
    train, cfg, lv, tables, mlp, x, t = small
    loss, dtab, dmlp, y = train.train_gradients(lv, cfg, tables, mlp, x, t)
    den = y[:, 5:8] ** 2 + 1e-2

    def L(tb, ml):
        return train.train_loss(lv, cfg, tb, ml, x, t, albedo_den=den)
    rng = np.random.default_rng(8)
    touched = np.nonzero(dtab)[0]
    assert len(touched) > 50
    h = 1e-6
    for i in rng.choice(touched, 40, replace=False):
        tp, tm = tables.copy(), tables.copy()
        tp[i] += h
        tm[i] -= h
        fd = (L(tp, mlp) - L(tm, mlp)) / (2 * h)
        assert abs(fd - dtab[i]) <= 1e-6 * max(1.0, abs(dtab[i])) + 1e-9, (i, fd, dtab[i])
    for i in rng.choice(len(mlp), 60, replace=False):
        mp, mm = mlp.copy(), mlp.copy()
        mp[i] += h
        mm[i] -= h
        fd = (L(tables, mp) - L(tables, mm)) / (2 * h)
        assert abs(fd - dmlp[i]) <= 1e-6 * max(1.0, abs(dmlp[i])) + 1e-9, (i, fd, dmlp[i])


def test_relative_l2_gradient_holds_the_denominator(small):
    train = small[0]
    y = np.full((1, 8), 0.5)
    y[0, 5] = 0.3
    t = np.ones((1, 8))
    _, g = train.composite_loss(y, t)
    h = 1e-7
    den = y[:, 5:8] ** 2 + 1e-2
    yp, ym = y.copy(), y.copy()
    yp[0, 5] += h
    ym[0, 5] -= h
    fd_frozen = (train.composite_loss(yp, t, den)[0][0, 5] - train.composite_loss(ym, t, den)[0][0, 5]) / (2 * h)
    fd_full = (train.composite_loss(yp, t)[0][0, 5] - train.composite_loss(ym, t)[0][0, 5]) / (2 * h)
    assert np.isclose(g[0, 5], fd_frozen, rtol=1e-6) and not np.isclose(g[0, 5], fd_full, rtol=1e-2)


def test_composite_loss_cases(small):
    train = small[0]
    y = np.array([[0.7, 0.4, 0.55, 0.5, 0.45, 0.2, 0.3, 0.9],
                  [0.2, 0.4, 0.55, 0.5, 0.45, 0.2, 0.3, 0.9]])
    t = np.array([[1.0, 0.5, 0.5, 0.5, 0.5, 0.1, 0.3, 0.5],
                  [0.0, 0.5, 0.5, 0.5, 0.5, 0.1, 0.3, 0.5]])
    loss, g = train.composite_loss(y, t)
    n8 = 16
    assert np.isclose(loss[0, 0] * n8, -np.log(0.7)) and np.isclose(g[0, 0] * n8, -(1 - 0.7) / (0.7 - 0.49))
    assert np.isclose(loss[0, 1] * n8, 0.1) and np.isclose(g[0, 1] * n8, -1.0)
    # normal: linearised ((y*2-1)/0.1)*0.5+0.5 -> 1.0 for y = 0.55, L1 against 0.5, gradient x 10
    assert np.isclose(loss[0, 2] * n8, 0.5) and np.isclose(g[0, 2] * n8, 10.0)
    assert np.isclose(loss[0, 3] * n8, 0.0)
    assert np.isclose(loss[0, 5] * n8, 0.01 / (0.04 + 1e-2)) and np.isclose(g[0, 5] * n8, 0.2 / 0.05)
    # target visibility < 0.5 masks every other output
    assert np.all(loss[1, 1:] == 0) and np.all(g[1, 1:] == 0) and np.isclose(loss[1, 0] * n8, -np.log(0.8))
    # BCE clamps the prediction to [1e-5, 1 - 1e-5]
    l2, g2 = train.composite_loss(np.array([[1.0] + [0.5] * 7]), np.array([[0.0] + [0.5] * 7]))
    assert np.isclose(l2[0, 0] * 8, -np.log(1e-5)) and np.isfinite(g2[0, 0])


def test_adam_closed_form(small):
    """tiny-cuda-nn's Adam in closed form: the first update of a parameter moves it by the rate
    against its gradient's sign; a constant gradient keeps doing so; L2 enters the matrices only;
    table entries with a zero gradient are skipped (value, moments and count), matrix entries are
    not; each parameter's bias correction uses its own count; epsilon sits outside the
    bias-corrected square root."""
    train = small[0]
    rng = np.random.default_rng(1)
    w, g = rng.normal(size=100), rng.normal(size=100)
    z = np.zeros(100)
    for matrix in (False, True):
        w1, m1, v1, s1 = train.adam_step(w, g, z, z, 0, lr=0.01, l2_reg=0.0, eps=0.0, matrix=matrix)
        assert np.allclose(w1, w - 0.01 * np.sign(g))  # first step: m^ = g, v^ = g^2
        assert np.allclose(m1, 0.1 * g) and np.allclose(v1, 0.001 * g * g) and np.all(s1 == 1)
        w2, _, _, s2 = train.adam_step(w1, g, m1, v1, s1, lr=0.01, l2_reg=0.0, eps=0.0, matrix=matrix)
        assert np.allclose(w2, w1 - 0.01 * np.sign(g)) and np.all(s2 == 2)  # constant gradient
    # L2 regularisation enters the matrices as g + l2 * w, and not the tables.
    tiny = np.full(100, 1e-30)
    w3, _, _, _ = train.adam_step(w, tiny, z, z, 0, lr=0.01, l2_reg=1.0, eps=0.0, matrix=True)
    assert np.allclose(w3, w - 0.01 * np.sign(w))
    w3t, _, _, _ = train.adam_step(w, tiny, z, z, 0, lr=0.01, l2_reg=1.0, eps=0.0, matrix=False)
    assert np.allclose(w3t, w - 0.01)  # sign of the tiny positive gradient only
    # Zero gradients: tables keep everything; matrices update from their decaying moments.
    g4 = g.copy()
    g4[::2] = 0.0
    w4, m4, v4, s4 = train.adam_step(w, g4, z, z, 0, lr=0.01)
    assert np.array_equal(w4[::2], w[::2]) and not m4[::2].any() and not v4[::2].any() and not s4[::2].any()
    assert np.allclose(w4[1::2], w[1::2] - 0.01 * np.sign(g[1::2]), atol=1e-6) and np.all(s4[1::2] == 1)
    m0, v0 = np.full(100, 0.5), np.full(100, 0.25)
    w5, m5, v5, s5 = train.adam_step(w, z, m0, v0, 3, lr=0.01, l2_reg=0.0, eps=0.0, matrix=True)
    rate = 0.01 * np.sqrt(1 - 0.999 ** 4) / (1 - 0.9 ** 4)
    assert np.allclose(m5, 0.45) and np.allclose(v5, 0.999 * 0.25) and np.all(s5 == 4)
    assert np.allclose(w5, w - rate * 0.45 / np.sqrt(0.999 * 0.25))
    # Per-parameter counts: an entry first touched at its own update 1 moves by the full rate,
    # whatever the optimiser's global step (a global count of 5 would move it by 0.01 *
    # sqrt(1 - 0.999^5) / (1 - 0.9^5) * 0.1 / sqrt(0.001), about 0.0049).
    w6, _, _, _ = train.adam_step(w, g, z, z, 0, lr=0.01, l2_reg=0.0, eps=0.0)
    assert np.allclose(np.abs(w6 - w), 0.01)
    # Epsilon: lr * sqrt(1 - b2) / (1 - b1) * m / (sqrt(v) + eps) at the first update, here with
    # eps = 1 and g = 1: 0.01 * sqrt(0.001) / (sqrt(0.001) + 1); the textbook form
    # m^ / (sqrt(v^) + eps) would give 0.01 * 1 / (1 + 1).
    w7, _, _, _ = train.adam_step(np.zeros(1), np.ones(1), np.zeros(1), np.zeros(1), 0, lr=0.01, l2_reg=0.0,
                                  eps=1.0)
    assert np.isclose(-w7[0], 0.01 * np.sqrt(0.001) / (np.sqrt(0.001) + 1.0))


def test_lr_decay_schedule():
    """The exponential decay's factor per update (from 1): x base after update decay_start and
    every decay_interval updates after it, worked by hand for small schedules and the
    reference's defaults (10000, 2000, 0.33)."""
    from oracle import train
    assert [train.lr_factor(t, 2, 2, 0.5) for t in range(1, 8)] == [1, 1, 0.5, 0.5, 0.25, 0.25, 0.125]
    assert [train.lr_factor(t, 0, 2, 0.5) for t in range(1, 7)] == [1, 1, 0.5, 0.5, 0.25, 0.25]
    assert [train.lr_factor(t, 10000, 2000, 0.33) for t in (1, 10000, 10001, 12000, 12001)] == \
        [1, 1, 0.33, 0.33, 0.33 ** 2]
    assert train.lr_factor(50, 1, 0, 0.5) == 1.0
