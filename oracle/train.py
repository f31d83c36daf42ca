"""FP64 restatement of one training step of the network (test infrastructure only).

Only tests/ use this module, as the checker of the GPU trainer. It restates, in numpy and double
precision:
  - the hash-grid encoding with its corner indices and trilinear weights (the oracle's grid
    convention, oracle/nbvh_oracle.cpp grid_point), grid positions rounded to FP32 as the FP32
    kernels form them, everything after them in FP64;
  - the MLP forward pass (no biases, ReLU hidden layers, Sigmoid on all outputs);
  - the reference's composite loss with its default configuration
    (include/cuda/neural/utils/pt_composite_loss.cuh:19-123, defaults
    include/cuda/neural/neural_bvh_module.h:250-266, loss functions
    include/cuda/neural/utils/loss_gradient.cuh): visibility BCE, depth L1, normal L1 on the
    linearised output, albedo relative L2, all but visibility masked where the target
    visibility is < 0.5, each element's loss and gradient divided by n * 8;
  - backpropagation to the MLP matrices and the grid tables;
  - one Adam step (Kingma and Ba 2015, Algorithm 1) with L2 regularisation added to the gradient,
    the reference's optimiser defaults (include/cuda/neural/utils/tcnn_config.h:153-166) and
    its exponential learning-rate decay (lr_factor); parameters with an exactly zero gradient
    are left alone (sparse hash-table updates).
The gradients are pinned by central finite differences of the loss (tests/test_oracle_train.py).
"""
import numpy as np

P1 = np.uint64(2654435761)
P2 = np.uint64(805459861)
M32 = np.uint64(0xFFFFFFFF)


def _corner_index(lv, l, g):
    """Corner indices [n, 8] and the offset of level l, g = integer cell [n, 3] (int64)."""
    res, size, dense = int(lv["res"][l]), int(lv["size"][l]), int(lv["dense"][l])
    out = np.zeros((g.shape[0], 8), np.int64)
    for c in range(8):
        gc = [((g[:, d] + ((c >> d) & 1)) & 0xFFFFFFFF).astype(np.uint64) for d in range(3)]
        if dense:
            i = (gc[0] + gc[1] * np.uint64(res) + gc[2] * np.uint64(res * res)) & M32
            out[:, c] = (i % np.uint64(size)).astype(np.int64)
        else:
            h = gc[0] ^ ((gc[1] * P1) & M32) ^ ((gc[2] * P2) & M32)
            out[:, c] = (h & np.uint64(size - 1)).astype(np.int64)
    return out


def encode(lv, x):
    """Per (point, level) corner indices [n, 8] (table entries, level offset added) and trilinear
    weights [n, 8] of the encoded points x [n, 3P]."""
    n, P, L = x.shape[0], x.shape[1] // 3, len(lv["res"])
    corners = []
    for p in range(P):
        pts = x[:, 3 * p:3 * p + 3].astype(np.float64)
        for l in range(L):
            # Grid position as the FP32 kernels form it, fmaf(scale, x, 0.5f) rounded once to
            # float (the product of two floats is exact in double): at the finest levels the
            # fraction keeps only ~11 bits, the same for every FP32 implementation, so the
            # restatement takes the positions in FP32 and does the rest in FP64.
            pos32 = (np.float64(np.float32(lv["scale"][l])) * pts.astype(np.float32).astype(np.float64)
                     + 0.5).astype(np.float32)
            g = np.floor(pos32).astype(np.float64)
            fr = (pos32 - np.floor(pos32)).astype(np.float64)
            idx = _corner_index(lv, l, g.astype(np.int64)) + int(lv["offset"][l])
            w = np.ones((n, 8))
            for c in range(8):
                for d in range(3):
                    w[:, c] *= fr[:, d] if (c >> d) & 1 else 1.0 - fr[:, d]
            corners.append((p, l, idx, w))
    return corners


def features(tables, F, corners, n, P, L):
    """Features [n, P*L*F], feature k = p*L*F + l*F + f."""
    t = tables.reshape(-1, F).astype(np.float64)
    feats = np.zeros((n, P * L * F))
    for p, l, idx, w in corners:
        k = p * L * F + l * F
        feats[:, k:k + F] = np.einsum("nc,ncf->nf", w, t[idx])
    return feats


def split_mlp(mlp, K0, H, n_hidden, O):
    mats, off, fan_in = [], 0, K0
    for layer in range(n_hidden + 1):
        out = O if layer == n_hidden else H
        mats.append(mlp[off:off + out * fan_in].astype(np.float64).reshape(out, fan_in))
        off += out * fan_in
        fan_in = H
    return mats


def composite_loss(y, t, albedo_den=None):
    """Per-element loss and gradient [n, 8] of the reference's default composite loss. The
    relative-L2 gradient treats its denominator y^2 + 1e-2 as a constant (loss_gradient.cuh), so
    the gradient is the derivative of the loss with the denominator frozen at its value;
    albedo_den, when given, freezes it at that value for the loss (the finite-difference pin)."""
    n = y.shape[0]
    loss, grad = np.zeros_like(y), np.zeros_like(y)
    p = np.clip(y[:, 0], 1e-5, 1.0 - 1e-5)
    loss[:, 0] = -(t[:, 0] * np.log(p) + (1.0 - t[:, 0]) * np.log(1.0 - p))
    grad[:, 0] = -(t[:, 0] - p) / (p - p * p)
    keep = t[:, 0] >= 0.5
    d = y[:, 1] - t[:, 1]
    loss[:, 1], grad[:, 1] = np.abs(d), np.copysign(1.0, d)
    lin = ((y[:, 2:5] * 2.0 - 1.0) / 0.1) * 0.5 + 0.5
    d = lin - t[:, 2:5]
    loss[:, 2:5], grad[:, 2:5] = np.abs(d), np.copysign(1.0, d) / 0.1
    d = y[:, 5:8] - t[:, 5:8]
    den = y[:, 5:8] * y[:, 5:8] + 1e-2 if albedo_den is None else albedo_den
    loss[:, 5:8], grad[:, 5:8] = d * d / den, 2.0 * d / den
    loss[~keep, 1:] = 0.0
    grad[~keep, 1:] = 0.0
    return loss / (n * 8), grad / (n * 8)


def forward(lv, cfg, tables, mlp, x):
    """Corners, layer activations [features, h1..h_H] and outputs y (FP64)."""
    F, L, P = cfg.grid.n_features_per_level, cfg.grid.n_levels, cfg.n_points
    corners = encode(lv, x)
    a = [features(tables, F, corners, x.shape[0], P, L)]
    W = split_mlp(mlp, P * L * F, cfg.n_neurons, cfg.n_hidden_layers, cfg.n_outputs)
    for layer in range(cfg.n_hidden_layers):
        a.append(np.maximum(a[-1] @ W[layer].T, 0.0))
    z = a[-1] @ W[cfg.n_hidden_layers].T
    return corners, a, W, 1.0 / (1.0 + np.exp(-z))


def train_loss(lv, cfg, tables, mlp, x, targets, albedo_den=None):
    """The loss alone (sum of the per-element values); albedo_den freezes the relative-L2
    denominators (see composite_loss)."""
    _, _, _, y = forward(lv, cfg, tables, mlp, x)
    return float(composite_loss(y, np.asarray(targets, np.float64), albedo_den)[0].sum())


def train_gradients(lv, cfg, tables, mlp, x, targets):
    """Loss (sum of the per-element values) and its gradients w.r.t. tables and MLP (FP64)."""
    F, L, P = cfg.grid.n_features_per_level, cfg.grid.n_levels, cfg.n_points
    H, nh, O = cfg.n_neurons, cfg.n_hidden_layers, cfg.n_outputs
    corners, a, W, y = forward(lv, cfg, tables, mlp, x)
    values, dy = composite_loss(y, np.asarray(targets, np.float64))
    delta = dy * y * (1.0 - y)
    dW = [None] * (nh + 1)
    for layer in range(nh, -1, -1):
        dW[layer] = delta.T @ a[layer]
        da = delta @ W[layer]
        if layer > 0:
            delta = da * (a[layer] > 0)
    dx = da
    dtab = np.zeros(tables.size)
    dt = dtab.reshape(-1, F)
    for p, l, idx, w in corners:
        k = p * L * F + l * F
        np.add.at(dt, idx.reshape(-1), (w[:, :, None] * dx[:, None, k:k + F]).reshape(-1, F))
    dmlp = np.concatenate([g.reshape(-1) for g in dW])
    return float(values.sum()), dtab, dmlp, y


def lr_factor(step, decay_start, decay_interval, decay_base):
    """Learning-rate factor of update number `step` (from 1) under the reference's exponential
    decay (src/cuda/neural/neural_bvh_module.cu:146-150, tcnn's ExponentialDecay wrapper around
    Adam): decay_base to the number of earlier updates s >= decay_start with
    (s - decay_start) % decay_interval == 0, counted one by one. Which update first uses a decayed
    rate is tcnn's convention, absent here: parity unpinned for that off-by-one."""
    if decay_interval == 0:
        return 1.0
    k = sum(1 for s in range(1, step) if s >= decay_start and (s - decay_start) % decay_interval == 0)
    return decay_base ** k


def adam_step(w, g, m, v, steps, lr=1e-2, beta1=0.9, beta2=0.999, eps=1e-8, l2_reg=1e-8, matrix=False):
    """One step of tiny-cuda-nn's AdamOptimizer, the reference's optimiser (public
    tiny-cuda-nn optimizers/adam.h; tcnn is an empty submodule in the reference, so this restates
    the published code): `steps` holds each parameter's count of earlier updates (array or
    scalar). matrix=True (the MLP's matrices): L2 regularisation is added to the gradient and the
    parameter is updated every step, also with a zero gradient. matrix=False (the grid tables):
    no L2, and parameters whose gradient is exactly zero keep their value, moments and count.
    The update is w -= lr * sqrt(1 - beta2^t) / (1 - beta1^t) * m / (sqrt(v) + eps) with t the
    parameter's own update number. Returns new (w, m, v, steps)."""
    steps = np.broadcast_to(np.asarray(steps, dtype=np.int64), np.shape(w))
    if matrix:
        live = np.ones(np.shape(w), dtype=bool)
        g2 = g + l2_reg * w
    else:
        live = g != 0
        g2 = g
    m2 = beta1 * m + (1.0 - beta1) * g2
    v2 = beta2 * v + (1.0 - beta2) * g2 * g2
    t = steps + 1
    rate = lr * np.sqrt(1.0 - beta2 ** t) / (1.0 - beta1 ** t)
    w2 = w - rate * m2 / (np.sqrt(v2) + eps)
    return np.where(live, w2, w), np.where(live, m2, m), np.where(live, v2, v), np.where(live, t, steps)
