"""FP64 CPU restatement of the learned minor-cycle proposal network.

TEST INFRASTRUCTURE ONLY: tests/ use it to check the HIP network's forward
pass, its gradients and its optimiser step.  The reference has no learned
component; this file defines the network the HIP kernels implement and is
itself pinned by finite differences (tests/test_cnn_oracle.py).

Network (encoder-decoder, `levels` levels, `base` channels at level 0,
doubling per level; tensors NHWC):

    encoder level l:  e_l = relu(conv(relu(conv(x_l, W_la) + b_la), W_lb) + b_lb)
                      x_{l+1} = avgpool2(e_l)            (l < levels - 1)
    decoder level l:  u_l = conv(upsample2(d_{l+1}), W_lu) + b_lu     (d_top = e_top)
                      d_l = relu(conv(relu(u_l + e_l), W_ld) + b_ld)
    head:             y = conv(d_0, W_h) + b_h

conv is a 3x3 cross-correlation with zero padding 1:
    out[n, y, x, o] = b[o] + sum_{ky, kx, i} in[n, y + ky - 1, x + kx - 1, i] W[ky, kx, i, o]
"""
from __future__ import annotations

import numpy as np


def conv3x3(x, w, b):
    n, h, wd, ci = x.shape
    co = w.shape[3]
    xp = np.zeros((n, h + 2, wd + 2, ci), x.dtype)
    xp[:, 1:-1, 1:-1] = x
    out = np.broadcast_to(b, (n, h, wd, co)).astype(np.float64).copy()
    for ky in range(3):
        for kx in range(3):
            out += (xp[:, ky:ky + h, kx:kx + wd].reshape(-1, ci) @ w[ky, kx]).reshape(n, h, wd, co)
    return out


def conv3x3_backward(x, w, gout):
    """Gradients of conv3x3 with respect to x, w and b."""
    n, h, wd, ci = x.shape
    xp = np.zeros((n, h + 2, wd + 2, ci), np.float64)
    xp[:, 1:-1, 1:-1] = x
    gw = np.zeros_like(w, dtype=np.float64)
    gxp = np.zeros_like(xp)
    co = gout.shape[3]
    g2 = gout.reshape(-1, co)
    for ky in range(3):
        for kx in range(3):
            gw[ky, kx] = xp[:, ky:ky + h, kx:kx + wd].reshape(-1, ci).T @ g2
            gxp[:, ky:ky + h, kx:kx + wd] += (g2 @ np.asarray(w[ky, kx], np.float64).T).reshape(n, h, wd, ci)
    gb = gout.sum(axis=(0, 1, 2))
    return gxp[:, 1:-1, 1:-1], gw, gb


def avgpool2(x):
    n, h, w, c = x.shape
    return x.reshape(n, h // 2, 2, w // 2, 2, c).mean(axis=(2, 4))


def avgpool2_backward(g):
    return np.repeat(np.repeat(g, 2, axis=1), 2, axis=2) / 4.0


def upsample2(x):
    return np.repeat(np.repeat(x, 2, axis=1), 2, axis=2)


def upsample2_backward(g):
    n, h, w, c = g.shape
    return g.reshape(n, h // 2, 2, w // 2, 2, c).sum(axis=(2, 4))


def relu(x):
    return np.maximum(x, 0.0)


def layer_shapes(levels, base, cin=2, cout=1):
    """Ordered (name, cin, cout) of every conv: the parameter layout."""
    ch = [base << l for l in range(levels)]
    shapes = []
    prev = cin
    for l in range(levels):
        shapes.append((f"enc{l}a", prev, ch[l]))
        shapes.append((f"enc{l}b", ch[l], ch[l]))
        prev = ch[l]
    for l in range(levels - 2, -1, -1):
        shapes.append((f"dec{l}u", ch[l + 1], ch[l]))
        shapes.append((f"dec{l}d", ch[l], ch[l]))
    shapes.append(("head", ch[0], cout))
    return shapes


def init_params(levels, base, seed, cin=2, cout=1):
    """He-normal weights (fan-in 9 cin), zero biases; float32."""
    rng = np.random.default_rng(seed)
    params = {}
    for name, ci, co in layer_shapes(levels, base, cin, cout):
        params[name + ".w"] = (rng.standard_normal((3, 3, ci, co)) * np.sqrt(2.0 / (9 * ci))).astype(np.float32)
        params[name + ".b"] = np.zeros(co, np.float32)
    return params


def forward(params, x, levels):
    """Returns the output and the saved activations for backward."""
    p = {k: v.astype(np.float64) for k, v in params.items()}
    act = {"x0": x.astype(np.float64)}
    xl = act["x0"]
    enc = []
    for l in range(levels):
        a = relu(conv3x3(xl, p[f"enc{l}a.w"], p[f"enc{l}a.b"]))
        act[f"enc{l}a"] = a
        e = relu(conv3x3(a, p[f"enc{l}b.w"], p[f"enc{l}b.b"]))
        act[f"enc{l}b"] = e
        enc.append(e)
        if l < levels - 1:
            xl = avgpool2(e)
            act[f"x{l + 1}"] = xl
    d = enc[-1]
    for l in range(levels - 2, -1, -1):
        up = upsample2(d)
        act[f"up{l}"] = up
        u = conv3x3(up, p[f"dec{l}u.w"], p[f"dec{l}u.b"])
        s = relu(u + enc[l])
        act[f"sum{l}"] = s
        d = relu(conv3x3(s, p[f"dec{l}d.w"], p[f"dec{l}d.b"]))
        act[f"dec{l}d"] = d
    y = conv3x3(d, p["head.w"], p["head.b"])
    act["d0"] = d
    return y, act


def backward(params, act, gy, levels):
    """Gradients of every parameter given dL/dy."""
    p = {k: v.astype(np.float64) for k, v in params.items()}
    grads = {}
    gd, grads["head.w"], grads["head.b"] = conv3x3_backward(act["d0"], p["head.w"], gy)
    genc = [None] * levels
    # decoder, from level 0 upwards
    gcur = gd
    for l in range(0, levels - 1):
        gcur = gcur * (act[f"dec{l}d"] > 0)
        gs, grads[f"dec{l}d.w"], grads[f"dec{l}d.b"] = conv3x3_backward(act[f"sum{l}"], p[f"dec{l}d.w"], gcur)
        gs = gs * (act[f"sum{l}"] > 0)
        genc[l] = gs                                   # skip branch
        gup, grads[f"dec{l}u.w"], grads[f"dec{l}u.b"] = conv3x3_backward(act[f"up{l}"], p[f"dec{l}u.w"], gs)
        gcur = upsample2_backward(gup)                 # gradient w.r.t. d_{l+1}
    # encoder, from the top down (the top level feeds only the decoder)
    ge = gcur
    gx_next = None
    for l in range(levels - 1, -1, -1):
        if l < levels - 1:
            ge = genc[l] + avgpool2_backward(gx_next)
        ge = ge * (act[f"enc{l}b"] > 0)
        ga, grads[f"enc{l}b.w"], grads[f"enc{l}b.b"] = conv3x3_backward(act[f"enc{l}a"], p[f"enc{l}b.w"], ge)
        ga = ga * (act[f"enc{l}a"] > 0)
        xin = act[f"x{l}"]
        gx_next, grads[f"enc{l}a.w"], grads[f"enc{l}a.b"] = conv3x3_backward(xin, p[f"enc{l}a.w"], ga)
    return grads


def mse(y, target):
    """Mean squared error and its gradient."""
    d = y - target
    return float(np.mean(d * d)), 2.0 * d / d.size


def adam_step(param, grad, m, v, t, lr, beta1=0.9, beta2=0.999, eps=1e-8):
    """One Adam update (Kingma and Ba 2015, Algorithm 1), in FP64."""
    m = beta1 * m + (1 - beta1) * grad
    v = beta2 * v + (1 - beta2) * grad * grad
    mhat = m / (1 - beta1 ** t)
    vhat = v / (1 - beta2 ** t)
    return param - lr * mhat / (np.sqrt(vhat) + eps), m, v
