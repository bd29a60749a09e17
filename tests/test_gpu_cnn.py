"""The learned proposal network's HIP kernels against its FP64 restatement
(oracle/cnn_oracle.py), through the C-ABI: forward pass, parameter
gradients and the Adam step.

Bars: FP32 mode within 1e-5 (forward) and 1e-4 (gradients) of the largest
magnitude; BF16 matrix-core mode within 3e-2 of the largest magnitude with
gradient cosine similarity above 0.999."""
import numpy as np
import pytest

import cnn_oracle as C

pytestmark = pytest.mark.gpu


def up(hip, a):
    return hip.DeviceBuffer.from_array(np.ascontiguousarray(a, dtype=np.float32))


def setup(hip, levels, base, H, B, precision, seed=3):
    rng = np.random.default_rng(seed)
    net = hip.Cnn(levels, base, H, H, B, precision=precision)
    params = C.init_params(levels, base, seed=seed)
    params = {k: (v + 0.05 * rng.standard_normal(v.shape)).astype(np.float32) for k, v in params.items()}
    net.set_params(params)
    x = rng.standard_normal((B, H, H, 2)).astype(np.float32)
    t = rng.standard_normal((B, H, H, 1)).astype(np.float32)
    return net, params, x, t


def run(hip, net, x, t):
    B, H = x.shape[0], x.shape[1]
    d_x, d_t = up(hip, x), up(hip, t)
    d_y, d_g = hip.DeviceBuffer(4 * B * H * H), hip.DeviceBuffer(4 * B * H * H)
    net.forward(d_x, d_y, B)
    loss = hip.cnn_mse(B * H * H, d_y, d_t, d_g)
    net.backward(d_g)
    hip.synchronize()
    return d_y.download((B, H, H, 1), np.float32), loss, net.unflatten(net.get_grads()), (d_x, d_t, d_y, d_g)


@pytest.mark.parametrize("levels,base,H", [(1, 4, 8), (2, 3, 16), (3, 4, 16)])
def test_fp32_forward_and_gradients_match_oracle(hip, levels, base, H):
    net, params, x, t = setup(hip, levels, base, H, 2, hip.CNN_FP32)
    y, loss, grads, _ = run(hip, net, x, t)
    yr, act = C.forward(params, x, levels)
    lr, gy = C.mse(yr, t)
    gr = C.backward(params, act, gy, levels)
    assert np.abs(y - yr).max() <= 1e-5 * np.abs(yr).max()
    assert loss == pytest.approx(lr, rel=1e-5)
    for k in gr:
        scale = np.abs(gr[k]).max()
        assert np.abs(grads[k] - gr[k]).max() <= 1e-4 * scale + 1e-9, k


def test_bf16_forward_and_gradients_track_oracle(hip):
    levels, base, H = 3, 8, 32
    net, params, x, t = setup(hip, levels, base, H, 2, hip.CNN_BF16)
    y, loss, grads, _ = run(hip, net, x, t)
    yr, act = C.forward(params, x, levels)
    lr, gy = C.mse(yr, t)
    gr = C.backward(params, act, gy, levels)
    assert np.abs(y - yr).max() <= 3e-2 * np.abs(yr).max()
    for k in gr:
        a, b = grads[k].ravel().astype(np.float64), gr[k].ravel()
        cos = a @ b / (np.linalg.norm(a) * np.linalg.norm(b) + 1e-30)
        assert cos > 0.999, (k, cos)


@pytest.mark.parametrize("levels,base", [(4, 32), (3, 16)])
def test_bf16_full_size_network_matches_fp32(hip, levels, base):
    """4 levels, 32 base channels (1.75M parameters), and the 3-level,
    16-channel network the learned cycle uses (whose full-resolution layers
    take the 16-channel forward kernel), on 64 x 64 patches: every GEMM
    shape of the real networks, BF16 against FP32."""
    outs, grads = [], []
    for prec in (hip.CNN_FP32, hip.CNN_BF16):
        net, params, x, t = setup(hip, levels, base, 64, 2, prec, seed=9)
        y, loss, g, _ = run(hip, net, x, t)
        outs.append(y)
        grads.append(net.get_grads())
        net.close()
    assert np.abs(outs[1] - outs[0]).max() <= 3e-2 * np.abs(outs[0]).max()
    a, b = grads[1].astype(np.float64), grads[0].astype(np.float64)
    assert a @ b / (np.linalg.norm(a) * np.linalg.norm(b)) > 0.999


def bf16_round(a):
    """Round to BF16 (nearest, ties to even), returned as float32."""
    u = np.ascontiguousarray(a, dtype=np.float32).view(np.uint32)
    r = (u + np.uint32(0x7FFF) + ((u >> 16) & np.uint32(1))) & np.uint32(0xFFFF0000)
    return r.view(np.float32)


def forward_bf16_operands(params, x, levels):
    """The oracle's network with every convolution's input and weights
    rounded to BF16 and activations kept in FP32, as the matrix-core kernels
    compute it (their BF16 products are exact; their sums FP32)."""
    f32 = lambda a: np.asarray(a, np.float32)
    conv = lambda a, name: f32(C.conv3x3(bf16_round(a), bf16_round(params[name + ".w"]),
                                         np.asarray(params[name + ".b"], np.float64)))
    xl, enc = f32(x), []
    for l in range(levels):
        e = C.relu(conv(C.relu(conv(xl, f"enc{l}a")), f"enc{l}b"))
        enc.append(e)
        if l < levels - 1:
            xl = f32(C.avgpool2(e))
    d = enc[-1]
    for l in range(levels - 2, -1, -1):
        s = C.relu(f32(conv(f32(C.upsample2(d)), f"dec{l}u") + enc[l]))
        d = C.relu(conv(s, f"dec{l}d"))
    return conv(d, "head")


@pytest.mark.parametrize("levels,base,bar", [(1, 32, 3e-3), (4, 32, 2e-2), (3, 16, 2e-2)])
def test_bf16_forward_matches_bf16_operand_restatement(hip, levels, base, bar):
    """The BF16 forward pass (the tiled matrix-core convolutions for layers
    of 32k input channels, the zero-padded input channels of the 2-channel
    first layer, the padded 16-channel tile for the 1-channel head) against
    the oracle with the same BF16
    rounding of every convolution's operands, at 128^2 (every level's tile
    fits).  Only summation order separates them, amplified where an
    activation lands on the other side of a BF16 rounding step in the next
    layer: the same restatement summed in FP32 instead of FP64 moves by
    9e-4 of the peak for the 3-layer network and by about 1e-2 for the deep
    ones, which sets the bars (an indexing error is of order 1)."""
    net, params, x, _ = setup(hip, levels, base, 128, 1, hip.CNN_BF16, seed=11)
    d_x, d_y = up(hip, x), hip.DeviceBuffer(4 * 128 * 128)
    net.forward(d_x, d_y, 1)
    y = d_y.download((1, 128, 128, 1), np.float32)
    ref = forward_bf16_operands(params, x, levels)
    err = np.abs(y - ref).max() / np.abs(ref).max()
    assert err <= bar, err
    net.close()


def test_bf16_forward_non_square_patch(hip):
    """A 24 x 32 patch: the full-resolution layers take the tiled kernel with
    8-row tiles (24 is not a multiple of 16), the half-resolution level (12 x 16)
    the general kernels; against the BF16-operand restatement."""
    levels, base, H, W = 2, 32, 24, 32
    rng = np.random.default_rng(21)
    net = hip.Cnn(levels, base, H, W, 2, precision=hip.CNN_BF16)
    params = C.init_params(levels, base, seed=21)
    params = {k: (v + 0.05 * rng.standard_normal(v.shape)).astype(np.float32) for k, v in params.items()}
    net.set_params(params)
    x = rng.standard_normal((2, H, W, 2)).astype(np.float32)
    d_x, d_y = up(hip, x), hip.DeviceBuffer(4 * 2 * H * W)
    net.forward(d_x, d_y, 2)
    y = d_y.download((2, H, W, 1), np.float32)
    ref = forward_bf16_operands(params, x, levels)
    err = np.abs(y - ref).max() / np.abs(ref).max()
    assert err <= 1e-2, err
    net.close()


def test_adam_step_matches_oracle(hip):
    net, params, x, t = setup(hip, 2, 4, 16, 2, hip.CNN_FP32)
    _, _, grads, _ = run(hip, net, x, t)
    p0 = net.get_params().astype(np.float64)
    g = net.get_grads().astype(np.float64)
    net.adam(1e-3)
    net.adam(1e-3)   # second step with the same gradients
    got = net.get_params()
    ref, m, v = C.adam_step(p0, g, np.zeros_like(p0), np.zeros_like(p0), 1, 1e-3)
    ref, m, v = C.adam_step(ref, g, m, v, 2, 1e-3)
    np.testing.assert_allclose(got, ref, rtol=1e-5, atol=1e-7)


def test_training_reduces_loss(hip):
    """A few hundred Adam steps on a fixed batch fit it: the loss falls by
    more than 10x (the kernels train, end to end, in BF16)."""
    net, params, x, t = setup(hip, 3, 8, 32, 4, hip.CNN_BF16)
    t = np.tanh(x[..., :1] + 0.5 * x[..., 1:2])   # a learnable target
    B, H = 4, 32
    d_x, d_t = up(hip, x), up(hip, t)
    d_y, d_g = hip.DeviceBuffer(4 * B * H * H), hip.DeviceBuffer(4 * B * H * H)
    losses = []
    for step in range(300):
        net.forward(d_x, d_y, B)
        losses.append(hip.cnn_mse(B * H * H, d_y, d_t, d_g))
        net.backward(d_g)
        net.adam(2e-3)
    assert losses[-1] < 0.1 * losses[0], (losses[0], losses[-1])


def test_fp32_configured_network_at_256_matches_oracle(hip):
    """The configured proposal network, 4 levels x 32 base channels
    (1,753,313 parameters), on one 256 x 256 (residual, PSF) patch in FP32
    mode: forward pass within 1e-5 and every parameter gradient within 1e-4
    of its largest magnitude of the FP64 oracle."""
    levels, base, H = 4, 32, 256
    net, params, x, t = setup(hip, levels, base, H, 1, hip.CNN_FP32, seed=17)
    assert sum(v.size for v in params.values()) == 1_753_313
    y, loss, grads, _ = run(hip, net, x, t)
    yr, act = C.forward(params, x, levels)
    lr, gy = C.mse(yr, t)
    gr = C.backward(params, act, gy, levels)
    e_fwd = np.abs(y - yr).max() / np.abs(yr).max()
    e_grad = max(np.abs(grads[k] - gr[k]).max() / (np.abs(gr[k]).max() + 1e-30) for k in gr)
    print(f"4x32 at 256^2, FP32: forward {e_fwd:.2e}, worst gradient {e_grad:.2e} of its largest magnitude")
    assert e_fwd <= 1e-5
    assert loss == pytest.approx(lr, rel=1e-5)
    for k in gr:
        assert np.abs(grads[k] - gr[k]).max() <= 1e-4 * np.abs(gr[k]).max() + 1e-9, k
    net.close()
