"""The array-factorised DFT predict on the matrix cores against the FP64
restatement of dft_actor (sdp_synth.dft_f64 on the baseline uvw, which are
the differences of the antenna uvw), for BF16 and FP16 operands with and
without the hi + lo split.

Bars (largest error over all visibilities relative to sum_s |g_s|, the
largest possible |V|), from the operand unit roundoff u (BF16 2^-9, FP16
2^-12): one product per term, 2u + FP32 accumulation; split, 3u^2 plus the
FP32 phasor and accumulation error."""
import numpy as np
import pytest

import sdp_synth as S

pytestmark = pytest.mark.gpu

BARS = {("bf16", True): 2e-5, ("fp16", True): 1e-6, ("bf16", False): 5e-3, ("fp16", False): 6e-4}


def array_case(n_ant, n_time, n_src, seed):
    rng = np.random.default_rng(seed)
    enu = S.log_radial_layout(n_ant, 20.0, 3000.0, seed=seed)
    ant = S.antenna_uvw(enu, -30.0, -30.0, np.linspace(-3.0, 3.0, n_time), 150e6)
    lm = rng.uniform(-0.03, 0.03, (n_src, 2))
    flux = rng.uniform(-0.2, 1.0, n_src)                  # model components may be negative
    src = np.concatenate([lm, flux[:, None]], axis=1).astype(np.float32)
    pairs = S.receiver_pairs(n_ant)
    uvw = (ant[:, pairs[:, 1]] - ant[:, pairs[:, 0]]).reshape(-1, 3)
    ref = S.dft_f64(uvw, src)
    l, m, f = (src[:, k].astype(np.float64) for k in range(3))
    gsum = np.abs(f / (1 - 0.5 * (l * l + m * m))).sum()
    return ant, src, ref, gsum


def run(hip, ant, src, flags):
    n_time, n_ant = ant.shape[0], ant.shape[1]
    n_vis = n_time * n_ant * (n_ant - 1) // 2
    d_ant = hip.DeviceBuffer.from_array(np.ascontiguousarray(ant, np.float64))
    d_src = hip.DeviceBuffer.from_array(src)
    d_out = hip.DeviceBuffer(8 * max(n_vis, 1))
    d_out.zero()
    hip.dft_predict_array(n_time, n_ant, d_ant, src.shape[0], d_src, d_out, flags)
    return d_out.download((n_vis, 2), np.float32)


@pytest.mark.parametrize("prec", ["bf16", "fp16"])
@pytest.mark.parametrize("split", [True, False])
def test_array_dft_matches_fp64(hip, prec, split):
    ant, src, ref, gsum = array_case(150, 5, 77, seed=31)   # 3 antenna tiles, ragged source chunk
    flags = (hip.DFT_ARRAY_FP16 if prec == "fp16" else hip.DFT_ARRAY_BF16) | (0 if split else hip.DFT_ARRAY_NO_SPLIT)
    got = run(hip, ant, src, flags)
    err = np.abs(got - ref).max() / gsum
    print(f"{prec} split={split}: max error {err:.3g} of sum|g|")
    assert err <= BARS[(prec, split)]
    if split:
        # not merely within the bar: the split must beat one product by far
        one = run(hip, ant, src, flags | hip.DFT_ARRAY_NO_SPLIT)
        assert err * 30 < np.abs(one - ref).max() / gsum


def test_array_dft_fp16_scale_keeps_bright_and_faint_sources(hip):
    """Fluxes spanning 1e-4 .. 1e3 Jy: the power-of-two scale keeps FP16
    operands finite and the split precise."""
    ant, src, _, _ = array_case(64, 2, 40, seed=5)
    src[:, 2] = np.geomspace(1e-4, 1e3, src.shape[0]).astype(np.float32)
    pairs = S.receiver_pairs(64)
    ref = S.dft_f64((ant[:, pairs[:, 1]] - ant[:, pairs[:, 0]]).reshape(-1, 3), src)
    l, m, f = (src[:, k].astype(np.float64) for k in range(3))
    gsum = np.abs(f / (1 - 0.5 * (l * l + m * m))).sum()
    got = run(hip, ant, src, hip.DFT_ARRAY_FP16)
    assert np.isfinite(got).all()
    assert np.abs(got - ref).max() / gsum <= BARS[("fp16", True)]


def test_array_dft_agrees_with_per_visibility_kernel(hip):
    """The factorised form and the per-visibility FP64-phase kernel compute
    the same sum (the latter from float32 baseline uvw)."""
    ant, src, ref, gsum = array_case(70, 3, 33, seed=8)
    pairs = S.receiver_pairs(70)
    uvw = (ant[:, pairs[:, 1]] - ant[:, pairs[:, 0]]).reshape(-1, 3).astype(np.float32)
    d_uvw, d_src = hip.DeviceBuffer.from_array(uvw), hip.DeviceBuffer.from_array(src)
    d_out = hip.DeviceBuffer(8 * uvw.shape[0])
    hip.dft_predict(uvw.shape[0], src.shape[0], d_src, d_uvw, d_out, hip.DFT_FP64_PHASE)
    per_vis = d_out.download((uvw.shape[0], 2), np.float32)
    got = run(hip, ant, src, hip.DFT_ARRAY_FP16)
    assert np.abs(got - per_vis).max() / gsum <= 2e-5


def test_array_dft_degenerate_inputs(hip):
    ant, src, _, _ = array_case(20, 2, 5, seed=3)
    assert np.all(run(hip, ant, src[:0], hip.DFT_ARRAY_BF16) == 0.0)      # no sources: zero visibilities
    d = hip.DeviceBuffer(8)
    hip.dft_predict_array(2, 1, hip.DeviceBuffer(48), 5, hip.DeviceBuffer.from_array(src), d)   # one antenna: nothing
    with pytest.raises(hip.SdpError):
        hip.dft_predict_array(2, 20, hip.DeviceBuffer(8), 5, hip.DeviceBuffer.from_array(src), d, flags=8)
