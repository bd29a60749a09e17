"""The synthetic workloads at their configured sizes, through the C-ABI,
against the oracle and its FP64 twins.

C1 (plain FFT imaging, 516,096 visibilities, 1024^2, h = 3, OS 16, 200
Hogbom iterations): gridding bit-exact with integer kernels and within 1e-6
of the peak with the real ones, the inverse FFT and the whole dirty image
against FP64, and 200 Hogbom iterations bit-exact.

C3 (precision variant: the C2 array's uvw with 4096 sources): the
matrix-core predict against the FP64 restatement, and the errors it leaves
in the dirty image and the residual image, each against the FP64 chain.
"""
import numpy as np
import pytest

import sdp_synth as S

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def c1():
    return S.config1()


def cfg_of(hip, d, **kw):
    return hip.make_config(d.grid_size, d.cell_size, oversampling=d.oversampling,
                           num_kernels=d.supports.shape[0], max_w=d.max_w, **kw)


def image_f64(oracle, G, grid64):
    """shift -> inverse FFT -> shift -> real part, in double"""
    yy, xx = np.mgrid[0:G, 0:G]
    sign = np.where((yy + xx) % 2 == 0, 1.0, -1.0).ravel()
    inv = oracle.fft2d_f64(G, grid64 * sign[:, None], +1)
    return inv[:, 0] * sign


def image_gpu(actors, G, grid, cfg):
    return actors.shift_c2r(G, actors.fft_inverse(G, actors.shift_c2c(G, grid, cfg)), cfg)


def test_config1_gridding_at_full_size(hip, actors, oracle, c1):
    """516,096 visibilities x 36 taps on the 1024^2 grid.  Integer-valued
    kernels and visibilities whose every partial sum stays below 2^24 make
    the float sums exact in any order, so the grid must be bit-identical
    to the oracle's; with the real kernels the grid is within 1e-6 of the
    peak of the FP64 sum and no further from it than the reference's
    sequential float sum (+1e-7 of the peak)."""
    d = c1
    G = d.grid_size
    cfg = cfg_of(hip, d)
    k = np.arange(d.kernels.shape[0])
    ikern = np.stack([(k % 5) - 2, (k % 3) - 1], axis=1).astype(np.float32)
    ivis = np.random.default_rng(11).integers(-1, 2, size=(d.n_vis, 2)).astype(np.float32)
    bound = oracle.std_gridding_f64(G, d.oversampling, np.abs(ikern).sum(1, keepdims=True) * [1, 0], d.supports,
                                    d.uvw, (np.abs(ivis).sum(1, keepdims=True) * [1, 0]).astype(np.float32),
                                    d.uv_scale, d.w_scale)
    assert bound.max() < 2 ** 24
    got = actors.std_gridding(G, d.oversampling, ikern, d.supports, d.uvw, ivis, cfg)
    ref = oracle.std_gridding(G, d.oversampling, ikern, d.supports, d.uvw, ivis, cfg)
    assert np.count_nonzero(ref) > 100_000
    np.testing.assert_array_equal(got, ref)

    got = actors.std_gridding(G, d.oversampling, d.kernels, d.supports, d.uvw, d.vis, cfg)
    ref32 = oracle.std_gridding(G, d.oversampling, d.kernels, d.supports, d.uvw, d.vis, cfg)
    ref64 = oracle.std_gridding_f64(G, d.oversampling, d.kernels, d.supports, d.uvw, d.vis, d.uv_scale, d.w_scale)
    peak = np.abs(ref64).max()
    err, err_ref = np.abs(got - ref64).max(), np.abs(ref32 - ref64).max()
    print(f"C1 grid: GPU {err / peak:.2e}, reference float order {err_ref / peak:.2e} of the peak")
    assert err <= 1e-6 * peak
    assert err <= err_ref + 1e-7 * peak


def test_config1_dirty_image_and_clean(hip, actors, oracle, c1):
    """The C1 dirty image (grid -> shift -> inverse FFT -> shift -> real ->
    taper correction) against the same chain in FP64: the inverse FFT of
    the GPU's own grid within 2e-6 of the image peak, the whole image
    within 1e-5 of the peak.  Then 200 Hogbom iterations at gain 0.1 on it
    with the array's own PSF: component list, model and residual bit-exact
    against the oracle."""
    d = c1
    G = d.grid_size
    cfg = cfg_of(hip, d)
    grid = actors.std_gridding(G, d.oversampling, d.kernels, d.supports, d.uvw, d.vis, cfg)
    img = image_gpu(actors, G, grid, cfg)
    ref_fft = image_f64(oracle, G, grid.astype(np.float64))
    peak = np.abs(ref_fft).max()
    e_fft = np.abs(img - ref_fft).max() / peak
    grid64 = oracle.std_gridding_f64(G, d.oversampling, d.kernels, d.supports, d.uvw, d.vis, d.uv_scale, d.w_scale)
    ref_img = image_f64(oracle, G, grid64)
    e_img = np.abs(img - ref_img).max() / np.abs(ref_img).max()
    print(f"C1 inverse FFT {e_fft:.2e}, dirty image {e_img:.2e} of the peak")
    assert e_fft <= 2e-6
    assert e_img <= 1e-5

    taper = S.taper_padded(G)
    dirty = actors.convolution_correction(G, img, taper[:-1].copy(), cfg)
    ones = np.tile(np.array([[1.0, 0.0]], np.float32), (d.n_vis, 1))
    psf_grid = actors.std_gridding(G, d.oversampling, d.kernels, d.supports, d.uvw, ones, cfg)
    psf_img = actors.convolution_correction(G, image_gpu(actors, G, psf_grid, cfg), taper[:-1].copy(), cfg)
    pmax = float(psf_img.max())
    psf = (psf_img.astype(np.float64) / pmax).astype(np.float32)
    ccfg = cfg_of(hip, d, psf_max_value=pmax)
    model = np.zeros(G * G, np.float32)
    ns, src, out_model = actors.hogbom_clean(G, 200, 400, dirty, psf, ccfg, model)
    rns, rsrc, rmodel, _ = oracle.hogbom_clean(G, 200, 400, dirty, np.concatenate([psf, [0.0]]).astype(np.float32),
                                               ccfg, model)
    assert rns == 200 and ns == rns
    np.testing.assert_array_equal(src, rsrc)
    np.testing.assert_array_equal(out_model, rmodel)


@pytest.fixture(scope="module")
def c3():
    """The first instant of the C2 array (32,640 baselines) and the 4096
    C3 sources; FP64 visibilities from the per-antenna phasors (checked
    against the direct restatement on a subset)."""
    d = S.config2()
    enu = d.extra["enu"]
    ant = S.antenna_uvw(enu, -30.0, -30.0, d.extra["ha_hours"][:1], 150e6)     # (1, 256, 3)
    src = S.config3_sources(4096)
    pairs = S.receiver_pairs(enu.shape[0])
    uvw = (ant[0, pairs[:, 1]] - ant[0, pairs[:, 0]])
    l, m, f = (src[:, k].astype(np.float64) for k in range(3))
    term = 0.5 * (l * l + m * m)
    g = f / (1.0 - term)
    ph = 2 * np.pi * (np.outer(ant[0, :, 0], l) + np.outer(ant[0, :, 1], m) - np.outer(ant[0, :, 2], term))
    E = np.exp(-1j * ph)                                    # (n_ant, n_src)
    V = (np.conj(E) * g) @ E.T                              # V[a, b] = sum_s g e^{-i (th_b - th_a)}
    vis64 = V[pairs[:, 0], pairs[:, 1]]
    vis64 = np.stack([vis64.real, vis64.imag], axis=1)
    sub = np.arange(0, uvw.shape[0], 97)
    direct = S.dft_f64(uvw[sub], src)
    assert np.abs(direct - vis64[sub]).max() <= 1e-9 * np.abs(g).sum()
    return d, ant, src, uvw, vis64, np.abs(g).sum()


@pytest.mark.parametrize("flags", ["fp16", "bf16"])
def test_config3_predict_dirty_and_residual_errors(hip, actors, oracle, c3, flags):
    """C3's three outputs against FP64, for the split matrix-core predict:
    visibilities (relative to sum |g|), the dirty image of the predicted
    visibilities (relative to its peak) and the residual image of
    measured - predicted (relative to the dirty peak), each through the
    GPU gridder and FFT against the FP64 gridding sum and FFT.

    The visibility bar is set by FP32 accumulation, not by the split: the
    4096 terms are summed in 256 chunks of 16 into FP32 accumulators whose
    value reaches sum |g| on short baselines, so each chunk's addition
    rounds by up to 2^-24 sum |g| and the max over 32,640 baselines of a
    random walk of 256 such roundings sits near 4 * 16 * 2^-24 = 3.8e-6 of
    sum |g| (measured 2.1e-6 FP16, 2.0e-6 BF16)."""
    d, ant, src, uvw, vis64, gsum = c3
    fl = hip.DFT_ARRAY_FP16 if flags == "fp16" else hip.DFT_ARRAY_BF16
    n_vis = uvw.shape[0]
    d_ant = hip.DeviceBuffer.from_array(np.ascontiguousarray(ant, np.float64))
    d_src = hip.DeviceBuffer.from_array(src)
    d_out = hip.DeviceBuffer(8 * n_vis)
    hip.dft_predict_array(1, ant.shape[1], d_ant, src.shape[0], d_src, d_out, fl)
    pred = d_out.download((n_vis, 2), np.float32)
    e_vis = np.abs(pred - vis64).max() / gsum
    G = d.grid_size
    cfg = cfg_of(hip, d)
    uvw32 = uvw.astype(np.float32)
    dirty = image_gpu(actors, G, actors.std_gridding(G, d.oversampling, d.kernels, d.supports, uvw32, pred, cfg), cfg)
    # the FP64 sum takes float inputs: vis64 rounded to float moves the image by ~6e-8 of its peak
    dirty64 = image_f64(oracle, G, oracle.std_gridding_f64(G, d.oversampling, d.kernels, d.supports, uvw32,
                                                           vis64.astype(np.float32), d.uv_scale, d.w_scale))
    dpeak = np.abs(dirty64).max()
    e_dirty = np.abs(dirty - dirty64).max() / dpeak
    noise = np.random.default_rng(4).normal(0.0, 0.01, vis64.shape)
    measured = (vis64 + noise).astype(np.float32)
    res_vis = actors.subtract_finegrid(measured, pred)
    res = image_gpu(actors, G, actors.std_gridding(G, d.oversampling, d.kernels, d.supports, uvw32, res_vis, cfg), cfg)
    # the exact residual rounded to float once (~1e-9 absolute at the noise level)
    res_exact = (measured.astype(np.float64) - vis64).astype(np.float32)
    res64 = image_f64(oracle, G, oracle.std_gridding_f64(G, d.oversampling, d.kernels, d.supports, uvw32,
                                                         res_exact, d.uv_scale, d.w_scale))
    e_res = np.abs(res - res64).max() / dpeak
    print(f"C3 {flags} split: vis {e_vis:.2e} of sum|g|, dirty image {e_dirty:.2e}, residual {e_res:.2e} "
          "of the dirty peak")
    assert e_vis <= 5e-6
    assert e_dirty <= 1e-5
    assert e_res <= (1e-5 if flags == "fp16" else 5e-5)
