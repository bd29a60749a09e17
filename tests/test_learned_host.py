"""Host-side logic of the learned proposal's training data (no GPU): the
random sky generator, the Gaussian smoothing and the network's parameter
count at the sizes used."""
import math

import numpy as np
import pytest


@pytest.fixture(scope="module")
def L():
    import sdp_learned
    return sdp_learned


def test_random_sky_is_seeded_and_stays_in_its_box(L):
    a = L.random_sky(256, np.random.default_rng(4), n_points=(30, 30), n_gauss=(0, 0), box=20)
    b = L.random_sky(256, np.random.default_rng(4), n_points=(30, 30), n_gauss=(0, 0), box=20)
    np.testing.assert_array_equal(a, b)
    ys, xs = np.nonzero(a)
    assert 0 < len(ys) <= 30
    assert ys.min() >= 128 - 20 and ys.max() < 128 + 20 and xs.min() >= 108 and xs.max() < 148
    assert a.min() >= 0 and a.max() <= 30 * 1.0 and a[a > 0].min() >= 1e-3 * (1 - 1e-6)


def test_random_sky_power_law_fluxes(L):
    """dN/dS ~ S^-1.5 on [1 mJy, 1 Jy]: the median of the inverse-CDF draw
    is ((lo^a + hi^a) / 2)^(1/a), a = -0.5."""
    rng = np.random.default_rng(9)
    flux = np.concatenate([L.random_sky(512, rng, n_points=(200, 200), n_gauss=(0, 0), margin=4).ravel()
                           for _ in range(20)])
    flux = flux[flux > 0]
    a = -0.5
    median = ((1e-3 ** a + 1.0 ** a) / 2) ** (1 / a)
    assert np.median(flux) == pytest.approx(median, rel=0.15)


def test_random_sky_gaussian_integral(L):
    """One Gaussian of peak 0.2 f and FWHMs (fa, fb): its integral is
    0.2 f * 2 pi sigma_a sigma_b, sigma = FWHM / sqrt(8 ln 2)."""
    rng = np.random.default_rng(2)
    sky = L.random_sky(512, rng, n_points=(0, 0), n_gauss=(1, 1), fwhm=(10.0, 10.0), box=4)
    rng = np.random.default_rng(2)
    rng.integers(0, 1)                       # the point-count draw
    rng.integers(252, 260, 0), rng.integers(252, 260, 0)
    rng.integers(1, 2)                       # the Gaussian-count draw
    u = rng.random(1)
    f = (1e-3 ** -0.5 + u * (1.0 ** -0.5 - 1e-3 ** -0.5)) ** (1 / -0.5)
    sigma = 10.0 / math.sqrt(8 * math.log(2))
    assert sky.sum() == pytest.approx(0.2 * f[0] * 2 * math.pi * sigma * sigma, rel=1e-3)


def test_smooth_keeps_the_sum_and_matches_a_gaussian(L):
    img = np.zeros((64, 64), np.float32)
    img[32, 32] = 1.0
    s = L.smooth(img, 2.0)
    assert s.sum() == pytest.approx(1.0, rel=1e-5)
    r2 = (np.arange(64) - 32)[:, None] ** 2 + (np.arange(64) - 32)[None, :] ** 2
    want = np.exp(-r2 / 8.0) / (8.0 * math.pi)
    assert np.abs(s - want).max() < 1e-4


def test_parameter_counts(L):
    import sdp_hip
    count = lambda lv, b: sum(9 * ci * co + co for _, ci, co in sdp_hip.cnn_layer_shapes(lv, b, 2, 1))
    assert count(4, 32) == 1_753_313         # the 4-level, 32-channel network (about 2M parameters)
    assert count(3, 16) == 106_737
