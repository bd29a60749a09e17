"""Pins of the FP64 imaging chain the pipeline tests compare against
(tests/imaging_ref.py): image orientation, PSF, flux scale and the predict
conventions, checked against what the physics fixes.  CPU only."""
import copy

import numpy as np
import pytest

import imaging_ref as R
import sdp_synth


@pytest.fixture(scope="module")
def sky():
    return sdp_synth.sky_case()


def rms(img):
    return float(np.sqrt(np.mean(img ** 2)))


def test_single_source_lands_on_its_pixel():
    """A point source at (l, m) = ((x - G/2) cell, (y - G/2) cell) peaks at
    row y, column x of the corrected dirty image, with peak / psf_max equal
    to its flux (the taper correction undoes the kernel's attenuation)."""
    for seed in (31, 32, 33):
        d = sdp_synth.sky_case(n_src=1, seed=seed)
        img = R.dirty(d, d.vis)
        _, peak = R.psf(d)
        x, y = d.extra["pixels"][0]
        assert np.unravel_index(np.argmax(img), img.shape) == (y, x)
        assert img[y, x] / peak == pytest.approx(float(d.sources[0, 2]), rel=0.05)


def test_psf_is_centred_with_unit_peak(sky):
    psf, peak = R.psf(sky)
    G = sky.grid_size
    assert np.unravel_index(np.argmax(psf), psf.shape) == (G // 2, G // 2)
    assert psf[G // 2, G // 2] == pytest.approx(1.0, abs=1e-7)
    assert peak == pytest.approx(sky.n_vis, rel=0.01)   # unit visibilities, kernels summing to 1


def test_predict_of_true_sky_cancels_the_visibilities(sky):
    """Subtracting the predicted visibilities of the true sky leaves a
    residual two orders below the dirty image; the same with the gridding
    table in place of its conjugate, or without the taper division, leaves
    several times more (planted mistakes)."""
    model = R.sky_model(sky)
    dirty_rms = rms(R.dirty(sky, sky.vis))
    res = R.dirty(sky, (sky.vis - R.predict(sky, model)).astype(np.float32))
    assert rms(res) < dirty_rms / 50

    wrong = copy.copy(sky)
    wrong.degrid_kernels = sky.kernels
    res_conj = R.dirty(sky, (sky.vis - R.predict(wrong, model)).astype(np.float32))
    assert rms(res_conj) > 3 * rms(res)

    saved = R.uncorrect_in
    try:
        R.uncorrect_in = lambda G, img: img
        res_taper = R.dirty(sky, (sky.vis - R.predict(sky, model)).astype(np.float32))
    finally:
        R.uncorrect_in = saved
    assert rms(res_taper) > 10 * rms(res)


def test_dirty_image_is_linear(sky, rng):
    a = rng.normal(size=(sky.n_vis, 2)).astype(np.float32)
    b = rng.normal(size=(sky.n_vis, 2)).astype(np.float32)
    lhs = R.dirty(sky, (a + 2 * b).astype(np.float32))
    rhs = R.dirty(sky, a) + 2 * R.dirty(sky, b)
    np.testing.assert_allclose(lhs, rhs, atol=1e-9 * np.abs(lhs).max() * sky.n_vis)
