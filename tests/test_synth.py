"""The synthetic-input generator: layouts and closed-form properties."""
import numpy as np
import pytest

import sdp_synth


def test_config1_shapes_and_ordering():
    d = sdp_synth.config1(n_time=4, with_vis=False)
    assert d.n_vis == 2016 * 4
    assert d.supports.tolist() == [[3, 0]] and d.kernels.shape == ((4 * 16) ** 2, 2)
    pairs = sdp_synth.receiver_pairs(64)
    assert pairs.shape == (2016, 2) and pairs[0].tolist() == [0, 1] and pairs[63].tolist() == [1, 2]
    # time-major: visibility t * 2016 + b is baseline b (pair order) at hour angle t
    enu = sdp_synth.disk_layout(64, 1000.0, seed=1)
    ha = np.linspace(-3.0, 3.0, 4)
    for t, b in ((0, 0), (2, 63), (3, 2015)):
        one = sdp_synth.uvw_tracks(enu[list(pairs[b])], -30.0, -30.0, ha[t:t + 1], 150e6)[0]
        np.testing.assert_allclose(d.uvw[t * 2016 + b], one.astype(np.float32), rtol=1e-6)
    assert np.abs(d.uvw[:, :2]).max() * d.uv_scale < d.grid_size / 2 - 4


def test_kernel_table_layout_and_normalisation():
    gk, dk, sup = sdp_synth.kernel_tables([3, 5, 8], 8, fov=0.1, w_scale=4 / 500.0)
    assert sup[:, 1].tolist() == [0, 32 * 32, 32 * 32 + 48 * 48]
    assert gk.shape[0] == 32 * 32 + 48 * 48 + 72 * 72
    np.testing.assert_array_equal(dk[:, 0], gk[:, 0])
    np.testing.assert_array_equal(dk[:, 1], -gk[:, 1])
    p0 = gk[:32 * 32].reshape(32, 32, 2)
    assert np.abs(p0[..., 1]).max() < 1e-6                      # w = 0 plane is real
    ints = p0[::8, ::8, 0]                                     # integer offsets 0..3
    full = ints[0, 0] + 2 * ints[0, 1:].sum() + 2 * ints[1:, 0].sum() + 4 * ints[1:, 1:].sum()
    assert full == pytest.approx(1.0, abs=2e-3)
    assert np.all(np.diff(p0[0, :, 0]) <= 1e-7)                 # decreasing away from the centre


def test_uvw_closed_form_at_transit():
    # HA = 0, dec = latitude: an east-west baseline gives pure u
    enu = np.array([[0.0, 0.0, 0.0], [100.0, 0.0, 0.0]])
    uvw = sdp_synth.uvw_tracks(enu, -30.0, -30.0, np.array([0.0]), sdp_synth.C_LIGHT)
    np.testing.assert_allclose(uvw[0], [100.0, 0.0, 0.0], atol=1e-9)


def test_dft_f64_matches_direct_sum():
    rng = np.random.default_rng(0)
    uvw = rng.uniform(-500, 500, (20, 3)).astype(np.float32)
    src = np.array([[0.01, 0.02, 1.0], [-0.03, 0.005, 0.5]], np.float32)
    got = sdp_synth.dft_f64(uvw, src)
    l, m, f = src[:, 0].astype(float), src[:, 1].astype(float), src[:, 2].astype(float)
    r2 = l * l + m * m
    ph = uvw[:, :1] * l + uvw[:, 1:2] * m - uvw[:, 2:3] * r2 / 2
    expect = (f / (1 - r2 / 2) * np.exp(-2j * np.pi * ph)).sum(1)
    np.testing.assert_allclose(got[:, 0] + 1j * got[:, 1], expect, atol=1e-9)
