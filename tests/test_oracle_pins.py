"""Pins the CPU oracle (oracle/sdp_oracle.c) without the reference binary.

The reference ships no tests, golden vectors or fixtures for this path and
its C actors cannot be compiled here (common.h:31 includes cblas.h, fft_run.c
includes fftw3.h; neither exists in this image).  Each oracle function is
therefore checked against something fixed independently of its code:
hand-derived known answers, closed forms, invariants, an independent FFT
(numpy) and brute force on tiny inputs.  A dropped term, swapped index or
wrong sign in the oracle fails at least one of these.
"""
import numpy as np
import pytest

import sdp_hip  # only for the Config record layout (no GPU calls)


def cfg_for(G, cell=None, n_kernels=1, max_w=1.0, **kw):
    cell = (1.0 / G) if cell is None else cell
    return sdp_hip.make_config(G, cell, num_kernels=n_kernels, max_w=max_w, **kw)


def index_kernel(h, OS):
    """kernel[k] = (k, 0.5): every tap reveals the table index it used."""
    n = ((h + 1) * OS) ** 2
    k = np.stack([np.arange(n, dtype=np.float32), np.full(n, 0.5, np.float32)], axis=1)
    return k, np.array([[h, 0]], np.int32)


# ---------------------------------------------------------------- gridding

def test_gridding_single_visibility_known_answer(oracle):
    G, OS, h = 64, 16, 3
    kern, sup = index_kernel(h, OS)
    cfg = cfg_for(G)                       # uv_scale = 1: grid position = u
    uvw = np.array([[10.25, -3.5, 0.0]], np.float32)
    vis = np.array([[1.0, 0.0]], np.float32)
    grid = oracle.std_gridding(G, OS, kern, sup, uvw, vis, cfg).reshape(G, G, 2)
    # u in [ceil(7.25), ceil(13.25)) = 8..13, |u - 10.25| * 16 = 36,20,4,12,28,44
    kx = {8: 36, 9: 20, 10: 4, 11: 12, 12: 28, 13: 44}
    # v in [ceil(-6.5), ceil(-0.5)) = -6..-1, |v + 3.5| * 16 = 40,24,8,8,24,40
    ky = {-6: 40, -5: 24, -4: 8, -3: 8, -2: 24, -1: 40}
    expect = np.zeros((G, G, 2), np.float32)
    for v, a in ky.items():
        for u, b in kx.items():
            expect[v + 32, u + 32] = (a * 64 + b, 0.5)
    np.testing.assert_array_equal(grid, expect)


def test_gridding_rounds_half_away_from_zero(oracle):
    G, OS, h = 64, 16, 3
    kern, sup = index_kernel(h, OS)
    cfg = cfg_for(G)
    # (10 - 10.03125) * 16 = -0.5 -> roundf gives -1 -> kx = 1 (banker's rounding would give 0)
    uvw = np.array([[10.03125, 0.0, 0.0]], np.float32)
    grid = oracle.std_gridding(G, OS, kern, sup, uvw, np.array([[1.0, 0.0]], np.float32), cfg).reshape(G, G, 2)
    assert grid[32, 42, 0] == 0 * 64 + 1


def test_gridding_conjugates_for_negative_w_and_skips_edge(oracle):
    G, OS, h = 64, 16, 3
    kern, sup = index_kernel(h, OS)
    cfg = cfg_for(G)
    uvw = np.array([[30.5, 0.0, -1.0]], np.float32)   # footprint 28..33, cells >= 32 dropped
    grid = oracle.std_gridding(G, OS, kern, sup, uvw, np.array([[1.0, 0.0]], np.float32), cfg).reshape(G, G, 2)
    cols = np.nonzero(grid[32, :, 0])[0] - 32
    assert list(cols) == [28, 29, 30, 31]
    assert np.all(grid[grid[..., 0] != 0][:, 1] == -0.5)


def test_gridding_plane_selection(oracle):
    G, OS = 64, 4
    sup = np.array([[1, 0], [2, 64], [3, 64 + 144]], np.int32)   # offsets sum ((h+1)*OS)^2
    n = 64 + 144 + 256
    kern = np.stack([np.arange(n, dtype=np.float32), np.zeros(n, np.float32)], axis=1)
    # w_scale = (3 - 1)^2 / max_w = 1 with max_w = 4
    cfg = cfg_for(G, n_kernels=3, max_w=4.0)
    for w, plane in ((2.25, 2), (2.2, 1), (0.2, 0), (0.3, 1)):
        g = oracle.std_gridding(G, OS, kern, sup, np.array([[0.0, 0.0, w]], np.float32),
                                np.array([[1.0, 0.0]], np.float32), cfg).reshape(G, G, 2)
        nz = g[..., 0][g[..., 0] != 0]
        lo, hi = sup[plane, 1], sup[plane, 1] + ((sup[plane, 0] + 1) * OS) ** 2
        assert np.all((nz >= lo) & (nz < hi)), (w, plane)
        assert np.count_nonzero(g[..., 0]) + (1 if g[32, 32, 0] == 0 else 0) == (2 * sup[plane, 0]) ** 2


def test_gridding_linear_and_bypass(oracle, rng):
    G, OS = 64, 16
    from sdp_synth import kernel_tables
    kern, _, sup = kernel_tables([3], OS, fov=1.0, w_scale=0.0)
    cfg = cfg_for(G)
    uvw = np.concatenate([rng.uniform(-25, 25, (50, 2)), rng.uniform(-1, 1, (50, 1))], axis=1).astype(np.float32)
    a = rng.normal(size=(50, 2)).astype(np.float32)
    b = rng.normal(size=(50, 2)).astype(np.float32)
    ga = oracle.std_gridding_f64(G, OS, kern, sup, uvw, a, 1.0, 0.0)
    gb = oracle.std_gridding_f64(G, OS, kern, sup, uvw, b, 1.0, 0.0)
    gab = oracle.std_gridding_f64(G, OS, kern, sup, uvw, (a + b).astype(np.float32), 1.0, 0.0)
    # the float a+b rounds, so compare against that rounding's size
    np.testing.assert_allclose(gab, ga + gb, atol=1e-5 * np.abs(ga + gb).max())
    f32 = oracle.std_gridding(G, OS, kern, sup, uvw, a, cfg)
    np.testing.assert_allclose(f32, ga, atol=1e-5 * np.abs(ga).max())
    prev = rng.normal(size=(G * G, 2)).astype(np.float32)
    out = oracle.std_gridding(G, OS, kern, sup, uvw, a, cfg, bypass=1, maj_iter=1, prev_grid=prev)
    np.testing.assert_array_equal(out, prev)


# ---------------------------------------------------------------- degridding

def test_degridding_nearest_known_answer(oracle, rng):
    G, OS = 64, 16
    kern, sup = index_kernel(3, OS)
    cfg = cfg_for(G)
    grid = rng.normal(size=(G * G, 2)).astype(np.float32)
    uvw = np.array([[5.4, -3.6, 0.0], [0.0, 0.0, 0.0]], np.float32)
    out = oracle.std_degridding(G, OS, kern, sup, grid, uvw, cfg, n=2, n_total=3)
    # (int)((5.4f + 32) + 0.5) = 37, (int)((-3.6f + 32) + 0.5) = 28
    np.testing.assert_array_equal(out[0], grid[37 + 28 * G])
    np.testing.assert_array_equal(out[1], grid[32 + 32 * G])
    np.testing.assert_array_equal(out[2], [0.0, 0.0])


def test_weighted_degridding_is_adjoint_of_gridding(oracle, rng):
    from sdp_synth import kernel_tables
    G, OS = 64, 8
    gk, dk, sup = kernel_tables([3, 4, 5], OS, fov=0.2, w_scale=(3 - 1) ** 2 / 300.0)
    n = 300
    uvw = np.concatenate([rng.uniform(-30, 30, (n, 2)) / 0.2, rng.uniform(-300, 300, (n, 1))],
                         axis=1).astype(np.float32)
    x = rng.normal(size=(n, 2)).astype(np.float32)
    y = rng.normal(size=(G * G, 2)).astype(np.float32)
    gx = oracle.std_gridding_f64(G, OS, gk, sup, uvw, x, 0.2, (3 - 1) ** 2 / 300.0)
    dy = oracle.degrid_weighted_f64(G, OS, dk, sup, y, uvw, 0.2, (3 - 1) ** 2 / 300.0)
    c = lambda a: a[..., 0] + 1j * a[..., 1]
    lhs = np.vdot(c(gx), c(y.astype(np.float64)))
    rhs = np.vdot(c(x.astype(np.float64)), c(dy))
    assert abs(lhs - rhs) <= 1e-10 * abs(lhs)
    # a wrong conjugation rule breaks it
    dy_bad = oracle.degrid_weighted_f64(G, OS, gk, sup, y, uvw, 0.2, (3 - 1) ** 2 / 300.0)
    assert abs(np.vdot(c(x.astype(np.float64)), c(dy_bad)) - lhs) > 1e-3 * abs(lhs)


# ---------------------------------------------------------------- fine grid

def test_finegrid_known_answer_and_simplification(oracle):
    G, OS = 64, 16
    cfg = cfg_for(G)
    uvw = np.array([[3.3, -3.3, 1.0], [3.31, -3.3, 1.0], [-7.0, 2.0, 1.0]], np.float32)
    vis = np.array([[1.0, 2.0], [0.5, 0.25], [3.0, 4.0]], np.float32)
    ov, ou, n = oracle.correct_to_finegrid(G, OS, 0, uvw, vis, cfg)
    assert n == 3
    np.testing.assert_array_equal(ou[0], [3.3125, -3.3125, 1.0])
    np.testing.assert_array_equal(ov, vis)
    ov, ou, n = oracle.correct_to_finegrid(G, OS, 1, uvw, vis, cfg)
    assert n == 2
    np.testing.assert_array_equal(ov[:2], [[1.5, 2.25], [3.0, 4.0]])
    np.testing.assert_array_equal(ou[:2], [[3.3125, -3.3125, 1.0], [-7.0, 2.0, 1.0]])
    np.testing.assert_array_equal(ov[2], [0.0, 0.0])


# ---------------------------------------------------------------- DFT

def test_dft_known_answers(oracle, rng):
    cfg = cfg_for(64)
    uvw = rng.uniform(-100, 100, (64, 3)).astype(np.float32)
    out = oracle.dft_actor(np.array([[0.0, 0.0, 0.7]], np.float32), uvw, cfg)
    np.testing.assert_array_equal(out, np.tile([0.7, 0.0], (64, 1)).astype(np.float32))
    # w = 0: V = f / (1 - r^2/2) exp(-2 pi i (u l + v m))
    l, m, f = 0.01, -0.02, 0.9
    uvw0 = uvw.copy()
    uvw0[:, 2] = 0
    expect = f / (1 - 0.5 * (l * l + m * m)) * np.exp(-2j * np.pi * (uvw0[:, 0] * l + uvw0[:, 1] * m))
    for fn in (lambda: oracle.dft_actor(np.array([[l, m, f]], np.float32), uvw0, cfg),
               lambda: oracle.dft_f64(np.array([[l, m, f]], np.float32), uvw0)):
        got = fn()
        np.testing.assert_allclose(got[:, 0] + 1j * got[:, 1], expect, atol=2e-5)
    # u = v = 0: V = f / (1 - r^2/2) exp(+i pi w r^2)
    uvww = np.zeros((16, 3), np.float32)
    uvww[:, 2] = np.linspace(-500, 500, 16)
    r2 = l * l + m * m
    expect = f / (1 - r2 / 2) * np.exp(1j * np.pi * uvww[:, 2] * r2)
    got = oracle.dft_f64(np.array([[l, m, f]], np.float32), uvww)
    np.testing.assert_allclose(got[:, 0] + 1j * got[:, 1], expect, atol=1e-6)
    # superposition
    s = np.array([[l, m, f], [0.003, 0.004, 0.2]], np.float32)
    both = oracle.dft_f64(s, uvw)
    sep = oracle.dft_f64(s[:1], uvw) + oracle.dft_f64(s[1:], uvw)
    np.testing.assert_allclose(both, sep, atol=1e-12)


# ---------------------------------------------------------------- Hogbom

def delta_psf(G, extra=None):
    psf = np.zeros(G * G + 1, np.float32)
    psf[(G // 2) * G + G // 2] = 1.0
    if extra:
        for (dy, dx), v in extra.items():
            psf[(G // 2 + dy) * G + G // 2 + dx] = v
    return psf


def test_hogbom_delta_psf_geometric_fluxes(oracle):
    G = 16
    cfg = cfg_for(G, cell=0.01, psf_max_value=1.0)
    res = np.zeros(G * G, np.float32)
    res[7 * G + 5] = 2.0
    ns, src, model, r = oracle.hogbom_clean(G, 3, 10, res, delta_psf(G), cfg, np.zeros(G * G, np.float32))
    assert ns == 3
    np.testing.assert_allclose(src[:, 2], [0.2, 0.18, 0.162], rtol=1e-6)
    np.testing.assert_allclose(src[:, 0], (5 - 8) * 0.01, rtol=1e-6)
    np.testing.assert_allclose(src[:, 1], (7 - 8) * 0.01, rtol=1e-6)
    np.testing.assert_allclose(model[7 * G + 5], 0.542, rtol=1e-6)
    np.testing.assert_allclose(r[7 * G + 5], 2.0 * 0.9 ** 3, rtol=1e-6)


def test_hogbom_psf_is_read_mirrored_in_x(oracle):
    G = 16
    cfg = cfg_for(G, psf_max_value=1.0)
    res = np.zeros(G * G, np.float32)
    res[7 * G + 5] = 2.0
    psf = delta_psf(G, {(0, 1): 0.5})         # lobe one pixel right of the centre
    _, _, _, r = oracle.hogbom_clean(G, 1, 10, res, psf, cfg, np.zeros(G * G, np.float32))
    # column index G - (x - mx + G/2) puts it one pixel LEFT of the peak
    np.testing.assert_allclose(r[7 * G + 4], -0.1, rtol=1e-6)
    assert r[7 * G + 6] == 0.0


def test_hogbom_tie_break_and_stops(oracle):
    G = 16
    cfg = cfg_for(G, psf_max_value=1.0)
    res = np.zeros(G * G, np.float32)
    res[20] = -1.0
    res[40] = 1.0
    ns, src, _, _ = oracle.hogbom_clean(G, 1, 4, res, delta_psf(G), cfg, np.zeros(G * G, np.float32))
    assert ns == 1 and src[0, 2] == pytest.approx(-0.1)
    # noise stop (noise_factor * gain > 1): peak 1.5 below 2 * mean(|r[1:]|) ~ 2.0
    cfgn = cfg_for(G, psf_max_value=1.0, noise_factor=20.0)
    res = np.ones(G * G, np.float32)
    res[0] = 0.0
    res[50] = 1.5
    ns, _, _, _ = oracle.hogbom_clean(G, 5, 4, res, delta_psf(G), cfgn, np.zeros(G * G, np.float32))
    assert ns == 0
    # weak-source stop: |max| < |first flux| * weak
    cfgw = cfg_for(G, psf_max_value=1.0, loop_gain=0.5, weak_source_percent_img=0.5)
    res = np.zeros(G * G, np.float32)
    res[30] = 4.0
    res[90] = 0.9
    ns, src, _, _ = oracle.hogbom_clean(G, 10, 10, res, delta_psf(G), cfgw, np.zeros(G * G, np.float32))
    assert ns == 3
    np.testing.assert_allclose(src[:, 2], [2.0, 1.0, 0.5])


# ---------------------------------------------------------------- FFT

@pytest.mark.parametrize("G", [8, 16])
@pytest.mark.parametrize("sign", [-1, 1])
def test_fft_matches_definition(oracle, rng, G, sign):
    x = rng.normal(size=(G * G, 2))
    np.testing.assert_allclose(oracle.fft2d_f64(G, x, sign), oracle.dft2d_direct_f64(G, x, sign), atol=1e-9)


def test_fft_matches_numpy_and_known_answers(oracle, rng):
    G = 64
    x = rng.normal(size=(G * G, 2))
    z = (x[:, 0] + 1j * x[:, 1]).reshape(G, G)
    fwd = oracle.fft2d_f64(G, x, -1)
    np.testing.assert_allclose(fwd[:, 0] + 1j * fwd[:, 1], np.fft.fft2(z).ravel(), atol=1e-9)
    inv = oracle.fft2d_f64(G, x, +1)
    np.testing.assert_allclose(inv[:, 0] + 1j * inv[:, 1], (np.fft.ifft2(z) * G * G).ravel(), atol=1e-9)
    # impulse -> ones; plane wave -> G^2 delta; Parseval; round trip
    imp = np.zeros((G * G, 2))
    imp[0, 0] = 1
    np.testing.assert_allclose(oracle.fft2d_f64(G, imp, -1), np.tile([1.0, 0.0], (G * G, 1)), atol=1e-12)
    yy, xx = np.mgrid[0:G, 0:G]
    pw = np.exp(2j * np.pi * (3 * yy + 5 * xx) / G).ravel()
    X = oracle.fft2d_f64(G, np.stack([pw.real, pw.imag], 1), -1)
    mag = np.hypot(X[:, 0], X[:, 1]).reshape(G, G)
    assert mag[3, 5] == pytest.approx(G * G) and np.sum(mag > 1e-6) == 1
    assert np.sum(fwd ** 2) == pytest.approx(G * G * np.sum(x ** 2), rel=1e-12)
    np.testing.assert_allclose(oracle.fft2d_f64(G, fwd, +1) / (G * G), x, atol=1e-12)


# ---------------------------------------------------------------- element-wise

def test_shifts_and_correction(oracle, rng):
    G = 32
    g = rng.normal(size=(G * G, 2)).astype(np.float32)
    s = oracle.shift_c2c(G, g)
    np.testing.assert_array_equal(oracle.shift_c2c(G, s), g)
    assert s[0, 0] == g[0, 0] and s[1, 0] == -g[1, 0] and s[G + 1, 0] == g[G + 1, 0] and s[G, 0] == -g[G, 0]
    np.testing.assert_array_equal(oracle.shift_c2r(G, g), s[:, 0])
    img = rng.normal(size=G * G).astype(np.float32)
    rc = oracle.shift_r2c(G, img)
    np.testing.assert_array_equal(rc[:, 0], oracle.shift_c2r(G, np.stack([img, np.zeros_like(img)], 1)))
    assert np.all(rc[:, 1] == 0)
    taper = np.concatenate([oracle.half_prolate(G), [0.0]]).astype(np.float32)
    out = oracle.convolution_correction(G, img, taper).reshape(G, G)
    t2 = np.outer(taper[np.abs(np.arange(G) - G // 2)], taper[np.abs(np.arange(G) - G // 2)])
    ok = t2 > 1e-10
    np.testing.assert_allclose(out[ok] * t2[ok], img.reshape(G, G)[ok], rtol=1e-6)
    assert np.all(out[0, :] == 0) and np.all(out[:, 0] == 0)


def test_correction_zeroes_distance_1229(oracle):
    G = 2460
    img = np.ones(G * G, np.float32)
    taper = np.concatenate([oracle.half_prolate(G), [0.0]]).astype(np.float32)
    out = oracle.convolution_correction(G, img, taper).reshape(G, G)
    assert np.all(out[G // 2 - 1229, :] == 0) and np.all(out[:, G // 2 + 1229] == 0)
    assert out[G // 2, G // 2] == pytest.approx(1.0 / taper[0] ** 2)


def test_prolate_published_values(oracle):
    # the approximation is normalised to psi(0) = 1 (to its published precision)
    assert oracle.prolate_spheroidal(0.0) == pytest.approx(1.0, abs=1e-6)
    assert oracle.prolate_spheroidal(1.0) == 0.0 and oracle.prolate_spheroidal(1.5) == 0.0
    # the two rational branches meet at nu = 0.75
    assert oracle.prolate_spheroidal(0.75 - 1e-9) == pytest.approx(oracle.prolate_spheroidal(0.75), rel=2e-3)
    # psi decreases monotonically on [0, 1)
    v = [oracle.prolate_spheroidal(x) for x in np.linspace(0, 0.999, 200)]
    assert np.all(np.diff(v) < 0)


def test_visibility_arithmetic(oracle, rng):
    m = rng.normal(size=(12, 2)).astype(np.float32)
    p = rng.normal(size=(12, 2)).astype(np.float32)
    pairs = np.array([[0, 1], [0, 2], [1, 2]], np.int32)
    ones = np.tile([1.0, 0.0], (3, 1)).astype(np.float32)
    np.testing.assert_array_equal(oracle.subtract_from_measurements(m, p, ones, pairs), m - p)
    gi = np.array([[0.0, 1.0], [1.0, 0.0], [1.0, 0.0]], np.float32)   # g0 = i
    got = oracle.apply_gains_only(m, gi, pairs)
    mc = m[:, 0] + 1j * m[:, 1]
    expect = np.where(np.arange(12) % 3 < 2, mc * 1j, mc)               # baselines 0,1 use antenna 0
    np.testing.assert_allclose(got[:, 0] + 1j * got[:, 1], expect, atol=1e-7)
    np.testing.assert_array_equal(oracle.add_visibilities(m, p), m + p)
    d = oracle.subtract_finegrid(m, p, n=5)
    np.testing.assert_array_equal(d[:5], m[:5] - p[:5])
    assert np.all(d[5:] == 0)


def test_reciprocal_transform(oracle, rng):
    """reciprocal_transform_actor returns 1 / z: known answers, z * (1 / z) = 1
    to float rounding, and the FP64 complex reciprocal."""
    z = np.array([[2.0, 0.0], [0.0, 1.0], [3.0, 4.0], [-1.0, -1.0]], np.float32)
    got = oracle.reciprocal_transform(z)
    np.testing.assert_array_equal(got, np.array([[0.5, 0.0], [0.0, -1.0], [np.float32(3) / np.float32(25),
                                                 -(np.float32(4) / np.float32(25))], [-0.5, 0.5]], np.float32))
    z = rng.normal(size=(1000, 2)).astype(np.float32)
    got = oracle.reciprocal_transform(z)
    zc = z[:, 0].astype(np.float64) + 1j * z[:, 1]
    gc = got[:, 0].astype(np.float64) + 1j * got[:, 1]
    assert np.abs(zc * gc - 1).max() < 1e-6
    np.testing.assert_allclose(gc, 1 / zc, rtol=1e-6)


def test_restore_conv2d_known_answers(oracle, rng):
    """conv2D: a model pixel a at (y0, x0) puts a * filter[c - fy][c - fx]
    (c = D/2 + 1) at (y0 + fy, x0 + fx) for fy, fx in [-h, h), clipped at the
    edges; pixels below 1e-5 (as a double) add nothing; disjoint sources add."""
    D, h = 32, 3
    c = D // 2 + 1
    f = rng.normal(size=(D, D)).astype(np.float32)
    for (y0, x0) in ((10, 12), (0, 1), (D - 1, D - 2)):
        img = np.zeros((D, D), np.float32)
        img[y0, x0] = 1.5
        got = oracle.conv2d(img, f, h, h)
        want = np.zeros((D, D), np.float32)
        for fy in range(-h, h):
            for fx in range(-h, h):
                if 0 <= y0 + fy < D and 0 <= x0 + fx < D:
                    want[y0 + fy, x0 + fx] = np.float32(1.5) * f[c - fy, c - fx]
        np.testing.assert_array_equal(got, want)
    img = np.zeros((D, D), np.float32)
    img[4, 4] = np.float32(1e-5)                                  # 9.99999975e-06 < 1e-5: skipped
    img[20, 20] = np.nextafter(np.float32(1e-5), np.float32(1))   # above 1e-5: kept
    got = oracle.conv2d(img, f, h, h)
    assert not got[:10, :10].any() and got[17:23, 17:23].any()
    a, b = np.zeros((D, D), np.float32), np.zeros((D, D), np.float32)
    a[5, 5], b[25, 25] = 2.0, 3.0
    np.testing.assert_array_equal(oracle.conv2d(a + b, f, h, h), oracle.conv2d(a, f, h, h) + oracle.conv2d(b, f, h, h))


def test_clean_psf_box_and_normalisation(oracle):
    """clean_psf_host_set_up: the box half size is the longest walk from
    (G/2 + 1, G/2 + 1) to its first non-decreasing step; the squared Gaussian
    over [-h, h)^2 sums to psf_max, is symmetric about the centre, and is
    zero outside the box."""
    G = 64
    c = G // 2 + 1
    # decreasing for 5 steps rightwards, 3 downwards, 4 leftwards, 2 upwards, then
    # rising: the walks stop at steps 6, 4, 5 and 3
    psf = np.zeros((G, G), np.float32)
    def walk(n):
        return np.concatenate([1.0 - 0.1 * np.arange(n + 1), 1.0 - 0.1 * n + 0.05 * np.arange(1, G)])
    psf[c, c:] = walk(5)[:G - c]
    psf[c:, c] = walk(3)[:G - c]
    psf[c, :c + 1] = walk(4)[:c + 1][::-1]
    psf[:c + 1, c] = walk(2)[:c + 1][::-1]
    cp, h = oracle.clean_psf(psf, 2.5)
    assert h == 6
    assert abs(float(cp.sum(dtype=np.float64)) - 2.5) < 1e-5
    box = cp[c - h:c + h, c - h:c + h]
    assert np.count_nonzero(cp) == np.count_nonzero(box) == (2 * h) ** 2
    inner = cp[c - h + 1:c + h, c - h + 1:c + h]
    np.testing.assert_array_equal(inner, inner[::-1, ::-1])
    assert cp[c, c] == cp.max()


def test_sliced_grid_degrid_matches_single_thread(oracle, rng):
    """The multi-core CPU baseline (splitmerge.c's slices on threads): with
    integer kernels and visibilities the merged grid equals the one-thread
    grid exactly, and every degridded visibility is the one-thread value."""
    import sdp_synth
    d = sdp_synth.small_case(n_vis=2000, grid_size=64, n_planes=3, oversampling=8, seed=4)
    cfg = sdp_hip.make_config(d.grid_size, d.cell_size, oversampling=d.oversampling,
                              num_kernels=d.supports.shape[0], max_w=d.max_w)
    k = np.arange(d.kernels.shape[0])
    kern = np.stack([(k % 251 + 1), (k % 13) - 6], axis=1).astype(np.float32)
    vis = rng.integers(-3, 4, size=(d.n_vis, 2)).astype(np.float32)
    one = oracle.std_gridding(d.grid_size, d.oversampling, kern, d.supports, d.uvw, vis, cfg)
    deg = oracle.std_degridding(d.grid_size, d.oversampling, kern, d.supports, one, d.uvw, cfg)
    for slices in (1, 3, 8):
        g, v, sec = oracle.sliced_grid_degrid(slices, d.grid_size, d.oversampling, kern, kern, d.supports, d.uvw, vis,
                                              cfg)
        np.testing.assert_array_equal(g, one)
        np.testing.assert_array_equal(v, deg)
        assert sec[0] >= 0 and sec[1] >= 0


def test_merge_and_split(oracle, rng):
    G, S = 8, 4
    grids = rng.normal(size=(S * G * G, 2)).astype(np.float32)
    np.testing.assert_allclose(oracle.merge(G, S, grids), grids.reshape(S, G * G, 2).sum(0), rtol=1e-6)
    assert list(oracle.split_counts(10, 4, 10)) == [2, 2, 2, 2]
    assert list(oracle.split_counts(12, 4, 7)) == [3, 3, 1, 0]


# ---------------------------------------------------------------- gridding_actor (gridding_run.c:63-133)

def _actor_taps(grid, G):
    """{(v, u): kernel index} of a grid made with index_kernel and vis = 1."""
    g = grid.reshape(G, G, 2)
    return {(v - G // 2, u - G // 2): int(g[v, u, 0]) for v, u in zip(*np.nonzero(g[..., 1]))}


def test_gridding_actor_snaps_to_sixteenths_and_takes_the_inclusive_range(oracle):
    """gx = 10.02 snaps to roundf(160.32) / 16 = 10.0 (gridding_run.c:79-82),
    and the inclusive range [ceil(s - h), floor(s + h)] (:84-92, :101-110)
    then holds 2h + 1 = 7 columns, 7..13, with |round((u - 10) * 16)| =
    48 32 16 0 16 32 48.  gy = -3.5 is already on a sixteenth: rows -6..-1
    (floor(-0.5) = -1), 6 of them.  Without the snap the columns would be
    8..13 with offsets 32 16 0 16 32 48; with std_gridding's half-open range
    7..12."""
    G, h = 64, 3
    kern, sup = index_kernel(h, 16)
    grid = oracle.gridding_actor(G, kern, sup, np.array([[10.02, -3.5, 0.0]], np.float32),
                                 np.array([[1.0, 0.0]], np.float32), cfg_for(G))
    kx = {7: 48, 8: 32, 9: 16, 10: 0, 11: 16, 12: 32, 13: 48}
    ky = {-6: 40, -5: 24, -4: 8, -3: 8, -2: 24, -1: 40}
    want = {(v, u): a * 64 + b for v, a in ky.items() for u, b in kx.items()}
    got = _actor_taps(grid, G)
    assert got == want
    no_snap = {(v, u): ky[v] * 64 + abs(round((u - 10.02) * 16)) for v in ky for u in range(8, 14)}
    half_open = {k: x for k, x in want.items() if k[1] <= 12}
    assert got != no_snap and got != half_open
    assert np.all(grid.reshape(G, G, 2)[..., 1][grid.reshape(G, G, 2)[..., 1] != 0] == 0.5)


def test_gridding_actor_plane_rounds_half_away_from_zero(oracle):
    """plane = (int)roundf(sqrtf(|w| * w_scale)) (gridding_run.c:67): with
    w_scale = 1, w = +-6.25 gives sqrtf = 2.5, plane 3 (banker's rounding
    or truncation would give 2); w = 2.2 gives 1.48, plane 1; w = -0.2
    gives 0.447, plane 0.  Negative w conjugates the kernel (:76)."""
    G, OS = 64, 16
    hs = [1, 1, 1, 1]
    sup = np.array([[1, p * 1024] for p in range(4)], np.int32)   # ((1 + 1) * 16)^2 = 1024 samples a plane
    n = 4 * 1024
    kern = np.stack([np.arange(n, dtype=np.float32), np.full(n, 0.25, np.float32)], axis=1)
    cfg = cfg_for(G, n_kernels=4, max_w=9.0)                      # w_scale = (4 - 1)^2 / 9 = 1
    assert cfg.w_scale == 1.0
    for w, plane in ((6.25, 3), (-6.25, 3), (2.2, 1), (-0.2, 0), (2.25, 2)):
        g = oracle.gridding_actor(G, kern, sup, np.array([[0.0, 0.0, w]], np.float32),
                                  np.array([[1.0, 0.0]], np.float32), cfg).reshape(G, G, 2)
        nz = g[..., 1] != 0
        assert np.count_nonzero(nz) == 9, w                        # s = 0: columns and rows -1..1
        idx = g[..., 0][nz]
        # tap (0, 0) reads sample 0 of its plane, whose real part is the plane offset
        assert np.all((idx >= plane * 1024) & (idx < (plane + 1) * 1024)), (w, plane)
        assert np.all(g[..., 1][nz] == (-0.25 if w < 0 else 0.25)), w


def test_gridding_actor_fixes_oversampling_at_16(oracle):
    """gridding_actor passes 16 to gridding_CPU whatever config->oversampling
    says (gridding_run.c:43): with oversampling = 8 in the Config, gx =
    10.0625 still snaps to 161 / 16 = 10.0625 (8 would give 10.125), the
    columns 8..13 take |round((u - 10.0625) * 16)| = 33 17 1 15 31 47 and
    rows step by (h + 1) * 16 = 64 samples."""
    G, h = 64, 3
    kern, sup = index_kernel(h, 16)
    cfg = cfg_for(G)
    cfg.oversampling = 8
    grid = oracle.gridding_actor(G, kern, sup, np.array([[10.0625, 0.0, 0.0]], np.float32),
                                 np.array([[1.0, 0.0]], np.float32), cfg)
    kx = {8: 33, 9: 17, 10: 1, 11: 15, 12: 31, 13: 47}
    ky = {-3: 48, -2: 32, -1: 16, 0: 0, 1: 16, 2: 32, 3: 48}
    assert _actor_taps(grid, G) == {(v, u): a * 64 + b for v, a in ky.items() for u, b in kx.items()}


def test_gridding_actor_float_sum_tracks_its_fp64_twin(oracle, rng):
    """The float grid against the same index arithmetic summed in double:
    the two differ by float rounding only (within 1e-5 of the peak); their
    tap sets are identical."""
    d = __import__("sdp_synth").small_case(n_vis=4000, grid_size=128, n_planes=4, oversampling=16, seed=21)
    cfg = cfg_for(d.grid_size, cell=d.cell_size, n_kernels=d.supports.shape[0], max_w=d.max_w)
    f32 = oracle.gridding_actor(d.grid_size, d.kernels, d.supports, d.uvw, d.vis, cfg)
    f64 = oracle.gridding_actor_f64(d.grid_size, d.kernels, d.supports, d.uvw, d.vis, d.uv_scale, d.w_scale)
    peak = np.abs(f64).max()
    assert peak > 0 and np.abs(f32 - f64).max() <= 1e-5 * peak
    np.testing.assert_array_equal(f32 != 0, f64 != 0)


# ---------------------------------------------------------------- Hogbom's "average"

def test_running_mean_recurrence_excludes_pixel_zero(oracle):
    """m = |r0|, m += (|r_i| - m) / i (deconvolution_run.c:226) is the mean
    of |r_1..r_{n-1}|: on [5, 2, 4, 6, 8] every step is exact, 5 -> 2 -> 3
    -> 4 -> 5 = mean(2, 4, 6, 8); the mean over all five would be 5 too, so
    also [100, 2, 4, 6, 8] -> 5, not 24."""
    assert oracle.running_mean(np.array([5, 2, 4, 6, 8], np.float32)) == 5.0
    assert oracle.running_mean(np.array([100, -2, 4, -6, 8], np.float32)) == 5.0
    assert oracle.running_mean(np.array([-3.5], np.float32)) == 3.5


def test_hogbom_noise_stop_excludes_pixel_zero(oracle):
    """Pixel 0 holds 100, every other pixel 1, noise_factor * gain = 80.
    The reference's average excludes pixel 0 (mean 1, threshold 80), so
    CLEAN takes pixel 0 three times (100, 90, 81) and stops at 72.9 < 80;
    a mean that counted pixel 0 (355 / 256 = 1.39, threshold 111) would stop
    at once with no component."""
    G = 16
    cfg = cfg_for(G, psf_max_value=1.0, noise_factor=800.0)
    res = np.ones(G * G, np.float32)
    res[0] = 100.0
    ns, src, model, r = oracle.hogbom_clean(G, 10, 10, res, delta_psf(G), cfg, np.zeros(G * G, np.float32))
    assert ns == 3
    np.testing.assert_allclose(src[:, 2], [10.0, 9.0, 8.1], rtol=1e-6)
    assert np.all(src[:, :2] == -8 * cfg.cell_size)            # pixel (0, 0) in radians
    assert r[0] == pytest.approx(72.9, rel=1e-6)


def test_running_mean_drift_at_4096(oracle):
    """The float recurrence drifts from the exact mean as the image grows:
    its increments (|r_i| - m) / i fall below half an ulp of m for i near
    2^24.  Measured on half-normal noise (three seeds): within 2e-6 of the
    exact mean at 128^2, 3e-5 at 1024^2, 1.2e-4 at 2048^2, and at 4096^2
    the float value sits 0.70% below it.  A stop decision within that band of the threshold is where
    a mean computed any other way can disagree with the reference."""
    rng = np.random.default_rng(0)
    small = rng.standard_normal(128 * 128).astype(np.float32)
    exact = np.abs(small[1:].astype(np.float64)).mean()
    assert oracle.running_mean(small) == pytest.approx(exact, rel=5e-6)
    big = rng.standard_normal(4096 * 4096).astype(np.float32)
    exact = np.abs(big[1:].astype(np.float64)).mean()
    drift = oracle.running_mean(big) / exact - 1.0
    assert -0.02 < drift < -0.001, drift


# ---------------------------------------------------------------- taper against the spheroidal function

def test_prolate_matches_the_spheroidal_function():
    """The taper's rational approximation (top.c:408-447, Schwab 1984) against
    the function it approximates, computed here from its differential
    equation: the alpha = 1, c = 3 pi (support 6) spheroidal function psi
    solves (1 - x^2) psi'' - 4 x psi' + (chi - c^2 x^2) psi = 0, whose
    Gegenbauer C_k^(3/2) expansion turns into a matrix eigenproblem
    (diag k (k + 3) + c^2 X^2) a = chi a, X the multiplication-by-x
    operator from the three-term recurrence; the smallest chi gives psi.
    The approximation's own error is 3.33e-7 (nu < 0.75) and 3.66e-7
    (nu >= 0.75) of psi; a change of the second-to-last printed digit of
    12 of its 14 coefficients, or of the last digit of 7 of them, pushes
    it past the bars."""
    import sdp_oracle
    from scipy.special import eval_gegenbauer
    c, lam, K = 3 * np.pi, 1.5, 60
    X = np.zeros((K + 1, K + 1))
    for k in range(K + 1):
        if k + 1 <= K:
            X[k + 1, k] = (k + 1) / (2 * (k + lam))
        if k >= 1:
            X[k - 1, k] = (k + 2 * lam - 1) / (2 * (k + lam))
    M = (np.diag([k * (k + 3.0) for k in range(K + 1)]) + c * c * X @ X)[:K - 2, :K - 2]
    w, v = np.linalg.eig(M)
    a = v[:, np.argmin(w.real)].real
    nu = np.linspace(0.0, 0.9999, 4000)
    psi = sum(a[k] * eval_gegenbauer(k, lam, nu) for k in range(len(a)))
    psi /= sum(a[k] * eval_gegenbauer(k, lam, 0.0) for k in range(len(a)))
    assert np.min(w.real) == pytest.approx(7.7154048, abs=1e-6)   # chi_0 of alpha = 1, c = 3 pi
    got = np.array([sdp_oracle.prolate_spheroidal(x) for x in nu])
    rel = np.abs(got - psi) / psi
    assert rel[nu < 0.75].max() <= 3.40e-7
    assert rel[nu >= 0.75].max() <= 3.70e-7
