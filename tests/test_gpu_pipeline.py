"""The device-resident major cycle (sdp_pipeline.h) and the helpers it is
built from, against the oracle and the FP64 imaging chain
(tests/imaging_ref.py), through the C-ABI.

Bars: copies, fills, divides and sign flips bit-exact; Hogbom bit-exact
against the oracle; images within 2e-5 of their peak before the taper
correction, i.e. |got - ref| * taper(r) * taper(c) <= 2e-5 * peak (float FFT
and float accumulation against FP64; the correction divides by a taper that
falls to 1e-5 at the edges and magnifies rounding there by as much); RMS
values within 1e-3 relative.
"""
import numpy as np
import pytest

import imaging_ref as R
import sdp_synth

pytestmark = pytest.mark.gpu

TOL_IMAGE = 2e-5


def up(hip, a):
    return hip.DeviceBuffer.from_array(np.ascontiguousarray(a))


@pytest.fixture(scope="module")
def sky():
    return sdp_synth.sky_case()


def make_pipeline(hip, d, n_minor, flags=0, vis=None, noise_factor=0.0, weak=0.0):
    G = d.grid_size
    bufs = [up(hip, d.uvw), up(hip, d.vis if vis is None else vis), up(hip, d.kernels),
            up(hip, d.degrid_kernels), up(hip, sdp_synth.taper_padded(G))]
    p = hip.Pipeline(G, d.oversampling, d.supports, d.uv_scale, d.w_scale, d.cell_size, bufs[0], bufs[1], d.n_vis,
                     bufs[2], bufs[3], bufs[4], n_minor, loop_gain=0.1, noise_factor=noise_factor,
                     weak_source_percent=weak, flags=flags)
    p._inputs = bufs
    return p


def assert_image_close(got, ref, peak):
    t = R.taper_2d(got.shape[0])
    err = float(np.max(np.abs(got.astype(np.float64) - ref) * t))
    assert err <= TOL_IMAGE * peak, (err, peak)


def rms(img):
    return float(np.sqrt(np.mean(np.asarray(img, np.float64) ** 2)))


# ---------------------------------------------------------------- helpers

def test_fft_correct_in_equals_separate_correction(hip, oracle, rng):
    """CORRECT_IN divides by the taper exactly as the correction actor does,
    before the shift and FFT: bit-exact against correcting first."""
    G = 256
    img = rng.normal(size=G * G).astype(np.float32)
    taper = sdp_synth.taper_padded(G)
    corrected = oracle.convolution_correction(G, img, taper)
    d_img, d_cor, d_tap = up(hip, img), up(hip, corrected), up(hip, taper)
    o1, o2 = hip.DeviceBuffer(8 * G * G), hip.DeviceBuffer(8 * G * G)
    w = hip.DeviceBuffer(hip.fft_work_bytes(G))
    base = hip.FFT_REAL_IN | hip.FFT_SHIFT_IN | hip.FFT_SHIFT_OUT
    hip.fft2d(d_img, o1, G, hip.FFT_FORWARD, base | hip.FFT_CORRECT_IN, d_tap, w)
    hip.fft2d(d_cor, o2, G, hip.FFT_FORWARD, base, None, w)
    np.testing.assert_array_equal(o1.download((G * G, 2), np.float32), o2.download((G * G, 2), np.float32))
    with pytest.raises(hip.SdpError):
        hip.fft2d(d_img, o1, G, hip.FFT_FORWARD, hip.FFT_CORRECT_IN, d_tap, w)   # needs REAL_IN


def test_image_stats_fill_divide(hip, rng):
    n = 300_001
    img = rng.normal(size=n).astype(np.float32)
    d = up(hip, img)
    sq, mx, amx = hip.image_stats(n, d)
    assert sq == pytest.approx(float(np.sum(img.astype(np.float64) ** 2)), rel=1e-12)
    assert mx == img.max() and amx == np.abs(img).max()
    hip.divide(n, d, 3.7)
    np.testing.assert_array_equal(d.download(n, np.float32), (img.astype(np.float64) / 3.7).astype(np.float32))
    c = hip.DeviceBuffer(8 * 1001)
    hip.fill_complex(1001, c, 1.5, -2.0)
    got = c.download((1001, 2), np.float32)
    assert (got[:, 0] == 1.5).all() and (got[:, 1] == -2.0).all()


def mirrored_layout(psf, G):
    """A PSF array that, read the reference's mirrored way
    (deconvolution_run.c:269), yields `psf` read in its own orientation."""
    out = np.zeros(G * G + 1, np.float32)
    p = psf.reshape(G, G)
    for r in range(G):
        out[r * G + 1:r * G + G] = p[r, G - 1:0:-1]     # column c -> index r*G + G - c, c >= 1
        out[r * G + G] = p[r, 0]                         # c = 0 wraps to the next row's first element
    return out


def test_hogbom_psf_as_stored_bit_exact(hip, oracle, rng):
    """The stored-orientation read equals the oracle's mirrored read of the
    re-laid-out PSF, bit for bit (residual, model, components)."""
    G = 64
    psf = np.exp(-((np.arange(G) - G // 2)[:, None] ** 2 / 8.0 + (np.arange(G) - G // 2)[None, :] ** 2 / 20.0))
    psf[G // 2, G // 2 + 3] = 0.7                          # break the x symmetry
    psf = psf.astype(np.float32).ravel()
    res = np.zeros(G * G, np.float32)
    for _ in range(6):
        y, x = rng.integers(4, G - 4, 2)
        res += rng.uniform(0.5, 2.0) * np.roll(np.roll(psf.reshape(G, G), y - G // 2, 0), x - G // 2, 1).ravel()
    res += rng.normal(0, 0.01, G * G).astype(np.float32)
    res = res.astype(np.float32)
    cfg = hip.make_config(G, 1e-4, psf_max_value=1.0)
    model = np.zeros(G * G, np.float32)
    rns, rsrc, rmodel, rres = oracle.hogbom_clean(G, 120, 200, res, mirrored_layout(psf, G), cfg, model)
    d_res = up(hip, res)
    d_psf = up(hip, np.concatenate([psf, np.zeros(1, np.float32)]))
    d_m, d_mo = up(hip, model), hip.DeviceBuffer(4 * G * G)
    d_src, d_n, d_w = hip.DeviceBuffer(12 * 200), hip.DeviceBuffer(16), hip.DeviceBuffer(hip.hogbom_work_bytes(G))
    hip.hogbom_ex(G, 120, 200, d_res, d_psf, 1.0, cfg.loop_gain, cfg.noise_factor, cfg.weak_source_percent_img,
                  cfg.cell_size, d_m, d_mo, d_src, d_n, d_w, hip.HOGBOM_PSF_AS_STORED)
    ns = int(d_n.download(1, np.int32)[0])
    assert ns == rns and ns > 20
    np.testing.assert_array_equal(d_src.download((200, 3), np.float32)[:ns], rsrc)
    np.testing.assert_array_equal(d_mo.download(G * G, np.float32), rmodel)
    np.testing.assert_array_equal(d_res.download(G * G, np.float32), rres)


# ---------------------------------------------------------------- pipeline

def test_pipeline_psf(hip, sky):
    p = make_pipeline(hip, sky, n_minor=0)
    ref, peak = R.psf(sky)
    assert p.psf_max == pytest.approx(peak, rel=1e-5)
    got = p.image(hip.IMAGE_PSF)
    assert_image_close(got, ref, 1.0)
    p.close()


def test_pipeline_dirty_image(hip, sky):
    """With no minor iterations a cycle returns the corrected dirty image."""
    p = make_pipeline(hip, sky, n_minor=0)
    before, after, ms = p.cycle()
    ref = R.dirty(sky, sky.vis)
    got = p.image(hip.IMAGE_RESIDUAL)
    assert_image_close(got, ref, np.abs(ref).max())
    _, peak = R.psf(sky)
    assert before == after
    assert before == pytest.approx(rms(ref) / peak, rel=1e-3)
    assert ms > 0
    p.close()


def test_pipeline_predict_of_loaded_model(hip, sky):
    """Loading the true sky as the model: the cycle's residual image equals
    the FP64 chain's dirty image of measured - predicted, and is two orders
    below the dirty image."""
    p = make_pipeline(hip, sky, n_minor=0)
    model = R.sky_model(sky)
    d_model = up(hip, model)
    p.load_model(d_model)
    before, _, _ = p.cycle()
    ref = R.dirty(sky, (sky.vis - R.predict(sky, model)).astype(np.float32))
    dirty = R.dirty(sky, sky.vis)
    got = p.image(hip.IMAGE_RESIDUAL)
    assert_image_close(got, ref, np.abs(dirty).max())
    assert rms(got) < rms(dirty) / 50
    np.testing.assert_array_equal(p.image(hip.IMAGE_MODEL), model)   # n_minor = 0 leaves the model alone
    p.close()


def test_pipeline_converges(hip, sky):
    """Residual RMS falls every cycle; the model's flux approaches the sky's;
    the brightest component sits on the brightest source."""
    p = make_pipeline(hip, sky, n_minor=200)
    rms_before = []
    for _ in range(4):
        before, after, _ = p.cycle()
        assert after <= before
        rms_before.append(before)
    assert all(b < a for a, b in zip(rms_before, rms_before[1:])), rms_before
    assert rms_before[-1] < 0.1 * rms_before[0], rms_before
    model = p.image(hip.IMAGE_MODEL)
    assert model.sum() == pytest.approx(float(sky.sources[:, 2].sum()), rel=0.15)
    x, y = sky.extra["pixels"][int(np.argmax(sky.sources[:, 2]))]
    assert np.unravel_index(np.argmax(model), model.shape) == (y, x)
    src = p.sources()
    assert 0 < src.shape[0] <= 200
    p.close()


def test_pipeline_graph_replay_equals_eager(hip, sky):
    """Cycles replayed from the captured graph give the same numbers and
    images, bit for bit, as eager cycles, replay after replay."""
    eager = make_pipeline(hip, sky, n_minor=50)
    graph = make_pipeline(hip, sky, n_minor=50, flags=hip.PIPE_GRAPH)
    for _ in range(4):
        e = eager.cycle()
        g = graph.cycle()
        assert e[:2] == g[:2]
        np.testing.assert_array_equal(eager.image(hip.IMAGE_RESIDUAL), graph.image(hip.IMAGE_RESIDUAL))
        np.testing.assert_array_equal(eager.image(hip.IMAGE_MODEL), graph.image(hip.IMAGE_MODEL))
    graph.reset()
    eager.reset()
    assert graph.cycle()[:2] == eager.cycle()[:2]
    eager.close()
    graph.close()


def test_pipeline_profile_stages(hip, sky):
    """Profiling records every stage and changes no result."""
    eager = make_pipeline(hip, sky, n_minor=30)
    prof = make_pipeline(hip, sky, n_minor=30, flags=hip.PIPE_PROFILE | hip.PIPE_GRAPH)
    for _ in range(2):
        e = eager.cycle()
        p = prof.cycle()
        assert e[:2] == p[:2]
        st = prof.stage_ms()
        assert set(st) == set(hip.STAGES) and all(v > 0 for v in st.values())
        assert sum(st.values()) <= p[2] * 1.01 + 0.05
    with pytest.raises(hip.SdpError):
        eager.stage_ms()
    eager.close()
    prof.close()


def test_pipeline_psf_orientation_matters(hip, sky):
    """Subtracting the PSF in its stored orientation cleans deeper than the
    reference's x-mirrored read (the w-term makes the PSF asymmetric)."""
    a = make_pipeline(hip, sky, n_minor=200)
    b = make_pipeline(hip, sky, n_minor=200, flags=hip.PIPE_PSF_MIRRORED)
    ra = [a.cycle()[0] for _ in range(3)]
    rb = [b.cycle()[0] for _ in range(3)]
    assert ra[0] == rb[0]
    assert ra[-1] < rb[-1]
    a.close()
    b.close()


def test_pipeline_rejects_bad_parameters(hip, sky):
    with pytest.raises(hip.SdpError):
        make_pipeline(hip, sky, n_minor=-1)
