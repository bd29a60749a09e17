"""Batched multi-channel residual imaging (sdp_batch_*) against the FP64
imaging chain (tests/imaging_ref.py), on a reduced multi-channel workload:
four channels of the config-4 array and sky with their own frequencies,
flagging and visibility counts.

Bar: every channel's dirty image within 1e-5 of its peak of the FP64
chain's, compared before the taper division (error times taper: the
division amplifies rounding near the field edge equally for any
arithmetic); the graph-replayed pass equals the eager one bit for bit."""
import types

import numpy as np
import pytest

import imaging_ref as R
import sdp_synth as S

pytestmark = pytest.mark.gpu

TOL = 1e-5


def up(hip, a):
    return hip.DeviceBuffer.from_array(np.ascontiguousarray(a))


@pytest.fixture(scope="module")
def mc():
    return S.config4(n_channels=4, n_time=48)


def channel_data(m, c):
    uvw = m.channel_uvw(c)
    vis = S.dft_f64(uvw, m.channel_sources(c)).astype(np.float32)
    d = types.SimpleNamespace(grid_size=m.grid_size, oversampling=m.oversampling, kernels=m.kernels,
                              degrid_kernels=m.degrid_kernels, supports=m.supports, uvw=uvw, uv_scale=m.uv_scale,
                              w_scale=m.w_scale, n_vis=uvw.shape[0])
    return d, vis


def make_batch(hip, m, data, flags):
    bufs = [(up(hip, d.uvw), up(hip, v)) for d, v in data]
    common = [up(hip, m.kernels), up(hip, m.degrid_kernels), up(hip, S.taper_padded(m.grid_size))]
    b = hip.Batch(m.grid_size, m.oversampling, m.supports, m.uv_scale, m.w_scale, [u for u, _ in bufs],
                  [v for _, v in bufs], [d.n_vis for d, _ in data], *common, flags=flags)
    b._inputs = (bufs, common)
    return b


def assert_image_close(got, ref):
    t = R.taper_2d(got.shape[0])
    err = float(np.max(np.abs(got.astype(np.float64) - ref) * t))
    peak = float(np.abs(ref).max())
    assert err <= TOL * peak, (err, peak)
    return err / peak


def test_batch_dirty_images_match_fp64_and_graph_replays(hip, mc):
    data = [channel_data(mc, c) for c in range(len(mc.freqs))]
    assert len({d.n_vis for d, _ in data}) == len(data)          # uneven counts
    b = make_batch(hip, mc, data, hip.BATCH_GRAPH)
    ms, _ = b.residuals()                                        # eager first pass
    first = [b.residual(c) for c in range(len(data))]
    for c, (d, vis) in enumerate(data):
        e = assert_image_close(first[c], R.dirty(d, vis))
        print(f"channel {c}: {d.n_vis} vis, dirty-image error {e:.2e} of peak")
    for _ in range(2):                                           # captured, then replayed
        ms2, _ = b.residuals()
        for c in range(len(data)):
            np.testing.assert_array_equal(b.residual(c), first[c])
    assert ms > 0 and ms2 > 0
    b.close()


def test_batch_predict_of_channel_models(hip, mc):
    """Per-channel models: each channel's residual is the FP64 chain's dirty
    image of its measured minus predicted visibilities."""
    data = [channel_data(mc, c) for c in range(2)]
    b = make_batch(hip, mc, data, 0)
    G = mc.grid_size
    rng = np.random.default_rng(4)
    models = []
    for c in range(2):
        model = np.zeros((G, G), np.float32)
        ys, xs = rng.integers(G // 2 - 40, G // 2 + 40, (2, 12))
        model[ys, xs] = rng.uniform(0.1, 1.0, 12)
        hip.check(hip.load().sdp_memcpy_htod(b.model_ptr(c), model.ctypes.data, model.nbytes, None))
        models.append(model)
    b.residuals()
    for c, (d, vis) in enumerate(data):
        ref = R.dirty(d, (vis - R.predict(d, models[c])).astype(np.float32))
        assert_image_close(b.residual(c), ref)
    b.close()


def test_batch_rejects_bad_input(hip, mc):
    d, vis = channel_data(mc, 0)
    with pytest.raises(hip.SdpError):
        hip.Batch(mc.grid_size, mc.oversampling, mc.supports, mc.uv_scale, mc.w_scale, [up(hip, d.uvw)],
                  [up(hip, vis)], [0], up(hip, mc.kernels), up(hip, mc.degrid_kernels),
                  up(hip, S.taper_padded(mc.grid_size)))


def test_multi_grid_plan_equals_per_grid_plans(hip, mc):
    """One plan over four channels (sdp_grid_plan_create_multi) against a
    plan per channel with the same tile and work-item size: the fused
    degrid + subtract + grid pass, the gridder alone and the degridder
    alone give every channel's grid and visibilities bit for bit (the
    channel is only a field of the sort key)."""
    data = [channel_data(mc, c) for c in range(len(mc.freqs))]
    G, nc = mc.grid_size, len(data)
    rng = np.random.default_rng(9)
    models = rng.normal(size=(nc, G * G, 2)).astype(np.float32)
    kw = dict(tile=8, chunk=2048)
    gk, dk = up(hip, mc.kernels), up(hip, mc.degrid_kernels)
    uvws = [up(hip, d.uvw) for d, _ in data]
    counts = [d.n_vis for d, _ in data]
    multi = hip.GridPlan.multi(G, mc.oversampling, mc.supports, mc.uv_scale, mc.w_scale, uvws, counts, **kw)
    assert multi.info()["chunks"] > 4 * nc
    n = sum(counts)
    meas = up(hip, np.concatenate([v for _, v in data]))
    d_models = up(hip, models)
    out_grid, out_vis = hip.DeviceBuffer(8 * G * G * nc), hip.DeviceBuffer(8 * n)
    multi.degrid_grid(dk, gk, d_models, meas, out_grid, d_residual_vis=out_vis)
    m_grid = out_grid.download((nc, G * G, 2), np.float32)
    m_vis = out_vis.download((n, 2), np.float32)
    multi.grid(gk, meas, out_grid)
    g_grid = out_grid.download((nc, G * G, 2), np.float32)
    multi.degrid(dk, d_models, out_vis)
    d_vis = out_vis.download((n, 2), np.float32)
    first = 0
    for c, (d, vis) in enumerate(data):
        one = hip.GridPlan(G, mc.oversampling, mc.supports, mc.uv_scale, mc.w_scale, uvws[c], d.n_vis, **kw)
        mv, mo = up(hip, vis), up(hip, models[c])
        og, ov = hip.DeviceBuffer(8 * G * G), hip.DeviceBuffer(8 * d.n_vis)
        one.degrid_grid(dk, gk, mo, mv, og, d_residual_vis=ov)
        part = slice(first, first + d.n_vis)
        assert np.count_nonzero(m_grid[c]) > 1000
        np.testing.assert_array_equal(og.download((G * G, 2), np.float32), m_grid[c])
        np.testing.assert_array_equal(ov.download((d.n_vis, 2), np.float32), m_vis[part])
        one.grid(gk, mv, og)
        np.testing.assert_array_equal(og.download((G * G, 2), np.float32), g_grid[c])
        one.degrid(dk, mo, ov)
        np.testing.assert_array_equal(ov.download((d.n_vis, 2), np.float32), d_vis[part])
        one.close()
        first += d.n_vis
    multi.close()


@pytest.mark.parametrize("G,nb", [(256, 3), (2048, 4)])
@pytest.mark.parametrize("direction,flags", [("forward", "real_in"), ("inverse", "real_out")])
def test_fft_batch_equals_one_grid_at_a_time(hip, direction, flags, G, nb):
    """sdp_fft2d_batch over several grids equals one sdp_fft2d call per grid
    bit for bit, with the major cycle's fused shift / real / taper flags.  At
    2048^2 the batch of four takes the 16-column blocks of the column pass and
    the single grids the 8-column blocks."""
    rng = np.random.default_rng(5)
    d_tap = up(hip, S.taper_padded(G))
    if flags == "real_in":
        fl = hip.FFT_REAL_IN | hip.FFT_SHIFT_IN | hip.FFT_SHIFT_OUT | hip.FFT_CORRECT_IN
        x = rng.normal(size=(nb, G * G)).astype(np.float32)
        out_shape, out_bytes = (nb, G * G, 2), 8
    else:
        fl = hip.FFT_SHIFT_IN | hip.FFT_SHIFT_OUT | hip.FFT_REAL_OUT | hip.FFT_CORRECT_OUT
        x = rng.normal(size=(nb, G * G, 2)).astype(np.float32)
        out_shape, out_bytes = (nb, G * G), 4
    sign = hip.FFT_FORWARD if direction == "forward" else hip.FFT_INVERSE
    d_in = up(hip, x)
    d_out = hip.DeviceBuffer(out_bytes * G * G * nb)
    work = hip.DeviceBuffer(8 * G * G * nb)
    hip.fft2d_batch(d_in, d_out, G, nb, sign, fl, d_tap, work)
    got = d_out.download(out_shape, np.float32)
    for b in range(nb):
        one_in = up(hip, x[b])
        one_out = hip.DeviceBuffer(out_bytes * G * G)
        hip.fft2d(one_in, one_out, G, sign, fl, d_tap, hip.DeviceBuffer(8 * G * G))
        np.testing.assert_array_equal(one_out.download(out_shape[1:], np.float32), got[b])
