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
