"""The learned minor cycle on the GPU: the patch kernels against numpy, the
training-pair convolution against numpy's FFT, the fallback that undoes a
harmful proposal, and what a network trained by the library's own kernels
gains over plain Hogbom on a small learned-proposal workload."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def up(hip, a, dtype=np.float32):
    return hip.DeviceBuffer.from_array(np.ascontiguousarray(a, dtype=dtype))


def candidates(G, T, margin):
    stride = T - 2 * margin
    npx = -(-G // stride)
    return [((q % npx) * stride - margin, (q // npx) * stride - margin) for q in range(npx * npx)]


def test_patch_peaks_gather_scatter_match_numpy(hip, rng):
    G, T, margin = 96, 32, 6
    lib = hip.load()
    res = rng.normal(size=(G, G)).astype(np.float32)
    psf = rng.normal(size=(G, G)).astype(np.float32)
    d_res, d_psf = up(hip, res), up(hip, psf)
    cand = candidates(G, T, margin)
    d_peaks = hip.DeviceBuffer(4 * len(cand))
    n = ctypes.c_int()
    hip.check(lib.sdp_cnn_patch_peaks(G, T, margin, d_res.ptr, d_peaks.ptr, ctypes.byref(n), None))
    assert n.value == len(cand)
    pad = np.zeros((G + 2 * T, G + 2 * T), np.float32)
    pad[T:T + G, T:T + G] = res
    tiles = [pad[T + y0:T + y0 + T, T + x0:T + x0 + T] for x0, y0 in cand]
    np.testing.assert_array_equal(d_peaks.download(len(cand), np.float32), [np.abs(t).max() for t in tiles])

    sel = [0, 4, 7, len(cand) - 1]                     # corners and an interior tile, centres disjoint
    origins = np.array([cand[q] for q in sel], np.int32)
    scales = rng.uniform(0.5, 2.0, len(sel)).astype(np.float32)
    d_org, d_sc = up(hip, origins, np.int32), up(hip, scales)
    d_batch = hip.DeviceBuffer(4 * len(sel) * T * T * 2)
    hip.check(lib.sdp_cnn_gather_patches(G, T, len(sel), d_org.ptr, d_sc.ptr, d_res.ptr, d_psf.ptr, d_batch.ptr,
                                         None))
    batch = d_batch.download((len(sel), T, T, 2), np.float32)
    c = G // 2 - T // 2
    for b, q in enumerate(sel):
        np.testing.assert_array_equal(batch[b, :, :, 0], tiles[q] / scales[b])
        np.testing.assert_array_equal(batch[b, :, :, 1], psf[c:c + T, c:c + T])

    out = rng.normal(size=(len(sel), T, T)).astype(np.float32)
    model = rng.normal(size=(G, G)).astype(np.float32)
    for in_floor, use_in in ((0.0, False), (0.3, True)):
        d_out, d_model = up(hip, out), up(hip, model)
        hip.check(lib.sdp_cnn_scatter_add(G, T, margin, len(sel), d_org.ptr, d_sc.ptr, d_out.ptr,
                                          ctypes.c_float(0.25), ctypes.c_float(0.1),
                                          d_batch.ptr if use_in else None, ctypes.c_float(in_floor), d_model.ptr,
                                          None))
        want = model.copy()
        for b, (x0, y0) in enumerate(origins):
            for i in range(margin, T - margin):
                for j in range(margin, T - margin):
                    gy, gx = y0 + i, x0 + j
                    v = out[b, i, j]
                    keep = v > 0.1 and (not use_in or batch[b, i, j, 0] > in_floor)
                    if 0 <= gy < G and 0 <= gx < G and keep:
                        want[gy, gx] += np.float32(np.float32(0.25) * scales[b]) * v
        np.testing.assert_allclose(d_model.download((G, G), np.float32), want, rtol=0, atol=1e-6)


def test_convolver_matches_numpy_circular_convolution(hip, rng):
    import sdp_learned as L
    T = 64
    psf = rng.normal(size=(T, T))
    sky = rng.normal(size=(T, T)).astype(np.float32)
    got = L.Convolver(psf)(sky)
    # the PSF's peak position T/2 is the convolution's origin
    want = np.fft.ifft2(np.fft.fft2(sky.astype(np.float64)) * np.fft.fft2(np.fft.ifftshift(psf))).real
    assert np.abs(got - want).max() <= 1e-5 * np.abs(want).max()


@pytest.fixture(scope="module")
def small_workload(hip):
    import sdp_learned as L
    return L.learned_workload(n_time=64, grid_size=512, seed=5)


def test_fallback_undoes_a_harmful_proposal(hip, small_workload):
    """An untrained network at a large gain makes the residual worse; the
    next cycle restores the model and from then on the cycles are plain
    Hogbom's, bit for bit."""
    import sdp_learned as L
    w = small_workload
    T, margin = 64, 8
    net = hip.Cnn(2, 4, T, T, 8, precision=hip.CNN_FP32)
    L.init_params(net, 11)
    plain = w.pipeline(100)
    ref = [plain.cycle() for _ in range(3)]
    lp = w.pipeline(100)
    lp.set_learned(net, T, margin, gain=50.0, threshold=0.0, floor=-1e30, max_cycles=0)
    assert lp.learned_active
    first = lp.cycle()
    assert first[0] == ref[0][0]
    assert first[1] > ref[0][0]                       # the proposal raised the residual
    rows = [lp.cycle() for _ in range(3)]
    assert not lp.learned_active
    assert rows[0][0] == ref[0][0]                    # proposal undone: the dirty residual again
    for a, b in zip(rows, ref):
        assert a[0] == b[0] and a[1] == b[1]
    plain.close()
    lp.close()
    net.close()


def test_trained_network_gains_over_hogbom(hip, small_workload):
    """A network trained by the library's kernels on pairs made from the
    pipeline's PSF reaches the residual RMS of 5 x 300 plain Hogbom
    iterations with at most half the Hogbom iterations, and cleans deeper
    than Hogbom in the first cycle."""
    import sdp_learned as L
    w = small_workload
    G, T, margin, n_minor = w.ds.grid_size, 128, 16, 300
    plain = w.pipeline(n_minor)
    psf = plain.image(hip.IMAGE_PSF)
    hog = [plain.cycle() for _ in range(5)]
    target = hog[-1][1]
    pool = L.make_pool(psf, 1024, seed=1, T=T, crops_per_sky=8, global_scale=True,
                       crop_range=(G // 2 - T - 32, G // 2 + 32), n_points=(20, 80), n_gauss=(0, 20), box=w.box)
    net = hip.Cnn(3, 16, T, T, 16, precision=hip.CNN_BF16)
    L.init_params(net, 0)
    losses = L.train(net, pool, 800, 16, lr=3e-4, lr_final=3e-5, data_weight=1.0, psf=L.psf_patch(psf, T))
    assert np.mean(losses[-100:]) < np.mean(losses[:100])
    lp = w.pipeline(n_minor // 2)
    lp.set_learned(net, T, margin, gain=1.0, threshold=0.05, floor=0.0, max_cycles=3,
                   scale=hip.LEARNED_GLOBAL_SCALE)
    ms, rows = L.cycles_to_target(lp, target, 5)
    print(f"plain: {[f'{b:.4g}' for _, b, _ in hog]} {sum(r[2] for r in hog):.1f} ms; "
          f"learned: {[f'{b:.4g}' for _, b, _ in rows]} {ms} ms")
    assert rows[0][1] < hog[0][1]
    assert ms is not None and len(rows) * (n_minor // 2) <= 0.5 * 5 * n_minor
    plain.close()
    lp.close()
    net.close()


def test_sky_render_matches_the_host_sky(hip):
    """sdp_sky_render against render_sky (numpy, double) for seeded skies
    with stamps clipped at the field edge: equal to 1e-6 of the peak (the
    two differ only by the last bits of exp, sin and cos in double)."""
    import sdp_learned as L
    G = 192
    rng = np.random.default_rng(7)
    d_sky = hip.DeviceBuffer(4 * G * G)
    for _ in range(4):
        pts, gs = L.sky_components(G, rng, n_points=(5, 40), n_gauss=(3, 12), margin=2)
        hip.sky_render(G, up(hip, pts, np.float64), len(pts), up(hip, gs, np.float64), len(gs), d_sky)
        got = d_sky.download((G, G), np.float32)
        want = L.render_sky(G, pts, gs)
        assert np.count_nonzero(want) > 1000
        assert np.abs(got - want).max() <= 1e-6 * np.abs(want).max()


def test_device_pool_matches_the_host_pool(hip, rng):
    """make_pool_device against make_pool (global scale) from the same seed:
    the same pairs, inputs, targets and dirty patches, to 1e-6 of each
    patch's scale."""
    import sdp_learned as L
    G, T = 256, 64
    psf = L.smooth(np.pad(np.ones((1, 1), np.float32), ((G // 2, G // 2 - 1), (G // 2, G // 2 - 1))), 3.0)
    psf = (psf / psf.max()).astype(np.float32)
    kw = dict(crops_per_sky=4, crop_range=(G // 2 - T - 16, G // 2 + 16), n_points=(5, 20), n_gauss=(0, 8), box=40)
    host = L.make_pool(psf, 22, seed=3, T=T, global_scale=True, **kw)
    dev = L.make_pool_device(psf, 22, seed=3, T=T, with_dirty=True, **kw)
    got = dev.download()
    d = dev.d_d.download((22, T, T), np.float32)
    assert np.abs(got.inputs - host.inputs).max() <= 1e-6
    assert np.abs(got.targets - host.targets).max() <= 1e-6
    np.testing.assert_array_equal(d, got.inputs[..., 0])
    assert np.count_nonzero(host.targets) > 100 and np.abs(host.inputs[..., 0]).max() <= 1.0
