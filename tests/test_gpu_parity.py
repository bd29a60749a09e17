"""HIP path against the CPU oracle, through the C-ABI (actor and device layers).

Bars: integer / index / copy work is bit-exact; floating-point work is within
the tolerance written in each test, relative to the data's peak.
"""
import dataclasses

import numpy as np
import pytest

import sdp_synth

pytestmark = pytest.mark.gpu


def cfg_of(hip, d, **kw):
    return hip.make_config(d.grid_size, d.cell_size, oversampling=d.oversampling,
                           num_kernels=d.supports.shape[0], max_w=d.max_w, **kw)


def integer_kernels(n):
    k = np.arange(n)
    return np.stack([(k % 251 + 1), (k % 13) - 6], axis=1).astype(np.float32)


@pytest.fixture(scope="module")
def case():
    return sdp_synth.small_case(n_vis=3000, grid_size=128, n_planes=4, oversampling=16, seed=7)


# ---------------------------------------------------------------- gridding

@pytest.mark.parametrize("snap", ["std", "actor"])
def test_grid_indices_bit_exact(hip, actors, oracle, case, rng, snap):
    """Integer-valued kernels and visibilities make every sum exact in float
    whatever the order, so any index difference shows as a mismatch."""
    d = case
    kern = integer_kernels(d.kernels.shape[0])
    vis = rng.integers(-3, 4, size=(d.n_vis, 2)).astype(np.float32)
    cfg = cfg_of(hip, d)
    if snap == "std":
        got = actors.std_gridding(d.grid_size, d.oversampling, kern, d.supports, d.uvw, vis, cfg)
        ref = oracle.std_gridding(d.grid_size, d.oversampling, kern, d.supports, d.uvw, vis, cfg)
    else:
        d16 = sdp_synth.small_case(n_vis=3000, grid_size=128, n_planes=4, oversampling=16, seed=8)
        kern = integer_kernels(d16.kernels.shape[0])
        got = actors.gridding_actor(d16.grid_size, kern, d16.supports, d16.uvw, vis, cfg_of(hip, d16))
        ref = oracle.gridding_actor(d16.grid_size, kern, d16.supports, d16.uvw, vis, cfg_of(hip, d16))
    assert np.count_nonzero(ref) > 1000
    np.testing.assert_array_equal(got, ref)


def test_grid_float_accuracy(hip, actors, oracle, case):
    """The GPU grid is at least as close to the FP64 grid as the reference's
    sequential float sum, and within 1e-6 of the peak."""
    d = case
    cfg = cfg_of(hip, d)
    got = actors.std_gridding(d.grid_size, d.oversampling, d.kernels, d.supports, d.uvw, d.vis, cfg)
    ref32 = oracle.std_gridding(d.grid_size, d.oversampling, d.kernels, d.supports, d.uvw, d.vis, cfg)
    ref64 = oracle.std_gridding_f64(d.grid_size, d.oversampling, d.kernels, d.supports, d.uvw, d.vis,
                                    d.uv_scale, d.w_scale)
    peak = np.abs(ref64).max()
    err_gpu = np.abs(got - ref64).max()
    err_ref = np.abs(ref32 - ref64).max()
    assert err_gpu <= 1e-6 * peak
    assert err_gpu <= err_ref + 1e-7 * peak
    np.testing.assert_allclose(got, ref32, atol=2e-6 * peak)


@pytest.mark.parametrize("G,tile", [(24, 4), (80, 4), (96, 8), (150, 16)])
def test_grid_sizes_off_the_merge_block(hip, actors, oracle, G, tile):
    """Grid sizes that are not multiples of the merge's 16- or 32-cell output
    blocks (and one that is not a power of two at all): every cell of the grid,
    including the partial blocks at its far edges, matches the FP64 grid."""
    d = sdp_synth.small_case(n_vis=4000, grid_size=G, n_planes=3, oversampling=16, seed=G)
    uvw = hip.DeviceBuffer.from_array(d.uvw)
    vis = hip.DeviceBuffer.from_array(d.vis)
    kern = hip.DeviceBuffer.from_array(d.kernels)
    plan = hip.GridPlan(G, d.oversampling, d.supports, d.uv_scale, d.w_scale, uvw, d.n_vis, tile=tile)
    g = hip.DeviceBuffer(8 * G * G)
    g.upload(np.full((G * G, 2), np.nan, np.float32))          # every cell must be written
    plan.grid(kern, vis, g)
    got = g.download((G * G, 2), np.float32)
    ref64 = oracle.std_gridding_f64(G, d.oversampling, d.kernels, d.supports, d.uvw, d.vis, d.uv_scale, d.w_scale)
    assert np.isfinite(got).all()
    assert np.abs(got - ref64).max() <= 1e-6 * np.abs(ref64).max()


@pytest.mark.parametrize("mode", ["blocked", "compensated"])
def test_grid_accuracy_dense_cells(hip, oracle, mode):
    """300,000 visibilities on a 128^2 grid: each cell sums thousands of
    taps.  Both accumulation modes stay within 1e-6 of the peak of the FP64
    grid and well inside the reference's own float error (measured about
    1.5e-6 of the peak here; the modes 2e-7 and 7e-8)."""
    d = sdp_synth.small_case(n_vis=300_000, grid_size=128, n_planes=4, oversampling=16, seed=8)
    G = d.grid_size
    acc = hip.ACC_BLOCKED if mode == "blocked" else hip.ACC_COMPENSATED
    uvw = hip.DeviceBuffer.from_array(d.uvw)
    vis = hip.DeviceBuffer.from_array(d.vis)
    kern = hip.DeviceBuffer.from_array(d.kernels)
    plan = hip.GridPlan(G, d.oversampling, d.supports, d.uv_scale, d.w_scale, uvw, d.n_vis, accumulate=acc)
    g = hip.DeviceBuffer(8 * G * G)
    plan.grid(kern, vis, g)
    got = g.download((G * G, 2), np.float32)
    ref64 = oracle.std_gridding_f64(G, d.oversampling, d.kernels, d.supports, d.uvw, d.vis, d.uv_scale, d.w_scale)
    cfg = cfg_of(hip, d)
    ref32 = oracle.std_gridding(G, d.oversampling, d.kernels, d.supports, d.uvw, d.vis, cfg)
    peak = np.abs(ref64).max()
    err = np.abs(got - ref64).max()
    assert err <= 1e-6 * peak
    assert err <= 0.5 * np.abs(ref32 - ref64).max()


def test_grid_keys_of_many_pieces(hip, oracle):
    """Work items of 512 taps on a dense grid: every (class, tile) key is cut
    into about 150 pieces, more than one reduce workgroup takes, so the merge
    sums them in groups and then sums the groups.  Gridding stays bit-exact
    with integer kernels and visibilities, and within 1e-6 of the peak of
    the FP64 grid with the real ones."""
    d = sdp_synth.small_case(n_vis=300_000, grid_size=128, n_planes=4, oversampling=16, seed=9)
    G = d.grid_size
    uvw = hip.DeviceBuffer.from_array(d.uvw)
    plan = hip.GridPlan(G, d.oversampling, d.supports, d.uv_scale, d.w_scale, uvw, d.n_vis, tile=8, chunk=512)
    info = plan.info()
    assert info["chunks"] > 64 * info["tiles"]   # so some key has more than 64 pieces
    g = hip.DeviceBuffer(8 * G * G)
    kern = integer_kernels(d.kernels.shape[0])
    vis = np.random.default_rng(9).integers(-3, 4, size=(d.n_vis, 2)).astype(np.float32)
    plan.grid(hip.DeviceBuffer.from_array(kern), hip.DeviceBuffer.from_array(vis), g)
    got = g.download((G * G, 2), np.float32)
    ref = oracle.std_gridding(G, d.oversampling, kern, d.supports, d.uvw, vis, cfg_of(hip, d))
    np.testing.assert_array_equal(got, ref)
    plan.grid(hip.DeviceBuffer.from_array(d.kernels), hip.DeviceBuffer.from_array(d.vis), g)
    got = g.download((G * G, 2), np.float32)
    ref64 = oracle.std_gridding_f64(G, d.oversampling, d.kernels, d.supports, d.uvw, d.vis, d.uv_scale, d.w_scale)
    assert np.abs(got - ref64).max() <= 1e-6 * np.abs(ref64).max()


def test_grid_device_api_plan_order_and_determinism(hip, case):
    d = case
    G = d.grid_size
    uvw = hip.DeviceBuffer.from_array(d.uvw)
    vis = hip.DeviceBuffer.from_array(d.vis)
    kern = hip.DeviceBuffer.from_array(d.kernels)
    plan = hip.GridPlan(G, d.oversampling, d.supports, d.uv_scale, d.w_scale, uvw, d.n_vis)
    info = plan.info()
    assert info["kept"] + info["dropped"] == d.n_vis
    g1 = hip.DeviceBuffer(8 * G * G)
    g2 = hip.DeviceBuffer(8 * G * G)
    plan.grid(kern, vis, g1)
    vis_p = hip.DeviceBuffer(8 * d.n_vis)
    plan.to_plan_order(vis, vis_p)
    plan.grid(kern, vis_p, g2, order=hip.ORDER_PLAN)
    a = g1.download((G * G, 2), np.float32)
    b = g2.download((G * G, 2), np.float32)
    np.testing.assert_array_equal(a, b)
    plan.grid(kern, vis, g2)
    np.testing.assert_array_equal(a, g2.download((G * G, 2), np.float32))
    back = hip.DeviceBuffer(8 * d.n_vis)
    plan.to_original_order(vis_p, back)
    np.testing.assert_array_equal(back.download((d.n_vis, 2), np.float32), d.vis)


@pytest.mark.parametrize("tile,chunk", [(8, 2000), (24, 1 << 16), (32, 5000)])
def test_grid_tilings_agree(hip, oracle, case, tile, chunk):
    d = case
    G = d.grid_size
    uvw = hip.DeviceBuffer.from_array(d.uvw)
    vis = hip.DeviceBuffer.from_array(d.vis)
    kern = hip.DeviceBuffer.from_array(d.kernels)
    plan = hip.GridPlan(G, d.oversampling, d.supports, d.uv_scale, d.w_scale, uvw, d.n_vis, tile=tile, chunk=chunk)
    g = hip.DeviceBuffer(8 * G * G)
    plan.grid(kern, vis, g)
    ref64 = oracle.std_gridding_f64(G, d.oversampling, d.kernels, d.supports, d.uvw, d.vis, d.uv_scale, d.w_scale)
    assert np.abs(g.download((G * G, 2), np.float32) - ref64).max() <= 1e-6 * np.abs(ref64).max()


# ---------------------------------------------------------------- degridding

def test_degrid_nearest_bit_exact(hip, actors, oracle, case, rng):
    d = case
    cfg = cfg_of(hip, d)
    grid = rng.normal(size=(d.grid_size ** 2, 2)).astype(np.float32)
    n = d.n_vis - 100
    actors.set_degrid_mode(hip.DEGRID_NEAREST)
    got = actors.std_degridding(d.grid_size, d.oversampling, d.kernels, d.supports, grid, d.uvw, cfg, n=n)
    ref = oracle.std_degridding(d.grid_size, d.oversampling, d.kernels, d.supports, grid, d.uvw, cfg, n=n)
    np.testing.assert_array_equal(got, ref)


def test_degrid_weighted_accuracy_and_adjointness(hip, actors, oracle, case, rng):
    d = case
    G = d.grid_size
    cfg = cfg_of(hip, d)
    grid = rng.normal(size=(G * G, 2)).astype(np.float32)
    actors.set_degrid_mode(hip.DEGRID_WEIGHTED)
    try:
        got = actors.std_degridding(G, d.oversampling, d.degrid_kernels, d.supports, grid, d.uvw, cfg)
    finally:
        actors.set_degrid_mode(hip.DEGRID_NEAREST)
    ref = oracle.degrid_weighted_f64(G, d.oversampling, d.degrid_kernels, d.supports, grid, d.uvw, d.uv_scale,
                                     d.w_scale)
    scale = np.abs(ref).max()
    assert np.abs(got - ref).max() <= 2e-6 * scale
    # <G x, y> == <x, D y> with D using the conjugate table
    x = d.vis
    gx = actors.std_gridding(G, d.oversampling, d.kernels, d.supports, d.uvw, x, cfg).astype(np.float64)
    c = lambda a: a[..., 0] + 1j * a[..., 1]
    lhs = np.vdot(c(gx), c(grid.astype(np.float64)))
    rhs = np.vdot(c(x.astype(np.float64)), c(got.astype(np.float64)))
    assert abs(lhs - rhs) <= 1e-5 * abs(lhs)


def wide_footprint_case(G=1024, seed=3):
    """Random visibilities plus ones whose float footprint spans 2h + 1 cells:
    gx - h, rounded to float just past a power of two, falls on the integer
    below, so std_gridding's [ceilf(gx - h), ceilf(gx + h)) (degridgrid.c:93,103)
    holds one cell more than 2h.  One visibility per power-of-two edge and
    per plane, at half supports 8, 12 and 16 (the largest of each support
    class); half of them are wide in both axes."""
    f32 = np.float32
    d = sdp_synth.small_case(n_vis=600, grid_size=G, h_list=[8, 12, 16], seed=seed)
    # uv_scale 1: a visibility's u is its float grid position
    rand = d.uvw.copy()
    rand[:, :2] *= np.float32(d.uv_scale)
    d = dataclasses.replace(d, cell_size=1.0 / G, uvw=rand)
    s = d.uv_scale

    def extent(u, h):
        g = f32(np.float64(f32(u)) * s)
        lo, hi = np.ceil(f32(g - f32(h))), np.ceil(f32(g + f32(h)))
        return int(hi - lo), int(lo) + G // 2, int(hi) + G // 2

    def find(m, h):
        # gx = -(m - h) + one ulp: gx - h = -(m - ulp) rounds (to even) onto -m
        u = f32(-(m - h)) + np.spacing(f32(m - h))
        n, lo, hi = extent(u, h)
        return u if n == 2 * h + 1 and lo >= 0 and hi <= G else None

    rows = []
    for p, h in enumerate([8, 12, 16]):
        w = f32(p * p / d.w_scale)
        e = 32
        while e + h < G // 2:
            for m in range(e + 1, e + h):
                u = find(m, h)
                if u is not None:
                    rows.append((u, u if len(rows) % 2 else f32(0.1 * e), w, h))
                    break
            e *= 2
    wide = np.array([r[:3] for r in rows], np.float32)
    for (u, v, _, h) in rows:
        assert extent(u, h)[0] == 2 * h + 1
    uvw = np.concatenate([wide, d.uvw]).astype(np.float32)
    rng = np.random.default_rng(seed)
    vis = rng.integers(-3, 4, size=(uvw.shape[0], 2)).astype(np.float32)
    return dataclasses.replace(d, uvw=uvw, vis=vis), len(rows)


def test_footprints_one_cell_wider_than_2h(hip, actors, oracle):
    """Footprints of 2h + 1 cells at each support class's largest h: gridding
    is bit-exact (integer kernels and visibilities) and kernel-weighted
    degridding within 2e-6 of the FP64 sum over the same cells."""
    d, n_wide = wide_footprint_case()
    assert n_wide >= 12
    G, cfg = d.grid_size, cfg_of(hip, d)
    kern = integer_kernels(d.kernels.shape[0])
    got = actors.std_gridding(G, d.oversampling, kern, d.supports, d.uvw, d.vis, cfg)
    ref = oracle.std_gridding(G, d.oversampling, kern, d.supports, d.uvw, d.vis, cfg)
    np.testing.assert_array_equal(got, ref)
    grid = np.random.default_rng(4).normal(size=(G * G, 2)).astype(np.float32)
    actors.set_degrid_mode(hip.DEGRID_WEIGHTED)
    try:
        deg = actors.std_degridding(G, d.oversampling, d.degrid_kernels, d.supports, grid, d.uvw, cfg)
    finally:
        actors.set_degrid_mode(hip.DEGRID_NEAREST)
    ref = oracle.degrid_weighted_f64(G, d.oversampling, d.degrid_kernels, d.supports, grid, d.uvw, d.uv_scale,
                                     d.w_scale)
    assert np.abs(deg - ref).max() <= 2e-6 * np.abs(ref).max()


def test_many_planes(hip, actors, oracle, rng):
    """300 w-planes: more than the 256 whose table entries the chunk kernel
    keeps in LDS, so it reads them from memory.  Gridding is bit-exact
    (integer kernels and visibilities), degridding within 2e-6 of the FP64 sum."""
    d = sdp_synth.small_case(n_vis=3000, grid_size=128, h_list=[3 + p % 2 for p in range(300)], oversampling=4,
                             seed=12)
    G, cfg = d.grid_size, cfg_of(hip, d)
    planes = np.floor(np.sqrt(np.abs(d.uvw[:, 2].astype(np.float64)) * d.w_scale) + 0.5)
    assert len(np.unique(planes)) > 256
    kern = integer_kernels(d.kernels.shape[0])
    vis = rng.integers(-3, 4, size=(d.n_vis, 2)).astype(np.float32)
    got = actors.std_gridding(G, d.oversampling, kern, d.supports, d.uvw, vis, cfg)
    ref = oracle.std_gridding(G, d.oversampling, kern, d.supports, d.uvw, vis, cfg)
    np.testing.assert_array_equal(got, ref)
    grid = rng.normal(size=(G * G, 2)).astype(np.float32)
    actors.set_degrid_mode(hip.DEGRID_WEIGHTED)
    try:
        deg = actors.std_degridding(G, d.oversampling, d.degrid_kernels, d.supports, grid, d.uvw, cfg)
    finally:
        actors.set_degrid_mode(hip.DEGRID_NEAREST)
    ref = oracle.degrid_weighted_f64(G, d.oversampling, d.degrid_kernels, d.supports, grid, d.uvw, d.uv_scale,
                                     d.w_scale)
    assert np.abs(deg - ref).max() <= 2e-6 * np.abs(ref).max()


@pytest.mark.parametrize("OS", [6, 10])
def test_oversampling_not_a_power_of_two(hip, actors, oracle, rng, OS):
    """Oversampling factors that are not powers of two take the kernel's
    division path for the phase-major table index: gridding bit-exact
    (integer kernels and visibilities), degridding within 2e-6 of the FP64 sum."""
    d = sdp_synth.small_case(n_vis=3000, grid_size=128, n_planes=4, oversampling=OS, seed=20 + OS)
    G, cfg = d.grid_size, cfg_of(hip, d)
    kern = integer_kernels(d.kernels.shape[0])
    vis = rng.integers(-3, 4, size=(d.n_vis, 2)).astype(np.float32)
    got = actors.std_gridding(G, OS, kern, d.supports, d.uvw, vis, cfg)
    ref = oracle.std_gridding(G, OS, kern, d.supports, d.uvw, vis, cfg)
    assert np.count_nonzero(ref) > 1000
    np.testing.assert_array_equal(got, ref)
    grid = rng.normal(size=(G * G, 2)).astype(np.float32)
    actors.set_degrid_mode(hip.DEGRID_WEIGHTED)
    try:
        deg = actors.std_degridding(G, OS, d.degrid_kernels, d.supports, grid, d.uvw, cfg)
    finally:
        actors.set_degrid_mode(hip.DEGRID_NEAREST)
    ref = oracle.degrid_weighted_f64(G, OS, d.degrid_kernels, d.supports, grid, d.uvw, d.uv_scale, d.w_scale)
    assert np.abs(deg - ref).max() <= 2e-6 * np.abs(ref).max()


def test_no_visibilities(hip, actors, oracle):
    """Zero visibilities: the grid is all zeros (degridgrid.c:73 clears it
    before the loop), degridding returns an empty set in both modes."""
    d = sdp_synth.small_case(n_vis=1, grid_size=64, n_planes=4, oversampling=16, seed=30)
    G, cfg = d.grid_size, cfg_of(hip, d)
    uvw, vis = np.ascontiguousarray(d.uvw[:0]), np.ascontiguousarray(d.vis[:0])
    got = actors.std_gridding(G, d.oversampling, d.kernels, d.supports, uvw, vis, cfg)
    ref = oracle.std_gridding(G, d.oversampling, d.kernels, d.supports, uvw, vis, cfg)
    assert not np.any(ref)
    np.testing.assert_array_equal(got, ref)
    grid = np.ones((G * G, 2), np.float32)
    for mode in (hip.DEGRID_NEAREST, hip.DEGRID_WEIGHTED):
        actors.set_degrid_mode(mode)
        try:
            deg = actors.std_degridding(G, d.oversampling, d.degrid_kernels, d.supports, grid, uvw, cfg)
        finally:
            actors.set_degrid_mode(hip.DEGRID_NEAREST)
        assert deg.shape == (0, 2)


def test_visibilities_off_the_grid(hip, actors, oracle, rng):
    """uv spread over 1.5x the grid: about half the footprints lie partly or
    wholly outside [-G/2, G/2), whose cells the reference skips
    (degridgrid.c:95-109).  Gridding bit-exact (integer kernels and
    visibilities); weighted degridding within 2e-6 of the FP64 sum, whose
    zero padding gives exactly 0 for a visibility that reaches no cell."""
    d = sdp_synth.small_case(n_vis=4000, grid_size=128, n_planes=4, oversampling=16, seed=31, fill=1.5)
    G, cfg = d.grid_size, cfg_of(hip, d)
    gx = np.abs(d.uvw[:, :2].astype(np.float64) * d.uv_scale).max(axis=1)
    outside = gx > G / 2 + 8
    assert outside.sum() > 500 and (~outside).sum() > 1500
    kern = integer_kernels(d.kernels.shape[0])
    vis = rng.integers(-3, 4, size=(d.n_vis, 2)).astype(np.float32)
    got = actors.std_gridding(G, d.oversampling, kern, d.supports, d.uvw, vis, cfg)
    ref = oracle.std_gridding(G, d.oversampling, kern, d.supports, d.uvw, vis, cfg)
    assert np.count_nonzero(ref) > 1000
    np.testing.assert_array_equal(got, ref)
    grid = rng.normal(size=(G * G, 2)).astype(np.float32)
    actors.set_degrid_mode(hip.DEGRID_WEIGHTED)
    try:
        deg = actors.std_degridding(G, d.oversampling, d.degrid_kernels, d.supports, grid, d.uvw, cfg)
    finally:
        actors.set_degrid_mode(hip.DEGRID_NEAREST)
    ref = oracle.degrid_weighted_f64(G, d.oversampling, d.degrid_kernels, d.supports, grid, d.uvw, d.uv_scale,
                                     d.w_scale)
    assert np.all(deg[outside] == 0.0) and np.all(ref[outside] == 0.0)
    assert np.abs(deg - ref).max() <= 2e-6 * np.abs(ref).max()


def test_reciprocal_transform_bit_exact(hip, actors, oracle, rng):
    """The correct_to_finegrid subgraph's first actor: each receiver gain's
    reciprocal, bit for bit (IEEE float division on both sides)."""
    g = rng.normal(size=(512, 2)).astype(np.float32)
    g[0] = (0.0, 0.0)                                     # 0 / 0: NaN on both sides
    cfg = hip.make_config(64, 1e-4)
    got = actors.reciprocal_transform_actor(g, cfg)
    ref = oracle.reciprocal_transform(g)
    np.testing.assert_array_equal(got, ref)


@pytest.mark.parametrize("D,h", [(64, 3), (128, 9)])
def test_restore_conv2d_bit_exact(hip, oracle, rng, D, h):
    """The restored image (conv2D, major_loop_iter.c:8-46) on the GPU against
    the reference's scatter loop: a sparse model with pixels on both sides of
    the 1e-5 threshold, every output pixel bit for bit."""
    img = np.zeros((D, D), np.float32)
    idx = rng.choice(D * D, size=D * D // 20, replace=False)
    img.flat[idx] = rng.uniform(-0.5, 1.0, idx.size).astype(np.float32)
    img.flat[idx[:20]] = np.float32(1e-5)                         # skipped (below 1e-5 as a double)
    img.flat[idx[20:40]] = np.nextafter(np.float32(1e-5), np.float32(1))
    beam = rng.uniform(0.0, 1.0, (D, D)).astype(np.float32)
    d_img, d_beam = hip.DeviceBuffer.from_array(img), hip.DeviceBuffer.from_array(beam)
    d_out = hip.DeviceBuffer(4 * D * D)
    hip.check(hip.load().sdp_restore_conv2d(D, d_img.ptr, d_beam.ptr, h, h, d_out.ptr, None))
    np.testing.assert_array_equal(d_out.download((D, D), np.float32), oracle.conv2d(img, beam, h, h))
    assert hip.load().sdp_restore_conv2d(D, d_img.ptr, d_beam.ptr, D // 2 - 1, h, d_out.ptr, None) == -2


def test_save_output_writes_restored_image(hip, actors, oracle, rng, tmp_path):
    """save_output (major_loop_iter.c:253-278): the five images as
    "%f, " text files named cycle_<n>_<name> under output_path; the restored
    image is the model convolved with the clean PSF (the GPU conv2D)."""
    import ctypes
    G = 64
    c = G // 2 + 1
    yy, xx = np.mgrid[0:G, 0:G]
    rr = np.hypot(yy - c, xx - c)
    dirty = (np.exp(-rr ** 2 / 40.0) * np.cos(rr / 2.0)).astype(np.float32)
    clean, h = oracle.clean_psf(dirty, 1.0)
    model = np.zeros((G, G), np.float32)
    model[20, 30], model[40, 12] = 0.8, 0.3
    resid = rng.normal(size=(G, G)).astype(np.float32)
    cfg = hip.make_config(G, 0.01)
    cfg.output_path = (str(tmp_path) + "/").encode()
    for f, name in (("residual_image_output", "res.csv"), ("psf_image_output", "psf.csv"),
                    ("clean_psf_image_output", "cpsf.csv"), ("model_image_output", "model.csv"),
                    ("final_image_output", "final.csv")):
        setattr(cfg, f, name.encode())
    hd = np.array([h, h], np.int32)
    cycle = ctypes.c_int(3)
    actors.lib.save_output(G, resid.ctypes.data, model.ctypes.data, clean.ctypes.data, hd.ctypes.data,
                           dirty.ctypes.data, ctypes.byref(cfg), ctypes.byref(cycle))

    def read(name):
        rows = (tmp_path / f"cycle_3_{name}").read_text().splitlines()
        return np.array([[float(v) for v in r.split(",") if v.strip()] for r in rows])

    np.testing.assert_allclose(read("model.csv"), model, atol=1e-6)
    np.testing.assert_allclose(read("res.csv"), resid, atol=1e-6)
    np.testing.assert_allclose(read("final.csv"), oracle.conv2d(model, clean, h, h), atol=1e-6)


def test_degrid_fused_subtract(hip, case, rng):
    d = case
    G = d.grid_size
    uvw = hip.DeviceBuffer.from_array(d.uvw)
    kern = hip.DeviceBuffer.from_array(d.degrid_kernels)
    grid = hip.DeviceBuffer.from_array(rng.normal(size=(G * G, 2)).astype(np.float32))
    meas = hip.DeviceBuffer.from_array(d.vis)
    plan = hip.GridPlan(G, d.oversampling, d.supports, d.uv_scale, d.w_scale, uvw, d.n_vis)
    pred = hip.DeviceBuffer(8 * d.n_vis)
    delta = hip.DeviceBuffer(8 * d.n_vis)
    plan.degrid(kern, grid, pred)
    plan.degrid(kern, grid, delta, d_measured=meas)
    p = pred.download((d.n_vis, 2), np.float32)
    np.testing.assert_array_equal(delta.download((d.n_vis, 2), np.float32), d.vis - p)


@pytest.mark.parametrize("tables", ["conjugate", "own"])
@pytest.mark.parametrize("order", ["original", "plan"])
def test_degrid_grid_fused_equals_separate(hip, case, rng, order, tables):
    """The fused major-cycle pass gives the residual visibilities and residual
    grid of degrid(+subtract) followed by grid.  The fused kernel may split a
    footprint's rows over more lanes than the separate kernels do, so sums
    are grouped differently: equal to float rounding (1e-6 of the peak).
    With the degridding table the conjugate of the gridding table each tap
    reads one sample for both; with a table of its own ("own": the
    conjugate table perturbed in one sample of every plane) it reads both."""
    d = case
    G = d.grid_size
    o = hip.ORDER_ORIGINAL if order == "original" else hip.ORDER_PLAN
    uvw = hip.DeviceBuffer.from_array(d.uvw)
    gk = hip.DeviceBuffer.from_array(d.kernels)
    dtab = d.degrid_kernels.copy()
    if tables == "own":
        dtab[d.supports[:, 1] + 5] *= np.float32(0.75)
        dtab *= np.float32(1.25)
    dk = hip.DeviceBuffer.from_array(dtab)
    model = hip.DeviceBuffer.from_array(rng.normal(size=(G * G, 2)).astype(np.float32))
    meas = hip.DeviceBuffer.from_array(d.vis)
    plan = hip.GridPlan(G, d.oversampling, d.supports, d.uv_scale, d.w_scale, uvw, d.n_vis)
    if o == hip.ORDER_PLAN:
        mp = hip.DeviceBuffer(8 * d.n_vis)
        plan.to_plan_order(meas, mp)
        meas = mp
    res_sep = hip.DeviceBuffer(8 * d.n_vis)
    grid_sep = hip.DeviceBuffer(8 * G * G)
    plan.degrid(dk, model, res_sep, d_measured=meas, order=o)
    plan.grid(gk, res_sep, grid_sep, order=o)
    res_f = hip.DeviceBuffer(8 * d.n_vis)
    grid_f = hip.DeviceBuffer(8 * G * G)
    plan.degrid_grid(dk, gk, model, meas, grid_f, d_residual_vis=res_f, order=o)
    rs = res_sep.download((d.n_vis, 2), np.float32)
    np.testing.assert_allclose(res_f.download((d.n_vis, 2), np.float32), rs, rtol=0, atol=1e-6 * np.abs(rs).max())
    gs = grid_sep.download((G * G, 2), np.float32)
    np.testing.assert_allclose(grid_f.download((G * G, 2), np.float32), gs, rtol=0, atol=1e-6 * np.abs(gs).max())


# ---------------------------------------------------------------- fine grid

def test_finegrid_bit_exact(hip, actors, oracle, case):
    d = case
    cfg = cfg_of(hip, d)
    uvw = np.concatenate([d.uvw, d.uvw[:500] + np.float32(1e-4), d.uvw[:50]], axis=0)   # near and exact repeats
    vis = np.concatenate([d.vis, d.vis[:500], d.vis[:50] * 2], axis=0)
    for simplify in (0, 1):
        gv, gu, gn = actors.correct_to_finegrid(d.grid_size, d.oversampling, simplify, uvw, vis, cfg)
        rv, ru, rn = oracle.correct_to_finegrid(d.grid_size, d.oversampling, simplify, uvw, vis, cfg)
        assert gn == rn
        if simplify:
            assert rn < uvw.shape[0]
        np.testing.assert_array_equal(gv, rv)
        np.testing.assert_array_equal(gu[:rn], ru[:rn])


# ---------------------------------------------------------------- FFT

@pytest.mark.parametrize("G", [16, 32, 64, 128, 256, 512, 1024, 2048, 4096, 8192])
def test_fft_against_fp64(actors, rng, G):
    x = rng.normal(size=(G * G, 2)).astype(np.float32)
    z = (x[:, 0] + 1j * x[:, 1]).astype(np.complex128).reshape(G, G)
    for got, ref in ((actors.fft_forward(G, x), np.fft.fft2(z)), (actors.fft_inverse(G, x), np.fft.ifft2(z) * G * G)):
        ref = ref.ravel()
        err = np.abs((got[:, 0] + 1j * got[:, 1]) - ref).max()
        assert err <= 3e-7 * np.log2(G) * np.abs(ref).max(), (G, err)


def test_fft_fused_equals_actor_chain(hip, actors, oracle, rng):
    G = 256
    cfg = hip.make_config(G, 1e-3)
    grid = rng.normal(size=(G * G, 2)).astype(np.float32)
    taper = sdp_synth.taper_padded(G)
    chain = actors.convolution_correction(
        G, actors.shift_c2r(G, actors.fft_inverse(G, actors.shift_c2c(G, grid, cfg)), cfg), taper[:-1], cfg)
    d_in = hip.DeviceBuffer.from_array(grid)
    d_out = hip.DeviceBuffer(4 * G * G)
    d_tap = hip.DeviceBuffer.from_array(taper)
    d_w = hip.DeviceBuffer(hip.fft_work_bytes(G))
    hip.fft2d(d_in, d_out, G, hip.FFT_INVERSE,
              hip.FFT_SHIFT_IN | hip.FFT_SHIFT_OUT | hip.FFT_REAL_OUT | hip.FFT_CORRECT_OUT, d_tap, d_w)
    np.testing.assert_array_equal(d_out.download(G * G, np.float32), chain)
    img = rng.normal(size=G * G).astype(np.float32)
    chain2 = actors.shift_c2c(G, actors.fft_forward(G, actors.shift_r2c(G, img, cfg)), cfg)
    d_img = hip.DeviceBuffer.from_array(img)
    d_g = hip.DeviceBuffer(8 * G * G)
    hip.fft2d(d_img, d_g, G, hip.FFT_FORWARD, hip.FFT_REAL_IN | hip.FFT_SHIFT_IN | hip.FFT_SHIFT_OUT, None, d_w)
    np.testing.assert_array_equal(d_g.download((G * G, 2), np.float32), chain2)


def test_fft_in_place_and_errors(hip, rng):
    G = 64
    x = rng.normal(size=(G * G, 2)).astype(np.float32)
    a = hip.DeviceBuffer.from_array(x)
    w = hip.DeviceBuffer(hip.fft_work_bytes(G))
    hip.fft2d(a, a, G, hip.FFT_FORWARD, 0, None, w)
    hip.fft2d(a, a, G, hip.FFT_INVERSE, 0, None, w)
    np.testing.assert_allclose(a.download((G * G, 2), np.float32) / (G * G), x, atol=1e-5)
    with pytest.raises(hip.SdpError, match="unsupported"):
        hip.fft2d(a, a, 96, hip.FFT_FORWARD, 0, None, w)


# ---------------------------------------------------------------- element-wise

def test_elementwise_bit_exact(hip, actors, oracle, rng):
    G = 128
    cfg = hip.make_config(G, 1e-3)
    grid = rng.normal(size=(G * G, 2)).astype(np.float32)
    img = rng.normal(size=G * G).astype(np.float32)
    np.testing.assert_array_equal(actors.shift_c2c(G, grid, cfg), oracle.shift_c2c(G, grid))
    np.testing.assert_array_equal(actors.shift_c2r(G, grid, cfg), oracle.shift_c2r(G, grid))
    np.testing.assert_array_equal(actors.shift_r2c(G, img, cfg), oracle.shift_r2c(G, img))
    taper = oracle.half_prolate(G)
    np.testing.assert_array_equal(actors.convolution_correction(G, img, taper, cfg),
                                  oracle.convolution_correction(G, img, np.concatenate([taper, [0]])))
    grids = rng.normal(size=(4 * G * G, 2)).astype(np.float32)
    np.testing.assert_array_equal(actors.merge(G, 4, grids), oracle.merge(G, 4, grids))
    m = rng.normal(size=(5000, 2)).astype(np.float32)
    p = rng.normal(size=(5000, 2)).astype(np.float32)
    gains = rng.normal(size=(10, 2)).astype(np.float32)
    pairs = sdp_synth.receiver_pairs(10)
    np.testing.assert_array_equal(actors.subtract_from_measurements(m, p, gains, pairs, cfg),
                                  oracle.subtract_from_measurements(m, p, gains, pairs))
    np.testing.assert_array_equal(actors.apply_gains_only(m, gains, pairs, cfg),
                                  oracle.apply_gains_only(m, gains, pairs))
    np.testing.assert_array_equal(actors.add_visibilities(m, p), oracle.add_visibilities(m, p))
    np.testing.assert_array_equal(actors.subtract_finegrid(m, p, n=4000), oracle.subtract_finegrid(m, p, n=4000))


# ---------------------------------------------------------------- Hogbom

def dirty_case(G, rng, n_src=6):
    yy, xx = np.mgrid[0:G, 0:G]
    c = G // 2
    r2 = (yy - c) ** 2 + (xx - c) ** 2
    psf = (np.exp(-r2 / 8.0) * (1 + 0.3 * np.cos(0.7 * (xx - c)))).astype(np.float32)
    psf /= psf.max()
    psf_padded = np.concatenate([psf.ravel(), [0.0]]).astype(np.float32)
    res = np.zeros((G, G), np.float64)
    for _ in range(n_src):
        y, x = rng.integers(G // 4, 3 * G // 4, 2)
        f = rng.uniform(0.2, 2.0)
        res += f * np.roll(np.roll(psf, y - c, 0), x - c, 1)
    res += rng.normal(0, 0.003, res.shape)
    return res.astype(np.float32).ravel(), psf_padded


@pytest.mark.parametrize("G,n_minor", [(64, 60), (128, 200)])
def test_hogbom_bit_exact(hip, actors, oracle, rng, G, n_minor):
    res, psf = dirty_case(G, rng)
    cfg = hip.make_config(G, 1e-4, psf_max_value=1.0)
    model = np.zeros(G * G, np.float32)
    ns, src, out_model = actors.hogbom_clean(G, n_minor, 400, res, psf[:-1].copy(), cfg, model)
    rns, rsrc, rmodel, _ = oracle.hogbom_clean(G, n_minor, 400, res, psf, cfg, model)
    assert ns == rns and ns > 10
    np.testing.assert_array_equal(src, rsrc)
    np.testing.assert_array_equal(out_model, rmodel)


def test_hogbom_quirks(hip, oracle, rng):
    """Mirrored PSF read, the padded PSF element, ties, both stop tests: the
    device path against the oracle, residual included."""
    G = 16
    cases = []
    psf = np.zeros(G * G + 1, np.float32)
    psf[(G // 2) * G + G // 2] = 1.0
    psf[(G // 2) * G + G // 2 + 1] = 0.5
    psf[(G - 1) * G + 3] = 0.25
    r = np.zeros(G * G, np.float32)
    r[7 * G + 12] = 2.0          # mx >= G/2 with my <= G/2 reads the padded element
    r[9 * G + 2] = -1.5
    cases.append((r, psf, dict(psf_max_value=1.0), 5))
    r = np.zeros(G * G, np.float32)
    r[20] = -1.0
    r[40] = 1.0
    cases.append((r, psf, dict(psf_max_value=2.0), 3))
    r = np.ones(G * G, np.float32)
    r[0] = 0.0
    r[50] = 1.5
    cases.append((r, psf, dict(psf_max_value=1.0, noise_factor=20.0), 5))
    r = np.zeros(G * G, np.float32)
    r[30] = 4.0
    r[90] = 0.9
    cases.append((r, psf, dict(psf_max_value=1.0, loop_gain=0.5, weak_source_percent_img=0.5), 10))
    cases.append((np.zeros(G * G, np.float32), psf, dict(psf_max_value=1.0), 3))
    for r, p, kw, n_minor in cases:
        cfg = hip.make_config(G, 1e-3, **kw)
        rns, rsrc, rmodel, rres = oracle.hogbom_clean(G, n_minor, 8, r, p, cfg, np.zeros(G * G, np.float32))
        d_res = hip.DeviceBuffer.from_array(r)
        d_psf = hip.DeviceBuffer.from_array(p)
        d_m = hip.DeviceBuffer.from_array(np.zeros(G * G, np.float32))
        d_mo = hip.DeviceBuffer(4 * G * G)
        d_src = hip.DeviceBuffer(12 * 8)
        d_n = hip.DeviceBuffer(16)
        d_w = hip.DeviceBuffer(hip.hogbom_work_bytes(G))
        hip.hogbom(G, n_minor, 8, d_res, d_psf, cfg.psf_max_value, cfg.loop_gain, cfg.noise_factor,
                   cfg.weak_source_percent_img, cfg.cell_size, d_m, d_mo, d_src, d_n, d_w)
        ns = int(d_n.download(1, np.int32)[0])
        assert ns == rns, kw
        np.testing.assert_array_equal(d_src.download((8, 3), np.float32)[:ns], rsrc)
        np.testing.assert_array_equal(d_mo.download(G * G, np.float32), rmodel)
        np.testing.assert_array_equal(d_res.download(G * G, np.float32), rres)


def run_device_hogbom(hip, G, n_minor, max_sources, r, psf_padded, cfg, flags=0):
    d_res = hip.DeviceBuffer.from_array(r)
    d_psf = hip.DeviceBuffer.from_array(psf_padded)
    d_m = hip.DeviceBuffer.from_array(np.zeros(G * G, np.float32))
    d_mo = hip.DeviceBuffer(4 * G * G)
    d_src = hip.DeviceBuffer(12 * max(max_sources, 1))
    d_n = hip.DeviceBuffer(16)
    d_w = hip.DeviceBuffer(hip.hogbom_work_bytes(G))
    hip.hogbom_ex(G, n_minor, max_sources, d_res, d_psf, cfg.psf_max_value, cfg.loop_gain, cfg.noise_factor,
                  cfg.weak_source_percent_img, cfg.cell_size, d_m, d_mo, d_src, d_n, d_w, flags)
    ns = int(d_n.download(1, np.int32)[0])
    return (ns, d_src.download((max(max_sources, 1), 3), np.float32)[:ns], d_mo.download(G * G, np.float32),
            d_res.download(G * G, np.float32))


@pytest.mark.parametrize("G", [256, 2048])
def test_hogbom_noise_stop_uses_the_reference_running_mean(hip, actors, oracle, G):
    """A noise image whose threshold noise_factor * gain * mean is set
    between the reference's float running average and the exact mean of
    |r_1..r_{N-1}|, so the two means disagree on whether to stop at once.
    The device path (its sequential recurrence) and the actor agree with
    the oracle bit for bit; the parallel double-sum option decides the
    other way, which shows that the case sits inside the drift band."""
    rng = np.random.default_rng(G)
    r = rng.standard_normal(G * G).astype(np.float32)
    r[(G // 3) * G + G // 5] = 6.0
    m_ref = oracle.running_mean(r)
    m_exact = np.float32(np.abs(r[1:].astype(np.float64)).sum() / (G * G - 1))
    assert m_ref != m_exact
    peak = float(np.abs(r).max())
    gain = 0.1
    nf = peak / (gain * 0.5 * (float(m_ref) + float(m_exact)))
    psf = np.zeros(G * G + 1, np.float32)
    psf[(G // 2) * G + G // 2] = 1.0
    cfg = hip.make_config(G, 1e-4, psf_max_value=1.0, noise_factor=nf, loop_gain=gain)
    n_minor = 3
    rns, rsrc, rmodel, rres = oracle.hogbom_clean(G, n_minor, 8, r, psf, cfg, np.zeros(G * G, np.float32))
    ns, src, model, res = run_device_hogbom(hip, G, n_minor, 8, r, psf, cfg)
    assert ns == rns
    np.testing.assert_array_equal(src, rsrc)
    np.testing.assert_array_equal(model, rmodel)
    np.testing.assert_array_equal(res, rres)
    ans, asrc, amodel = actors.hogbom_clean(G, n_minor, 8, r, psf[:-1].copy(), cfg, np.zeros(G * G, np.float32))
    assert ans == rns
    np.testing.assert_array_equal(asrc, rsrc)
    np.testing.assert_array_equal(amodel, rmodel)
    dns, _, _, _ = run_device_hogbom(hip, G, n_minor, 8, r, psf, cfg, flags=hip.HOGBOM_DOUBLE_MEAN)
    assert (dns == 0) != (rns == 0)


def test_hogbom_stops_after_max_sources(hip, oracle, rng):
    """n_minor > max_sources: the reference would keep cleaning and write
    past its sources array; the library stops after max_sources components,
    which are the reference's first max_sources iterations bit for bit."""
    G = 64
    res, psf = dirty_case(G, rng)
    cfg = hip.make_config(G, 1e-4, psf_max_value=1.0)
    ns, src, model, r = run_device_hogbom(hip, G, 40, 12, res, psf, cfg)
    rns, rsrc, rmodel, rres = oracle.hogbom_clean(G, 12, 12, res, psf, cfg, np.zeros(G * G, np.float32))
    assert ns == rns == 12
    np.testing.assert_array_equal(src, rsrc)
    np.testing.assert_array_equal(model, rmodel)
    np.testing.assert_array_equal(r, rres)
    full = oracle.hogbom_clean(G, 40, 12, res, psf, cfg, np.zeros(G * G, np.float32))
    assert full[0] == 40 and np.array_equal(full[1][:12], src)


# ---------------------------------------------------------------- DFT

def test_dft_modes(hip, actors, oracle, rng):
    d = sdp_synth.config1(n_time=8, with_vis=False)
    cfg = cfg_of(hip, d)
    src = d.sources
    ref32 = oracle.dft_actor(src, d.uvw, cfg)
    ref64 = oracle.dft_f64(src, d.uvw)
    scale = np.abs(src[:, 2]).sum()
    actors.set_dft_mode(hip.DFT_REFERENCE_FP32)
    try:
        got_ref = actors.dft_actor(src, d.uvw, cfg)
    finally:
        actors.set_dft_mode(hip.DFT_FP64_PHASE)
    got64 = actors.dft_actor(src, d.uvw, cfg)
    # same float phase as the reference: only sinf/cosf ulps and summation differ
    assert np.abs(got_ref - ref32).max() <= 2e-6 * scale
    # FP64 phase: float rounding of the reduced phase and the float sum only
    assert np.abs(got64 - ref64).max() <= 2e-6 * scale
    # the reference's float phase is itself this far from FP64 (stated, not a bar)
    assert np.abs(ref32 - ref64).max() < 1e-2 * scale


def test_autotune_picks_a_working_candidate(hip, oracle, case):
    d = case
    G = d.grid_size
    uvw = hip.DeviceBuffer.from_array(d.uvw)
    gk = hip.DeviceBuffer.from_array(d.kernels)
    dk = hip.DeviceBuffer.from_array(d.degrid_kernels)
    tile, chunk, ms = hip.autotune(G, d.oversampling, d.supports, d.uv_scale, d.w_scale, uvw, d.n_vis, gk, dk,
                                   tiles=(4, 8, 16), chunks=(4096, 65536))
    assert tile in (4, 8, 16) and chunk in (4096, 65536) and ms > 0
    plan = hip.GridPlan(G, d.oversampling, d.supports, d.uv_scale, d.w_scale, uvw, d.n_vis, tile=tile, chunk=chunk)
    vis = hip.DeviceBuffer.from_array(d.vis)
    g = hip.DeviceBuffer(8 * G * G)
    plan.grid(gk, vis, g)
    ref64 = oracle.std_gridding_f64(G, d.oversampling, d.kernels, d.supports, d.uvw, d.vis, d.uv_scale, d.w_scale)
    assert np.abs(g.download((G * G, 2), np.float32) - ref64).max() <= 1e-6 * np.abs(ref64).max()


@pytest.mark.parametrize("G,n_minor", [(16, 30), (256, 300), (1024, 200), (2048, 100), (4096, 40)])
def test_hogbom_one_launch_equals_per_iteration_launches(hip, rng, G, n_minor):
    """All iterations in one cooperative launch (residual held in registers,
    one word per workgroup and iteration as the exchange) against one launch
    per iteration: residual, model and components bit for bit, at sizes that
    take one to sixteen 16-byte groups per thread."""
    res, psf = dirty_case(G, rng, n_src=12)
    cfg = hip.make_config(G, 1e-4, psf_max_value=1.0)
    a = run_device_hogbom(hip, G, n_minor, n_minor, res, psf, cfg)
    b = run_device_hogbom(hip, G, n_minor, n_minor, res, psf, cfg, flags=hip.HOGBOM_PER_ITERATION)
    assert a[0] == b[0] == n_minor
    for x, y in zip(a[1:], b[1:]):
        np.testing.assert_array_equal(x, y)


def test_hogbom_one_launch_exit_path_ends_every_workgroup(hip, rng):
    """The one-launch form's bounded-wait exit, raised from the host before the
    launch: every workgroup leaves at its first exchange, the run ends (no
    hang) and reports -1 sources; a normal run afterwards on the same work
    buffer is unaffected."""
    G = 1024
    res, psf = dirty_case(G, rng)
    cfg = hip.make_config(G, 1e-4, psf_max_value=1.0)
    ns, _, _, _ = run_device_hogbom(hip, G, 50, 50, res, psf, cfg, flags=hip.HOGBOM_TEST_ABORT)
    assert ns == -1
    a = run_device_hogbom(hip, G, 50, 50, res, psf, cfg)
    b = run_device_hogbom(hip, G, 50, 50, res, psf, cfg, flags=hip.HOGBOM_PER_ITERATION)
    assert a[0] == b[0] == 50
    for x, y in zip(a[1:], b[1:]):
        np.testing.assert_array_equal(x, y)
