"""Training of the learned minor-cycle proposal (sdp_cnn_* kernels).

A training pair is made from a seeded random sky patch S (Jy per pixel):
power-law point sources plus elliptical Gaussians.  Its dirty patch is
R = P (*) S, the circular convolution with the T x T centre of the imaging
pipeline's own PSF P (peak 1), computed with the library's FFT on the GPU.
The network input is (R / s, P) with s = max |R|, the target S / s: the same
normalisation the pipeline applies to residual tiles (sdp_pipeline.h), so
the network returns the sky in the units of its input.

The network, its gradients and the optimiser all run in the HIP kernels;
this module only draws the random skies and drives the training loop.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

import sdp_hip as hip


def psf_patch(psf_image: np.ndarray, T: int) -> np.ndarray:
    """Centre T x T of a G x G PSF image whose peak is at (G/2, G/2)."""
    G = psf_image.shape[0]
    c = G // 2 - T // 2
    return np.ascontiguousarray(psf_image[c:c + T, c:c + T], dtype=np.float32)


def sky_components(T: int, rng: np.random.Generator, n_points=(1, 30), n_gauss=(0, 20), fwhm=(2.0, 30.0),
                   flux=(1e-3, 1.0), index=-1.5, margin=8, box: int | None = None) -> tuple[np.ndarray, np.ndarray]:
    """The random components of a T x T sky: point sources with
    dN/dS ~ S^index on [flux] and elliptical Gaussians (FWHM per axis
    uniform in [fwhm], random orientation, peak 0.2 times a flux drawn the
    same way), centred `margin` pixels away from the edges, or, with `box`,
    within `box` pixels of the centre pixel (T/2, T/2) on either axis.
    Returns (points: n x (y, x, flux), gaussians: m x (cy, cx, fwhm_a,
    fwhm_b, theta, peak)), float64."""
    a = index + 1.0
    lo, hi = (margin, T - margin) if box is None else (T // 2 - box, T // 2 + box)

    def draw(n):
        u = rng.random(n)
        return (flux[0] ** a + u * (flux[1] ** a - flux[0] ** a)) ** (1.0 / a)

    n = int(rng.integers(n_points[0], n_points[1] + 1))
    ys = rng.integers(lo, hi, n)
    xs = rng.integers(lo, hi, n)
    points = np.stack([ys, xs, draw(n)], axis=1).astype(np.float64)
    ng = int(rng.integers(n_gauss[0], n_gauss[1] + 1))
    f = draw(ng)
    cy, cx = rng.uniform(lo, hi, ng), rng.uniform(lo, hi, ng)
    fa, fb = rng.uniform(*fwhm, ng), rng.uniform(*fwhm, ng)
    th = rng.uniform(0, np.pi, ng)
    gauss = np.stack([cy, cx, fa, fb, th, 0.2 * f], axis=1).astype(np.float64)
    return points.reshape(-1, 3), gauss.reshape(-1, 6)


def render_sky(T: int, points: np.ndarray, gauss: np.ndarray) -> np.ndarray:
    """The T x T sky (float32) of sky_components' output, summed in double:
    the points, then each Gaussian over the square stamp of half side
    ceil(3 max FWHM) around its (floor) centre pixel."""
    sky = np.zeros((T, T), np.float64)
    np.add.at(sky, (points[:, 0].astype(int), points[:, 1].astype(int)), points[:, 2])
    k = 1.0 / (8.0 * math.log(2.0))   # sigma^2 = FWHM^2 k
    for cy, cx, fa, fb, th, peak in gauss:
        r = int(math.ceil(3.0 * max(fa, fb)))   # stamp to 3 FWHM
        y0, y1 = max(0, int(cy) - r), min(T, int(cy) + r + 1)
        x0, x1 = max(0, int(cx) - r), min(T, int(cx) + r + 1)
        yy, xx = np.mgrid[y0:y1, x0:x1].astype(np.float64)
        dy, dx = yy - cy, xx - cx
        u = dx * math.cos(th) + dy * math.sin(th)
        v = -dx * math.sin(th) + dy * math.cos(th)
        sky[y0:y1, x0:x1] += peak * np.exp(-0.5 * (u * u / (fa * fa * k) + v * v / (fb * fb * k)))
    return sky.astype(np.float32)


def random_sky(T: int, rng: np.random.Generator, **kw) -> np.ndarray:
    """render_sky of sky_components (same keywords)."""
    return render_sky(T, *sky_components(T, rng, **kw))


def init_params(net: "hip.Cnn", seed: int, head_scale: float = 1.0) -> np.ndarray:
    """He-normal weights (standard deviation sqrt(2 / (9 cin))), zero biases,
    as one parameter vector in the network's layer order; the output
    layer's weights are scaled by head_scale (0: the untrained network
    proposes nothing)."""
    rng = np.random.default_rng(seed)
    parts = []
    shapes = hip.cnn_layer_shapes(net.levels, net.base, net.in_ch, net.out_ch)
    for li, (name, ci, co) in enumerate(shapes):
        w = rng.standard_normal(9 * ci * co) * math.sqrt(2.0 / (9 * ci))
        if li == len(shapes) - 1:
            w = w * head_scale
        parts.append(w.astype(np.float32))
        parts.append(np.zeros(co, np.float32))
    flat = np.concatenate(parts)
    net.set_params(flat)
    return flat


class Convolver:
    """Circular convolution of T x T real patches with the PSF patch, on the
    GPU: FFT (library kernels), multiplication by the PSF's transform,
    inverse FFT, real part."""

    def __init__(self, psf: np.ndarray):
        self.T = T = psf.shape[0]
        # transform of the PSF with its peak moved to (0, 0), divided by T^2
        # (the library's inverse FFT is unnormalised)
        ft = np.fft.fft2(np.fft.ifftshift(psf.astype(np.float64))) / (T * T)
        self.d_ft = hip.DeviceBuffer.from_array(np.stack([ft.real, ft.imag], -1).astype(np.float32))
        self.d_in = hip.DeviceBuffer(4 * T * T)
        self.d_c = hip.DeviceBuffer(8 * T * T)
        self.d_out = hip.DeviceBuffer(4 * T * T)
        self.d_work = hip.DeviceBuffer(hip.fft_work_bytes(T))

    def __call__(self, sky: np.ndarray) -> np.ndarray:
        T = self.T
        self.d_in.upload(np.ascontiguousarray(sky, np.float32))
        self.run()
        return self.d_out.download((T, T), np.float32)

    def run(self):
        """d_out = the convolution of d_in (both device buffers of T x T floats)."""
        T = self.T
        hip.fft2d(self.d_in, self.d_c, T, hip.FFT_FORWARD, hip.FFT_REAL_IN, None, self.d_work)
        hip.complex_multiply(T * T, self.d_c, self.d_ft, T * T, self.d_c)
        hip.fft2d(self.d_c, self.d_out, T, hip.FFT_INVERSE, hip.FFT_REAL_OUT, None, self.d_work)


class DataTerm:
    """Residual of a proposal in the data domain: r = P (*) out - dirty for a
    batch of T x T patches (circular convolution by FFT), its mean square,
    and its gradient 2 P^T (*) r / n added to a gradient buffer: the part of
    the loss that scores a proposal the way the next major cycle will."""

    def __init__(self, psf: np.ndarray, batch: int):
        self.T = T = psf.shape[0]
        self.batch = batch
        ft = np.fft.fft2(np.fft.ifftshift(psf.astype(np.float64))) / (T * T)
        self.d_ft = hip.DeviceBuffer.from_array(np.stack([ft.real, ft.imag], -1).astype(np.float32))
        self.d_ftc = hip.DeviceBuffer.from_array(np.stack([ft.real, -ft.imag], -1).astype(np.float32))
        self.d_c = hip.DeviceBuffer(8 * T * T * batch)
        self.d_r = hip.DeviceBuffer(4 * T * T * batch)
        self.d_work = hip.DeviceBuffer(hip.fft_work_bytes(T))

    def _conv(self, d_src, d_dst, d_filter, n):
        T = self.T
        for b in range(n):
            hip.fft2d(d_src.ptr + 4 * b * T * T, self.d_c.ptr + 8 * b * T * T, T, hip.FFT_FORWARD, hip.FFT_REAL_IN,
                      None, self.d_work)
        hip.complex_multiply(n * T * T, self.d_c, d_filter, T * T, self.d_c)
        for b in range(n):
            hip.fft2d(self.d_c.ptr + 8 * b * T * T, d_dst.ptr + 4 * b * T * T, T, hip.FFT_INVERSE, hip.FFT_REAL_OUT,
                      None, self.d_work)

    def add_grad(self, d_out, d_dirty, d_grad, n, weight) -> float:
        """d_grad += weight * 2 P^T (*) r / count; returns mean(r^2)."""
        count = n * self.T * self.T
        self._conv(d_out, self.d_r, self.d_ft, n)
        hip.axpby(count, 1.0, self.d_r, -1.0, d_dirty, self.d_r)
        sq, _, _ = hip.image_stats(count, self.d_r)
        self._conv(self.d_r, self.d_r, self.d_ftc, n)
        hip.axpby(count, 1.0, d_grad, 2.0 * weight / count, self.d_r, d_grad)
        return sq / count


@dataclass
class Pool:
    inputs: np.ndarray    # (n, T, T, 2)
    targets: np.ndarray   # (n, T, T, 1)


def smooth(image: np.ndarray, sigma: float) -> np.ndarray:
    """Circular convolution with a unit-sum Gaussian of standard deviation
    sigma pixels (sigma 0: unchanged)."""
    if sigma <= 0:
        return image
    G = image.shape[0]
    f = np.fft.fftfreq(G)
    g = np.exp(-2.0 * (math.pi * sigma) ** 2 * (f[:, None] ** 2 + f[None, :] ** 2))
    return np.fft.ifft2(np.fft.fft2(image) * g).real.astype(np.float32)


def make_pool(psf: np.ndarray, n: int, seed: int, T: int | None = None, crops_per_sky: int = 4,
              target_sigma: float = 0.0, global_scale: bool = False, crop_range: tuple[int, int] | None = None,
              **sky_kw) -> Pool:
    """n training pairs of T x T patches.  psf is the pipeline's PSF image
    (G x G, peak 1 at G/2; pass a T x T PSF patch to draw skies of the patch
    size).  Each random sky covers the PSF's full G x G field; its dirty
    image is the circular convolution with the PSF; crops_per_sky patches
    are cut from it at random origins (uniform on [crop_range) per axis,
    default the whole field), so patches carry the sidelobes of sources
    outside them, as residual tiles in the major cycle do.  Input
    (dirty / s, central PSF patch), target sky / s, with s = max |dirty| over
    the patch, or over the whole field with global_scale (the pipeline's
    SDP_LEARNED_TILE_SCALE and SDP_LEARNED_GLOBAL_SCALE); with target_sigma
    > 0 the target sky is smoothed by a Gaussian of that many pixels."""
    G = psf.shape[0]
    T = T or G
    lo, hi = crop_range if crop_range is not None else (0, G - T + 1)
    if lo < 0 or hi > G - T + 1 or lo >= hi:
        raise ValueError("crop_range outside the field")
    conv = Convolver(psf)
    rng = np.random.default_rng(seed)
    pp = psf_patch(psf, T)
    x = np.empty((n, T, T, 2), np.float32)
    y = np.empty((n, T, T, 1), np.float32)
    i = 0
    while i < n:
        sky = random_sky(G, rng, **sky_kw)
        dirty = conv(sky)
        target = smooth(sky, target_sigma)
        s_field = float(np.abs(dirty).max()) or 1.0
        for _ in range(crops_per_sky):
            if i >= n:
                break
            oy, ox = (int(v) for v in rng.integers(lo, hi, 2))
            d = dirty[oy:oy + T, ox:ox + T]
            s = s_field if global_scale else (float(np.abs(d).max()) or 1.0)
            x[i, :, :, 0] = d / s
            x[i, :, :, 1] = pp
            y[i, :, :, 0] = target[oy:oy + T, ox:ox + T] / s
            i += 1
    return Pool(x, y)


@dataclass
class DevicePool:
    """Training pairs in device memory: inputs d_x (n, T, T, 2), targets
    d_t (n, T, T) and, when made with_dirty, the scaled dirty patches d_d
    (n, T, T) the data-domain loss compares against."""
    n: int
    T: int
    d_x: "hip.DeviceBuffer"
    d_t: "hip.DeviceBuffer"
    d_d: "hip.DeviceBuffer | None"

    def download(self) -> Pool:
        T = self.T
        return Pool(self.d_x.download((self.n, T, T, 2), np.float32),
                    self.d_t.download((self.n, T, T, 1), np.float32))


def make_pool_device(psf: np.ndarray, n: int, seed: int, T: int, crops_per_sky: int = 4,
                     crop_range: tuple[int, int] | None = None, with_dirty: bool = False, **sky_kw) -> DevicePool:
    """make_pool with global_scale and no target smoothing, built on the
    GPU: each sky's components are drawn on the host (the same draws, in
    the same order, as make_pool's), rendered by sdp_sky_render, convolved
    with the PSF by the library's FFT, and its crops cut by
    sdp_cnn_gather_patches / sdp_cnn_crop_patches straight into the
    device pool; nothing of size G^2 or n T^2 passes through the host."""
    G = psf.shape[0]
    lo, hi = crop_range if crop_range is not None else (0, G - T + 1)
    if lo < 0 or hi > G - T + 1 or lo >= hi:
        raise ValueError("crop_range outside the field")
    conv = Convolver(psf)
    rng = np.random.default_rng(seed)
    d_psf = hip.DeviceBuffer.from_array(np.ascontiguousarray(psf, np.float32))
    pool = DevicePool(n, T, hip.DeviceBuffer(8 * n * T * T), hip.DeviceBuffer(4 * n * T * T),
                      hip.DeviceBuffer(4 * n * T * T) if with_dirty else None)
    d_pts, d_gs = hip.DeviceBuffer(8 * 3 * 1024), hip.DeviceBuffer(8 * 6 * 512)
    d_orig = hip.DeviceBuffer(8 * crops_per_sky)
    d_peak = hip.DeviceBuffer(4 * crops_per_sky)
    i = 0
    while i < n:
        pts, gs = sky_components(G, rng, **sky_kw)
        if len(pts):
            d_pts.upload(np.ascontiguousarray(pts))
        if len(gs):
            d_gs.upload(np.ascontiguousarray(gs))
        hip.sky_render(G, d_pts, len(pts), d_gs, len(gs), conv.d_in)
        conv.run()
        s = hip.image_stats(G * G, conv.d_out)[2] or 1.0
        k = min(crops_per_sky, n - i)
        orig = np.empty((k, 2), np.int32)
        for j in range(k):
            oy, ox = (int(v) for v in rng.integers(lo, hi, 2))
            orig[j] = (ox, oy)
        d_orig.upload(orig)
        d_peak.upload(np.full(k, s, np.float32))
        hip.cnn_gather_patches(G, T, k, d_orig, d_peak, conv.d_out, d_psf, pool.d_x.ptr + 8 * i * T * T)
        hip.cnn_crop_patches(G, T, k, d_orig, d_peak, conv.d_in, pool.d_t.ptr + 4 * i * T * T)
        if with_dirty:
            hip.cnn_crop_patches(G, T, k, d_orig, d_peak, conv.d_out, pool.d_d.ptr + 4 * i * T * T)
        i += k
    hip.synchronize()
    return pool


def train(net: "hip.Cnn", pool: "Pool | DevicePool", steps: int, batch: int, lr: float = 1e-3, seed: int = 0,
          lr_final: float = 1e-4, log_every: int = 0, sky_weight: float = 1.0, data_weight: float = 0.0,
          psf: np.ndarray | None = None) -> list[float]:
    """Adam on batches drawn from the pool; learning rate decays
    geometrically from lr to lr_final.  The loss is sky_weight times the
    mean square error against the true sky plus data_weight times the mean
    square of the proposal's data-domain residual (DataTerm, needs the PSF
    patch).  Returns the loss per step."""
    rng = np.random.default_rng(seed)
    # the pool stays in device memory; batches are gathered there
    if isinstance(pool, DevicePool):
        n, T = pool.n, pool.T
        d_pool_x, d_pool_t, d_pool_d = pool.d_x, pool.d_t, pool.d_d
        if data_weight > 0 and d_pool_d is None:
            raise ValueError("the data-domain loss needs a pool made with_dirty")
    else:
        n, T = pool.inputs.shape[0], pool.inputs.shape[1]
        d_pool_x = hip.DeviceBuffer.from_array(pool.inputs)
        d_pool_t = hip.DeviceBuffer.from_array(pool.targets)
        d_pool_d = (hip.DeviceBuffer.from_array(np.ascontiguousarray(pool.inputs[..., 0])) if data_weight > 0
                    else None)
    d_x = hip.DeviceBuffer(4 * batch * T * T * 2)
    d_t = hip.DeviceBuffer(4 * batch * T * T)
    d_y = hip.DeviceBuffer(4 * batch * T * T)
    d_g = hip.DeviceBuffer(4 * batch * T * T)
    d_dirty = hip.DeviceBuffer(4 * batch * T * T)
    data = DataTerm(psf, batch) if data_weight > 0 else None
    lib = hip.load()
    px, pt = 4 * T * T * 2, 4 * T * T
    losses = []
    for step in range(steps):
        idx = rng.choice(n, batch, replace=False)
        for b, k in enumerate(idx):
            hip.check(lib.sdp_memcpy_dtod(d_x.ptr + b * px, d_pool_x.ptr + int(k) * px, px, None))
            hip.check(lib.sdp_memcpy_dtod(d_t.ptr + b * pt, d_pool_t.ptr + int(k) * pt, pt, None))
            if d_pool_d is not None:
                hip.check(lib.sdp_memcpy_dtod(d_dirty.ptr + b * pt, d_pool_d.ptr + int(k) * pt, pt, None))
        net.forward(d_x, d_y, batch)
        loss = sky_weight * hip.cnn_mse(batch * T * T, d_y, d_t, d_g)
        if sky_weight != 1.0:
            hip.axpby(batch * T * T, sky_weight, d_g, 0.0, d_g, d_g)
        if data is not None:
            loss += data_weight * data.add_grad(d_y, d_dirty, d_g, batch, data_weight)
        losses.append(loss)
        net.backward(d_g)
        frac = step / max(steps - 1, 1)
        net.adam(lr * (lr_final / lr) ** frac)
        if log_every and step % log_every == 0:
            print(f"step {step}: loss {losses[-1]:.4g}", flush=True)
    return losses


def evaluate(net: "hip.Cnn", pool: Pool, batch: int) -> float:
    """Mean squared error of the network on the pool."""
    n, T = pool.inputs.shape[0], pool.inputs.shape[1]
    d_x = hip.DeviceBuffer(4 * batch * T * T * 2)
    d_t = hip.DeviceBuffer(4 * batch * T * T)
    d_y = hip.DeviceBuffer(4 * batch * T * T)
    d_g = hip.DeviceBuffer(4 * batch * T * T)
    tot, cnt = 0.0, 0
    for b0 in range(0, n - batch + 1, batch):
        d_x.upload(np.ascontiguousarray(pool.inputs[b0:b0 + batch]))
        d_t.upload(np.ascontiguousarray(pool.targets[b0:b0 + batch]))
        net.forward(d_x, d_y, batch)
        tot += hip.cnn_mse(batch * T * T, d_y, d_t, d_g) * batch
        cnt += batch
    return tot / max(cnt, 1)


# ------------------------------------------------------- learned workload

def sky_visibilities(sky: np.ndarray, cell: float, d_uvw: "hip.DeviceBuffer", n_vis: int, noise: float,
                     seed: int) -> "hip.DeviceBuffer":
    """Visibilities of a pixel sky (Jy per pixel; pixel (y, x) sits at
    l = (x - G/2) cell, m = (y - G/2) cell) from the library's FP64-phase DFT
    over its non-zero pixels, plus Gaussian noise of standard deviation
    `noise` per component (seeded)."""
    G = sky.shape[0]
    yy, xx = np.nonzero(sky)
    src = np.stack([(xx - G // 2) * cell, (yy - G // 2) * cell, sky[yy, xx]], 1).astype(np.float32)
    d_src = hip.DeviceBuffer.from_array(src)
    d_vis = hip.DeviceBuffer(8 * n_vis)
    hip.dft_predict(n_vis, src.shape[0], d_src, d_uvw, d_vis)
    if noise > 0:
        vis = d_vis.download((n_vis, 2), np.float32)
        vis += np.random.default_rng(seed).normal(0.0, noise, vis.shape).astype(np.float32)
        d_vis.upload(vis)
    d_src.close()
    return d_vis


@dataclass
class Workload:
    """A sky, its measured visibilities and the imaging set-up, in device memory."""
    ds: object
    sky: np.ndarray
    box: int
    d_uvw: "hip.DeviceBuffer"
    d_vis: "hip.DeviceBuffer"
    d_gk: "hip.DeviceBuffer"
    d_dk: "hip.DeviceBuffer"
    d_taper: "hip.DeviceBuffer"

    def pipeline(self, n_minor: int, **kw) -> "hip.Pipeline":
        d = self.ds
        return hip.Pipeline(d.grid_size, d.oversampling, d.supports, d.uv_scale, d.w_scale, d.cell_size, self.d_uvw,
                            self.d_vis, d.n_vis, self.d_gk, self.d_dk, self.d_taper, n_minor, **kw)


def learned_workload(n_time: int = 256, grid_size: int = 1024, seed: int = 5, noise: float = 0.01,
                     n_points: int = 50, n_gauss: int = 20, cell_size: float = 4.0e-4) -> Workload:
    """The learned-proposal workload: the single-channel array and imaging
    set-up of sdp_synth.config1 (64 antennas, 2 km disk, 2016 n_time
    visibilities, PSWF support 7, OS 16, no w-term), observing a random sky
    (random_sky, seed) of n_points power-law point sources and n_gauss
    elliptical Gaussians (FWHM 2-30 pixels) centred in the 2 x 2 degree box
    of config1's sources, with noise sigma 0.01 Jy per component.  The
    cell stays 4e-4 rad at every grid size, so that the uv grid (1250
    wavelengths across half of it) holds the longest baselines (1000)."""
    import sdp_synth
    d = sdp_synth.config1(n_time=n_time, grid_size=grid_size, with_vis=False, cell_size=cell_size)
    box = int(round(math.radians(1.0) / d.cell_size))
    sky = random_sky(grid_size, np.random.default_rng(seed), n_points=(n_points, n_points),
                     n_gauss=(n_gauss, n_gauss), box=box)
    up = hip.DeviceBuffer.from_array
    d_uvw = up(d.uvw)
    d_vis = sky_visibilities(sky, d.cell_size, d_uvw, d.n_vis, noise, seed + 1000)
    return Workload(d, sky, box, d_uvw, d_vis, up(d.kernels), up(d.degrid_kernels),
                    up(sdp_synth.taper_padded(grid_size)))


def cycles_to_target(pipe: "hip.Pipeline", target: float, max_cycles: int) -> tuple[float | None, list]:
    """Runs major cycles until the residual RMS after the minor cycle is at
    most `target`; returns (summed cycle ms when reached or None, [(rms
    before, rms after, ms) per cycle])."""
    rows, total = [], 0.0
    for _ in range(max_cycles):
        a, b, ms = pipe.cycle()
        total += ms
        rows.append((a, b, ms))
        if b <= target:
            return total, rows
    return None, rows
