"""Seeded synthetic inputs for the imaging major cycle.

The reference reads GLEAM visibilities, kernel CSVs and a PSF CSV that are not
shipped with it (README "Download GLEAM" step, top.c:19-60).  This module
builds inputs of the same shapes and layouts from seeded random numbers
(numpy PCG64):

* antenna layouts (uniform disk, log-radial), Earth-rotation uvw tracks in
  wavelengths, ordered time-major (visibility index = time * n_baselines +
  baseline, the order subtract_from_measurements assumes,
  gains_apply_run.c:81), baselines in calculate_receiver_pairs order
  (top.c:119-137);
* point-source skies and their visibilities (exact FP64 DFT plus noise);
* gridding-kernel tables in the kernel_host_set_up layout (top.c:252-257):
  per w-plane a ((h+1)*OS)^2 quadrant indexed [ky*(h+1)*OS + kx], the
  support table holding (h, sample offset).  Plane 0 is the prolate
  spheroidal kernel; plane p>0 is the w-projection kernel for
  w_p = p^2 / w_scale (the sqrt spacing implied by the plane index
  round(sqrt(|w| w_scale)), degridgrid.c:84).

The generator is input preparation, run once per configuration; it is not
part of the major-cycle hot path.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

C_LIGHT = 299792458.0


def prolate_spheroidal(nu: np.ndarray) -> np.ndarray:
    """Schwab's rational approximation used by the reference taper
    (top.c:408-447), vectorised; 0 outside [0, 1)."""
    p = np.array([0.08203343, -0.3644705, 0.627866, -0.5335581, 0.2312756,
                  0.004028559, -0.03697768, 0.1021332, -0.1201436, 0.06412774])
    q = np.array([1.0, 0.8212018, 0.2078043, 1.0, 0.9599102, 0.2918724])
    nu = np.abs(np.asarray(nu, dtype=np.float64))
    out = np.zeros_like(nu)
    for part, lo, hi in ((0, 0.0, 0.75), (1, 0.75, 1.0)):
        m = (nu >= lo) & (nu < hi)
        d = nu[m] ** 2 - hi ** 2
        top = sum(p[part * 5 + i] * d ** i for i in range(5))
        bot = sum(q[part * 3 + i] * d ** i for i in range(3))
        out[m] = top / bot
    return out


def taper(grid_size: int) -> np.ndarray:
    """create_1D_half_prolate (top.c:395-404): G/2 floats."""
    half = grid_size // 2
    return prolate_spheroidal(np.arange(half) / half).astype(np.float32)


def taper_padded(grid_size: int) -> np.ndarray:
    """The taper with the trailing 0 the correction reads for row/column 0."""
    return np.concatenate([taper(grid_size), np.zeros(1, np.float32)])


# ------------------------------------------------------------------ geometry

def disk_layout(n_ant: int, radius_m: float, seed: int) -> np.ndarray:
    rng = np.random.default_rng(seed)
    r = radius_m * np.sqrt(rng.random(n_ant))
    th = 2 * np.pi * rng.random(n_ant)
    return np.stack([r * np.cos(th), r * np.sin(th), np.zeros(n_ant)], axis=1)


def log_radial_layout(n_ant: int, r_min: float, r_max: float, seed: int) -> np.ndarray:
    rng = np.random.default_rng(seed)
    r = np.exp(rng.uniform(np.log(r_min), np.log(r_max), n_ant))
    th = 2 * np.pi * rng.random(n_ant)
    return np.stack([r * np.cos(th), r * np.sin(th), np.zeros(n_ant)], axis=1)


def receiver_pairs(n_ant: int) -> np.ndarray:
    """calculate_receiver_pairs order (top.c:119-137)."""
    a, b = np.triu_indices(n_ant, k=1)
    return np.stack([a, b], axis=1).astype(np.int32)


def uvw_tracks(enu: np.ndarray, lat_deg: float, dec_deg: float, ha_hours: np.ndarray,
               freq_hz: float) -> np.ndarray:
    """uvw (wavelengths, float64) for every (time, baseline), time-major."""
    pairs = receiver_pairs(enu.shape[0])
    d = enu[pairs[:, 1]] - enu[pairs[:, 0]]
    lat, dec = np.radians(lat_deg), np.radians(dec_deg)
    # ENU -> local equatorial XYZ
    X = -np.sin(lat) * d[:, 1] + np.cos(lat) * d[:, 2]
    Y = d[:, 0]
    Z = np.cos(lat) * d[:, 1] + np.sin(lat) * d[:, 2]
    H = np.radians(15.0 * np.asarray(ha_hours))[:, None]
    u = np.sin(H) * X + np.cos(H) * Y
    v = -np.sin(dec) * np.cos(H) * X + np.sin(dec) * np.sin(H) * Y + np.cos(dec) * Z
    w = np.cos(dec) * np.cos(H) * X - np.cos(dec) * np.sin(H) * Y + np.sin(dec) * Z
    lam = C_LIGHT / freq_hz
    return np.stack([u, v, w], axis=-1).reshape(-1, 3) / lam


def antenna_uvw(enu: np.ndarray, lat_deg: float, dec_deg: float, ha_hours: np.ndarray,
                freq_hz: float) -> np.ndarray:
    """Per-antenna uvw (wavelengths, float64), shape (n_time, n_ant, 3), about
    the array origin: baseline (a, b) of uvw_tracks is [:, b] - [:, a]."""
    lat, dec = np.radians(lat_deg), np.radians(dec_deg)
    X = -np.sin(lat) * enu[:, 1] + np.cos(lat) * enu[:, 2]
    Y = enu[:, 0]
    Z = np.cos(lat) * enu[:, 1] + np.sin(lat) * enu[:, 2]
    H = np.radians(15.0 * np.asarray(ha_hours))[:, None]
    u = np.sin(H) * X + np.cos(H) * Y
    v = -np.sin(dec) * np.cos(H) * X + np.sin(dec) * np.sin(H) * Y + np.cos(dec) * Z
    w = np.cos(dec) * np.cos(H) * X - np.cos(dec) * np.sin(H) * Y + np.sin(dec) * Z
    return np.stack([u, v, w], axis=-1) / (C_LIGHT / freq_hz)


def dft_f64(uvw: np.ndarray, sources: np.ndarray, block: int = 65536) -> np.ndarray:
    """Exact FP64 predict with the reference's formula (dft_run.c:21-32)."""
    out = np.zeros((uvw.shape[0], 2), np.float64)
    l, m, f = (sources[:, k].astype(np.float64) for k in range(3))
    term = 0.5 * (l * l + m * m)
    corr = f / (1.0 - term)
    for s in range(0, uvw.shape[0], block):
        u = uvw[s:s + block].astype(np.float64)
        ph = u[:, 0:1] * l + u[:, 1:2] * m - u[:, 2:3] * term
        ph -= np.rint(ph)
        out[s:s + block, 0] = (np.cos(2 * np.pi * ph) * corr).sum(axis=1)
        out[s:s + block, 1] = -(np.sin(2 * np.pi * ph) * corr).sum(axis=1)
    return out


# ------------------------------------------------------------------ kernels

def kernel_tables(supports_h: list[int], oversampling: int, fov: float, w_scale: float,
                  q_min: int = 0) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
    """Gridding kernel table, its conjugate (degridding) table and the
    support table (h, offset) for planes with half supports supports_h.

    Plane p images the w-term of w_p = p^2 / w_scale: the uv kernel is the
    Fourier transform of psi(l)psi(m) exp(-i pi w_p (l^2 + m^2)) over the
    field of view (l, m in units of the half field), sampled at 1/OS cell.
    All planes are scaled by the sum of plane 0 over integer offsets, so the
    w = 0 kernel sums to 1."""
    n_planes = len(supports_h)
    h_max = max(supports_h)
    q = max(q_min, 2 * (h_max + 4))
    q += q % 2
    n_pad = q * oversampling
    x = (np.arange(q) - q // 2) / (q / 2)          # image coordinate in half fields, [-1, 1)
    xx, yy = np.meshgrid(x, x, indexing="ij")
    base = prolate_spheroidal(xx) * prolate_spheroidal(yy)
    l = xx * fov / 2
    m = yy * fov / 2
    tables, offsets = [], []
    norm = None
    off = 0
    for p, h in enumerate(supports_h):
        w_p = (p * p) / w_scale if w_scale > 0 else 0.0
        screen = base * np.exp(-1j * np.pi * w_p * (l * l + m * m))
        pad = np.zeros((n_pad, n_pad), np.complex128)
        c = n_pad // 2
        pad[c - q // 2:c + q // 2, c - q // 2:c + q // 2] = screen
        k = np.fft.fftshift(np.fft.fft2(np.fft.ifftshift(pad)))   # uv kernel, 1/OS-cell samples, centred
        if norm is None:
            ints = k[c % oversampling::oversampling, c % oversampling::oversampling]
            norm = ints.sum().real
        k = k / norm
        width = (h + 1) * oversampling
        quad = k[c:c + width, c:c + width]                 # offsets (ky, kx) >= 0
        tables.append(quad.reshape(-1))
        offsets.append(off)
        off += width * width
    kern = np.concatenate(tables)
    grid_k = np.stack([kern.real, kern.imag], axis=1).astype(np.float32)
    degrid_k = np.stack([kern.real, -kern.imag], axis=1).astype(np.float32)
    sup = np.stack([np.asarray(supports_h), np.asarray(offsets)], axis=1).astype(np.int32)
    return grid_k, degrid_k, sup


# ------------------------------------------------------------------ configs

@dataclass
class Dataset:
    name: str
    grid_size: int
    cell_size: float
    oversampling: int
    uvw: np.ndarray            # (N, 3) float32 wavelengths
    vis: np.ndarray            # (N, 2) float32
    kernels: np.ndarray        # (S, 2) float32 gridding table
    degrid_kernels: np.ndarray # (S, 2) float32 conjugate table
    supports: np.ndarray       # (P, 2) int32 (h, offset)
    sources: np.ndarray        # (n_src, 3) float32 (l, m, flux)
    max_w: float
    n_minor: int
    n_major: int
    extra: dict = field(default_factory=dict)

    @property
    def uv_scale(self) -> float:
        return self.cell_size * self.grid_size

    @property
    def w_scale(self) -> float:
        n = self.supports.shape[0]
        return float((n - 1) ** 2) / self.max_w if n > 1 else 0.0

    @property
    def n_vis(self) -> int:
        return self.uvw.shape[0]


def config1(n_time: int = 256, grid_size: int = 1024, with_vis: bool = True,
            cell_size: float | None = None) -> Dataset:
    """Plain FFT imaging, single channel: 64 antennas uniform in a 2 km
    diameter disk (seed 1), latitude -30 deg, dec -30 deg, HA -3..+3 h in
    n_time steps at 150 MHz -> 2016 * n_time visibilities; 50 sources with
    flux U[0.1, 1] Jy in a 2 x 2 deg box (seed 2), noise sigma 0.01 Jy per
    component (seed 3); PSWF support 7 (h = 3), OS 16, one plane.  The cell
    defaults to 4e-4 rad at 1024^2 (field of view 23.5 deg whatever the
    grid size); pass cell_size to keep the uv extent instead."""
    enu = disk_layout(64, 1000.0, seed=1)
    ha = np.linspace(-3.0, 3.0, n_time)
    uvw = uvw_tracks(enu, -30.0, -30.0, ha, 150e6)
    rng = np.random.default_rng(2)
    half_box = np.radians(1.0)
    src = np.stack([rng.uniform(-half_box, half_box, 50), rng.uniform(-half_box, half_box, 50),
                    rng.uniform(0.1, 1.0, 50)], axis=1).astype(np.float32)
    cell = cell_size if cell_size is not None else 4.0e-4 * 1024 / grid_size
    gk, dk, sup = kernel_tables([3], 16, fov=cell * grid_size, w_scale=0.0)
    uvw32 = uvw.astype(np.float32)
    if with_vis:
        vis = dft_f64(uvw32, src)
        vis += np.random.default_rng(3).normal(0.0, 0.01, vis.shape)
        vis32 = vis.astype(np.float32)
    else:
        vis32 = np.zeros((uvw32.shape[0], 2), np.float32)
    return Dataset("config1", grid_size, cell, 16, uvw32, vis32, gk, dk, sup, src, max_w=1.0,
                   n_minor=200, n_major=1)


def config2(n_ant: int = 256, n_time: int = 512, grid_size: int = 4096, n_planes: int = 32) -> Dataset:
    """Plain full major cycle at real size: 256 antennas log-radial out to
    5 km radius (10 km across, seed 11), 512 time steps (HA -4..+4 h) at
    150 MHz -> 32,640 * 512 = 16,711,680 visibilities; 1000 sources with
    dN/dS ~ S^-1.5 on [1 mJy, 1 Jy] (seed 12) in the central half of a
    3.5 deg field; 4096^2 grid with w-projection, 32 planes, half support
    growing 4 -> 16 with |w|, OS 8.  Measured visibilities are left zero
    here: the major-cycle driver predicts them on the GPU."""
    enu = log_radial_layout(n_ant, 50.0, 5000.0, seed=11)
    ha = np.linspace(-4.0, 4.0, n_time)
    uvw = uvw_tracks(enu, -30.0, -30.0, ha, 150e6).astype(np.float32)
    cell = 1.5e-5
    fov = cell * grid_size
    max_w = float(np.abs(uvw[:, 2]).max()) * 1.0001
    w_scale = (n_planes - 1) ** 2 / max_w
    # full support 8 -> 32 cells across the planes; every plane's w-term
    # spread (about 3 + w_p fov^2 / 2 cells) stays inside its support
    sup_h = [4 + int(round(12 * p / (n_planes - 1))) for p in range(n_planes)]
    assert all(3 + (p * p / w_scale) * fov * fov / 2 <= h for p, h in enumerate(sup_h))
    gk, dk, sup = kernel_tables(sup_h, 8, fov=fov, w_scale=w_scale)
    rng = np.random.default_rng(12)
    n_src = 1000
    a = -1.5
    u01 = rng.random(n_src)
    s_lo, s_hi = 1e-3, 1.0
    flux = (s_lo ** (a + 1) + u01 * (s_hi ** (a + 1) - s_lo ** (a + 1))) ** (1.0 / (a + 1))
    lm = rng.uniform(-fov / 4, fov / 4, (n_src, 2))
    src = np.concatenate([lm, flux[:, None]], axis=1).astype(np.float32)
    vis = np.zeros((uvw.shape[0], 2), np.float32)
    ds = Dataset("config2", grid_size, cell, 8, uvw, vis, gk, dk, sup, src, max_w=max_w,
                 n_minor=1000, n_major=5)
    ds.extra["enu"] = enu
    ds.extra["ha_hours"] = ha
    return ds


def config3_sources(n_src: int = 4096, grid_size: int = 4096, seed: int = 13) -> np.ndarray:
    """Precision-variant sky: n_src sources with dN/dS ~ S^-1.5 on
    [1 mJy, 1 Jy] (seed) in the central half of config2's field."""
    rng = np.random.default_rng(seed)
    fov = 1.5e-5 * grid_size
    a = -1.5
    u01 = rng.random(n_src)
    flux = (1e-3 ** (a + 1) + u01 * (1.0 - 1e-3 ** (a + 1))) ** (1.0 / (a + 1))
    lm = rng.uniform(-fov / 4, fov / 4, (n_src, 2))
    return np.concatenate([lm, flux[:, None]], axis=1).astype(np.float32)


@dataclass
class MultiChannel:
    """Batched multi-channel workload: one array observation at n_channels
    frequencies.  base_uvw holds every (time, baseline) uvw in metres, in a
    seeded random order; channel c keeps the first n_vis[c] rows (its
    unflagged visibilities: a uniformly random subset of the fraction
    1 - flag[c]) and scales them by freqs[c] / c."""
    grid_size: int
    cell_size: float
    oversampling: int
    base_uvw: np.ndarray       # (N, 3) float32 metres
    freqs: np.ndarray          # (n_channels,) Hz
    n_vis: np.ndarray          # (n_channels,) int
    flag: np.ndarray           # (n_channels,) flagged fraction
    kernels: np.ndarray
    degrid_kernels: np.ndarray
    supports: np.ndarray
    sources: np.ndarray        # (n_src, 3) l, m, flux at ref_freq
    spectral_index: float
    ref_freq: float
    max_w: float

    @property
    def uv_scale(self) -> float:
        return self.cell_size * self.grid_size

    @property
    def w_scale(self) -> float:
        return float((self.supports.shape[0] - 1) ** 2) / self.max_w

    def channel_uvw(self, c: int) -> np.ndarray:
        return (self.base_uvw[:self.n_vis[c]] * np.float32(self.freqs[c] / C_LIGHT)).astype(np.float32)

    def channel_sources(self, c: int) -> np.ndarray:
        s = self.sources.copy()
        s[:, 2] *= (self.freqs[c] / self.ref_freq) ** self.spectral_index
        return s


def config4(n_channels: int = 64, n_time: int = 8192, grid_size: int = 2048, n_planes: int = 16,
            f_lo: float = 100e6, f_hi: float = 200e6) -> MultiChannel:
    """Batched multi-channel: config1's array (64 antennas, 2 km disk, seed 1)
    and sky (50 sources, U[0.1, 1] Jy at 150 MHz, seed 2) with spectral
    index -0.7, n_time steps over HA -3..+3 h (8192: 2016 * 8192 =
    16,515,072 rows per channel), n_channels frequencies on [f_lo, f_hi],
    per-channel flagged fraction U[0, 0.6] (seed 3).  grid_size^2 grids with
    cell 3e-5 rad (the 3.5 degree field of config2, so that w-projection
    supports stay small at the longest w), n_planes w-planes, half support
    4 -> 7 with |w|, OS 16."""
    enu = disk_layout(64, 1000.0, seed=1)
    ha = np.linspace(-3.0, 3.0, n_time)
    base = uvw_tracks(enu, -30.0, -30.0, ha, C_LIGHT)           # metres (lambda = 1 m)
    perm = np.random.default_rng(3).permutation(base.shape[0])
    base = base[perm].astype(np.float32)
    freqs = np.linspace(f_lo, f_hi, n_channels)
    flag = np.random.default_rng(3).uniform(0.0, 0.6, n_channels)
    n_vis = np.rint(base.shape[0] * (1.0 - flag)).astype(np.int64)
    cell = 3.0e-5
    fov = cell * grid_size
    max_w = float(np.abs(base[:, 2]).max()) * f_hi / C_LIGHT * 1.0001
    w_scale = (n_planes - 1) ** 2 / max_w
    sup_h = [4 + int(round(3 * p / (n_planes - 1))) for p in range(n_planes)]
    assert all(3 + (p * p / w_scale) * fov * fov / 2 <= h for p, h in enumerate(sup_h))
    gk, dk, sup = kernel_tables(sup_h, 16, fov=fov, w_scale=w_scale)
    rng = np.random.default_rng(2)
    half_box = np.radians(1.0)
    src = np.stack([rng.uniform(-half_box, half_box, 50), rng.uniform(-half_box, half_box, 50),
                    rng.uniform(0.1, 1.0, 50)], axis=1).astype(np.float32)
    return MultiChannel(grid_size, cell, 16, base, freqs, n_vis, flag, gk, dk, sup, src, -0.7, 150e6, max_w)


def small_case(n_vis: int = 4000, grid_size: int = 128, n_planes: int = 4, oversampling: int = 16,
               seed: int = 7, h_list=None, fill: float = 0.98, w_max: float = 200.0) -> Dataset:
    """A small random case for parity tests: uvw uniform so that footprints
    cover most of the grid (some reach past the edge), w uniform in
    [-w_max, w_max], complex Gaussian visibilities."""
    rng = np.random.default_rng(seed)
    cell = 1.0 / grid_size / 10.0
    uv_scale = cell * grid_size
    half = grid_size / 2
    u = rng.uniform(-half * fill, half * fill, n_vis) / uv_scale
    v = rng.uniform(-half * fill, half * fill, n_vis) / uv_scale
    w = rng.uniform(-w_max, w_max, n_vis)
    uvw = np.stack([u, v, w], axis=1).astype(np.float32)
    vis = rng.normal(size=(n_vis, 2)).astype(np.float32)
    if h_list is None:
        h_list = [3 + (p % 3) for p in range(n_planes)]
    max_w = w_max * 1.0001
    w_scale = (len(h_list) - 1) ** 2 / max_w if len(h_list) > 1 else 0.0
    gk, dk, sup = kernel_tables(h_list, oversampling, fov=cell * grid_size, w_scale=w_scale if w_scale else 0.0)
    src = np.zeros((0, 3), np.float32)
    return Dataset("small", grid_size, cell, oversampling, uvw, vis, gk, dk, sup, src, max_w=max_w,
                   n_minor=50, n_major=1)


def sky_case(grid_size: int = 256, n_ant: int = 48, n_time: int = 64, n_src: int = 8, oversampling: int = 8,
             seed: int = 21, noise: float = 0.0) -> Dataset:
    """A small end-to-end imaging case: n_ant antennas log-radial 20 m to
    1.5 km (seed), HA -3..+3 h in n_time steps at 150 MHz, 4 w-planes
    (half supports 4, 4, 6, 8, each wide enough for its w-term), n_src point
    sources on pixel centres inside the central half of the field with flux
    U[0.2, 1] Jy, visibilities from the FP64 DFT plus optional Gaussian
    noise (sigma `noise` per component)."""
    rng = np.random.default_rng(seed)
    enu = log_radial_layout(n_ant, 20.0, 1500.0, seed=seed)
    uvw = uvw_tracks(enu, -30.0, -30.0, np.linspace(-3.0, 3.0, n_time), 150e6).astype(np.float32)
    sup_h = [4, 4, 6, 8]
    u_max = float(np.abs(uvw[:, :2]).max())
    cell = (grid_size / 2 - max(sup_h) - 2) / (u_max * grid_size)
    fov = cell * grid_size
    max_w = float(np.abs(uvw[:, 2]).max()) * 1.0001
    w_scale = (len(sup_h) - 1) ** 2 / max_w
    assert all(3 + (p * p / w_scale) * fov * fov / 2 <= h for p, h in enumerate(sup_h))
    gk, dk, sup = kernel_tables(sup_h, oversampling, fov=fov, w_scale=w_scale)
    px = rng.choice(np.arange(grid_size // 4, 3 * grid_size // 4), size=(n_src, 2))
    lm = (px - grid_size // 2) * cell
    flux = rng.uniform(0.2, 1.0, n_src)
    src = np.concatenate([lm, flux[:, None]], axis=1).astype(np.float32)
    vis = dft_f64(uvw, src)
    if noise > 0:
        vis += rng.normal(0.0, noise, vis.shape)
    ds = Dataset("sky", grid_size, cell, oversampling, uvw, vis.astype(np.float32), gk, dk, sup, src,
                 max_w=max_w, n_minor=200, n_major=4)
    ds.extra["pixels"] = px
    return ds
