"""ctypes binding of the CPU oracle (libsdp_oracle.so).

TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() (as the
checker) and bench.py's cpu_baseline leg may import this module.  Every
function takes and returns numpy arrays and calls the C restatement of the
reference actor named in its docstring (sdp_oracle.c has the file:line
citations).
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libsdp_oracle.so")
_lib = None

P = ctypes.c_void_p
I = ctypes.c_int
D = ctypes.c_double


def build() -> None:
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = ctypes.CDLL(LIB_PATH)
        sigs = {
            "oracle_correct_to_finegrid": [I, I, I, I, P, P, P, P, P, P],
            "oracle_std_gridding": [I, I, I, I, I, I, P, P, P, P, P, P, P, P, P],
            "oracle_std_gridding_f64": [I, I, I, P, P, P, P, D, D, P],
            "oracle_std_degridding": [I, I, I, I, I, P, P, P, P, P, P, P],
            "oracle_degrid_weighted_f64": [I, I, I, P, P, P, P, D, D, P],
            "oracle_gridding_actor": [I, I, I, I, P, P, P, P, P, P],
            "oracle_gridding_actor_f64": [I, I, P, P, P, P, D, D, P],
            "oracle_running_mean": [P, ctypes.c_long],
            "oracle_dft_actor": [I, I, P, P, P, P, P],
            "oracle_dft_f64": [I, I, P, P, P],
            "oracle_hogbom_clean": [I, I, I, P, P, P, P, P, P, P, P],
            "oracle_shift_c2c": [I, P, P], "oracle_shift_c2r": [I, P, P], "oracle_shift_r2c": [I, P, P],
            "oracle_convolution_correction": [I, P, P, P],
            "oracle_inv_convolution_correction": [I, P, P],
            "oracle_fft2d_f64": [I, P, P, I],
            "oracle_dft2d_direct_f64": [I, P, P, I],
            "oracle_subtract_from_measurements_finegrid": [I, P, P, P, P],
            "oracle_subtract_from_measurements": [I, I, I, P, P, P, P, P],
            "oracle_apply_gains_only": [I, I, I, P, P, P, P],
            "oracle_reciprocal_transform": [I, P, P],
            "oracle_conv2d": [P, P, P, I, I, I],
            "oracle_sliced_grid_degrid": [I, I, I, I, I, I, P, P, P, P, P, P, P, P, P],
            "oracle_clean_psf": [I, D, P, P, P],
            "oracle_add_visibilities": [I, P, P, P],
            "oracle_merge_gridded_visibilities": [I, I, P, P],
            "oracle_split_fftdegrid_counts": [I, I, I, P],
            "oracle_prolate_spheroidal": [D],
            "oracle_create_1D_half_prolate": [P, I],
        }
        for name, args in sigs.items():
            f = getattr(_lib, name)
            f.argtypes = args
            f.restype = (D if name == "oracle_prolate_spheroidal" else
                         ctypes.c_float if name == "oracle_running_mean" else
                         I if name in ("oracle_fft2d_f64", "oracle_sliced_grid_degrid") else None)
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data


def _c(a, dt):
    return np.ascontiguousarray(a, dtype=dt)


def _iref(v):
    return ctypes.byref(ctypes.c_int(int(v)))


def config_bytes(cfg) -> ctypes.c_void_p:
    """Address of a sdp_hip.Config instance (same layout as Config)."""
    return ctypes.cast(ctypes.pointer(cfg), P)


def correct_to_finegrid(G, OS, simplify, uvw, vis, cfg):
    """correct_to_finegrid, degridgrid.c:12-65"""
    uvw, vis = _c(uvw, np.float32), _c(vis, np.float32)
    n = uvw.shape[0]
    ov = np.zeros((n, 2), np.float32)
    ou = np.zeros((n, 3), np.float32)
    kept = ctypes.c_int(0)
    load().oracle_correct_to_finegrid(n, G, OS, simplify, _p(uvw), _p(vis), config_bytes(cfg), _p(ov), _p(ou),
                                      ctypes.byref(kept))
    return ov, ou, kept.value


def std_gridding(G, OS, kernels, supports, uvw, vis, cfg, n=None, bypass=0, maj_iter=0, prev_grid=None):
    """std_gridding, degridgrid.c:69-125"""
    kernels, supports = _c(kernels, np.float32), _c(supports, np.int32)
    uvw, vis = _c(uvw, np.float32), _c(vis, np.float32)
    n = uvw.shape[0] if n is None else n
    out = np.empty((G * G, 2), np.float32)
    prev = _c(prev_grid, np.float32) if prev_grid is not None else out
    load().oracle_std_gridding(G, uvw.shape[0], supports.shape[0], kernels.shape[0], OS, bypass, _iref(maj_iter),
                               _iref(n), _p(kernels), _p(supports), _p(uvw), _p(vis), config_bytes(cfg), _p(prev),
                               _p(out))
    return out


def std_gridding_f64(G, OS, kernels, supports, uvw, vis, uv_scale, w_scale):
    kernels, supports = _c(kernels, np.float32), _c(supports, np.int32)
    uvw, vis = _c(uvw, np.float32), _c(vis, np.float32)
    out = np.empty((G * G, 2), np.float64)
    load().oracle_std_gridding_f64(G, uvw.shape[0], OS, _p(kernels), _p(supports), _p(uvw), _p(vis), uv_scale,
                                   w_scale, _p(out))
    return out


def std_degridding(G, OS, kernels, supports, grid, uvw, cfg, n=None, n_total=None):
    """std_degridding, degridgrid.c:127-202 (nearest-cell output)"""
    kernels, supports = _c(kernels, np.float32), _c(supports, np.int32)
    grid, uvw = _c(grid, np.float32), _c(uvw, np.float32)
    n = uvw.shape[0] if n is None else n
    n_total = uvw.shape[0] if n_total is None else n_total
    out = np.empty((n_total, 2), np.float32)
    load().oracle_std_degridding(G, n_total, supports.shape[0], kernels.shape[0], OS, _p(kernels), _p(supports),
                                 _p(grid), _p(uvw), _iref(n), config_bytes(cfg), _p(out))
    return out


def degrid_weighted_f64(G, OS, kernels, supports, grid, uvw, uv_scale, w_scale):
    kernels, supports = _c(kernels, np.float32), _c(supports, np.int32)
    grid, uvw = _c(grid, np.float32), _c(uvw, np.float32)
    out = np.empty((uvw.shape[0], 2), np.float64)
    load().oracle_degrid_weighted_f64(G, uvw.shape[0], OS, _p(kernels), _p(supports), _p(grid), _p(uvw), uv_scale,
                                      w_scale, _p(out))
    return out


def gridding_actor(G, kernels, supports, uvw, vis, cfg):
    """gridding_actor / gridding_CPU, gridding_run.c:5-137"""
    kernels, supports = _c(kernels, np.float32), _c(supports, np.int32)
    uvw, vis = _c(uvw, np.float32), _c(vis, np.float32)
    out = np.empty((G * G, 2), np.float32)
    load().oracle_gridding_actor(G, uvw.shape[0], supports.shape[0], kernels.shape[0], _p(kernels), _p(supports),
                                 _p(uvw), _p(vis), config_bytes(cfg), _p(out))
    return out


def gridding_actor_f64(G, kernels, supports, uvw, vis, uv_scale, w_scale):
    """gridding_actor's index arithmetic with double products and sums"""
    kernels, supports = _c(kernels, np.float32), _c(supports, np.int32)
    uvw, vis = _c(uvw, np.float32), _c(vis, np.float32)
    out = np.empty((G * G, 2), np.float64)
    load().oracle_gridding_actor_f64(G, uvw.shape[0], _p(kernels), _p(supports), _p(uvw), _p(vis), uv_scale, w_scale,
                                     _p(out))
    return out


def running_mean(r):
    """hogbom_clean's float "average" recurrence, deconvolution_run.c:223-227"""
    r = _c(r, np.float32).ravel()
    return float(load().oracle_running_mean(_p(r), r.size))


def dft_actor(sources, uvw, cfg):
    """dft_actor, dft_run.c:3-47"""
    sources, uvw = _c(sources, np.float32), _c(uvw, np.float32)
    out = np.empty((uvw.shape[0], 2), np.float32)
    load().oracle_dft_actor(uvw.shape[0], sources.shape[0], _p(sources), _p(uvw), _iref(sources.shape[0]),
                            config_bytes(cfg), _p(out))
    return out


def dft_f64(sources, uvw):
    sources, uvw = _c(sources, np.float32), _c(uvw, np.float32)
    out = np.empty((uvw.shape[0], 2), np.float64)
    load().oracle_dft_f64(uvw.shape[0], sources.shape[0], _p(sources), _p(uvw), _p(out))
    return out


def hogbom_clean(G, n_minor, max_sources, residual, psf_padded, cfg, model):
    """hogbom_clean, deconvolution_run.c:197-312; psf_padded has G*G+1 floats"""
    residual, psf_padded, model = _c(residual, np.float32), _c(psf_padded, np.float32), _c(model, np.float32)
    out_model = np.empty(G * G, np.float32)
    out_res = np.empty(G * G, np.float32)
    # the reference writes one source per iteration whatever MAX_SOURCES is
    sources = np.zeros((max(max_sources, n_minor, 1), 3), np.float32)
    ns = ctypes.c_int(0)
    load().oracle_hogbom_clean(G, n_minor, max_sources, _p(residual), _p(psf_padded), config_bytes(cfg), _p(model),
                               ctypes.byref(ns), _p(sources), _p(out_model), _p(out_res))
    return ns.value, sources[: ns.value], out_model, out_res


def shift_c2c(G, grid):
    grid = _c(grid, np.float32)
    out = np.empty_like(grid)
    load().oracle_shift_c2c(G, _p(grid), _p(out))
    return out


def shift_c2r(G, grid):
    grid = _c(grid, np.float32)
    out = np.empty(G * G, np.float32)
    load().oracle_shift_c2r(G, _p(grid), _p(out))
    return out


def shift_r2c(G, image):
    image = _c(image, np.float32)
    out = np.empty((G * G, 2), np.float32)
    load().oracle_shift_r2c(G, _p(image), _p(out))
    return out


def convolution_correction(G, image, prolate_padded):
    image, prolate_padded = _c(image, np.float32), _c(prolate_padded, np.float32)
    out = np.empty_like(image)
    load().oracle_convolution_correction(G, _p(image), _p(prolate_padded), _p(out))
    return out


def fft2d_f64(G, x, sign):
    x = _c(x, np.float64)
    out = np.empty_like(x)
    rc = load().oracle_fft2d_f64(G, _p(x), _p(out), sign)
    assert rc == 0
    return out


def dft2d_direct_f64(G, x, sign):
    x = _c(x, np.float64)
    out = np.empty_like(x)
    load().oracle_dft2d_direct_f64(G, _p(x), _p(out), sign)
    return out


def subtract_finegrid(measured, pred, n=None):
    measured, pred = _c(measured, np.float32), _c(pred, np.float32)
    out = np.zeros_like(measured)
    n = measured.shape[0] if n is None else n
    load().oracle_subtract_from_measurements_finegrid(measured.shape[0], _p(measured), _p(pred), _iref(n), _p(out))
    return out


def subtract_from_measurements(measured, pred, gains, pairs):
    measured, pred, gains, pairs = (_c(measured, np.float32), _c(pred, np.float32), _c(gains, np.float32),
                                    _c(pairs, np.int32))
    out = np.empty_like(measured)
    load().oracle_subtract_from_measurements(gains.shape[0], pairs.shape[0], measured.shape[0], _p(measured),
                                             _p(pred), _p(gains), _p(pairs), _p(out))
    return out


def reciprocal_transform(gains):
    """reciprocal_transform_actor, gains_apply_run.c:4-17."""
    gains = _c(gains, np.float32)
    out = np.empty_like(gains)
    load().oracle_reciprocal_transform(gains.shape[0], _p(gains), _p(out))
    return out


def apply_gains_only(measured, gains, pairs):
    measured, gains, pairs = _c(measured, np.float32), _c(gains, np.float32), _c(pairs, np.int32)
    out = np.empty_like(measured)
    load().oracle_apply_gains_only(gains.shape[0], pairs.shape[0], measured.shape[0], _p(measured), _p(gains),
                                   _p(pairs), _p(out))
    return out


def add_visibilities(a, b):
    a, b = _c(a, np.float32), _c(b, np.float32)
    out = np.empty_like(a)
    load().oracle_add_visibilities(a.shape[0], _p(a), _p(b), _p(out))
    return out


def merge(G, slices, grids):
    grids = _c(grids, np.float32)
    out = np.empty((G * G, 2), np.float32)
    load().oracle_merge_gridded_visibilities(G, slices, _p(grids), _p(out))
    return out


def split_counts(n_vis, slices, n_corrected):
    out = np.zeros(slices, np.int32)
    load().oracle_split_fftdegrid_counts(n_vis, slices, n_corrected, _p(out))
    return out


def prolate_spheroidal(nu: float) -> float:
    return float(load().oracle_prolate_spheroidal(nu))


def half_prolate(G):
    out = np.empty(G // 2, np.float32)
    load().oracle_create_1D_half_prolate(_p(out), G)
    return out


def conv2d(image, clean_psf, hx, hy):
    """conv2D, major_loop_iter.c:8-46: the model convolved with the clean PSF."""
    image, clean_psf = _c(image, np.float32), _c(clean_psf, np.float32)
    D = image.shape[0]
    out = np.empty_like(image)
    load().oracle_conv2d(_p(image), _p(clean_psf), _p(out), D, int(hx), int(hy))
    return out


def clean_psf(dirty_psf, psf_max_value):
    """clean_psf_host_set_up, top.c:527-592: (clean PSF, half size)."""
    dirty_psf = _c(dirty_psf, np.float32)
    G = dirty_psf.shape[0]
    out = np.empty_like(dirty_psf)
    h = ctypes.c_int(0)
    load().oracle_clean_psf(G, float(psf_max_value), _p(dirty_psf), _p(out), ctypes.byref(h))
    return out, h.value


def sliced_grid_degrid(slices, G, OS, grid_kernels, degrid_kernels, supports, uvw, vis, cfg):
    """std_gridding then std_degridding of the gridded result over `slices`
    threads, as splitmerge.c splits and merges: (grid, visibilities,
    (grid seconds, degrid seconds))."""
    gk, dk = _c(grid_kernels, np.float32), _c(degrid_kernels, np.float32)
    supports, uvw, vis = _c(supports, np.int32), _c(uvw, np.float32), _c(vis, np.float32)
    grid = np.empty((G * G, 2), np.float32)
    out = np.empty((uvw.shape[0], 2), np.float32)
    sec = np.zeros(2, np.float64)
    rc = load().oracle_sliced_grid_degrid(int(slices), G, uvw.shape[0], supports.shape[0], gk.shape[0], OS, _p(gk),
                                          _p(dk), _p(supports), _p(uvw), _p(vis), config_bytes(cfg), _p(grid),
                                          _p(out), _p(sec))
    if rc != 0:
        raise MemoryError("oracle_sliced_grid_degrid failed")
    return grid, out, (float(sec[0]), float(sec[1]))
