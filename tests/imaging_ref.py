"""FP64 restatement of one imaging major cycle, built from the oracle's
restated actors, for the pipeline tests.

The chain follows the reference's FFT imaging graph
(preesm_pipelines/simsdp_fft_imaging_pipeline/Algo/fft_delta.pi):
grid (degridgrid.c:69-125) -> fft_shift_complex_to_complex -> inverse FFT
(FFTW_BACKWARD, unnormalised, fft_run.c:23-48) -> fft_shift_complex_to_real
-> execute_convolution_correction_actor (convolution_correction_run.c:17-31);
and for the predict: taper division -> fft_shift_real_to_complex -> forward
FFT (FFTW_FORWARD, fft_run.c:50-73) -> fft_shift_complex_to_complex ->
kernel-weighted degridding with the conjugate table.

The grid / degrid sums come from the oracle's FP64 restatements; the FFTs are
numpy's (unnormalised forward, inverse times G^2), which the oracle pins
check against the FFT definition.
"""
import numpy as np

import sdp_oracle
import sdp_synth


def checkerboard(G):
    return (-1.0) ** np.add.outer(np.arange(G), np.arange(G))


def taper_2d(G):
    """p[|r - G/2|] * p[|c - G/2|] with the padded taper (element G/2 = 0)."""
    p = sdp_synth.taper_padded(G).astype(np.float64)
    idx = np.abs(np.arange(G) - G // 2)
    return np.outer(p[idx], p[idx])


def correct(G, img):
    t = taper_2d(G)
    out = np.zeros_like(img)
    ok = np.abs(t) > 1e-10
    out[ok] = img[ok] / t[ok]
    return out


def uncorrect_in(G, img):
    """The predict's taper division (the inverse of correct on its support)."""
    return correct(G, img)


def dirty(d, vis):
    """Corrected dirty image (G x G float64) of visibilities vis."""
    G = d.grid_size
    g = sdp_oracle.std_gridding_f64(G, d.oversampling, d.kernels, d.supports, d.uvw, vis, d.uv_scale, d.w_scale)
    gc = (g[:, 0] + 1j * g[:, 1]).reshape(G, G)
    s = checkerboard(G)
    img = (np.fft.ifft2(gc * s) * (G * G) * s).real
    return correct(G, img)


def predict(d, model):
    """Kernel-weighted predicted visibilities (N x 2 float64) of a model image."""
    G = d.grid_size
    s = checkerboard(G)
    grid = np.fft.fft2(uncorrect_in(G, model.astype(np.float64)) * s) * s
    g32 = np.stack([grid.real.ravel(), grid.imag.ravel()], axis=1).astype(np.float32)
    return sdp_oracle.degrid_weighted_f64(G, d.oversampling, d.degrid_kernels, d.supports, g32, d.uvw, d.uv_scale,
                                          d.w_scale)


def psf(d):
    """(peak-normalised PSF, its peak) from unit visibilities."""
    ones = np.zeros((d.n_vis, 2), np.float32)
    ones[:, 0] = 1.0
    img = dirty(d, ones)
    peak = float(np.float32(img.max()))
    return img / peak, peak


def sky_model(d):
    """The true sky of a sky_case as a model image (flux on source pixels)."""
    G = d.grid_size
    m = np.zeros((G, G), np.float32)
    for (x, y), f in zip(d.extra["pixels"], d.sources[:, 2]):
        m[y, x] += f
    return m
