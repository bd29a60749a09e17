"""The generated single-core schedule of the FFT imaging pipeline
(param_code/code_fft_1core/generated/Core0.c:2588-4045) run as a C program,
sdp_major_cycle, linked against libsdp_actors.so, against the same actor
sequence on the oracle.

The case is C1 at full size (516,096 visibilities, 1024^2, h = 3, OS 16)
with per-antenna complex gains applied to the measured visibilities, four
slices as the schedule splits them (split_fftdegrid_visibilities,
split_uvgrid, merge_gridded_visibilities), and three major cycles of 200
Hogbom iterations, each cycle's model the next one's image estimate.  The
oracle runs every actor of the sequence as its C restatement, with the two
FFTs in double rounded to float.
"""
import os
import subprocess

import numpy as np
import pytest

import sdp_synth as S

pytestmark = pytest.mark.gpu

PROGRAM = os.path.join(os.path.dirname(S.__file__), "sdp_major_cycle")
SLICES, N_MAJOR, N_MINOR, MAX_SOURCES = 4, 3, 200, 200


def oracle_fft(oracle, G, x, sign):
    return oracle.fft2d_f64(G, x.astype(np.float64), sign).astype(np.float32)


def oracle_schedule(oracle, d, cfg, measured, gains, pairs, psf, prolate):
    """The Core0.c actor sequence on the oracle: (residual, model, sources)
    per major cycle."""
    G, N = d.grid_size, d.n_vis
    per = N // SLICES
    psf_padded = np.concatenate([psf, [0.0]]).astype(np.float32)
    estimate = np.zeros(G * G, np.float32)
    out = []
    for cycle in range(N_MAJOR):
        image = estimate.copy()                                   # inverse correction: the identity
        gains_out = oracle.reciprocal_transform(gains)
        grid_in = oracle.shift_c2c(G, oracle_fft(oracle, G, oracle.shift_r2c(G, image), -1))
        corrected = oracle.apply_gains_only(measured, gains_out, pairs)
        fine_vis, fine_uvw, n_corr = oracle.correct_to_finegrid(G, d.oversampling, 0, d.uvw, corrected, cfg)
        counts = oracle.split_counts(N, SLICES, n_corr)
        grids = []
        for s in range(SLICES):
            part = slice(s * per, (s + 1) * per)
            pred = oracle.std_degridding(G, d.oversampling, d.degrid_kernels, d.supports, grid_in, fine_uvw[part],
                                         cfg, n=int(counts[s]), n_total=per)
            delta = oracle.subtract_finegrid(fine_vis[part], pred, n=int(counts[s]))
            grids.append(oracle.std_gridding(G, d.oversampling, d.kernels, d.supports, fine_uvw[part], delta, cfg,
                                             n=int(counts[s]), maj_iter=cycle))
        merged = oracle.merge(G, SLICES, np.stack(grids))
        dirty = oracle.shift_c2r(G, oracle_fft(oracle, G, oracle.shift_c2c(G, merged), +1))
        residual = oracle.convolution_correction(G, dirty, prolate)
        ns, src, model, _ = oracle.hogbom_clean(G, N_MINOR, MAX_SOURCES, residual, psf_padded, cfg, estimate)
        out.append((residual, model, src[:ns]))
        estimate = model
    return out


def write_case(path, d, cfg, measured, gains, pairs, psf):
    G = d.grid_size
    values = dict(grid_size=G, num_vis=d.n_vis, num_kernels=d.supports.shape[0],
                  total_kernel_samples=d.kernels.shape[0], num_receivers=gains.shape[0],
                  num_baselines=pairs.shape[0], slices=SLICES, num_major=N_MAJOR, num_minor=N_MINOR,
                  max_sources=MAX_SOURCES, oversampling=d.oversampling)
    with open(os.path.join(path, "case.txt"), "w") as f:
        for k, v in values.items():
            f.write(f"{k} {v}\n")
    with open(os.path.join(path, "config.bin"), "wb") as f:
        f.write(bytes(cfg))
    arrays = dict(uvw=(d.uvw, np.float32), vis=(measured, np.float32), gains=(gains, np.float32),
                  pairs=(pairs, np.int32), kernels=(d.kernels, np.float32), supports=(d.supports, np.int32),
                  dkernels=(d.degrid_kernels, np.float32), dsupports=(d.supports, np.int32),
                  psf=(psf, np.float32))
    for name, (a, dt) in arrays.items():
        np.ascontiguousarray(a, dt).tofile(os.path.join(path, name + ".bin"))


@pytest.fixture(scope="module")
def case(hip, oracle):
    d = S.config1()
    G = d.grid_size
    enu = S.disk_layout(64, 1000.0, seed=1)
    pairs = S.receiver_pairs(enu.shape[0])
    rng = np.random.default_rng(21)
    g = rng.uniform(0.8, 1.2, 64) * np.exp(1j * rng.uniform(-0.5, 0.5, 64))
    gains = np.stack([g.real, g.imag], axis=1).astype(np.float32)
    # measured = V g_a conj(g_b): what apply_gains_only's reciprocal gains undo
    b = np.arange(d.n_vis) % pairs.shape[0]
    v = d.vis[:, 0].astype(np.float64) + 1j * d.vis[:, 1]
    m = v * g[pairs[b, 0]] * np.conj(g[pairs[b, 1]])
    measured = np.stack([m.real, m.imag], axis=1).astype(np.float32)
    cfg0 = hip.make_config(G, d.cell_size, oversampling=d.oversampling, num_kernels=d.supports.shape[0],
                           max_w=d.max_w)
    prolate = np.concatenate([oracle.half_prolate(G), [0.0]]).astype(np.float32)
    # the dirty PSF (oracle chain), normalised to a unit peak
    ones = np.tile(np.array([[1.0, 0.0]], np.float32), (d.n_vis, 1))
    pgrid = oracle.std_gridding(G, d.oversampling, d.kernels, d.supports, d.uvw, ones, cfg0)
    pimg = oracle.convolution_correction(
        G, oracle.shift_c2r(G, oracle_fft(oracle, G, oracle.shift_c2c(G, pgrid), +1)), prolate)
    pmax = float(pimg.max())
    psf = (pimg.astype(np.float64) / pmax).astype(np.float32)
    cfg = hip.make_config(G, d.cell_size, oversampling=d.oversampling, num_kernels=d.supports.shape[0],
                          max_w=d.max_w, psf_max_value=pmax)
    return d, cfg, measured, gains, pairs, psf, prolate


def test_core0_schedule_program_matches_the_oracle_chain(oracle, case, tmp_path):
    """Per major cycle: the residual image fed to hogbom_clean, times the
    taper (the image before the correction divides by it, which magnifies
    rounding near the edges by up to 1/taper), within 1e-5 of the first
    cycle's peak; the component list with the same pixels in the same order
    and fluxes within 1e-5 of the largest; the model image within 1e-5 of
    its peak.  The bars are the float gridding's and FFT's rounding carried
    through three cycles (about 1e-6 of the peak per pass); the Hogbom
    choices themselves must agree."""
    d, cfg, measured, gains, pairs, psf, prolate = case
    G = d.grid_size
    write_case(str(tmp_path), d, cfg, measured, gains, pairs, psf)
    print(f"running {PROGRAM} on C1 (G {G}, {d.n_vis} vis, {SLICES} slices, {N_MAJOR} cycles)", flush=True)
    run = subprocess.run([PROGRAM, str(tmp_path)], capture_output=True, text=True, timeout=300)
    print(run.stdout, run.stderr, flush=True)
    assert run.returncode == 0, run.stderr
    assert run.stdout.count("cycle ") == N_MAJOR
    ref = oracle_schedule(oracle, d, cfg, measured, gains, pairs, psf, prolate)
    half = np.abs(np.arange(G) - G // 2)
    taper = np.outer(prolate[half], prolate[half]).ravel()
    peak0 = np.abs(ref[0][0] * taper).max()
    for cycle, (r_res, r_model, r_src) in enumerate(ref):
        res = np.fromfile(tmp_path / f"residual_{cycle}.bin", np.float32)
        model = np.fromfile(tmp_path / f"model_{cycle}.bin", np.float32)
        raw = np.fromfile(tmp_path / f"sources_{cycle}.bin", np.int32, count=1)
        src = np.fromfile(tmp_path / f"sources_{cycle}.bin", np.float32, offset=4).reshape(-1, 3)
        e_res = np.abs((res - r_res) * taper).max() / peak0
        mpeak = np.abs(r_model).max()
        e_model = np.abs(model - r_model).max() / mpeak
        print(f"cycle {cycle}: residual rms {np.sqrt(np.mean(r_res.astype(np.float64) ** 2)):.4g}, "
              f"residual {e_res:.2e} of the first peak, model {e_model:.2e} of its peak, {len(r_src)} components")
        assert raw[0] == len(r_src) == src.shape[0]
        np.testing.assert_array_equal(src[:, :2], r_src[:, :2])
        assert np.abs(src[:, 2] - r_src[:, 2]).max() <= 1e-5 * np.abs(r_src[:, 2]).max()
        assert e_res <= 1e-5
        assert e_model <= 1e-5
    rms = [np.sqrt(np.mean(r[0].astype(np.float64) ** 2)) for r in ref]
    assert rms[-1] < rms[0]
