"""The drop-in boundary: the shared library loads on a host without a GPU
and exports every function the public headers declare (no compute calls)."""
import ctypes
import os
import subprocess

import pytest

import sdp_hip


def test_library_exports_every_declared_function():
    lib = ctypes.CDLL(sdp_hip.LIB_PATH)
    missing = []
    for header in ("sdp_actors.h", "sdp_device.h", "sdp_pipeline.h"):
        names = sdp_hip.header_functions(header)
        assert len(names) > 5, header
        missing += [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_python_binding_covers_device_and_pipeline_headers():
    """Every device-layer and pipeline function has a ctypes signature."""
    lib = sdp_hip.load()
    unbound = []
    for header in ("sdp_device.h", "sdp_pipeline.h"):
        for n in sdp_hip.header_functions(header):
            if getattr(lib, n).argtypes is None:
                unbound.append(n)
    assert not unbound, unbound


def test_pipeline_argument_errors_need_no_gpu():
    """Argument checks run before any device call and report a message."""
    lib = sdp_hip.load()
    h = ctypes.c_void_p()
    rc = lib.sdp_pipeline_create(ctypes.byref(h), None, None, None, None, None, None, None, None)
    assert rc == -2 and h.value is None
    assert b"null argument" in lib.sdp_last_error()
    assert lib.sdp_pipeline_cycle(None, None, None, None) == -2
    assert lib.sdp_pipeline_image(None, 0) is None


def test_actor_prototypes_cover_reference_symbols():
    names = set(sdp_hip.header_functions("sdp_actors.h"))
    for n in ("correct_to_finegrid", "std_gridding", "std_degridding", "gridding_actor", "add_visibilities",
              "dft_actor", "hogbom_clean", "fft_shift_complex_to_complex_actor", "CUFFT_EXECUTE_FORWARD_C2C_actor",
              "CUFFT_EXECUTE_INVERSE_C2C_actor", "fft_shift_complex_to_real_actor",
              "fft_shift_real_to_complex_actor", "execute_convolution_correction_actor",
              "execute_inv_convolution_correction_actor", "subtract_from_measurements_finegrid",
              "subtract_from_measurements", "apply_gains_only", "reciprocal_transform_actor",
              "split_fftdegrid_visibilities", "split_uvgrid",
              "merge_gridded_visibilities", "correction_set_up", "psf_host_set_up", "config_struct_set_up_sequel",
              "kernel_host_set_up"):
        assert n in names, n


def test_config_layout_matches_reference():
    C = sdp_hip.Config
    assert ctypes.sizeof(C) == 360
    off = {f: getattr(C, f).offset for f, _ in C._fields_}
    assert off["cell_size"] == 24 and off["frequency_hz"] == 128 and off["oversampling"] == 136
    assert off["uv_scale"] == 144 and off["max_w"] == 168 and off["w_scale"] == 176
    assert off["loop_gain"] == 248 and off["weak_source_percent_img"] == 264
    assert off["noise_factor"] == 272 and off["psf_max_value"] == 280


def test_header_compiles_as_c(tmp_path):
    src = tmp_path / "t.c"
    src.write_text('#include "sdp_actors.h"\n#include "sdp_device.h"\n#include "sdp_pipeline.h"\n'
                   'int main(void){return (int)sizeof(Config) - 360;}\n')
    exe = tmp_path / "t"
    subprocess.run(["gcc", "-std=c99", "-Wall", "-Werror", "-I", sdp_hip.INCLUDE_DIR, str(src), "-o", str(exe)],
                   check=True)
    assert subprocess.run([str(exe)]).returncode == 0


def test_setup_actors_on_host(tmp_path):
    """Set-up actors that stay on the host (top.c): no GPU needed."""
    lib = ctypes.CDLL(sdp_hip.LIB_PATH)
    import numpy as np
    import sdp_oracle
    G = 64
    out = np.zeros(G // 2, np.float32)
    lib.correction_set_up.argtypes = [ctypes.c_int, ctypes.c_void_p]
    lib.correction_set_up(G, out.ctypes.data)
    np.testing.assert_array_equal(out, sdp_oracle.half_prolate(G))
    # psf_host_set_up: crop a 6x6 file to 4x4 and normalise by the file's max
    psf = np.arange(36, dtype=np.float64).reshape(6, 6) / 7.0
    f = tmp_path / "psf.csv"
    f.write_text(" ".join(f"{v:.9f}" for v in psf.ravel()))
    cfg = sdp_hip.make_config(4, 0.01)
    cfg.psf_input_file = str(f).encode()
    cfg.psf_max_value = 0.0
    lib.psf_host_set_up.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.POINTER(sdp_hip.Config), ctypes.c_void_p,
                                    ctypes.c_void_p]
    got = np.zeros(16, np.float32)
    mx = ctypes.c_double(0.0)
    lib.psf_host_set_up(4, 6, ctypes.byref(cfg), got.ctypes.data, ctypes.byref(mx))
    assert mx.value == pytest.approx(35 / 7.0, rel=1e-6)
    np.testing.assert_allclose(got.reshape(4, 4), (psf[1:5, 1:5] / (35 / 7.0)).astype(np.float32), rtol=1e-6)
    # clean_psf_host_set_up on the host against the oracle, bit for bit
    G = 64
    c = G // 2 + 1
    yy, xx = np.mgrid[0:G, 0:G]
    rr = np.hypot(yy - c, xx - c)
    dirty = (np.exp(-rr ** 2 / 40.0) * np.cos(rr / 2.0)).astype(np.float32)
    cfg = sdp_hip.make_config(G, 0.01, psf_max_value=1.75)
    lib.clean_psf_host_set_up.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.POINTER(sdp_hip.Config),
                                          ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
    clean = np.zeros((G, G), np.float32)
    hd = np.zeros(2, np.int32)
    lib.clean_psf_host_set_up(G, 1, ctypes.byref(cfg), dirty.ctypes.data, clean.ctypes.data, hd.ctypes.data)
    want, h = sdp_oracle.clean_psf(dirty, 1.75)
    assert hd[0] == hd[1] == h and h > 1
    np.testing.assert_array_equal(clean, want)
    # visibility_host_set_up: the count line, then "u v w re im weight" in metres
    rows = np.array([[100.0, -50.5, 3.25, 1.5, -2.0, 0.5], [-7.0, 8.0, -9.5, 0.25, 4.0, 2.0],
                     [1e3, 2e3, -3e3, -1.0, 1.0, 1.0]])
    f = tmp_path / "vis.csv"
    f.write_text("3\n" + "\n".join(" ".join(f"{x:.6f}" for x in r) for r in rows) + "\n")
    lib.visibility_host_set_up.argtypes = [ctypes.c_int, ctypes.POINTER(sdp_hip.Config), ctypes.c_void_p,
                                           ctypes.c_void_p]
    for ra, unit_weight in ((True, True), (False, False)):
        cfg = sdp_hip.make_config(64, 0.01)
        cfg.visibility_source_file = str(f).encode()
        cfg.frequency_hz = 150e6
        cfg.right_ascension = ra
        cfg.force_weight_to_one = unit_weight
        uvw = np.zeros((3, 3), np.float32)
        vis = np.zeros((3, 2), np.float32)
        lib.visibility_host_set_up(3, ctypes.byref(cfg), uvw.ctypes.data, vis.ctypes.data)
        sign = np.array([-1.0, 1.0, -1.0]) if ra else np.ones(3)
        np.testing.assert_array_equal(uvw, (rows[:, :3] * sign * (150e6 / 299792458.0)).astype(np.float32))
        wts = np.ones(3) if unit_weight else rows[:, 5]
        np.testing.assert_array_equal(vis, (rows[:, 3:5] * wts[:, None]).astype(np.float32))


def _split_lib():
    lib = ctypes.CDLL(sdp_hip.LIB_PATH)
    P, I, CP = ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(sdp_hip.Config)
    lib.split_visibilities.argtypes = [CP, I, I, P, P, P, P, P, P]
    lib.split_fftdegrid_visibilities.argtypes = [I, I, I, CP, P, P, P, P, P, P, P, P, P]
    lib.split_uvgrid.argtypes = [I, I, P, P]
    lib.merge_configs.argtypes = [P, P]
    return lib


def _config_bytes(c):
    return bytes(memoryview(c))


@pytest.mark.parametrize("n_vis,slices,n_corrected", [(1000, 4, 1000), (1003, 4, 1003), (1003, 4, 777),
                                                      (10, 3, 0), (7, 8, 7)])
def test_split_actors_on_host(n_vis, slices, n_corrected):
    """split_visibilities, split_fftdegrid_visibilities, split_uvgrid and
    merge_configs (host copies, no GPU) against a restatement of
    splitmerge.c:3-56: copies of the inputs, one Config and one major-cycle
    counter per slice, and NUM_VISIBILITIES / SLICES corrected visibilities
    per slice while they last (a short last slice, then empty ones)."""
    import numpy as np
    lib = _split_lib()
    rng = np.random.default_rng(n_vis + slices)
    vis = rng.normal(size=(n_vis, 2)).astype(np.float32)
    uvw = rng.normal(size=(n_vis, 3)).astype(np.float32)
    cfg = sdp_hip.make_config(32, 0.01, num_kernels=3)
    cfg.num_visibilities = n_vis
    # split_visibilities: splitmerge.c:3-12
    out_vis, out_uvw = np.zeros_like(vis), np.zeros_like(uvw)
    cfgs = (sdp_hip.Config * slices)()
    one = sdp_hip.Config()
    lib.split_visibilities(ctypes.byref(cfg), n_vis, slices, vis.ctypes.data, uvw.ctypes.data, out_vis.ctypes.data,
                           out_uvw.ctypes.data, ctypes.addressof(cfgs), ctypes.byref(one))
    np.testing.assert_array_equal(out_vis, vis)
    np.testing.assert_array_equal(out_uvw, uvw)
    assert all(_config_bytes(c) == _config_bytes(cfg) for c in cfgs) and _config_bytes(one) == _config_bytes(cfg)
    # split_fftdegrid_visibilities: splitmerge.c:30-48
    out_vis, out_uvw = np.zeros_like(vis), np.zeros_like(uvw)
    cfgs = (sdp_hip.Config * slices)()
    num_vis = np.full(slices, -1, np.int32)
    maj_out = np.full(slices, -1, np.int32)
    lib.split_fftdegrid_visibilities(n_vis, slices, 32, ctypes.byref(cfg), uvw.ctypes.data, vis.ctypes.data,
                                     ctypes.byref(ctypes.c_int(n_corrected)), ctypes.byref(ctypes.c_int(3)),
                                     out_vis.ctypes.data, out_uvw.ctypes.data, ctypes.addressof(cfgs),
                                     num_vis.ctypes.data, maj_out.ctypes.data)
    np.testing.assert_array_equal(out_vis, vis)
    np.testing.assert_array_equal(out_uvw, uvw)
    assert all(_config_bytes(c) == _config_bytes(cfg) for c in cfgs)
    assert list(maj_out) == [3] * slices
    ideal, left, want = n_vis // slices, n_corrected, []
    for _ in range(slices):
        cur = min(ideal, left)
        want.append(cur)
        left -= cur
    assert list(num_vis) == want
    assert sum(want) == min(n_corrected, ideal * slices)
    # split_uvgrid: splitmerge.c:50-56, one copy of the grid per slice
    G = 16
    grid = rng.normal(size=(G * G, 2)).astype(np.float32)
    grids = np.zeros((slices, G * G, 2), np.float32)
    lib.split_uvgrid(G, slices, grid.ctypes.data, grids.ctypes.data)
    for s in range(slices):
        np.testing.assert_array_equal(grids[s], grid)
    # merge_configs: splitmerge.c:26-28, the first slice's Config
    cfgs = (sdp_hip.Config * 2)(sdp_hip.make_config(64, 0.5), sdp_hip.make_config(32, 0.25))
    out = sdp_hip.Config()
    lib.merge_configs(ctypes.addressof(cfgs), ctypes.byref(out))
    assert _config_bytes(out) == _config_bytes(cfgs[0]) and out.grid_size == 64


def test_kernel_csv_loaders(tmp_path):
    """kernel_host_set_up / degridding_kernel_host_set_up read the files the
    Config names (top.c:235-379): one half support per line, plane p's
    samples at offset sum_{q<p} ((h_q + 1) * OS)^2, the real and imaginary
    files holding whitespace-separated samples plane after plane, each
    value parsed as a double and stored as float."""
    import numpy as np
    lib = ctypes.CDLL(sdp_hip.LIB_PATH)
    CP = ctypes.POINTER(sdp_hip.Config)
    OS = 4
    hs = [2, 3, 1]
    counts = [((h + 1) * OS) ** 2 for h in hs]          # 144, 256, 64
    total = sum(counts)
    rng = np.random.default_rng(5)
    re = rng.normal(size=total) * 10.0 ** rng.integers(-6, 3, total)
    im = rng.normal(size=total)
    for name, which in (("kernel_host_set_up", ""), ("degridding_kernel_host_set_up", "degridding_")):
        d = tmp_path / name
        d.mkdir()
        (d / "sup.csv").write_text("".join(f"{h}\n" for h in hs))
        # the real file one sample per line, the imaginary one space-separated rows of 16
        (d / "re.csv").write_text("\n".join(repr(float(v)) for v in re) + "\n")
        rows = [" ".join(f"{v:.12e}" for v in im[i:i + 16]) for i in range(0, total, 16)]
        (d / "im.csv").write_text("\n".join(rows) + "\n")
        cfg = sdp_hip.make_config(64, 0.01, oversampling=OS, num_kernels=len(hs))
        setattr(cfg, which + "kernel_support_file", str(d / "sup.csv").encode())
        setattr(cfg, which + "kernel_real_file", str(d / "re.csv").encode())
        setattr(cfg, which + "kernel_imag_file", str(d / "im.csv").encode())
        sup = np.full((len(hs), 2), -1, np.int32)
        ker = np.full((total + 8, 2), np.nan, np.float32)
        f = getattr(lib, name)
        f.argtypes = [ctypes.c_int, ctypes.c_int, CP, ctypes.c_void_p, ctypes.c_void_p]
        f(len(hs), total, ctypes.byref(cfg), sup.ctypes.data, ker.ctypes.data)
        np.testing.assert_array_equal(sup[:, 0], hs)
        np.testing.assert_array_equal(sup[:, 1], [0, counts[0], counts[0] + counts[1]])
        np.testing.assert_array_equal(ker[:total, 0], re.astype(np.float32))
        np.testing.assert_array_equal(ker[:total, 1], np.array([float(f"{v:.12e}") for v in im], np.float32))
        assert np.isnan(ker[total:]).all()                 # nothing written past the last plane


def test_schedule_program_checks_its_case_directory(tmp_path):
    """sdp_major_cycle (the generated single-core schedule as a C program
    over libsdp_actors.so) is built, links, and rejects a missing argument,
    a missing case file and an incomplete case before any actor runs."""
    prog = os.path.join(os.path.dirname(sdp_hip.LIB_PATH), "sdp_major_cycle")
    run = subprocess.run([prog], capture_output=True, text=True, timeout=60)
    assert run.returncode == 2 and "usage" in run.stderr
    run = subprocess.run([prog, str(tmp_path)], capture_output=True, text=True, timeout=60)
    assert run.returncode == 2 and "case.txt" in run.stderr
    (tmp_path / "case.txt").write_text("grid_size 64\nnum_vis 10\n")
    run = subprocess.run([prog, str(tmp_path)], capture_output=True, text=True, timeout=60)
    assert run.returncode == 2 and "does not set every key" in run.stderr
    keys = dict(grid_size=64, num_vis=10, num_kernels=1, total_kernel_samples=16, num_receivers=5, num_baselines=10,
                slices=4, num_major=1, num_minor=1, max_sources=1, oversampling=16)
    (tmp_path / "case.txt").write_text("".join(f"{k} {v}\n" for k, v in keys.items()))
    run = subprocess.run([prog, str(tmp_path)], capture_output=True, text=True, timeout=60)
    assert run.returncode == 2 and "multiple of slices" in run.stderr
    keys["slices"] = 5
    (tmp_path / "case.txt").write_text("".join(f"{k} {v}\n" for k, v in keys.items()))
    (tmp_path / "config.bin").write_bytes(bytes(100))
    run = subprocess.run([prog, str(tmp_path)], capture_output=True, text=True, timeout=60)
    assert run.returncode == 2 and "fewer than the 360 bytes" in run.stderr
