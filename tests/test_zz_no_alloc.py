"""The allocation-free entry point, sdp_dft_predict_array_noalloc, through
the C-ABI: its output equals sdp_dft_predict_array's bit for bit on two
calls in a row (the second smaller, so the scratch array still holds the
first call's terms), for both operand types, and it rejects more sources
than its scratch array holds."""
import ctypes

import numpy as np
import pytest

import sdp_synth as S

pytestmark = pytest.mark.gpu


def array_case(n_ant, n_time, n_src, seed):
    rng = np.random.default_rng(seed)
    enu = S.log_radial_layout(n_ant, 20.0, 3000.0, seed=seed)
    ant = S.antenna_uvw(enu, -30.0, -30.0, np.linspace(-3.0, 3.0, n_time), 150e6)
    src = np.concatenate([rng.uniform(-0.03, 0.03, (n_src, 2)), rng.uniform(-0.2, 1.0, (n_src, 1))],
                         axis=1).astype(np.float32)
    return np.ascontiguousarray(ant, np.float64), src


def test_dft_noalloc_matches_the_allocating_entry_point(hip):
    lib = hip.load()
    for flags in (hip.DFT_ARRAY_BF16, hip.DFT_ARRAY_FP16):
        for n_ant, n_time, n_src, seed in ((150, 4, 300, 1), (40, 2, 17, 2)):
            ant, src = array_case(n_ant, n_time, n_src, seed)
            n_vis = n_time * n_ant * (n_ant - 1) // 2
            d_ant, d_src = hip.DeviceBuffer.from_array(ant), hip.DeviceBuffer.from_array(src)
            d_ref, d_out = hip.DeviceBuffer(8 * n_vis), hip.DeviceBuffer(8 * n_vis)
            d_out.upload(np.full((n_vis, 2), np.nan, np.float32))      # every output must be written
            hip.dft_predict_array(n_time, n_ant, d_ant, n_src, d_src, d_ref, flags)
            hip.check(lib.sdp_dft_predict_array_noalloc(n_time, n_ant, d_ant.ptr, n_src, d_src.ptr, d_out.ptr,
                                                         flags, None))
            hip.synchronize()
            ref = d_ref.download((n_vis, 2), np.float32)
            got = d_out.download((n_vis, 2), np.float32)
            assert np.isfinite(ref).all()
            np.testing.assert_array_equal(got, ref)


def test_dft_noalloc_rejects_more_sources_than_its_scratch_holds(hip):
    lib = hip.load()
    ant, src = array_case(20, 1, 4, 3)
    d_ant, d_src, d_out = hip.DeviceBuffer.from_array(ant), hip.DeviceBuffer.from_array(src), hip.DeviceBuffer(8 * 190)
    rc = lib.sdp_dft_predict_array_noalloc(1, 20, d_ant.ptr, 65537, d_src.ptr, d_out.ptr, 0, None)
    assert rc != 0
    assert b"at most 65536" in lib.sdp_last_error()
