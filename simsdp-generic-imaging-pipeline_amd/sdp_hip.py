"""ctypes binding of libsdp_actors.so (the MI355X imaging library).

Two layers, mirroring include/sdp_actors.h and include/sdp_device.h:

* ``Actors``: the reference actor ABI (host numpy arrays in, host arrays
  out), the same names and argument meaning as the C actors in
  preesm_pipelines/simsdp_fft_imaging_pipeline/Code/include/*.h, so tests read
  like calls made by the PREESM schedule.
* ``Device`` / ``DeviceBuffer``: device-resident entry points for callers
  that keep data in HBM (bench, major-cycle driver).

The library must be built (``__graft_entry__.build()``); importing this module
never falls back to any other implementation.  Failing C calls raise
``SdpError`` with the library's message.
"""
from __future__ import annotations

import ctypes
import os
import re

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libsdp_actors.so")
REPO = os.path.dirname(HERE)
INCLUDE_DIR = os.path.join(REPO, "include")

c_int, c_double, c_float, c_size_t = ctypes.c_int, ctypes.c_double, ctypes.c_float, ctypes.c_size_t
c_void_p, c_char_p = ctypes.c_void_p, ctypes.c_char_p

# device-layer enums (sdp_device.h)
DEGRID_NEAREST, DEGRID_WEIGHTED = 0, 1
DFT_REFERENCE_FP32, DFT_FP64_PHASE = 0, 1
DFT_ARRAY_BF16, DFT_ARRAY_FP16, DFT_ARRAY_NO_SPLIT = 0, 1, 4
SNAP_STD, SNAP_GRIDDING_ACTOR = 0, 1
ORDER_ORIGINAL, ORDER_PLAN = 0, 1
ACC_BLOCKED, ACC_COMPENSATED = 0, 1
FFT_FORWARD, FFT_INVERSE = -1, 1
FFT_SHIFT_IN, FFT_SHIFT_OUT, FFT_REAL_IN, FFT_REAL_OUT, FFT_CORRECT_OUT, FFT_CORRECT_IN = 1, 2, 4, 8, 16, 32
HOGBOM_PSF_AS_STORED = 1
HOGBOM_DOUBLE_MEAN = 2
HOGBOM_PER_ITERATION = 4
HOGBOM_TEST_ABORT = 8
# major-cycle pipeline (sdp_pipeline.h)
PIPE_GRAPH, PIPE_REFERENCE_PREDICT, PIPE_PSF_MIRRORED, PIPE_PROFILE = 1, 2, 4, 8
STAGES = ("predict_fft", "degrid_subtract_grid", "grid_merge", "inverse_fft", "clean")
CNN_FP32, CNN_BF16 = 0, 1
LEARNED_TILE_SCALE, LEARNED_GLOBAL_SCALE = 0, 1
IMAGE_RESIDUAL, IMAGE_MODEL, IMAGE_PSF = 0, 1, 2


class SdpError(RuntimeError):
    pass


class Config(ctypes.Structure):
    """The reference ``Config`` record (common.h:187-265), 360 bytes."""

    _fields_ = [
        ("retain_device_mem", ctypes.c_bool), ("num_major_cycles", c_int), ("num_recievers", c_int),
        ("num_baselines", c_int), ("grid_size", c_int), ("cell_size", c_double), ("output_path", c_char_p),
        ("dirty_image_output", c_char_p), ("gpu_max_threads_per_block", c_int),
        ("gpu_max_threads_per_block_dimension", c_int), ("right_ascension", ctypes.c_bool),
        ("save_dirty_image", ctypes.c_bool), ("save_residual_image", ctypes.c_bool),
        ("save_extracted_sources", ctypes.c_bool), ("save_predicted_visibilities", ctypes.c_bool),
        ("save_estimated_gains", ctypes.c_bool), ("perform_system_test", ctypes.c_bool),
        ("system_test_image", c_char_p), ("system_test_sources", c_char_p),
        ("system_test_visibilities", c_char_p), ("system_test_gains", c_char_p),
        ("default_gains_file", c_char_p), ("output_gains_file", c_char_p),
        ("perform_gain_calibration", ctypes.c_bool), ("use_default_gains", ctypes.c_bool),
        ("max_calibration_cycles", c_int), ("number_cal_major_cycles", c_int),
        ("force_weight_to_one", ctypes.c_bool), ("frequency_hz", c_double), ("oversampling", c_int),
        ("uv_scale", c_double), ("num_visibilities", c_int), ("num_kernels", c_int),
        ("total_kernel_samples", c_int), ("max_w", c_double), ("w_scale", c_double),
        ("kernel_real_file", c_char_p), ("kernel_imag_file", c_char_p), ("kernel_support_file", c_char_p),
        ("degridding_kernel_real_file", c_char_p), ("degridding_kernel_imag_file", c_char_p),
        ("degridding_kernel_support_file", c_char_p), ("visibility_source_file", c_char_p),
        ("number_minor_cycles_cal", ctypes.c_uint), ("number_minor_cycles_img", ctypes.c_uint),
        ("loop_gain", c_double), ("weak_source_percent_gc", c_double), ("weak_source_percent_img", c_double),
        ("noise_factor", c_double), ("psf_max_value", c_double), ("model_sources_output", c_char_p),
        ("residual_image_output", c_char_p), ("psf_input_file", c_char_p), ("psf_image_output", c_char_p),
        ("clean_psf_image_output", c_char_p), ("model_image_output", c_char_p),
        ("final_image_output", c_char_p), ("num_sources", c_int), ("predicted_vis_output", c_char_p),
    ]


assert ctypes.sizeof(Config) == 360


def make_config(grid_size: int, cell_size: float, oversampling: int = 16, num_kernels: int = 1,
                max_w: float = 1895.410847844, loop_gain: float = 0.1, weak_source_percent_img: float = 2e-4,
                noise_factor: float = 1.5, psf_max_value: float = 1.0) -> Config:
    """A Config filled like config_struct_set_up (top.c:9-69) for a given
    grid, cell size and w-plane count."""
    c = Config()
    c.grid_size = grid_size
    c.cell_size = cell_size
    c.frequency_hz = 299792458.0
    c.oversampling = oversampling
    c.uv_scale = cell_size * grid_size
    c.num_kernels = num_kernels
    c.max_w = max_w
    c.w_scale = float((num_kernels - 1) ** 2) / max_w
    c.loop_gain = loop_gain
    c.weak_source_percent_gc = 0.005
    c.weak_source_percent_img = weak_source_percent_img
    c.noise_factor = noise_factor
    c.psf_max_value = psf_max_value
    c.use_default_gains = True
    c.force_weight_to_one = True
    return c


_lib = None


def header_functions(header: str) -> list[str]:
    """Names of the functions a public header declares."""
    text = open(os.path.join(INCLUDE_DIR, header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?[A-Za-z_][A-Za-z0-9_]*\s*\**\s*([A-Za-z_][A-Za-z0-9_]*)\s*\(", text,
                       flags=re.M)
    return [n for n in names if n not in ("if", "while", "for", "sizeof", "return")]


def load() -> ctypes.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise SdpError(f"{LIB_PATH} is not built; run __graft_entry__.build()")
    lib = ctypes.CDLL(LIB_PATH, mode=ctypes.RTLD_GLOBAL)
    P = c_void_p
    sig = {
        "sdp_init": (c_int, []), "sdp_last_error": (c_char_p, []),
        "sdp_device_malloc": (c_int, [ctypes.POINTER(c_void_p), c_size_t]), "sdp_device_free": (c_int, [P]),
        "sdp_memcpy_htod": (c_int, [P, P, c_size_t, P]), "sdp_memcpy_dtoh": (c_int, [P, P, c_size_t, P]),
        "sdp_memcpy_dtod": (c_int, [P, P, c_size_t, P]), "sdp_memset_zero": (c_int, [P, c_size_t, P]),
        "sdp_device_copy": (c_int, [P, P, c_size_t, P]),
        "sdp_stream_create": (c_int, [ctypes.POINTER(c_void_p)]), "sdp_stream_destroy": (c_int, [P]),
        "sdp_stream_sync": (c_int, [P]), "sdp_device_sync": (c_int, []),
        "sdp_event_create": (c_int, [ctypes.POINTER(c_void_p)]), "sdp_event_destroy": (c_int, [P]),
        "sdp_event_record": (c_int, [P, P]), "sdp_event_sync": (c_int, [P]),
        "sdp_event_elapsed_ms": (c_int, [P, P, ctypes.POINTER(c_float)]),
        "sdp_set_degrid_mode": (None, [c_int]), "sdp_get_degrid_mode": (c_int, []),
        "sdp_set_dft_mode": (None, [c_int]), "sdp_get_dft_mode": (c_int, []),
        "sdp_finegrid_correct": (c_int, [c_int, c_int, c_int, c_double, P, P, P]),
        "sdp_finegrid_simplify": (c_int, [c_int, c_int, c_int, c_double, c_double, P, P, P, P,
                                          ctypes.POINTER(c_int), P]),
        "sdp_grid_plan_create": (c_int, [ctypes.POINTER(c_void_p), P, P, P, c_int, P]),
        "sdp_grid_plan_create_multi": (c_int, [ctypes.POINTER(c_void_p), P, P, c_int, P, P, P]),
        "sdp_grid_plan_destroy": (c_int, [P]), "sdp_grid_plan_info": (c_int, [P, P]),
        "sdp_grid_plan_to_plan_order": (c_int, [P, P, P, P]),
        "sdp_grid_plan_to_original_order": (c_int, [P, P, P, P]),
        "sdp_grid": (c_int, [P, P, P, c_int, P, P]),
        "sdp_grid_accumulate": (c_int, [P, P, P, c_int, P]),
        "sdp_grid_merge": (c_int, [P, P, P]),
        "sdp_degrid": (c_int, [P, P, P, P, c_int, P, P]),
        "sdp_degrid_grid": (c_int, [P, P, P, P, P, c_int, P, P, P]),
        "sdp_degrid_grid_accumulate": (c_int, [P, P, P, P, P, c_int, P, P]),
        "sdp_grid_autotune": (c_int, [P, P, P, c_int, P, P, P, c_int, P, c_int, ctypes.POINTER(c_int),
                                      ctypes.POINTER(c_int), ctypes.POINTER(c_float), P]),
        "sdp_degrid_nearest": (c_int, [c_int, c_int, c_int, c_double, P, P, P, P]),
        "sdp_fft_work_bytes": (c_size_t, [c_int]),
        "sdp_fft2d": (c_int, [P, P, c_int, c_int, c_int, P, P, P]),
        "sdp_fft2d_batch": (c_int, [P, P, c_int, c_int, c_int, c_int, P, P, P]),
        "sdp_shift_c2c": (c_int, [c_int, P, P, P]), "sdp_shift_c2r": (c_int, [c_int, P, P, P]),
        "sdp_shift_r2c": (c_int, [c_int, P, P, P]),
        "sdp_convolution_correct": (c_int, [c_int, P, P, P, P]),
        "sdp_vis_subtract": (c_int, [c_int, P, P, P, P]), "sdp_vis_add": (c_int, [c_int, P, P, P, P]),
        "sdp_vis_subtract_gains": (c_int, [c_int, c_int, P, P, P, P, c_int, P, P]),
        "sdp_gains_reciprocal": (c_int, [c_int, P, P, P]),
        "sdp_restore_conv2d": (c_int, [c_int, P, P, c_int, c_int, P, P]),
        "sdp_merge_grids": (c_int, [c_int, c_int, P, P, P]),
        "sdp_hogbom_work_bytes": (c_size_t, [c_int]),
        "sdp_hogbom": (c_int, [c_int, c_int, c_int, P, P, c_double, c_double, c_double, c_double, c_double,
                               P, P, P, P, P, P]),
        "sdp_hogbom_ex": (c_int, [c_int, c_int, c_int, P, P, c_double, c_double, c_double, c_double, c_double,
                                  P, P, P, P, P, c_int, P]),
        "sdp_dft_predict": (c_int, [c_int, c_int, P, P, P, c_int, P]),
        "sdp_set_last_error": (c_int, [c_int, c_char_p]),
        "sdp_image_stats": (c_int, [c_size_t, P, ctypes.POINTER(c_double), ctypes.POINTER(c_float),
                                    ctypes.POINTER(c_float), P]),
        "sdp_image_stats_work_bytes": (c_size_t, []),
        "sdp_image_stats_async": (c_int, [c_size_t, P, P, P, P]),
        "sdp_fill_complex": (c_int, [c_size_t, P, c_float, c_float, P]),
        "sdp_divide": (c_int, [c_size_t, P, c_double, P]),
        "sdp_graph_begin_capture": (c_int, [P]),
        "sdp_graph_end_capture": (c_int, [P, ctypes.POINTER(c_void_p)]),
        "sdp_graph_launch": (c_int, [P, P]), "sdp_graph_destroy": (c_int, [P]),
        "sdp_pipeline_create": (c_int, [ctypes.POINTER(c_void_p), P, P, P, P, P, P, P, P]),
        "sdp_pipeline_destroy": (c_int, [P]), "sdp_pipeline_reset": (c_int, [P]),
        "sdp_pipeline_load_model": (c_int, [P, P]),
        "sdp_pipeline_cycle": (c_int, [P, ctypes.POINTER(c_double), ctypes.POINTER(c_double),
                                       ctypes.POINTER(c_float)]),
        "sdp_pipeline_image": (c_void_p, [P, c_int]),
        "sdp_pipeline_psf_max": (c_double, [P]),
        "sdp_pipeline_stage_ms": (c_int, [P, P]),
        "sdp_pipeline_sources": (c_int, [P, P, ctypes.POINTER(c_int)]),
        "sdp_pipeline_plan": (c_void_p, [P]),
        "sdp_batch_create": (c_int, [ctypes.POINTER(c_void_p), P, P, P, P, P, P, P, P, P]),
        "sdp_batch_destroy": (c_int, [P]),
        "sdp_batch_residuals": (c_int, [P, P, P]),
        "sdp_batch_model": (c_void_p, [P, c_int]),
        "sdp_batch_residual": (c_void_p, [P, c_int]),
        "sdp_batch_plan": (c_void_p, [P, c_int]),
        "sdp_cnn_create": (c_int, [ctypes.POINTER(c_void_p), P]), "sdp_cnn_destroy": (c_int, [P]),
        "sdp_cnn_param_count": (c_size_t, [P]),
        "sdp_cnn_set_params": (c_int, [P, P]), "sdp_cnn_get_params": (c_int, [P, P]),
        "sdp_cnn_get_grads": (c_int, [P, P]),
        "sdp_cnn_forward": (c_int, [P, P, P, c_int, P]), "sdp_cnn_backward": (c_int, [P, P, P]),
        "sdp_cnn_mse": (c_int, [c_size_t, P, P, P, ctypes.POINTER(c_double), P]),
        "sdp_cnn_adam": (c_int, [P, c_float, c_float, c_float, c_float, P]),
        "sdp_cnn_describe": (c_int, [P, P]),
        "sdp_dft_array_work_bytes": (c_size_t, [c_int]),
        "sdp_dft_predict_array": (c_int, [c_int, c_int, P, c_int, P, P, c_int, P, P]),
        "sdp_dft_predict_array_noalloc": (c_int, [c_int, c_int, P, c_int, P, P, c_int, P]),
        "sdp_cnn_patch_peaks": (c_int, [c_int, c_int, c_int, P, P, ctypes.POINTER(c_int), P]),
        "sdp_cnn_gather_patches": (c_int, [c_int, c_int, c_int, P, P, P, P, P, P]),
        "sdp_cnn_scatter_add": (c_int, [c_int, c_int, c_int, c_int, P, P, P, c_float, c_float, P, c_float, P, P]),
        "sdp_cnn_crop_patches": (c_int, [c_int, c_int, c_int, P, P, P, P, P]),
        "sdp_sky_render": (c_int, [c_int, c_int, P, c_int, P, P, P]),
        "sdp_complex_multiply": (c_int, [c_size_t, P, P, c_size_t, P, P]),
        "sdp_axpby": (c_int, [c_size_t, c_float, P, c_float, P, P, P]),
        "sdp_pipeline_set_learned": (c_int, [P, P, P]),
        "sdp_pipeline_learned_active": (c_int, [P]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _lib = lib
    return lib


def check(rc: int) -> None:
    if rc != 0:
        raise SdpError(load().sdp_last_error().decode(errors="replace") or f"error {rc}")


def _ptr(a):
    if a is None:
        return None
    if isinstance(a, DeviceBuffer):
        return a.ptr
    if isinstance(a, int):
        return a
    if not (a.flags["C_CONTIGUOUS"]):
        raise SdpError("host arrays must be C-contiguous")
    return a.ctypes.data


# ---------------------------------------------------------------- device layer

class DeviceBuffer:
    """A device allocation owned by Python (freed on close / garbage collection)."""

    def __init__(self, nbytes: int):
        lib = load()
        check(lib.sdp_init())
        p = c_void_p()
        check(lib.sdp_device_malloc(ctypes.byref(p), c_size_t(max(int(nbytes), 16))))
        self.ptr = p.value
        self.nbytes = int(nbytes)

    @classmethod
    def from_array(cls, a: np.ndarray, stream=None) -> "DeviceBuffer":
        a = np.ascontiguousarray(a)
        b = cls(a.nbytes)
        b.upload(a, stream)
        return b

    def upload(self, a: np.ndarray, stream=None) -> None:
        a = np.ascontiguousarray(a)
        if a.nbytes > self.nbytes:
            raise SdpError("upload larger than the buffer")
        check(load().sdp_memcpy_htod(self.ptr, a.ctypes.data, a.nbytes, stream))

    def download(self, shape, dtype, stream=None) -> np.ndarray:
        out = np.empty(shape, dtype=dtype)
        if out.nbytes > self.nbytes:
            raise SdpError("download larger than the buffer")
        check(load().sdp_memcpy_dtoh(out.ctypes.data, self.ptr, out.nbytes, stream))
        return out

    def zero(self, stream=None) -> None:
        check(load().sdp_memset_zero(self.ptr, self.nbytes, stream))

    def close(self) -> None:
        if getattr(self, "ptr", None):
            load().sdp_device_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Event:
    def __init__(self):
        e = c_void_p()
        check(load().sdp_event_create(ctypes.byref(e)))
        self.ptr = e.value

    def record(self, stream=None):
        check(load().sdp_event_record(self.ptr, stream))

    def sync(self):
        check(load().sdp_event_sync(self.ptr))

    def elapsed_ms(self, later: "Event") -> float:
        ms = c_float()
        check(load().sdp_event_elapsed_ms(self.ptr, later.ptr, ctypes.byref(ms)))
        return float(ms.value)

    def __del__(self):
        try:
            load().sdp_event_destroy(self.ptr)
        except Exception:
            pass


class _Params(ctypes.Structure):
    _fields_ = [("grid_size", c_int), ("oversampling", c_int), ("n_kernels", c_int), ("snap_mode", c_int),
                ("uv_scale", c_double), ("w_scale", c_double), ("tile", c_int), ("chunk", c_int),
                ("accumulate", c_int)]


def autotune(grid_size, oversampling, supports, uv_scale, w_scale, d_uvw, n_vis, d_grid_kernels, d_degrid_kernels,
             tiles=(4, 8, 16), chunks=(16384, 32768, 65536, 131072), snap_mode=SNAP_STD, accumulate=ACC_BLOCKED,
             stream=None):
    """Fastest (tile, taps per work item) for these coordinates: (tile, chunk, ms)."""
    sup = np.ascontiguousarray(np.asarray(supports, dtype=np.int32).reshape(-1, 2))
    p = _Params(grid_size, oversampling, sup.shape[0], snap_mode, uv_scale, w_scale, 0, 0, accumulate)
    t = np.asarray(tiles, np.int32)
    c = np.asarray(chunks, np.int32)
    bt, bc, ms = c_int(0), c_int(0), c_float(0)
    check(load().sdp_grid_autotune(ctypes.byref(p), sup.ctypes.data, _ptr(d_uvw), int(n_vis), _ptr(d_grid_kernels),
                                   _ptr(d_degrid_kernels), t.ctypes.data, t.size, c.ctypes.data, c.size,
                                   ctypes.byref(bt), ctypes.byref(bc), ctypes.byref(ms), stream))
    return bt.value, bc.value, ms.value


class GridPlan:
    """Owner of an ``sdp_grid_plan`` (tile-sorted visibility records)."""

    def __init__(self, grid_size, oversampling, supports, uv_scale, w_scale, d_uvw, n_vis,
                 snap_mode=SNAP_STD, tile=0, chunk=0, accumulate=ACC_BLOCKED, stream=None):
        lib = load()
        check(lib.sdp_init())
        sup = np.ascontiguousarray(np.asarray(supports, dtype=np.int32).reshape(-1, 2))
        p = _Params(grid_size, oversampling, sup.shape[0], snap_mode, uv_scale, w_scale, tile, chunk, accumulate)
        h = c_void_p()
        check(lib.sdp_grid_plan_create(ctypes.byref(h), ctypes.byref(p), sup.ctypes.data, _ptr(d_uvw),
                                       int(n_vis), stream))
        self.ptr = h.value
        self.grid_size = grid_size
        self.n_vis = int(n_vis)
        self.n_grids = 1

    @classmethod
    def multi(cls, grid_size, oversampling, supports, uv_scale, w_scale, d_uvws, n_vis, snap_mode=SNAP_STD,
              tile=0, chunk=0, accumulate=ACC_BLOCKED, stream=None):
        """One plan over len(d_uvws) grids (sdp_grid_plan_create_multi):
        visibilities in grid order, grid arguments hold that many grids."""
        lib = load()
        check(lib.sdp_init())
        sup = np.ascontiguousarray(np.asarray(supports, dtype=np.int32).reshape(-1, 2))
        p = _Params(grid_size, oversampling, sup.shape[0], snap_mode, uv_scale, w_scale, tile, chunk, accumulate)
        ng = len(d_uvws)
        ptrs = (c_void_p * ng)(*[_ptr(d) for d in d_uvws])
        counts = np.ascontiguousarray(np.asarray(n_vis, dtype=np.int32))
        assert counts.size == ng
        h = c_void_p()
        check(lib.sdp_grid_plan_create_multi(ctypes.byref(h), ctypes.byref(p), sup.ctypes.data, ng,
                                             ctypes.cast(ptrs, c_void_p), counts.ctypes.data, stream))
        self = cls.__new__(cls)
        self.ptr = h.value
        self.grid_size = grid_size
        self.n_vis = int(counts.sum())
        self.n_grids = ng
        return self

    def info(self) -> dict:
        a = np.zeros(10, dtype=np.int64)
        check(load().sdp_grid_plan_info(self.ptr, a.ctypes.data))
        keys = ["kept", "dropped", "tiles", "chunks", "tile", "chunk_taps", "max_half_support", "scratch_bytes",
                "class_launches", "taps"]
        return dict(zip(keys, (int(x) for x in a)))

    def class_launches(self) -> int:
        return self.info()["class_launches"]

    def total_taps(self) -> int:
        return self.info()["taps"]

    def grid(self, d_kernels, d_vis, d_grid, order=ORDER_ORIGINAL, stream=None):
        check(load().sdp_grid(self.ptr, _ptr(d_kernels), _ptr(d_vis), order, _ptr(d_grid), stream))

    def accumulate(self, d_kernels, d_vis, order=ORDER_ORIGINAL, stream=None):
        check(load().sdp_grid_accumulate(self.ptr, _ptr(d_kernels), _ptr(d_vis), order, stream))

    def merge(self, d_grid, stream=None):
        check(load().sdp_grid_merge(self.ptr, _ptr(d_grid), stream))

    def degrid(self, d_kernels, d_grid, d_vis_out, d_measured=None, order=ORDER_ORIGINAL, stream=None):
        check(load().sdp_degrid(self.ptr, _ptr(d_kernels), _ptr(d_grid), _ptr(d_measured), order,
                                _ptr(d_vis_out), stream))

    def degrid_grid(self, d_degrid_kernels, d_grid_kernels, d_model, d_measured, d_residual_grid,
                    d_residual_vis=None, order=ORDER_ORIGINAL, stream=None):
        check(load().sdp_degrid_grid(self.ptr, _ptr(d_degrid_kernels), _ptr(d_grid_kernels), _ptr(d_model),
                                     _ptr(d_measured), order, _ptr(d_residual_vis), _ptr(d_residual_grid), stream))

    def degrid_grid_accumulate(self, d_degrid_kernels, d_grid_kernels, d_model, d_measured, d_residual_vis=None,
                               order=ORDER_ORIGINAL, stream=None):
        check(load().sdp_degrid_grid_accumulate(self.ptr, _ptr(d_degrid_kernels), _ptr(d_grid_kernels),
                                                _ptr(d_model), _ptr(d_measured), order, _ptr(d_residual_vis),
                                                stream))

    def to_plan_order(self, d_in, d_out, stream=None):
        check(load().sdp_grid_plan_to_plan_order(self.ptr, _ptr(d_in), _ptr(d_out), stream))

    def to_original_order(self, d_in, d_out, stream=None):
        check(load().sdp_grid_plan_to_original_order(self.ptr, _ptr(d_in), _ptr(d_out), stream))

    def close(self):
        if getattr(self, "ptr", None):
            load().sdp_grid_plan_destroy(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def synchronize(stream=None):
    check(load().sdp_stream_sync(stream))


def fft2d(d_in, d_out, grid_size, direction, flags=0, d_taper=None, d_work=None, stream=None):
    check(load().sdp_fft2d(_ptr(d_in), _ptr(d_out), grid_size, direction, flags, _ptr(d_taper), _ptr(d_work),
                           stream))


def fft2d_batch(d_in, d_out, grid_size, n_batch, direction, flags=0, d_taper=None, d_work=None, stream=None):
    check(load().sdp_fft2d_batch(_ptr(d_in), _ptr(d_out), grid_size, n_batch, direction, flags, _ptr(d_taper),
                                 _ptr(d_work), stream))


def fft_work_bytes(grid_size: int) -> int:
    return int(load().sdp_fft_work_bytes(grid_size))


def hogbom_work_bytes(grid_size: int) -> int:
    return int(load().sdp_hogbom_work_bytes(grid_size))


def hogbom(grid_size, n_minor, max_sources, d_residual, d_psf, psf_max, loop_gain, noise_factor, weak,
           cell_size, d_model_in, d_model_out, d_sources, d_num_sources, d_work, stream=None):
    check(load().sdp_hogbom(grid_size, n_minor, max_sources, _ptr(d_residual), _ptr(d_psf), psf_max, loop_gain,
                            noise_factor, weak, cell_size, _ptr(d_model_in), _ptr(d_model_out), _ptr(d_sources),
                            _ptr(d_num_sources), _ptr(d_work), stream))


def hogbom_ex(grid_size, n_minor, max_sources, d_residual, d_psf, psf_max, loop_gain, noise_factor, weak,
              cell_size, d_model_in, d_model_out, d_sources, d_num_sources, d_work, flags=0, stream=None):
    check(load().sdp_hogbom_ex(grid_size, n_minor, max_sources, _ptr(d_residual), _ptr(d_psf), psf_max, loop_gain,
                               noise_factor, weak, cell_size, _ptr(d_model_in), _ptr(d_model_out), _ptr(d_sources),
                               _ptr(d_num_sources), _ptr(d_work), flags, stream))


def image_stats(n, d_img, stream=None) -> tuple[float, float, float]:
    """(sum of squares, max, max |value|) of n floats on the device."""
    sq, mx, amx = c_double(), c_float(), c_float()
    check(load().sdp_image_stats(n, _ptr(d_img), ctypes.byref(sq), ctypes.byref(mx), ctypes.byref(amx), stream))
    return sq.value, mx.value, amx.value


def fill_complex(n, d, re, im, stream=None):
    check(load().sdp_fill_complex(n, _ptr(d), re, im, stream))


def divide(n, d, by, stream=None):
    check(load().sdp_divide(n, _ptr(d), by, stream))


class Graph:
    """A captured stream sequence (sdp_graph_begin_capture .. end_capture)."""

    def __init__(self, handle):
        self.handle = handle

    @staticmethod
    def begin(stream=None):
        check(load().sdp_graph_begin_capture(stream))

    @classmethod
    def end(cls, stream=None) -> "Graph":
        h = c_void_p()
        check(load().sdp_graph_end_capture(stream, ctypes.byref(h)))
        return cls(h.value)

    def launch(self, stream=None):
        check(load().sdp_graph_launch(self.handle, stream))

    def close(self):
        if self.handle:
            check(load().sdp_graph_destroy(self.handle))
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class _PipelineParams(ctypes.Structure):
    _fields_ = [("grid_size", c_int), ("oversampling", c_int), ("n_kernels", c_int), ("n_vis", c_int),
                ("uv_scale", c_double), ("w_scale", c_double), ("cell_size", c_double),
                ("n_minor", c_int), ("max_sources", c_int), ("loop_gain", c_double), ("noise_factor", c_double),
                ("weak_source_percent", c_double), ("flags", c_int), ("tile", c_int), ("chunk", c_int)]


BATCH_GRAPH, BATCH_PROFILE = 1, 2


class _BatchParams(ctypes.Structure):
    _fields_ = [("n_channels", c_int), ("grid_size", c_int), ("oversampling", c_int), ("n_kernels", c_int),
                ("uv_scale", c_double), ("w_scale", c_double), ("tile", c_int), ("chunk", c_int),
                ("accumulate", c_int), ("flags", c_int)]


class Batch:
    """Batched multi-channel residual imaging (sdp_batch_* in sdp_pipeline.h)."""

    def __init__(self, grid_size, oversampling, supports, uv_scale, w_scale, d_uvws, d_viss, n_vis, d_grid_kernels,
                 d_degrid_kernels, d_taper, tile=0, chunk=0, accumulate=0, flags=0, stream=None):
        self._sup = np.ascontiguousarray(supports, dtype=np.int32)
        self.n_channels = len(n_vis)
        self.grid_size = grid_size
        self.params = _BatchParams(self.n_channels, grid_size, oversampling, self._sup.shape[0], uv_scale, w_scale,
                                   tile, chunk, accumulate, flags)
        self._n = np.ascontiguousarray(n_vis, dtype=np.int32)
        self._uvw = (c_void_p * self.n_channels)(*[_ptr(b) for b in d_uvws])
        self._vis = (c_void_p * self.n_channels)(*[_ptr(b) for b in d_viss])
        self._keep = (d_grid_kernels, d_degrid_kernels, d_taper)
        h = c_void_p()
        check(load().sdp_batch_create(ctypes.byref(h), ctypes.byref(self.params), _ptr(self._sup), _ptr(self._n),
                                      self._uvw, self._vis, _ptr(d_grid_kernels), _ptr(d_degrid_kernels),
                                      _ptr(d_taper), stream))
        self.handle = h.value

    def residuals(self) -> tuple[float, list]:
        """One multi-channel pass: (ms, per-stage ms of the pass when profiling)."""
        ms = c_float()
        st = (c_float * len(STAGES))()
        check(load().sdp_batch_residuals(self.handle, ctypes.byref(ms), st))
        return ms.value, list(st)

    def model_ptr(self, k) -> int:
        return load().sdp_batch_model(self.handle, k)

    def residual(self, k) -> np.ndarray:
        G = self.grid_size
        out = np.empty((G, G), np.float32)
        check(load().sdp_memcpy_dtoh(out.ctypes.data, load().sdp_batch_residual(self.handle, k), out.nbytes, None))
        return out

    def plan_info(self, k) -> dict:
        info = np.zeros(10, np.int64)
        check(load().sdp_grid_plan_info(load().sdp_batch_plan(self.handle, k), info.ctypes.data))
        keys = ["kept", "dropped", "tiles", "chunks", "tile", "chunk_taps", "max_half_support", "scratch_bytes",
                "class_launches", "taps"]
        return dict(zip(keys, (int(v) for v in info)))

    def close(self):
        if getattr(self, "handle", None):
            load().sdp_batch_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class _LearnedParams(ctypes.Structure):
    _fields_ = [("patch", c_int), ("margin", c_int), ("gain", c_double), ("threshold", c_double),
                ("floor", c_double), ("input_floor", c_double), ("max_cycles", c_int), ("scale", c_int)]


class Pipeline:
    """The device-resident major cycle (sdp_pipeline.h)."""

    def __init__(self, grid_size, oversampling, supports, uv_scale, w_scale, cell_size, d_uvw, d_vis, n_vis,
                 d_grid_kernels, d_degrid_kernels, d_taper, n_minor, max_sources=None, loop_gain=0.1,
                 noise_factor=0.0, weak_source_percent=0.0, flags=0, tile=0, chunk=0, stream=None):
        self._sup = np.ascontiguousarray(supports, dtype=np.int32)
        self.grid_size = grid_size
        self.params = _PipelineParams(grid_size, oversampling, self._sup.shape[0], n_vis, uv_scale, w_scale,
                                      cell_size, n_minor, max_sources or max(n_minor, 1), loop_gain, noise_factor,
                                      weak_source_percent, flags, tile, chunk)
        # the tables and taper are used by every cycle: keep them alive
        self._keep = (d_grid_kernels, d_degrid_kernels, d_taper)
        h = c_void_p()
        check(load().sdp_pipeline_create(ctypes.byref(h), ctypes.byref(self.params), _ptr(self._sup), _ptr(d_uvw),
                                         _ptr(d_vis), _ptr(d_grid_kernels), _ptr(d_degrid_kernels), _ptr(d_taper),
                                         stream))
        self.handle = h.value

    def cycle(self) -> tuple[float, float, float]:
        """One major cycle: (residual RMS before CLEAN, after CLEAN, ms)."""
        a, b, t = c_double(), c_double(), c_float()
        check(load().sdp_pipeline_cycle(self.handle, ctypes.byref(a), ctypes.byref(b), ctypes.byref(t)))
        return a.value, b.value, t.value

    def reset(self):
        check(load().sdp_pipeline_reset(self.handle))

    def set_learned(self, net, patch=0, margin=0, gain=0.5, threshold=0.02, floor=0.01, input_floor=-1.0,
                    max_cycles=0, scale=LEARNED_TILE_SCALE):
        """Learned minor cycle with network `net` (a Cnn), or Hogbom when None."""
        self._net = net
        lp = _LearnedParams(patch, margin, gain, threshold, floor, input_floor, max_cycles, scale)
        check(load().sdp_pipeline_set_learned(self.handle, net.handle if net is not None else None,
                                              ctypes.byref(lp)))

    @property
    def learned_active(self) -> bool:
        return bool(load().sdp_pipeline_learned_active(self.handle))

    def load_model(self, d_model):
        check(load().sdp_pipeline_load_model(self.handle, _ptr(d_model)))

    def image(self, which) -> np.ndarray:
        G = self.grid_size
        ptr = load().sdp_pipeline_image(self.handle, which)
        out = np.empty((G, G), np.float32)
        check(load().sdp_memcpy_dtoh(out.ctypes.data, ptr, out.nbytes, None))
        return out

    def stage_ms(self) -> dict:
        """Per-stage milliseconds of the last cycle (pipeline made with PIPE_PROFILE)."""
        ms = np.zeros(len(STAGES), np.float32)
        check(load().sdp_pipeline_stage_ms(self.handle, ms.ctypes.data))
        return dict(zip(STAGES, (float(v) for v in ms)))

    @property
    def psf_max(self) -> float:
        return float(load().sdp_pipeline_psf_max(self.handle))

    def sources(self) -> np.ndarray:
        buf = np.zeros((self.params.max_sources, 3), np.float32)
        n = c_int()
        check(load().sdp_pipeline_sources(self.handle, buf.ctypes.data, ctypes.byref(n)))
        return buf[:n.value].copy()

    def plan_info(self) -> dict:
        info = np.zeros(10, np.int64)
        check(load().sdp_grid_plan_info(load().sdp_pipeline_plan(self.handle), info.ctypes.data))
        keys = ["kept", "dropped", "tiles", "chunks", "tile", "chunk_taps", "max_half_support", "scratch_bytes",
                "class_launches", "taps"]
        return dict(zip(keys, (int(v) for v in info)))

    def close(self):
        if getattr(self, "handle", None):
            check(load().sdp_pipeline_destroy(self.handle))
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class _CnnParams(ctypes.Structure):
    _fields_ = [("levels", c_int), ("base", c_int), ("in_ch", c_int), ("out_ch", c_int), ("height", c_int),
                ("width", c_int), ("max_batch", c_int), ("precision", c_int)]


def cnn_layer_shapes(levels, base, in_ch=2, out_ch=1):
    """(name, cin, cout) of every convolution, in parameter order."""
    ch = [base << l for l in range(levels)]
    out, prev = [], in_ch
    for l in range(levels):
        out += [(f"enc{l}a", prev, ch[l]), (f"enc{l}b", ch[l], ch[l])]
        prev = ch[l]
    for l in range(levels - 2, -1, -1):
        out += [(f"dec{l}u", ch[l + 1], ch[l]), (f"dec{l}d", ch[l], ch[l])]
    out.append(("head", ch[0], out_ch))
    return out


class Cnn:
    """The learned minor-cycle proposal network (sdp_cnn_* in sdp_device.h)."""

    def __init__(self, levels, base, height, width, max_batch, in_ch=2, out_ch=1, precision=CNN_BF16):
        self.levels, self.base, self.in_ch, self.out_ch = levels, base, in_ch, out_ch
        self.p = _CnnParams(levels, base, in_ch, out_ch, height, width, max_batch, precision)
        h = c_void_p()
        check(load().sdp_cnn_create(ctypes.byref(h), ctypes.byref(self.p)))
        self.handle = h.value
        self.n_params = int(load().sdp_cnn_param_count(self.handle))

    def flatten(self, params: dict) -> np.ndarray:
        parts = []
        for name, ci, co in cnn_layer_shapes(self.levels, self.base, self.in_ch, self.out_ch):
            parts += [np.asarray(params[name + ".w"], np.float32).reshape(-1),
                      np.asarray(params[name + ".b"], np.float32).reshape(-1)]
        flat = np.concatenate(parts)
        assert flat.size == self.n_params
        return flat

    def unflatten(self, flat: np.ndarray) -> dict:
        out, off = {}, 0
        for name, ci, co in cnn_layer_shapes(self.levels, self.base, self.in_ch, self.out_ch):
            n = 9 * ci * co
            out[name + ".w"] = flat[off:off + n].reshape(3, 3, ci, co)
            out[name + ".b"] = flat[off + n:off + n + co]
            off += n + co
        return out

    def set_params(self, params):
        flat = np.ascontiguousarray(params if isinstance(params, np.ndarray) else self.flatten(params),
                                    dtype=np.float32)
        check(load().sdp_cnn_set_params(self.handle, flat.ctypes.data))

    def get_params(self) -> np.ndarray:
        out = np.empty(self.n_params, np.float32)
        check(load().sdp_cnn_get_params(self.handle, out.ctypes.data))
        return out

    def get_grads(self) -> np.ndarray:
        out = np.empty(self.n_params, np.float32)
        check(load().sdp_cnn_get_grads(self.handle, out.ctypes.data))
        return out

    def forward(self, d_in, d_out, batch, stream=None):
        check(load().sdp_cnn_forward(self.handle, _ptr(d_in), _ptr(d_out), int(batch), stream))

    def backward(self, d_grad_out, stream=None):
        check(load().sdp_cnn_backward(self.handle, _ptr(d_grad_out), stream))

    def adam(self, lr, beta1=0.9, beta2=0.999, eps=1e-8, stream=None):
        check(load().sdp_cnn_adam(self.handle, lr, beta1, beta2, eps, stream))

    def close(self):
        if getattr(self, "handle", None):
            check(load().sdp_cnn_destroy(self.handle))
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def axpby(n, a, d_x, b, d_y, d_out, stream=None):
    check(load().sdp_axpby(n, a, _ptr(d_x), b, _ptr(d_y), _ptr(d_out), stream))


def complex_multiply(n, d_a, d_b, nb, d_out, stream=None):
    check(load().sdp_complex_multiply(n, _ptr(d_a), _ptr(d_b), nb, _ptr(d_out), stream))


def cnn_gather_patches(grid_size, patch, n, d_origins, d_peaks, d_image, d_psf, d_batch, stream=None):
    check(load().sdp_cnn_gather_patches(grid_size, patch, n, _ptr(d_origins), _ptr(d_peaks), _ptr(d_image),
                                        _ptr(d_psf), _ptr(d_batch), stream))


def cnn_crop_patches(grid_size, patch, n, d_origins, d_peaks, d_image, d_out, stream=None):
    check(load().sdp_cnn_crop_patches(grid_size, patch, n, _ptr(d_origins), _ptr(d_peaks), _ptr(d_image),
                                      _ptr(d_out), stream))


def sky_render(grid_size, d_points, n_points, d_gauss, n_gauss, d_sky, stream=None):
    check(load().sdp_sky_render(grid_size, n_points, _ptr(d_points), n_gauss, _ptr(d_gauss), _ptr(d_sky), stream))


def cnn_mse(count, d_out, d_target, d_grad, stream=None) -> float:
    loss = c_double()
    check(load().sdp_cnn_mse(count, _ptr(d_out), _ptr(d_target), _ptr(d_grad), ctypes.byref(loss), stream))
    return loss.value


def dft_predict(n_vis, n_sources, d_sources, d_uvw, d_out, mode=DFT_FP64_PHASE, stream=None):
    check(load().sdp_dft_predict(n_vis, n_sources, _ptr(d_sources), _ptr(d_uvw), _ptr(d_out), mode, stream))


def dft_array_work_bytes(n_sources) -> int:
    return load().sdp_dft_array_work_bytes(n_sources)


def dft_predict_array(n_time, n_ant, d_ant_uvw, n_sources, d_sources, d_out, flags=DFT_ARRAY_BF16, d_work=None,
                      stream=None):
    """Array-factorised predict on the matrix cores (sdp_dft_predict_array);
    d_ant_uvw holds n_time * n_ant * 3 doubles."""
    work = d_work if d_work is not None else DeviceBuffer(dft_array_work_bytes(n_sources))
    check(load().sdp_dft_predict_array(n_time, n_ant, _ptr(d_ant_uvw), n_sources, _ptr(d_sources), _ptr(d_out),
                                       flags, _ptr(work), stream))
    return work


def degrid_nearest(n_vis, n_out, grid_size, uv_scale, d_grid, d_uvw, d_out, stream=None):
    check(load().sdp_degrid_nearest(n_vis, n_out, grid_size, uv_scale, _ptr(d_grid), _ptr(d_uvw), _ptr(d_out),
                                    stream))


# ---------------------------------------------------------------- actor layer

class Actors:
    """The reference actor ABI, called with numpy host arrays."""

    def __init__(self):
        self.lib = load()
        P = c_void_p
        cfg = ctypes.POINTER(Config)
        ip = ctypes.POINTER(c_int)
        sigs = {
            "correct_to_finegrid": [c_int, c_int, c_int, c_int, P, P, cfg, P, P, ip],
            "std_gridding": [c_int, c_int, c_int, c_int, c_int, c_int, ip, ip, P, P, P, P, cfg, P, P],
            "std_degridding": [c_int, c_int, c_int, c_int, c_int, P, P, P, P, ip, cfg, P],
            "gridding_actor": [c_int, c_int, c_int, c_int, P, P, P, P, cfg, P],
            "add_visibilities": [c_int, P, P, P],
            "dft_actor": [c_int, c_int, P, P, ip, cfg, P],
            "hogbom_clean": [c_int, c_int, c_int, P, P, P, cfg, P, ip, P, P],
            "fft_shift_complex_to_complex_actor": [c_int, P, cfg, P],
            "CUFFT_EXECUTE_FORWARD_C2C_actor": [c_int, P, P],
            "CUFFT_EXECUTE_INVERSE_C2C_actor": [c_int, P, P],
            "fft_shift_complex_to_real_actor": [c_int, P, cfg, P],
            "fft_shift_real_to_complex_actor": [c_int, P, cfg, P],
            "execute_convolution_correction_actor": [c_int, P, P, cfg, P],
            "execute_inv_convolution_correction_actor": [c_int, P, P, cfg, P],
            "subtract_from_measurements": [c_int, c_int, c_int, P, P, P, P, cfg, P],
            "apply_gains_only": [c_int, c_int, c_int, P, P, P, cfg, P],
            "reciprocal_transform_actor": [c_int, P, cfg, P],
            "save_output": [c_int, P, P, P, P, P, cfg, ip],
            "subtract_from_measurements_finegrid": [c_int, P, P, ip, P],
            "merge_gridded_visibilities": [c_int, c_int, P, P],
            "split_uvgrid": [c_int, c_int, P, P],
            "correction_set_up": [c_int, P],
            "prolate_spheroidal": [c_double],
        }
        for name, args in sigs.items():
            f = getattr(self.lib, name)
            f.argtypes = args
            f.restype = c_double if name == "prolate_spheroidal" else None

    @staticmethod
    def _i(v):
        return ctypes.byref(c_int(int(v)))

    def correct_to_finegrid(self, G, OS, simplify, uvw, vis, cfg):
        n = uvw.shape[0]
        out_vis = np.zeros((n, 2), np.float32)
        out_uvw = np.zeros((n, 3), np.float32)
        kept = c_int(0)
        self.lib.correct_to_finegrid(n, G, OS, simplify, _ptr(uvw), _ptr(vis), ctypes.byref(cfg), _ptr(out_vis),
                                     _ptr(out_uvw), ctypes.byref(kept))
        return out_vis, out_uvw, kept.value

    def std_gridding(self, G, OS, kernels, supports, uvw, vis, cfg, n=None, bypass=0, maj_iter=0, prev_grid=None):
        n = uvw.shape[0] if n is None else n
        out = np.empty((G * G, 2), np.float32)
        prev = prev_grid if prev_grid is not None else out
        self.lib.std_gridding(G, uvw.shape[0], supports.shape[0], kernels.shape[0], OS, bypass, self._i(maj_iter),
                              self._i(n), _ptr(kernels), _ptr(supports), _ptr(uvw), _ptr(vis), ctypes.byref(cfg),
                              _ptr(prev), _ptr(out))
        return out

    def std_degridding(self, G, OS, kernels, supports, grid, uvw, cfg, n=None, n_total=None):
        n = uvw.shape[0] if n is None else n
        n_total = uvw.shape[0] if n_total is None else n_total
        out = np.empty((n_total, 2), np.float32)
        self.lib.std_degridding(G, n_total, supports.shape[0], kernels.shape[0], OS, _ptr(kernels), _ptr(supports),
                                _ptr(grid), _ptr(uvw), self._i(n), ctypes.byref(cfg), _ptr(out))
        return out

    def gridding_actor(self, G, kernels, supports, uvw, vis, cfg):
        out = np.empty((G * G, 2), np.float32)
        self.lib.gridding_actor(G, uvw.shape[0], supports.shape[0], kernels.shape[0], _ptr(kernels),
                                _ptr(supports), _ptr(uvw), _ptr(vis), ctypes.byref(cfg), _ptr(out))
        return out

    def add_visibilities(self, a, b):
        out = np.empty_like(a)
        self.lib.add_visibilities(a.shape[0], _ptr(a), _ptr(b), _ptr(out))
        return out

    def dft_actor(self, sources, uvw, cfg, max_sources=None):
        n = uvw.shape[0]
        out = np.empty((n, 2), np.float32)
        ns = sources.shape[0]
        self.lib.dft_actor(n, max_sources or ns, _ptr(sources), _ptr(uvw), self._i(ns), ctypes.byref(cfg),
                           _ptr(out))
        return out

    def hogbom_clean(self, G, n_minor, max_sources, residual, psf, cfg, model):
        out_model = np.empty(G * G, np.float32)
        sources = np.zeros((max_sources, 3), np.float32)
        ns = c_int(0)
        halfdims = np.zeros(2, np.int32)
        self.lib.hogbom_clean(G, n_minor, max_sources, _ptr(residual), _ptr(psf), _ptr(halfdims),
                              ctypes.byref(cfg), _ptr(model), ctypes.byref(ns), _ptr(sources), _ptr(out_model))
        return ns.value, sources[: ns.value], out_model

    def shift_c2c(self, G, grid, cfg):
        out = np.empty_like(grid)
        self.lib.fft_shift_complex_to_complex_actor(G, _ptr(grid), ctypes.byref(cfg), _ptr(out))
        return out

    def shift_c2r(self, G, grid, cfg):
        out = np.empty(G * G, np.float32)
        self.lib.fft_shift_complex_to_real_actor(G, _ptr(grid), ctypes.byref(cfg), _ptr(out))
        return out

    def shift_r2c(self, G, image, cfg):
        out = np.empty((G * G, 2), np.float32)
        self.lib.fft_shift_real_to_complex_actor(G, _ptr(image), ctypes.byref(cfg), _ptr(out))
        return out

    def fft_forward(self, G, grid):
        out = np.empty_like(grid)
        self.lib.CUFFT_EXECUTE_FORWARD_C2C_actor(G, _ptr(grid), _ptr(out))
        return out

    def fft_inverse(self, G, grid):
        out = np.empty_like(grid)
        self.lib.CUFFT_EXECUTE_INVERSE_C2C_actor(G, _ptr(grid), _ptr(out))
        return out

    def convolution_correction(self, G, image, prolate, cfg):
        out = np.empty_like(image)
        self.lib.execute_convolution_correction_actor(G, _ptr(image), _ptr(prolate), ctypes.byref(cfg), _ptr(out))
        return out

    def subtract_from_measurements(self, measured, pred, gains, pairs, cfg):
        out = np.empty_like(measured)
        self.lib.subtract_from_measurements(gains.shape[0], pairs.shape[0], measured.shape[0], _ptr(measured),
                                            _ptr(pred), _ptr(gains), _ptr(pairs), ctypes.byref(cfg), _ptr(out))
        return out

    def reciprocal_transform_actor(self, gains, cfg):
        gains = np.ascontiguousarray(gains, dtype=np.float32)
        out = np.empty_like(gains)
        self.lib.reciprocal_transform_actor(gains.shape[0], _ptr(gains), ctypes.byref(cfg), _ptr(out))
        return out

    def apply_gains_only(self, measured, gains, pairs, cfg):
        out = np.empty_like(measured)
        self.lib.apply_gains_only(gains.shape[0], pairs.shape[0], measured.shape[0], _ptr(measured), _ptr(gains),
                                  _ptr(pairs), ctypes.byref(cfg), _ptr(out))
        return out

    def subtract_finegrid(self, measured, pred, n=None):
        out = np.zeros_like(measured)
        n = measured.shape[0] if n is None else n
        self.lib.subtract_from_measurements_finegrid(measured.shape[0], _ptr(measured), _ptr(pred), self._i(n),
                                                     _ptr(out))
        return out

    def merge(self, G, slices, grids):
        out = np.empty((G * G, 2), np.float32)
        self.lib.merge_gridded_visibilities(G, slices, _ptr(grids), _ptr(out))
        return out

    def correction_set_up(self, G):
        out = np.empty(G // 2, np.float32)
        self.lib.correction_set_up(G, _ptr(out))
        return out

    def set_degrid_mode(self, mode):
        self.lib.sdp_set_degrid_mode(mode)

    def set_dft_mode(self, mode):
        self.lib.sdp_set_dft_mode(mode)
