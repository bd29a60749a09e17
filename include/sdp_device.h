/*
 * sdp_device.h -- device-resident C-ABI of the MI355X imaging library.
 *
 * The actor ABI (sdp_actors.h) takes host pointers and copies on every call,
 * like the reference's GPU actors (src/GPU/dft_run.cu:25-73).  A caller that
 * keeps data in HBM across a major cycle uses these entry points instead:
 * plain device pointers, sizes, and an optional HIP stream (NULL = the
 * library's default stream).  No function here allocates per call except
 * the plan constructors, which allocate once.
 *
 * Complex data are interleaved float pairs (re, im), row-major G x G for
 * grids and images, index (v + G/2) * G + (u + G/2) for grid cell (u, v).
 * Coordinates are float triples (u, v, w) in wavelengths.
 *
 * Every function returns 0 on success and a negative SDP_E* code on error;
 * sdp_last_error() gives the message.  These have no counterpart in the
 * reference (which only has host-pointer actors); each cites the actor whose
 * arithmetic it implements.
 */
#ifndef SDP_DEVICE_H
#define SDP_DEVICE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef void *sdp_stream;   /* hipStream_t */
typedef void *sdp_event;    /* hipEvent_t  */

enum {
    SDP_OK = 0,
    SDP_E_HIP = -1,        /* a HIP runtime call failed */
    SDP_E_ARG = -2,        /* invalid argument (size, pointer, mode) */
    SDP_E_NOMEM = -3,      /* device allocation failed */
    SDP_E_UNSUPPORTED = -4 /* shape outside what the kernels support */
};

/* ---- runtime ------------------------------------------------------------ */
int sdp_init(void);
const char *sdp_last_error(void);
/* records `message` as this thread's last error and returns `code` (for
 * host layers built on this ABI) */
int sdp_set_last_error(int code, const char *message);
int sdp_device_malloc(void **ptr, size_t bytes);
int sdp_device_free(void *ptr);
int sdp_memcpy_htod(void *dst, const void *src, size_t bytes, sdp_stream s);
int sdp_memcpy_dtoh(void *dst, const void *src, size_t bytes, sdp_stream s);
int sdp_memcpy_dtod(void *dst, const void *src, size_t bytes, sdp_stream s);
int sdp_memset_zero(void *dst, size_t bytes, sdp_stream s);   /* fill kernel, graph safe */
/* copy kernel (graph safe); pointers and size multiples of 16 bytes */
int sdp_device_copy(void *dst, const void *src, size_t bytes, sdp_stream s);
int sdp_stream_create(sdp_stream *s);
int sdp_stream_destroy(sdp_stream s);
int sdp_stream_sync(sdp_stream s);
int sdp_device_sync(void);
int sdp_event_create(sdp_event *e);
int sdp_event_destroy(sdp_event e);
int sdp_event_record(sdp_event e, sdp_stream s);
int sdp_event_sync(sdp_event e);
int sdp_event_elapsed_ms(sdp_event start, sdp_event stop, float *ms);

/* stream capture into an executable graph (replayed with sdp_graph_launch);
 * every call between begin and end must be enqueue-only on that stream */
typedef void *sdp_graph;
int sdp_graph_begin_capture(sdp_stream s);
int sdp_graph_end_capture(sdp_stream s, sdp_graph *graph);
int sdp_graph_launch(sdp_graph graph, sdp_stream s);
int sdp_graph_destroy(sdp_graph graph);

/* ---- actor-level switches ---------------------------------------------- */
enum { SDP_DEGRID_NEAREST = 0, SDP_DEGRID_WEIGHTED = 1 };
/* std_degridding output: nearest grid cell (reference default,
 * degridgrid.c:194-200) or kernel-weighted sum (pre-override loop
 * degridgrid.c:147-186 with zero padding, no comm_norm division). */
void sdp_set_degrid_mode(int mode);
int sdp_get_degrid_mode(void);

enum { SDP_DFT_REFERENCE_FP32 = 0, SDP_DFT_FP64_PHASE = 1 };
/* dft_actor arithmetic: the reference's FP32 phase (dft_run.c:21-32), or the
 * same formula with the phase formed and range-reduced in FP64.  (The
 * matrix-core form needs per-antenna tracks: sdp_dft_predict_array.) */
void sdp_set_dft_mode(int mode);
int sdp_get_dft_mode(void);

/* ---- fine-grid snapping (degridgrid.c:12-65, PERFORM_SIMPLIFICATION = 0) */
/* uvw_out[i] = uvw snapped to the oversampled fine grid; w copied. */
int sdp_finegrid_correct(int n_vis, int grid_size, int oversampling, double uv_scale,
                         const float *d_uvw_in, float *d_uvw_out, sdp_stream s);
/* PERFORM_SIMPLIFICATION = 1: visibilities that fall in the same fine-grid
 * cell and w-plane are summed into the first occurrence and compacted in
 * order of first occurrence.  *n_out receives the number kept (host int). */
int sdp_finegrid_simplify(int n_vis, int grid_size, int oversampling, double uv_scale, double w_scale,
                          const float *d_uvw_in, const float *d_vis_in,
                          float *d_uvw_out, float *d_vis_out, int *n_out, sdp_stream s);

/* ---- gridding / degridding plans --------------------------------------- */
enum { SDP_SNAP_STD = 0,        /* std_gridding: half-open [ceil(p-h), ceil(p+h)) */
       SDP_SNAP_GRIDDING_ACTOR  /* gridding_actor: snapped, inclusive [ceil(s-h), floor(s+h)],
                                   plane = round(sqrt(|w| w_scale)) */ };
enum { SDP_ORDER_ORIGINAL = 0, SDP_ORDER_PLAN = 1 };

typedef struct sdp_grid_params {
    int grid_size;        /* G, even, <= 32768 */
    int oversampling;     /* OS */
    int n_kernels;        /* number of w-planes in the support table, 1 to 8192 */
    int snap_mode;        /* SDP_SNAP_* */
    double uv_scale;      /* cell_size * G */
    double w_scale;       /* (n_kernels - 1)^2 / max_w */
    int tile;             /* uv tile edge in cells; 0: 8, or the largest power of two
                             within 2h + 1 of the smallest support when narrower */
    int chunk;            /* taps per work item; 0: about 8 work items per CU, within
                             4096 and 131072 taps */
    int accumulate;       /* SDP_ACC_*: how a work item sums into its sub-grid */
} sdp_grid_params;
/* Grid-cell accumulation.  Every cell's value is a TwoSum-compensated sum of
 * the work items' partial sums, merged in a fixed order.  Inside a work item
 * (a few hundred visibilities at most per cell) the partial is a float sum
 * (SDP_ACC_BLOCKED, the default) or itself TwoSum-compensated
 * (SDP_ACC_COMPENSATED: about twice the arithmetic and LDS). */
enum { SDP_ACC_BLOCKED = 0, SDP_ACC_COMPENSATED = 1 };

typedef struct sdp_grid_plan sdp_grid_plan;

/* Builds the per-visibility records (exact reference index arithmetic) and
 * the tile-sorted order for n_vis coordinates.  supports is a HOST array of
 * n_kernels (half_support, sample_offset) int pairs (kernel_host_set_up
 * layout, top.c:252-257); half supports up to 31 (a footprint, up to 2h + 1
 * cells across, must fit a 64-lane wavefront), else SDP_E_UNSUPPORTED. */
int sdp_grid_plan_create(sdp_grid_plan **plan, const sdp_grid_params *p, const int *supports,
                         const float *d_uvw, int n_vis, sdp_stream s);
/* One plan over n_grids grids (e.g. frequency channels) that share the
 * parameters and kernel tables: grid g's n_vis[g] coordinates are at
 * d_uvw[g] (host array of device pointers).  Visibilities are indexed over
 * all grids in grid order (grid g's first is sum_{k<g} n_vis[k]): the
 * caller-order arrays the plan functions take hold them in that order, and
 * every grid argument of the functions below then holds n_grids
 * consecutive G x G grids.  Each pass covers all grids in the launches of
 * one grid; a visibility's footprint is clipped to its own grid, and grid
 * g's cells and visibilities come out as with a plan of grid g alone (same
 * tile and work-item size, bit for bit while the sort key keeps the same
 * low field: with very many grids and tiles the plane part of that field
 * is dropped, which only reorders float sums inside a work item). */
int sdp_grid_plan_create_multi(sdp_grid_plan **plan, const sdp_grid_params *p, const int *supports, int n_grids,
                               const float *const *d_uvw, const int *n_vis, sdp_stream s);
int sdp_grid_plan_destroy(sdp_grid_plan *plan);
/* plan statistics (10 values): [0] visibilities kept, [1] dropped (outside
 * grid or beyond the support table), [2] tiles used, [3] work items, [4] tile
 * edge, [5] taps per work item, [6] max half support, [7] scratch bytes,
 * [8] kernel launches per pass (support classes present), [9] total taps */
int sdp_grid_plan_info(const sdp_grid_plan *plan, int64_t *info10);
/* reorder n_vis complex values between caller order and plan order */
int sdp_grid_plan_to_plan_order(const sdp_grid_plan *plan, const float *d_in, float *d_out, sdp_stream s);
int sdp_grid_plan_to_original_order(const sdp_grid_plan *plan, const float *d_in, float *d_out, sdp_stream s);

/* grid[G*G] = sum_vis vis * kernel (std_gridding, degridgrid.c:83-124);
 * the whole grid is written.  vis is in `order`. */
int sdp_grid(const sdp_grid_plan *plan, const float *d_kernels, const float *d_vis, int order,
             float *d_grid, sdp_stream s);
/* the two phases of sdp_grid: per-work-item LDS accumulation into the
 * plan's slabs, then the fixed-order merge of the slabs into the grid */
int sdp_grid_accumulate(const sdp_grid_plan *plan, const float *d_kernels, const float *d_vis, int order,
                        sdp_stream s);
int sdp_grid_merge(const sdp_grid_plan *plan, float *d_grid, sdp_stream s);
/* kernel-weighted degridding (adjoint of sdp_grid when d_kernels holds the
 * conjugate table): vis_out = sum_taps grid * kernel.  If d_vis_measured is
 * not NULL the output is measured - predicted (subtract_from_measurements_finegrid
 * fused, gains_apply_run.c:133-141).  Input and output in `order`. */
int sdp_degrid(const sdp_grid_plan *plan, const float *d_kernels, const float *d_grid,
               const float *d_vis_measured, int order, float *d_vis_out, sdp_stream s);

/* the major cycle's visibility-domain step in one pass: predict each
 * visibility from d_model_grid (as sdp_degrid with d_degrid_kernels),
 * subtract it from d_vis_measured, and grid the residual into
 * d_residual_grid (as sdp_grid with d_grid_kernels); the residual
 * visibilities are also written to d_residual_vis when it is not NULL.
 * Arithmetic of std_degridding + subtract_from_measurements_finegrid +
 * std_gridding (degridgrid.c:69-186, gains_apply_run.c:133-141). */
int sdp_degrid_grid(const sdp_grid_plan *plan, const float *d_degrid_kernels, const float *d_grid_kernels,
                    const float *d_model_grid, const float *d_vis_measured, int order, float *d_residual_vis,
                    float *d_residual_grid, sdp_stream s);
/* sdp_degrid_grid without the final sdp_grid_merge */
int sdp_degrid_grid_accumulate(const sdp_grid_plan *plan, const float *d_degrid_kernels,
                               const float *d_grid_kernels, const float *d_model_grid, const float *d_vis_measured,
                               int order, float *d_residual_vis, sdp_stream s);

/* Autotuner: builds a plan for every (tile, taps-per-work-item) candidate
 * on these coordinates, times one sdp_degrid_grid_accumulate + merge pass of each with
 * HIP events and returns the fastest pair (and its time).  Candidates whose
 * windows do not fit LDS are skipped.  Allocates scratch for the timing. */
int sdp_grid_autotune(const sdp_grid_params *p, const int *supports, const float *d_uvw, int n_vis,
                      const float *d_grid_kernels, const float *d_degrid_kernels, const int *tiles, int n_tiles,
                      const int *chunks, int n_chunk, int *best_tile, int *best_chunk, float *best_ms, sdp_stream s);

/* nearest-cell degridding (std_degridding output, degridgrid.c:194-200);
 * writes n_out (>= n_vis) values, zero beyond n_vis. */
int sdp_degrid_nearest(int n_vis, int n_out, int grid_size, double uv_scale, const float *d_grid,
                       const float *d_uvw, float *d_vis_out, sdp_stream s);

/* ---- FFT (fft_run.c:5-111) ---------------------------------------------- */
enum { SDP_FFT_FORWARD = -1, SDP_FFT_INVERSE = +1 };
enum { SDP_FFT_SHIFT_IN = 1,      /* multiply input by (-1)^(r+c) */
       SDP_FFT_SHIFT_OUT = 2,     /* multiply output by (-1)^(r+c) */
       SDP_FFT_REAL_IN = 4,       /* input is a real G x G image */
       SDP_FFT_REAL_OUT = 8,      /* output only the real part as a G x G image */
       SDP_FFT_CORRECT_OUT = 16,  /* divide the real output by the PSWF taper */
       SDP_FFT_CORRECT_IN = 32    /* divide the real input by the taper (the inverse correction
                                     the reference leaves as a copy, convolution_correction_run.c:65) */ };
/* bytes of device work memory sdp_fft2d needs for grid size G */
size_t sdp_fft_work_bytes(int grid_size);
/* unnormalised 2D C2C DFT with sign `direction`; in == out allowed.
 * d_taper: G/2+1 floats (taper with a trailing 0) when SDP_FFT_CORRECT_OUT. */
int sdp_fft2d(const void *d_in, void *d_out, int grid_size, int direction, int flags,
              const float *d_taper, void *d_work, sdp_stream s);
/* sdp_fft2d of n_batch consecutive grids (input, output and the work
 * buffer, n_batch * sdp_fft_work_bytes(G) bytes, each hold n_batch grids)
 * in the same three launches: results identical to n_batch sdp_fft2d calls.
 * No reference counterpart (the reference transforms one grid per actor
 * call); it serves the batched multi-channel pass. */
int sdp_fft2d_batch(const void *d_in, void *d_out, int grid_size, int n_batch, int direction, int flags,
                    const float *d_taper, void *d_work, sdp_stream s);

/* ---- element-wise actors ------------------------------------------------ */
int sdp_shift_c2c(int grid_size, const float *d_in, float *d_out, sdp_stream s);
int sdp_shift_c2r(int grid_size, const float *d_in, float *d_out, sdp_stream s);
int sdp_shift_r2c(int grid_size, const float *d_in, float *d_out, sdp_stream s);
/* d_taper: G/2+1 floats, element G/2 must be 0 (convolution_correction_run.c:23) */
int sdp_convolution_correct(int grid_size, const float *d_in, const float *d_taper, float *d_out, sdp_stream s);
int sdp_vis_subtract(int n, const float *d_a, const float *d_b, float *d_out, sdp_stream s);
int sdp_vis_add(int n, const float *d_a, const float *d_b, float *d_out, sdp_stream s);
int sdp_vis_subtract_gains(int n_vis, int n_baselines, const float *d_measured, const float *d_pred,
                           const float *d_gains, const int *d_pairs, int apply_only,
                           float *d_out, sdp_stream s);
/* the restored image (conv2D, major_loop_iter.c:8-46): out = the model
 * convolved with the clean beam over the [-hx, hx) x [-hy, hy) box, model
 * pixels below 1e-5 skipped, the beam read flipped about (D/2 + 1, D/2 + 1)
 * and zero padding at the edges; bit-exact with the reference's float sums.
 * Requires D/2 + 1 + h < D. */
int sdp_restore_conv2d(int D, const float *d_image, const float *d_filter, int hx, int hy, float *d_out,
                       sdp_stream s);
/* out[i] = 1 / in[i] for n complex gains, as conj(z) / |z|^2 in float
 * (reciprocal_transform_actor, gains_apply_run.c:4-17); in == out allowed */
int sdp_gains_reciprocal(int n, const float *d_in, float *d_out, sdp_stream s);
/* out[i] = sum_{j<slices} grids[j*G*G + i] in slice order */
int sdp_merge_grids(int grid_size, int slices, const float *d_grids, float *d_out, sdp_stream s);

/* image statistics: sum of squares (double), largest value, largest |value|;
 * synchronous (results on the host) */
int sdp_image_stats(size_t n, const float *d_img, double *sum_sq, float *max_value, float *max_abs, sdp_stream s);
/* the same, enqueued only: d_stats[0..2] = {sum of squares, max, max |value|}
 * as doubles, d_work of sdp_image_stats_work_bytes() bytes */
size_t sdp_image_stats_work_bytes(void);
int sdp_image_stats_async(size_t n, const float *d_img, double *d_stats, void *d_work, sdp_stream s);
int sdp_fill_complex(size_t n, float *d, float re, float im, sdp_stream s);
/* out[i] = a x[i] + b y[i] for n floats (out may alias x or y) */
int sdp_axpby(size_t n, float a, const float *d_x, float b, const float *d_y, float *d_out, sdp_stream s);
/* out[i] = a[i] * b[i % nb] for n complex values (b repeats: one filter for a batch) */
int sdp_complex_multiply(size_t n, const float *d_a, const float *d_b, size_t nb, float *d_out, sdp_stream s);
/* d[i] = (float)((double)d[i] / by): psf_host_set_up's normalisation (top.c:505-507) */
int sdp_divide(size_t n, float *d, double by, sdp_stream s);

/* ---- Hogbom CLEAN (deconvolution_run.c:197-312) ------------------------ */
/* d_psf: G*G+1 floats (peak-normalised PSF followed by one 0).
 * d_residual is updated in place (it ends as the residual after the last
 * subtraction); d_model_out = d_model_in + components; d_sources holds
 * max_sources (x, y, flux) triples in radians after the conversion
 * (deconvolution_run.c:297-306); *d_num_sources is a device int.
 * d_work: sdp_hogbom_work_bytes(G) bytes. */
size_t sdp_hogbom_work_bytes(int grid_size);
int sdp_hogbom(int grid_size, int n_minor, int max_sources, float *d_residual, const float *d_psf,
               double psf_max, double loop_gain, double noise_factor, double weak_source_percent,
               double cell_size, const float *d_model_in, float *d_model_out, float *d_sources,
               int *d_num_sources, void *d_work, sdp_stream s);
/* flags for sdp_hogbom_ex.  By default the PSF is read x-mirrored about its
 * centre exactly as the reference indexes it (deconvolution_run.c:269);
 * SDP_HOGBOM_PSF_AS_STORED subtracts it in its stored orientation.
 * The noise stop (deconvolution_run.c:238) compares the peak with
 * noise_factor * gain * the reference's float running "average" (:226),
 * which drifts from the true mean of |r_1..r_{N-1}| as N grows (0.7% at
 * 4096^2).  When noise_factor * gain > 1/2 (below that the stop cannot
 * fire) sdp_hogbom runs that recurrence sequentially on one wavefront before
 * every iteration, so the stop matches the reference bit for bit at the
 * cost of about 0.1 s per iteration at 2048^2; SDP_HOGBOM_DOUBLE_MEAN
 * uses the parallel double sum / (N - 1) instead, which can decide
 * differently only when the peak is within that drift of the threshold.
 * Both entry points also stop after max_sources components (the reference
 * keeps going and writes past its sources array); the first max_sources
 * iterations are the reference's.
 * When the side is a power of two up to 4096, the noise stop cannot fire
 * and the stream is not being captured, all iterations run in one
 * cooperative launch that keeps the residual in registers (the same results
 * bit for bit); SDP_HOGBOM_PER_ITERATION forces one launch per iteration.
 * A run that cannot complete its cross-workgroup exchange reports -1
 * sources; SDP_HOGBOM_TEST_ABORT raises that exit at the first exchange (a
 * test hook: the one-launch form then ends at once with -1 sources). */
enum {
    SDP_HOGBOM_PSF_AS_STORED = 1,
    SDP_HOGBOM_DOUBLE_MEAN = 2,
    SDP_HOGBOM_PER_ITERATION = 4,
    SDP_HOGBOM_TEST_ABORT = 8
};
int sdp_hogbom_ex(int grid_size, int n_minor, int max_sources, float *d_residual, const float *d_psf,
                  double psf_max, double loop_gain, double noise_factor, double weak_source_percent,
                  double cell_size, const float *d_model_in, float *d_model_out, float *d_sources,
                  int *d_num_sources, void *d_work, int flags, sdp_stream s);

/* ---- DFT predict (dft_run.c:3-47) --------------------------------------- */
/* vis[j] = sum_s flux_s / (1 - r^2/2) * exp(-2 pi i (u l + v m - w r^2/2)),
 * sources (l, m, flux) in radians.  mode: SDP_DFT_*. */
int sdp_dft_predict(int n_vis, int n_sources, const float *d_sources, const float *d_uvw,
                    float *d_vis_out, int mode, sdp_stream s);

/* The same predict for a whole array, factorised per antenna and contracted
 * on the matrix cores.  For n_ant antennas at n_time instants and every
 * baseline (a, b), a < b, in calculate_receiver_pairs order
 * (top.c:119-137), visibility index t * n_ant (n_ant - 1) / 2 + k:
 *   V_ab(t) = sum_s g_s E_bs(t) conj(E_as(t)),  g_s = f_s / (1 - r_s^2/2),
 *   E_xs(t) = exp(-i 2 pi (u_x l_s + v_x m_s - w_x r_s^2/2)),
 * which is dft_actor's sum (dft_run.c:21-41) for uvw_ab = uvw_b - uvw_a.
 * d_ant_uvw: n_time * n_ant * 3 doubles, antenna uvw in wavelengths about
 * any common origin.  Each phase is formed and range-reduced in FP64 per
 * antenna and source; the contraction over sources is a complex matrix
 * product per instant on the matrix cores (BF16 or FP16 operands, FP32
 * accumulation), each operand split into hi + lo parts with the three
 * products hi*hi + hi*lo + lo*hi, or hi only with SDP_DFT_ARRAY_NO_SPLIT.
 * d_work: sdp_dft_array_work_bytes(n_sources) bytes. */
enum { SDP_DFT_ARRAY_BF16 = 0, SDP_DFT_ARRAY_FP16 = 1, SDP_DFT_ARRAY_NO_SPLIT = 4 };
size_t sdp_dft_array_work_bytes(int n_sources);
int sdp_dft_predict_array(int n_time, int n_ant, const double *d_ant_uvw, int n_sources, const float *d_sources,
                          float *d_vis_out, int flags, void *d_work, sdp_stream s);
/* No counterpart in the reference: sdp_dft_predict_array for callers that
 * cannot allocate (code under graph capture, tight serving loops, harnesses
 * that forbid hipMalloc).  Its work buffer is a fixed 2.0 MiB device array
 * of the library, one per GPU, held until the process ends; the call itself
 * allocates nothing.  At most SDP_DFT_NOALLOC_MAX_SOURCES sources (larger
 * counts return SDP_E_ARG).  All calls on one GPU share that array: calls
 * that overlap in time on one GPU (different streams or threads) race, so
 * serialise them.  Results are those of sdp_dft_predict_array bit for bit. */
#define SDP_DFT_NOALLOC_MAX_SOURCES 65536
int sdp_dft_predict_array_noalloc(int n_time, int n_ant, const double *d_ant_uvw, int n_sources,
                                  const float *d_sources, float *d_vis_out, int flags, sdp_stream s);

/* ---- learned minor-cycle proposal ---------------------------------------
 * No reference counterpart (the reference has no learned component).  An
 * encoder-decoder CNN over NHWC float tensors: `levels` levels of two 3x3
 * convolutions + ReLU with `base` << l channels, 2x2 average pooling down,
 * nearest upsampling + 3x3 convolution + additive skip + ReLU + 3x3
 * convolution up, and a 3x3 head to out_ch channels.  Parameters are one
 * float vector, layer by layer (enc0a, enc0b, ..., dec{L-2}u, dec{L-2}d, ...,
 * dec0u, dec0d, head), each layer its weights [3][3][cin][cout] then its
 * biases [cout]. */
enum { SDP_CNN_FP32 = 0,   /* FP32 FMA on the vector ALUs */
       SDP_CNN_BF16 = 1 }; /* BF16 operands on the matrix cores, FP32 accumulation */
typedef struct sdp_cnn_params {
    int levels, base, in_ch, out_ch;
    int height, width, max_batch;   /* activation buffers; height, width divisible by 2^(levels-1) */
    int precision;                  /* SDP_CNN_* */
} sdp_cnn_params;
typedef struct sdp_cnn sdp_cnn;
int sdp_cnn_create(sdp_cnn **net, const sdp_cnn_params *params);
int sdp_cnn_destroy(sdp_cnn *net);
size_t sdp_cnn_param_count(const sdp_cnn *net);
int sdp_cnn_describe(const sdp_cnn *net, sdp_cnn_params *params);
/* host copies of the parameter / gradient vectors (set resets Adam) */
int sdp_cnn_set_params(sdp_cnn *net, const float *h_params);
int sdp_cnn_get_params(const sdp_cnn *net, float *h_params);
int sdp_cnn_get_grads(const sdp_cnn *net, float *h_grads);
/* d_in: batch x height x width x in_ch, d_out: batch x height x width x out_ch */
int sdp_cnn_forward(sdp_cnn *net, const float *d_in, float *d_out, int batch, sdp_stream s);
/* gradients of all parameters (overwritten) for the last forward pass, given dL/d(out) */
int sdp_cnn_backward(sdp_cnn *net, const float *d_grad_out, sdp_stream s);
/* mean squared error over count values; d_grad = 2 (out - target) / count; synchronous */
int sdp_cnn_mse(size_t count, const float *d_out, const float *d_target, float *d_grad, double *loss, sdp_stream s);
/* one Adam step on the current gradients */
int sdp_cnn_adam(sdp_cnn *net, float lr, float beta1, float beta2, float eps, sdp_stream s);
/* Patches of a G x G image for the network.  Candidate patch p of the grid
 * of stride T - 2 margin has origin (x0, y0) = ((p % npx) stride - margin,
 * (p / npx) stride - margin), npx = ceil(G / stride); the centres
 * [margin, T - margin)^2 of the candidates tile the image.
 * sdp_cnn_patch_peaks: largest |residual| over each candidate's T x T tile
 * (d_peaks may be NULL to only count the candidates).
 * sdp_cnn_gather_patches: for n chosen patches (origins: int pairs x0, y0;
 * peaks: each patch's scale, e.g. its tile's largest |residual|),
 * batch[p][i][j] = {residual[y0+i][x0+j] / peak[p],
 * psf[G/2 - T/2 + i][G/2 - T/2 + j]}, zero outside the image.
 * sdp_cnn_scatter_add: model[y0+i][x0+j] += scale * peak[p] * out[p][i][j]
 * over each centre [margin, T - margin)^2 inside the image, where
 * out > floor and, when d_batch_in (the gathered batch) is given, where
 * batch_in[p][i][j][0] > in_floor.  The centres of the n patches must not
 * overlap (no atomics). */
int sdp_cnn_patch_peaks(int grid_size, int patch, int margin, const float *d_residual, float *d_peaks,
                        int *n_patches, sdp_stream s);
int sdp_cnn_gather_patches(int grid_size, int patch, int n, const int *d_origins, const float *d_peaks,
                           const float *d_residual, const float *d_psf, float *d_batch, sdp_stream s);
int sdp_cnn_scatter_add(int grid_size, int patch, int margin, int n, const int *d_origins, const float *d_peaks,
                        const float *d_out, float scale, float floor, const float *d_batch_in, float in_floor,
                        float *d_model, sdp_stream s);
/* sdp_cnn_crop_patches: one channel, out[p][i][j] = image[y0+i][x0+j] / peak[p]
 * (zero outside the image): the training target and the dirty patch of a
 * pair, beside sdp_cnn_gather_patches's network input. */
int sdp_cnn_crop_patches(int grid_size, int patch, int n, const int *d_origins, const float *d_peaks,
                         const float *d_image, float *d_out, sdp_stream s);
/* A G x G sky image (overwritten) for training pairs: n_points point sources
 * (d_points: y, x, flux doubles; pixel row and column) and n_gauss
 * elliptical Gaussians (d_gauss: cy, cx, fwhm_a, fwhm_b, theta, peak
 * doubles), each Gaussian over the square stamp of half side
 * ceil(3 max(fwhm_a, fwhm_b)) pixels around (floor cy, floor cx), with
 * peak exp(-(u^2 / (fwhm_a^2 k) + v^2 / (fwhm_b^2 k)) / 2), k = 1 / (8 ln 2),
 * (u, v) the pixel offset rotated by theta.  Each pixel is summed in double
 * (points first, then Gaussians, each in order) and rounded once.  At most
 * 1024 points and 512 Gaussians.  No reference counterpart: it serves the
 * learned proposal's training. */
int sdp_sky_render(int grid_size, int n_points, const double *d_points, int n_gauss, const double *d_gauss,
                   float *d_sky, sdp_stream s);

#ifdef __cplusplus
}
#endif

#endif /* SDP_DEVICE_H */
