/*
 * sdp_actors.c -- the reference actor ABI (include/sdp_actors.h) on top of
 * the device C-ABI (include/sdp_device.h).
 *
 * Each actor copies its host inputs to device buffers, runs the HIP
 * kernels, and copies the outputs back before returning, as the reference's
 * GPU actors do (src/GPU/dft_run.cu:25-73), but without per-call device
 * allocation: device buffers are kept in a grow-only pool.  A mutex
 * serialises actors because they share that pool (PREESM may call actors
 * from several core threads, P1/generated/main.c:41-79).
 *
 * Set-up actors that only read files or fill small host tables (top.c)
 * stay on the host, as in the reference.
 */
#define _GNU_SOURCE
#include <math.h>
#include <pthread.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "sdp_actors.h"
#include "sdp_device.h"

enum {
    B_IN0, B_IN1, B_IN2, B_IN3, B_OUT0, B_OUT1, B_WORK, B_COUNT
};

static pthread_mutex_t g_lock = PTHREAD_MUTEX_INITIALIZER;
static void *g_buf[B_COUNT];
static size_t g_buf_bytes[B_COUNT];

static void die(const char *what)
{
    fprintf(stderr, "sdp_actors: %s failed: %s\n", what, sdp_last_error());
    abort();
}

#define CHECK(x) do { if ((x) != 0) die(#x); } while (0)

static void lock(void)
{
    pthread_mutex_lock(&g_lock);
    CHECK(sdp_init());
}

static void unlock(void) { pthread_mutex_unlock(&g_lock); }

/* device buffer `slot` of at least `bytes` bytes (contents undefined) */
static void *dev(int slot, size_t bytes)
{
    if (bytes == 0) bytes = 16;
    if (g_buf_bytes[slot] < bytes) {
        if (g_buf[slot]) CHECK(sdp_device_free(g_buf[slot]));
        g_buf[slot] = NULL;
        size_t want = bytes + bytes / 8;
        CHECK(sdp_device_malloc(&g_buf[slot], want));
        g_buf_bytes[slot] = want;
    }
    return g_buf[slot];
}

static void *upload(int slot, const void *host, size_t bytes)
{
    void *d = dev(slot, bytes);
    CHECK(sdp_memcpy_htod(d, host, bytes, NULL));
    return d;
}

static void download(void *host, const void *d, size_t bytes)
{
    CHECK(sdp_memcpy_dtoh(host, d, bytes, NULL));
}

static size_t cells(int G) { return (size_t)G * (size_t)G; }

/* ---- degridgrid.h --------------------------------------------------------- */

void correct_to_finegrid(int NUM_VISIBILITIES, int GRID_SIZE, int OVERSAMPLING_FACTOR,
        int PERFORM_SIMPLIFICATION, PRECISION3 *vis_uvw_coords, PRECISION2 *input_visibilities,
        Config *config, PRECISION2 *output_visibilities, PRECISION3 *output_finegrid_vis_coords,
        int *num_output_visibilities)
{
    const int n = NUM_VISIBILITIES;
    lock();
    float *d_uvw = upload(B_IN0, vis_uvw_coords, sizeof(float3) * (size_t)n);
    float *d_out = dev(B_OUT0, sizeof(float3) * (size_t)n);
    if (PERFORM_SIMPLIFICATION == 1) {
        float *d_vis = upload(B_IN1, input_visibilities, sizeof(float2) * (size_t)n);
        float *d_vout = dev(B_OUT1, sizeof(float2) * (size_t)n);
        int kept = 0;
        CHECK(sdp_finegrid_simplify(n, GRID_SIZE, OVERSAMPLING_FACTOR, config->uv_scale, config->w_scale, d_uvw,
                                    d_vis, d_out, d_vout, &kept, NULL));
        memset(output_visibilities, 0, sizeof(float2) * (size_t)n);
        download(output_visibilities, d_vout, sizeof(float2) * (size_t)kept);
        download(output_finegrid_vis_coords, d_out, sizeof(float3) * (size_t)kept);
        *num_output_visibilities = kept;
    } else {
        CHECK(sdp_finegrid_correct(n, GRID_SIZE, OVERSAMPLING_FACTOR, config->uv_scale, d_uvw, d_out, NULL));
        download(output_finegrid_vis_coords, d_out, sizeof(float3) * (size_t)n);
        if (output_visibilities != input_visibilities)
            memmove(output_visibilities, input_visibilities, sizeof(float2) * (size_t)n);
        *num_output_visibilities = n;
    }
    unlock();
}

static sdp_grid_plan *make_plan(int G, int OS, int n_kernels, int snap, const int2 *supports, const Config *config,
                                const float *d_uvw, int n)
{
    sdp_grid_params p;
    memset(&p, 0, sizeof p);
    p.grid_size = G;
    p.oversampling = OS;
    p.n_kernels = n_kernels;
    p.snap_mode = snap;
    p.uv_scale = config->uv_scale;
    p.w_scale = config->w_scale;
    sdp_grid_plan *plan = NULL;
    CHECK(sdp_grid_plan_create(&plan, &p, (const int *)supports, d_uvw, n, NULL));
    return plan;
}

void std_gridding(int GRID_SIZE, int NUM_VISIBILITIES, int NUM_KERNELS, int TOTAL_KERNEL_SAMPLES,
        int OVERSAMPLING_FACTOR, int BYPASS, int *maj_iter, int *num_corrected_visibilities,
        PRECISION2 *kernels, int2 *kernel_supports, PRECISION3 *corrected_vis_uvw_coords,
        PRECISION2 *visibilities, Config *config, PRECISION2 *prev_grid, PRECISION2 *output_grid)
{
    (void)NUM_VISIBILITIES;
    if (BYPASS && *maj_iter > 0) {   /* degridgrid.c:75-79 */
        memmove(output_grid, prev_grid, sizeof(float2) * cells(GRID_SIZE));
        return;
    }
    const int n = *num_corrected_visibilities;
    lock();
    float *d_uvw = upload(B_IN0, corrected_vis_uvw_coords, sizeof(float3) * (size_t)n);
    float *d_vis = upload(B_IN1, visibilities, sizeof(float2) * (size_t)n);
    float *d_k = upload(B_IN2, kernels, sizeof(float2) * (size_t)TOTAL_KERNEL_SAMPLES);
    float *d_grid = dev(B_OUT0, sizeof(float2) * cells(GRID_SIZE));
    sdp_grid_plan *plan = make_plan(GRID_SIZE, OVERSAMPLING_FACTOR, NUM_KERNELS, SDP_SNAP_STD, kernel_supports,
                                    config, d_uvw, n);
    CHECK(sdp_grid(plan, d_k, d_vis, SDP_ORDER_ORIGINAL, d_grid, NULL));
    download(output_grid, d_grid, sizeof(float2) * cells(GRID_SIZE));
    sdp_grid_plan_destroy(plan);
    unlock();
}

void std_degridding(int GRID_SIZE, int NUM_VISIBILITIES, int NUM_KERNELS, int TOTAL_KERNEL_SAMPLES,
        int OVERSAMPLING_FACTOR, PRECISION2 *kernels, int2 *kernel_supports, PRECISION2 *input_grid,
        PRECISION3 *corrected_vis_uvw_coords, int *num_corrected_visibilities, Config *config,
        PRECISION2 *output_visibilities)
{
    const int n = *num_corrected_visibilities;
    const int n_out = NUM_VISIBILITIES > n ? NUM_VISIBILITIES : n;
    lock();
    float *d_grid = upload(B_IN0, input_grid, sizeof(float2) * cells(GRID_SIZE));
    float *d_uvw = upload(B_IN1, corrected_vis_uvw_coords, sizeof(float3) * (size_t)n);
    float *d_out = dev(B_OUT0, sizeof(float2) * (size_t)n_out);
    if (sdp_get_degrid_mode() == SDP_DEGRID_WEIGHTED) {
        float *d_k = upload(B_IN2, kernels, sizeof(float2) * (size_t)TOTAL_KERNEL_SAMPLES);
        sdp_grid_plan *plan = make_plan(GRID_SIZE, OVERSAMPLING_FACTOR, NUM_KERNELS, SDP_SNAP_STD, kernel_supports,
                                        config, d_uvw, n);
        CHECK(sdp_degrid(plan, d_k, d_grid, NULL, SDP_ORDER_ORIGINAL, d_out, NULL));
        sdp_grid_plan_destroy(plan);
        download(output_visibilities, d_out, sizeof(float2) * (size_t)(n < NUM_VISIBILITIES ? n : NUM_VISIBILITIES));
        if (NUM_VISIBILITIES > n)   /* degridgrid.c:132: entries past the corrected count stay 0 */
            memset(output_visibilities + n, 0, sizeof(float2) * (size_t)(NUM_VISIBILITIES - n));
        unlock();
        return;
    }
    CHECK(sdp_degrid_nearest(n, n_out, GRID_SIZE, config->uv_scale, d_grid, d_uvw, d_out, NULL));
    download(output_visibilities, d_out, sizeof(float2) * (size_t)NUM_VISIBILITIES);
    unlock();
}

/* ---- gridding_run.h --------------------------------------------------------- */

void gridding_actor(int GRID_SIZE, int NUM_VISIBILITIES, int NUM_KERNELS, int TOTAL_KERNEL_SAMPLES,
        PRECISION2 *kernels, int2 *kernel_supports, PRECISION3 *vis_uvw_coords,
        PRECISION2 *visibilities, Config *config, PRECISION2 *uv_grid)
{
    const int n = NUM_VISIBILITIES;
    lock();
    float *d_uvw = upload(B_IN0, vis_uvw_coords, sizeof(float3) * (size_t)n);
    float *d_vis = upload(B_IN1, visibilities, sizeof(float2) * (size_t)n);
    float *d_k = upload(B_IN2, kernels, sizeof(float2) * (size_t)TOTAL_KERNEL_SAMPLES);
    float *d_grid = dev(B_OUT0, sizeof(float2) * cells(GRID_SIZE));
    /* gridding_run.c:43 passes an oversampling of 16 regardless of config */
    sdp_grid_plan *plan = make_plan(GRID_SIZE, 16, NUM_KERNELS, SDP_SNAP_GRIDDING_ACTOR, kernel_supports, config,
                                    d_uvw, n);
    CHECK(sdp_grid(plan, d_k, d_vis, SDP_ORDER_ORIGINAL, d_grid, NULL));
    download(uv_grid, d_grid, sizeof(float2) * cells(GRID_SIZE));
    sdp_grid_plan_destroy(plan);
    unlock();
}

void add_visibilities(int NUM_VISIBILITIES, PRECISION2 *v1, PRECISION2 *v2, PRECISION2 *output)
{
    const size_t b = sizeof(float2) * (size_t)NUM_VISIBILITIES;
    lock();
    float *a = upload(B_IN0, v1, b), *c = upload(B_IN1, v2, b), *o = dev(B_OUT0, b);
    CHECK(sdp_vis_add(NUM_VISIBILITIES, a, c, o, NULL));
    download(output, o, b);
    unlock();
}

/* ---- dft_run.h ---------------------------------------------------------------- */

void dft_actor(int NUM_VISIBILITIES, int NUM_MAX_SOURCES, PRECISION3 *sources, PRECISION3 *vis_uvw_coords,
        int *num_sources, Config *config, PRECISION2 *visibilities)
{
    (void)NUM_MAX_SOURCES; (void)config;
    const int n = NUM_VISIBILITIES, ns = *num_sources;
    lock();
    float *d_src = upload(B_IN0, sources, sizeof(float3) * (size_t)ns);
    float *d_uvw = upload(B_IN1, vis_uvw_coords, sizeof(float3) * (size_t)n);
    float *d_out = dev(B_OUT0, sizeof(float2) * (size_t)n);
    CHECK(sdp_dft_predict(n, ns, d_src, d_uvw, d_out, sdp_get_dft_mode(), NULL));
    download(visibilities, d_out, sizeof(float2) * (size_t)n);
    unlock();
}

/* ---- deconvolution_run.h ------------------------------------------------------ */

void hogbom_clean(int GRID_SIZE, int NUM_MINOR_CYCLES, int MAX_SOURCES, PRECISION *residual,
        PRECISION *partial_psf, int2 *partial_psf_halfdims, Config *config, PRECISION *current_model,
        int *num_sources_out, PRECISION3 *sources_out, PRECISION *output_model)
{
    (void)partial_psf_halfdims;   /* ignored by the reference too, deconvolution_run.c:259-264 */
    const size_t n = cells(GRID_SIZE);
    lock();
    float *d_res = upload(B_IN0, residual, sizeof(float) * n);
    float *d_psf = dev(B_IN1, sizeof(float) * (n + 1));
    CHECK(sdp_memcpy_htod(d_psf, partial_psf, sizeof(float) * n, NULL));
    /* the reference reads psf[G*G] for some peaks (deconvolution_run.c:269); it is 0 here */
    {
        const float zero = 0.0f;
        CHECK(sdp_memcpy_htod(d_psf + n, &zero, sizeof(float), NULL));
    }
    float *d_model_in = upload(B_IN2, current_model, sizeof(float) * n);
    float *d_model = dev(B_OUT0, sizeof(float) * n);
    float *d_src = dev(B_OUT1, sizeof(float3) * (size_t)(MAX_SOURCES > 0 ? MAX_SOURCES : 1) + 16);
    size_t wb = sdp_hogbom_work_bytes(GRID_SIZE);
    char *d_work = dev(B_WORK, wb + 64);
    int *d_num = (int *)(d_work + ((wb + 15) / 16) * 16);
    CHECK(sdp_hogbom(GRID_SIZE, NUM_MINOR_CYCLES, MAX_SOURCES, d_res, d_psf, config->psf_max_value,
                     config->loop_gain, config->noise_factor, config->weak_source_percent_img, config->cell_size,
                     d_model_in, d_model, d_src, d_num, d_work, NULL));
    int ns = 0;
    download(&ns, d_num, sizeof(int));
    download(output_model, d_model, sizeof(float) * n);
    if (ns > 0) download(sources_out, d_src, sizeof(float3) * (size_t)ns);
    *num_sources_out = ns;
    unlock();
}

/* ---- fft_run.h ------------------------------------------------------------------ */

void fft_shift_complex_to_complex_actor(int GRID_SIZE, PRECISION2 *uv_grid_in, Config *config, PRECISION2 *uv_grid_out)
{
    (void)config;
    const size_t b = sizeof(float2) * cells(GRID_SIZE);
    lock();
    float *d = upload(B_IN0, uv_grid_in, b);
    CHECK(sdp_shift_c2c(GRID_SIZE, d, d, NULL));
    download(uv_grid_out, d, b);
    unlock();
}

static void fft_actor(int G, const float2 *in, float2 *out, int direction)
{
    const size_t b = sizeof(float2) * cells(G);
    lock();
    float *d = upload(B_IN0, in, b);
    void *w = dev(B_WORK, sdp_fft_work_bytes(G));
    CHECK(sdp_fft2d(d, d, G, direction, 0, NULL, w, NULL));
    download(out, d, b);
    unlock();
}

void CUFFT_EXECUTE_FORWARD_C2C_actor(int GRID_SIZE, PRECISION2 *uv_grid_in, PRECISION2 *uv_grid_out)
{
    fft_actor(GRID_SIZE, uv_grid_in, uv_grid_out, SDP_FFT_FORWARD);
}

void CUFFT_EXECUTE_INVERSE_C2C_actor(int GRID_SIZE, PRECISION2 *uv_grid_in, PRECISION2 *uv_grid_out)
{
    fft_actor(GRID_SIZE, uv_grid_in, uv_grid_out, SDP_FFT_INVERSE);
}

void fft_shift_complex_to_real_actor(int GRID_SIZE, PRECISION2 *uv_grid, Config *config, PRECISION *dirty_image)
{
    (void)config;
    lock();
    float *d = upload(B_IN0, uv_grid, sizeof(float2) * cells(GRID_SIZE));
    float *o = dev(B_OUT0, sizeof(float) * cells(GRID_SIZE));
    CHECK(sdp_shift_c2r(GRID_SIZE, d, o, NULL));
    download(dirty_image, o, sizeof(float) * cells(GRID_SIZE));
    unlock();
}

void fft_shift_real_to_complex_actor(int GRID_SIZE, PRECISION *image, Config *config, PRECISION2 *fourier)
{
    (void)config;
    lock();
    float *d = upload(B_IN0, image, sizeof(float) * cells(GRID_SIZE));
    float *o = dev(B_OUT0, sizeof(float2) * cells(GRID_SIZE));
    CHECK(sdp_shift_r2c(GRID_SIZE, d, o, NULL));
    download(fourier, o, sizeof(float2) * cells(GRID_SIZE));
    unlock();
}

/* ---- convolution_correction_run.h ---------------------------------------------- */

void execute_convolution_correction_actor(int GRID_SIZE, PRECISION *dirty_image_in, PRECISION *prolate,
        Config *config, PRECISION *dirty_image_out)
{
    (void)config;
    const size_t n = cells(GRID_SIZE);
    const int half = GRID_SIZE / 2;
    lock();
    float *d_in = upload(B_IN0, dirty_image_in, sizeof(float) * n);
    float *d_taper = dev(B_IN1, sizeof(float) * (size_t)(half + 1));
    float *h_taper = (float *)malloc(sizeof(float) * (size_t)(half + 1));
    memcpy(h_taper, prolate, sizeof(float) * (size_t)half);
    h_taper[half] = 0.0f;   /* the element read one past the end (convolution_correction_run.c:23) */
    CHECK(sdp_memcpy_htod(d_taper, h_taper, sizeof(float) * (size_t)(half + 1), NULL));
    free(h_taper);
    float *d_out = dev(B_OUT0, sizeof(float) * n);
    CHECK(sdp_convolution_correct(GRID_SIZE, d_in, d_taper, d_out, NULL));
    download(dirty_image_out, d_out, sizeof(float) * n);
    unlock();
}

void execute_inv_convolution_correction_actor(int GRID_SIZE, PRECISION *dirty_image_in, PRECISION *prolate,
        Config *config, PRECISION *dirty_image_out)
{
    (void)prolate; (void)config;
    /* convolution_correction_run.c:65: an identity copy */
    memmove(dirty_image_out, dirty_image_in, sizeof(float) * cells(GRID_SIZE));
}

/* ---- gains_apply_run.h ---------------------------------------------------------- */

static void gains_actor(int NUM_RECEIVERS, int NUM_BASELINES, int NUM_VISIBILITIES, const float2 *measured,
        const float2 *pred, const float2 *gains, const int2 *pairs, int apply_only, float2 *out)
{
    const size_t b = sizeof(float2) * (size_t)NUM_VISIBILITIES;
    lock();
    float *d_m = upload(B_IN0, measured, b);
    float *d_p = apply_only ? d_m : upload(B_IN1, pred, b);
    float *d_g = upload(B_IN2, gains, sizeof(float2) * (size_t)NUM_RECEIVERS);
    int *d_pairs = upload(B_IN3, pairs, sizeof(int2) * (size_t)NUM_BASELINES);
    float *d_o = dev(B_OUT0, b);
    CHECK(sdp_vis_subtract_gains(NUM_VISIBILITIES, NUM_BASELINES, d_m, d_p, d_g, d_pairs, apply_only, d_o, NULL));
    download(out, d_o, b);
    unlock();
}

void reciprocal_transform_actor(int NUM_RECEIVERS, PRECISION2 *gains_in, Config *config, PRECISION2 *gains_out)
{
    (void)config;
    const size_t b = sizeof(float2) * (size_t)NUM_RECEIVERS;
    lock();
    float *d_i = upload(B_IN0, gains_in, b);
    float *d_o = dev(B_OUT0, b);
    CHECK(sdp_gains_reciprocal(NUM_RECEIVERS, d_i, d_o, NULL));
    download(gains_out, d_o, b);
    unlock();
}

void subtract_from_measurements(int NUM_RECEIVERS, int NUM_BASELINES, int NUM_VISIBILITIES,
        PRECISION2 *measured_vis, PRECISION2 *visibilities_in, PRECISION2 *gains, int2 *receiver_pairs,
        Config *config, PRECISION2 *visibilities_out)
{
    (void)config;
    gains_actor(NUM_RECEIVERS, NUM_BASELINES, NUM_VISIBILITIES, measured_vis, visibilities_in, gains,
                receiver_pairs, 0, visibilities_out);
}

void apply_gains_only(int NUM_RECEIVERS, int NUM_BASELINES, int NUM_VISIBILITIES, PRECISION2 *measured_vis,
        PRECISION2 *gains, int2 *receiver_pairs, Config *config, PRECISION2 *visibilities_out)
{
    (void)config;
    gains_actor(NUM_RECEIVERS, NUM_BASELINES, NUM_VISIBILITIES, measured_vis, NULL, gains, receiver_pairs, 1,
                visibilities_out);
}

void subtract_from_measurements_finegrid(int NUM_VISIBILITIES, PRECISION2 *gridded_measured_vis,
        PRECISION2 *input_vis, int *num_finegrid_vis, PRECISION2 *delta_vis)
{
    (void)NUM_VISIBILITIES;
    const int n = *num_finegrid_vis;
    if (n <= 0) return;
    const size_t b = sizeof(float2) * (size_t)n;
    lock();
    float *a = upload(B_IN0, gridded_measured_vis, b), *c = upload(B_IN1, input_vis, b), *o = dev(B_OUT0, b);
    CHECK(sdp_vis_subtract(n, a, c, o, NULL));
    download(delta_vis, o, b);
    unlock();
}

/* ---- splitmerge.h --------------------------------------------------------------- */

void split_visibilities(Config *config, int NUM_VISIBILITIES, int SLICES, PRECISION2 *measured_visibilities,
        PRECISION3 *vis_coords, PRECISION2 *visibilities, PRECISION3 *out_vis_coords, Config *output_configs,
        Config *output_config)
{
    memmove(visibilities, measured_visibilities, sizeof(float2) * (size_t)NUM_VISIBILITIES);
    memmove(out_vis_coords, vis_coords, sizeof(float3) * (size_t)NUM_VISIBILITIES);
    for (int i = 0; i < SLICES; ++i) output_configs[i] = *config;
    *output_config = *config;
}

void merge_gridded_visibilities(int GRID_SIZE, int SLICES, PRECISION2 *grids, PRECISION2 *output_grid)
{
    const size_t b = sizeof(float2) * cells(GRID_SIZE);
    lock();
    float *d = upload(B_IN0, grids, b * (size_t)SLICES);
    float *o = dev(B_OUT0, b);
    CHECK(sdp_merge_grids(GRID_SIZE, SLICES, d, o, NULL));
    download(output_grid, o, b);
    unlock();
}

void merge_configs(Config *in_configs, Config *out_config) { *out_config = in_configs[0]; }

void split_fftdegrid_visibilities(int NUM_VISIBILITIES, int SLICES, int GRID_SIZE, Config *config,
        PRECISION3 *vis_coords, PRECISION2 *measured_visibilities, int *num_corrected_vis, int *maj_iter,
        PRECISION2 *output_visibilities, PRECISION3 *out_vis_coords, Config *output_configs, int *num_vis,
        int *maj_iter_out)
{
    (void)GRID_SIZE;
    memmove(out_vis_coords, vis_coords, sizeof(float3) * (size_t)NUM_VISIBILITIES);
    memmove(output_visibilities, measured_visibilities, sizeof(float2) * (size_t)NUM_VISIBILITIES);
    for (int i = 0; i < SLICES; ++i) {
        output_configs[i] = *config;
        maj_iter_out[i] = *maj_iter;
    }
    /* splitmerge.c:40-47: NUM_VISIBILITIES / SLICES per slice, the tail short */
    int left = *num_corrected_vis;
    const int ideal = NUM_VISIBILITIES / SLICES;
    for (int i = 0; i < SLICES; ++i) {
        const int cur = left > ideal ? ideal : left;
        left -= cur;
        num_vis[i] = cur;
    }
}

void split_uvgrid(int GRID_SIZE, int SLICES, PRECISION2 *input_grid, PRECISION2 *output_grid)
{
    for (int i = 0; i < SLICES; ++i)
        memmove(output_grid + (size_t)i * cells(GRID_SIZE), input_grid, sizeof(float2) * cells(GRID_SIZE));
}

/* ---- top.h ----------------------------------------------------------------------- */

void config_struct_set_up(int GRID_SIZE, int NUM_KERNELS, Config *config_struct)
{
    /* the constants of top.c:9-69; file names are those of the reference layout */
    config_struct->gpu_max_threads_per_block = 1024;
    config_struct->gpu_max_threads_per_block_dimension = 32;
    config_struct->cell_size = 8.52211548825356E-06;
    config_struct->frequency_hz = 299792458.0;
    config_struct->dirty_image_output = "gleam_small_dirty_image.csv";
    config_struct->right_ascension = true;
    config_struct->visibility_source_file = "../data/input/GLEAM_small_visibilities_corrupted.csv";
    config_struct->output_path = "../data/output/small/";
    config_struct->perform_system_test = false;
    config_struct->system_test_image = "";
    config_struct->system_test_sources = "";
    config_struct->system_test_visibilities = "";
    config_struct->default_gains_file = "../data/input/TrueGainsNotRotated.csv";
    config_struct->output_gains_file = "estimated_gains.csv";
    config_struct->use_default_gains = true;
    config_struct->max_w = 1895.410847844;
    config_struct->w_scale = pow(NUM_KERNELS - 1, 2.0) / config_struct->max_w;
    config_struct->oversampling = 16;
    config_struct->uv_scale = config_struct->cell_size * GRID_SIZE;
    config_struct->kernel_real_file = "../data/input/kernels/new/wproj_manualconj_gridding_kernels_real_x16.csv";
    config_struct->kernel_imag_file = "../data/input/kernels/new/wproj_manualconj_gridding_kernels_imag_x16.csv";
    config_struct->kernel_support_file = "../data/input/kernels/new/wproj_manualconj_gridding_kernel_supports_x16.csv";
    config_struct->degridding_kernel_real_file = "../data/input/kernels/new/wproj_manualconj_degridding_kernels_real_x16.csv";
    config_struct->degridding_kernel_imag_file = "../data/input/kernels/new/wproj_manualconj_degridding_kernels_imag_x16.csv";
    config_struct->degridding_kernel_support_file = "../data/input/kernels/new/wproj_manualconj_degridding_kernel_supports_x16.csv";
    config_struct->force_weight_to_one = true;
    config_struct->loop_gain = 0.1;
    config_struct->weak_source_percent_gc = 0.005;
    config_struct->weak_source_percent_img = 0.0002;
    config_struct->psf_max_value = 0.0;
    config_struct->noise_factor = 1.5;
    config_struct->model_sources_output = "sample_model_sources.csv";
    config_struct->residual_image_output = "residual_image.csv";
    config_struct->psf_input_file = "../data/input/gleam_small_psf.csv";
    config_struct->psf_image_output = "dirty_psf.csv";
    config_struct->clean_psf_image_output = "clean_psf.csv";
    config_struct->model_image_output = "model.csv";
    config_struct->final_image_output = "deconvolved.csv";
    config_struct->predicted_vis_output = "predicted_visibilities.csv";
}

static void load_kernels(int NUM_KERNELS, const Config *config, const char *sup_file, const char *re_file,
                         const char *im_file, int2 *supports, float2 *kernels)
{
    /* top.c:235-315: supports file holds one half support per plane; the
     * sample offset of plane p is the sum of ((h_q + 1) * OS)^2 over q < p;
     * the real and imaginary files hold the samples plane after plane. */
    FILE *fs = fopen(sup_file, "r");
    if (!fs) {
        fprintf(stderr, "sdp_actors: cannot open kernel support file %s\n", sup_file);
        exit(EXIT_FAILURE);
    }
    int total = 0;
    for (int p = 0; p < NUM_KERNELS; ++p) {
        if (fscanf(fs, "%d\n", &supports[p].x) != 1) supports[p].x = 0;
        supports[p].y = total;
        total += (int)pow((supports[p].x + 1) * config->oversampling, 2.0);
    }
    fclose(fs);
    FILE *fr = fopen(re_file, "r"), *fi = fopen(im_file, "r");
    if (!fr || !fi) {
        if (fr) fclose(fr);
        if (fi) fclose(fi);
        fprintf(stderr, "sdp_actors: cannot open kernel files %s / %s\n", re_file, im_file);
        exit(EXIT_FAILURE);
    }
    int k = 0;
    for (int p = 0; p < NUM_KERNELS; ++p) {
        const int ns = (int)pow((supports[p].x + 1) * config->oversampling, 2.0);
        for (int q = 0; q < ns; ++q, ++k) {
            double re = 0.0, im = 0.0;
            if (fscanf(fr, "%lf ", &re) != 1) re = 0.0;
            if (fscanf(fi, "%lf ", &im) != 1) im = 0.0;
            kernels[k].x = (float)re;
            kernels[k].y = (float)im;
        }
    }
    fclose(fr);
    fclose(fi);
}

void kernel_host_set_up(int NUM_KERNELS, int TOTAL_KERNEL_SAMPLES, Config *config, int2 *kernel_supports,
        PRECISION2 *kernels)
{
    (void)TOTAL_KERNEL_SAMPLES;
    load_kernels(NUM_KERNELS, config, config->kernel_support_file, config->kernel_real_file,
                 config->kernel_imag_file, kernel_supports, kernels);
}

void degridding_kernel_host_set_up(int NUM_DEGRIDDING_KERNELS, int TOTAL_DEGRIDDING_KERNEL_SAMPLES, Config *config,
        int2 *degridding_kernel_supports, PRECISION2 *degridding_kernels)
{
    (void)TOTAL_DEGRIDDING_KERNEL_SAMPLES;
    load_kernels(NUM_DEGRIDDING_KERNELS, config, config->degridding_kernel_support_file,
                 config->degridding_kernel_real_file, config->degridding_kernel_imag_file,
                 degridding_kernel_supports, degridding_kernels);
}

double prolate_spheroidal(double nu)
{
    /* Schwab's rational approximation of the zero-order prolate spheroidal
     * wave function (top.c:408-447); the coefficient table is the published
     * one and is reproduced as such. */
    static const double p[] = {0.08203343, -0.3644705, 0.627866, -0.5335581, 0.2312756,
                               0.004028559, -0.03697768, 0.1021332, -0.1201436, 0.06412774};
    static const double q[] = {1.0, 0.8212018, 0.2078043, 1.0, 0.9599102, 0.2918724};
    int part;
    double nu_end;
    if (nu >= 0.0 && nu < 0.75) { part = 0; nu_end = 0.75; }
    else if (nu >= 0.75 && nu < 1.0) { part = 1; nu_end = 1.0; }
    else return 0.0;
    const double d = nu * nu - nu_end * nu_end;
    double num = p[5 * part], den = q[3 * part];
    for (int i = 1; i < 5; i++) num += p[5 * part + i] * pow(d, i);
    for (int i = 1; i < 3; i++) den += q[3 * part + i] * pow(d, i);
    return den == 0.0 ? 0.0 : num / den;
}

void create_1D_half_prolate(PRECISION *prolate, int grid_size)
{
    const int half = grid_size / 2;
    for (int i = 0; i < half; ++i) prolate[i] = (float)prolate_spheroidal((double)i / (double)half);
}

void correction_set_up(int GRID_SIZE, PRECISION *prolate)
{
    if (prolate == NULL) exit(EXIT_FAILURE);
    create_1D_half_prolate(prolate, GRID_SIZE);
}

void psf_host_set_up(int GRID_SIZE, int PSF_GRID_SIZE, Config *config, PRECISION *psf, double *psf_max_value)
{
    /* top.c:449-512: read a PSF_GRID_SIZE^2 text file, track the largest
     * sample, centre-crop (or pad) to GRID_SIZE^2, divide by the maximum. */
    const size_t full_n = (size_t)PSF_GRID_SIZE * PSF_GRID_SIZE;
    float *full = (float *)calloc(full_n, sizeof(float));
    if (psf == NULL || full == NULL) exit(EXIT_FAILURE);
    memset(psf, 0, sizeof(float) * cells(GRID_SIZE));
    FILE *f = fopen(config->psf_input_file, "r");
    if (!f) {
        perror("PSF: ");
        exit(EXIT_FAILURE);
    }
    for (size_t i = 0; i < full_n; i++) {
        double v = 0.0;
        if (fscanf(f, "%lf ", &v) != 1) v = 0.0;
        full[i] = (float)v;
        if (v > *psf_max_value) *psf_max_value = (float)v;
    }
    fclose(f);
    const int small = PSF_GRID_SIZE < GRID_SIZE ? PSF_GRID_SIZE : GRID_SIZE;
    const int psf_off = (PSF_GRID_SIZE - GRID_SIZE) / 2 > 0 ? (PSF_GRID_SIZE - GRID_SIZE) / 2 : 0;
    const int grid_off = (GRID_SIZE - PSF_GRID_SIZE) / 2 > 0 ? (GRID_SIZE - PSF_GRID_SIZE) / 2 : 0;
    for (int y = 0; y < small; ++y)
        for (int x = 0; x < small; ++x)
            psf[(size_t)(y + grid_off) * GRID_SIZE + (x + grid_off)] =
                full[(size_t)(y + psf_off) * PSF_GRID_SIZE + (x + psf_off)];
    free(full);
    for (size_t i = 0; i < cells(GRID_SIZE); i++) psf[i] = (float)((double)psf[i] / *psf_max_value);
}

void visibility_host_set_up(int NUM_VISIBILITIES, Config *config, PRECISION3 *vis_uvw_coords,
        PRECISION2 *measured_vis)
{
    /* top.c:140-232: the leading count is read and not used; each line's
     * values are doubles, scaled and weighted in double, stored as float */
    if (vis_uvw_coords == NULL || measured_vis == NULL) exit(EXIT_FAILURE);
    FILE *f = fopen(config->visibility_source_file, "r");
    if (!f) {
        fprintf(stderr, "ERROR >>> Unable to open visibility file %s\n", config->visibility_source_file);
        exit(EXIT_FAILURE);
    }
    int count = 0;
    if (fscanf(f, "%d", &count) != 1) count = 0;
    (void)count;
    const double to_wavelengths = config->frequency_hz / 299792458.0;   /* SPEED_OF_LIGHT, common.h:59 */
    for (int i = 0; i < NUM_VISIBILITIES; ++i) {
        double u = 0.0, v = 0.0, w = 0.0, re = 0.0, im = 0.0, weight = 0.0;
        if (fscanf(f, "%lf %lf %lf %lf %lf %lf", &u, &v, &w, &re, &im, &weight) != 6) {
            fprintf(stderr, "ERROR >>> visibility file %s: line %d unreadable\n", config->visibility_source_file,
                    i + 2);
            fclose(f);
            exit(EXIT_FAILURE);
        }
        if (config->right_ascension) {
            u = -u;
            w = -w;
        }
        if (!config->force_weight_to_one) {
            re *= weight;
            im *= weight;
        }
        vis_uvw_coords[i].x = (float)(u * to_wavelengths);
        vis_uvw_coords[i].y = (float)(v * to_wavelengths);
        vis_uvw_coords[i].z = (float)(w * to_wavelengths);
        measured_vis[i].x = (float)re;
        measured_vis[i].y = (float)im;
    }
    fclose(f);
}

/* conv2D, major_loop_iter.c:8-46, on the GPU */
static void restore_image(int G, const float *model, const float *clean_psf, const int2 *halfdims, float *out)
{
    const size_t b = sizeof(float) * cells(G);
    lock();
    float *d_m = upload(B_IN0, model, b);
    float *d_f = upload(B_IN1, clean_psf, b);
    float *d_o = dev(B_OUT0, b);
    CHECK(sdp_restore_conv2d(G, d_m, d_f, halfdims->x, halfdims->y, d_o, NULL));
    download(out, d_o, b);
    unlock();
}

/* gaussian2D, top.c:520-523: float arithmetic, a double exp, PI = 3.141592654f */
static float clean_gaussian(float x, float y, float x_sigma, float y_sigma)
{
    const float pi = 3.141592654f;
    return (float)(exp(-(x * x) / (2 * x_sigma * x_sigma) - (y * y) / (2 * y_sigma * y_sigma)) /
                   (2 * pi * x_sigma * y_sigma));
}

void clean_psf_host_set_up(int GRID_SIZE, int GAUSSIAN_CLEAN_PSF, Config *config, PRECISION *dirty_psf,
        PRECISION *clean_psf, int2 *partial_psf_halfdims)
{
    /* top.c:527-592.  From (G/2 + 1, G/2 + 1) the dirty PSF is walked right,
     * down, left and up; each walk stops at the first step that does not
     * decrease (the reference compares with a sign it never updates), and a
     * walk that never stops keeps its initial extent (G/2 rightwards, 0
     * otherwise).  The largest extent h is both half sizes; over the box
     * [-h, h)^2 the clean PSF is a squared Gaussian (sigma h / 2),
     * normalised to sum to psf_max_value, zero elsewhere. */
    (void)GAUSSIAN_CLEAN_PSF;
    if (dirty_psf == NULL || clean_psf == NULL || partial_psf_halfdims == NULL) exit(EXIT_FAILURE);
    const int G = GRID_SIZE, cx = G / 2 + 1, cy = G / 2 + 1;
    const int dirs[4][2] = {{1, 0}, {0, 1}, {-1, 0}, {0, -1}};
    int box[4] = {G / 2, 0, 0, 0};
    for (int i = 0; i < 4; ++i) {
        for (int j = 1; j < G / 2; ++j) {
            const int cur = (cx + dirs[i][0] * j) + (cy + dirs[i][1] * j) * G;
            const int last = (cx + dirs[i][0] * (j - 1)) + (cy + dirs[i][1] * (j - 1)) * G;
            if (dirty_psf[cur] - dirty_psf[last] >= 0.f) {
                box[i] = j;
                break;
            }
        }
    }
    int h = box[0];
    for (int i = 1; i < 4; ++i) h = box[i] > h ? box[i] : h;
    /* a walk that never stops leaves h = G/2, and the box (and conv2D's
     * flipped reads) would run past the image: h is capped at G/2 - 2 */
    if (h > G / 2 - 2) h = G / 2 - 2;
    partial_psf_halfdims->x = partial_psf_halfdims->y = h;
    memset(clean_psf, 0, sizeof(float) * cells(G));
    const float psf_max = (float)config->psf_max_value;
    const float sigma = h / 2.f;
    float sum = 0.f;
    for (int y = -h; y < h; ++y)
        for (int x = -h; x < h; ++x) {
            const int idx = cx + x + (cy + y) * G;
            clean_psf[idx] = clean_gaussian((float)x, (float)y, sigma, sigma);
            clean_psf[idx] *= clean_psf[idx];
            sum += clean_psf[idx];
        }
    for (int y = -h; y < h; ++y)
        for (int x = -h; x < h; ++x) {
            const int idx = cx + x + (cy + y) * G;
            clean_psf[idx] = clean_psf[idx] / sum * psf_max;
        }
}

/* save_image_to_file, major_loop_iter.c:168-199: "<output_path>cycle_<n>_<name>",
 * "%f, " per pixel, one row per line; a file that cannot be opened is reported
 * and skipped */
static void save_image(const Config *config, const float *image, const char *name, int G, int cycle)
{
    char path[256];
    snprintf(path, sizeof(path), "%scycle_%d_%s", config->output_path ? config->output_path : "", cycle,
             name ? name : "image.csv");
    FILE *f = fopen(path, "w");
    if (f == NULL) {
        fprintf(stderr, ">>> ERROR: Unable to save image to file %s\n", path);
        return;
    }
    for (int row = 0; row < G; ++row) {
        for (int col = 0; col < G; ++col) fprintf(f, "%f, ", image[(size_t)row * G + col]);
        fprintf(f, "\n");
    }
    fclose(f);
}

void save_output(int GRID_SIZE, PRECISION *residual, PRECISION *model, PRECISION *clean_psf,
        int2 *clean_psf_halfdims, PRECISION *psf, Config *config, int *cycle)
{
    /* major_loop_iter.c:253-278: the residual, dirty PSF, clean PSF and model
     * images, then the restored image (conv2D of the model with the clean
     * PSF, on the GPU) */
    const int G = GRID_SIZE;
    save_image(config, residual, config->residual_image_output, G, *cycle);
    save_image(config, psf, config->psf_image_output, G, *cycle);
    save_image(config, clean_psf, config->clean_psf_image_output, G, *cycle);
    save_image(config, model, config->model_image_output, G, *cycle);
    float *final_image = (float *)malloc(sizeof(float) * cells(G));
    if (final_image == NULL) exit(EXIT_FAILURE);
    restore_image(G, model, clean_psf, clean_psf_halfdims, final_image);
    save_image(config, final_image, config->final_image_output, G, *cycle);
    free(final_image);
}

void config_struct_set_up_sequel(Config *config_in, double *psf_max_value, Config *config_out)
{
    memmove(config_out, config_in, sizeof(Config));
    config_out->psf_max_value = *psf_max_value;
}
