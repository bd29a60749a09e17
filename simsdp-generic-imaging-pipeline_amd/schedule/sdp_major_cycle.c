/*
 * sdp_major_cycle -- the FFT imaging pipeline's generated single-core
 * schedule, restated as a plain C program over the actor C-ABI
 * (include/sdp_actors.h) and linked against libsdp_actors.so.
 *
 * The actor sequence and its arguments follow the PREESM-generated schedule
 * param_code/code_fft_1core/generated/Core0.c:
 *   set-up      correction_set_up (:1856), clean_psf_host_set_up (:2358)
 *   per cycle   execute_inv_convolution_correction_actor (:2588)
 *               reciprocal_transform_actor (:2866)
 *               fft_shift_real_to_complex_actor -> CUFFT_EXECUTE_FORWARD_C2C_actor
 *                 -> fft_shift_complex_to_complex_actor -> split_uvgrid (:3026-3034)
 *               apply_gains_only -> correct_to_finegrid -> split_fftdegrid_visibilities (:3080-3125)
 *               std_degridding x SLICES (:3955-3973)
 *               subtract_from_measurements_finegrid x SLICES (:3979-3988)
 *               std_gridding x SLICES, BYPASS 0 (:3991-4009)
 *               merge_gridded_visibilities (:4024)
 *               fft_shift_complex_to_complex_actor -> CUFFT_EXECUTE_INVERSE_C2C_actor
 *                 -> fft_shift_complex_to_real_actor (:4025-4030)
 *               execute_convolution_correction_actor (:4033)
 *               hogbom_clean (:4045); its model is the next cycle's image
 *               estimate (the FIFO of :4810), the first estimate is zero.
 * Each slice actor works on its own contiguous NUM_VISIBILITIES / SLICES
 * part of the split buffers, as the schedule's explode forks hand them out.
 *
 * Usage: sdp_major_cycle CASE_DIR
 *   CASE_DIR/case.txt     "key value" lines: grid_size num_vis num_kernels
 *                         total_kernel_samples num_receivers num_baselines
 *                         slices num_major num_minor max_sources oversampling
 *   CASE_DIR/config.bin   the Config record (sizeof(Config) bytes)
 *   CASE_DIR/{uvw,vis,gains,pairs,kernels,supports,dkernels,dsupports,psf}.bin
 *                         raw arrays in the actors' types
 * Writes CASE_DIR/residual_<k>.bin, model_<k>.bin and sources_<k>.bin (an
 * int count then that many float3) for every major cycle k, and prints one
 * line per cycle.  Exit status 0 on success, 2 on a bad case directory; an
 * actor error aborts, as the reference's actors do.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "sdp_actors.h"

typedef struct {
    int grid_size, num_vis, num_kernels, total_kernel_samples, num_receivers, num_baselines;
    int slices, num_major, num_minor, max_sources, oversampling;
} Case;

static const char *g_dir;

static void *xalloc(size_t bytes)
{
    void *p = calloc(1, bytes ? bytes : 1);
    if (!p) {
        fprintf(stderr, "sdp_major_cycle: out of host memory (%zu bytes)\n", bytes);
        exit(2);
    }
    return p;
}

static void *read_file(const char *name, size_t bytes)
{
    char path[4096];
    snprintf(path, sizeof path, "%s/%s", g_dir, name);
    FILE *f = fopen(path, "rb");
    if (!f) {
        fprintf(stderr, "sdp_major_cycle: cannot open %s\n", path);
        exit(2);
    }
    void *p = xalloc(bytes);
    size_t got = fread(p, 1, bytes, f);
    int extra = fgetc(f) != EOF;
    fclose(f);
    if (got != bytes || extra) {
        fprintf(stderr, "sdp_major_cycle: %s holds %s than the %zu bytes the case needs\n", path,
                got != bytes ? "fewer" : "more", bytes);
        exit(2);
    }
    return p;
}

static void write_file(const char *name, int cycle, const void *head, size_t head_bytes, const void *p,
                       size_t bytes)
{
    char path[4096];
    snprintf(path, sizeof path, "%s/%s_%d.bin", g_dir, name, cycle);
    FILE *f = fopen(path, "wb");
    if (!f || (head_bytes && fwrite(head, 1, head_bytes, f) != head_bytes) ||
        (bytes && fwrite(p, 1, bytes, f) != bytes)) {
        fprintf(stderr, "sdp_major_cycle: cannot write %s\n", path);
        exit(2);
    }
    fclose(f);
}

static int read_case(Case *c)
{
    char path[4096];
    snprintf(path, sizeof path, "%s/case.txt", g_dir);
    FILE *f = fopen(path, "r");
    if (!f) {
        fprintf(stderr, "sdp_major_cycle: cannot open %s\n", path);
        return -1;
    }
    struct { const char *key; int *val; } keys[] = {
        {"grid_size", &c->grid_size}, {"num_vis", &c->num_vis}, {"num_kernels", &c->num_kernels},
        {"total_kernel_samples", &c->total_kernel_samples}, {"num_receivers", &c->num_receivers},
        {"num_baselines", &c->num_baselines}, {"slices", &c->slices}, {"num_major", &c->num_major},
        {"num_minor", &c->num_minor}, {"max_sources", &c->max_sources}, {"oversampling", &c->oversampling},
    };
    const int nkeys = (int)(sizeof keys / sizeof keys[0]);
    int seen = 0;
    char key[64];
    int val;
    memset(c, 0, sizeof *c);
    while (fscanf(f, "%63s %d", key, &val) == 2) {
        int k = 0;
        while (k < nkeys && strcmp(key, keys[k].key)) k++;
        if (k == nkeys) {
            fprintf(stderr, "sdp_major_cycle: unknown key %s in %s\n", key, path);
            fclose(f);
            return -1;
        }
        *keys[k].val = val;
        seen |= 1 << k;
    }
    fclose(f);
    if (seen != (1 << nkeys) - 1) {
        fprintf(stderr, "sdp_major_cycle: %s does not set every key\n", path);
        return -1;
    }
    if (c->grid_size <= 0 || c->grid_size % 2 || c->num_vis <= 0 || c->slices <= 0 ||
        c->num_vis % c->slices || c->num_kernels <= 0 || c->total_kernel_samples <= 0 ||
        c->num_receivers <= 0 || c->num_baselines <= 0 || c->num_major <= 0 || c->num_minor < 0 ||
        c->max_sources <= 0 || c->oversampling <= 0) {
        fprintf(stderr, "sdp_major_cycle: out-of-range value in %s (num_vis must be a multiple of slices)\n",
                path);
        return -1;
    }
    return 0;
}

int main(int argc, char **argv)
{
    if (argc != 2) {
        fprintf(stderr, "usage: %s CASE_DIR\n", argv[0]);
        return 2;
    }
    g_dir = argv[1];
    Case c;
    if (read_case(&c)) return 2;

    const int G = c.grid_size, N = c.num_vis, S = c.slices, per = N / S;
    const size_t cells = (size_t)G * G;
    Config *config = read_file("config.bin", sizeof(Config));
    PRECISION3 *uvw = read_file("uvw.bin", sizeof(PRECISION3) * N);
    PRECISION2 *measured = read_file("vis.bin", sizeof(PRECISION2) * N);
    PRECISION2 *gains = read_file("gains.bin", sizeof(PRECISION2) * c.num_receivers);
    int2 *pairs = read_file("pairs.bin", sizeof(int2) * c.num_baselines);
    PRECISION2 *kernels = read_file("kernels.bin", sizeof(PRECISION2) * c.total_kernel_samples);
    int2 *supports = read_file("supports.bin", sizeof(int2) * c.num_kernels);
    PRECISION2 *dkernels = read_file("dkernels.bin", sizeof(PRECISION2) * c.total_kernel_samples);
    int2 *dsupports = read_file("dsupports.bin", sizeof(int2) * c.num_kernels);
    PRECISION *psf = read_file("psf.bin", sizeof(PRECISION) * cells);

    /* set-up: the half prolate taper (one spare zero element: the correction
     * actor's own bound is G/2) and the clean beam's half size */
    PRECISION *prolate = xalloc(sizeof(PRECISION) * (G / 2 + 1));
    correction_set_up(G, prolate);
    PRECISION *clean_psf = xalloc(sizeof(PRECISION) * cells);
    int2 halfdims = {0, 0};
    clean_psf_host_set_up(G, 0 /* GAUSSIAN_CLEAN_PSF */, config, psf, clean_psf, &halfdims);

    PRECISION *estimate = xalloc(sizeof(PRECISION) * cells);      /* first estimate: zero */
    PRECISION *image = xalloc(sizeof(PRECISION) * cells);
    PRECISION *dirty = xalloc(sizeof(PRECISION) * cells);
    PRECISION *residual = xalloc(sizeof(PRECISION) * cells);
    PRECISION *model = xalloc(sizeof(PRECISION) * cells);
    PRECISION2 *gains_out = xalloc(sizeof(PRECISION2) * c.num_receivers);
    PRECISION2 *fourier = xalloc(sizeof(PRECISION2) * cells);
    PRECISION2 *uv = xalloc(sizeof(PRECISION2) * cells);
    PRECISION2 *input_grid = xalloc(sizeof(PRECISION2) * cells);
    PRECISION2 *slice_grids = xalloc(sizeof(PRECISION2) * cells * S);
    PRECISION2 *out_grids = xalloc(sizeof(PRECISION2) * cells * S);
    PRECISION2 *prev_grid = xalloc(sizeof(PRECISION2) * cells);   /* read only when BYPASS is set */
    PRECISION2 *merged = xalloc(sizeof(PRECISION2) * cells);
    PRECISION2 *corrected = xalloc(sizeof(PRECISION2) * N);
    PRECISION2 *fine_vis = xalloc(sizeof(PRECISION2) * N);
    PRECISION3 *fine_uvw = xalloc(sizeof(PRECISION3) * N);
    PRECISION2 *split_vis = xalloc(sizeof(PRECISION2) * N);
    PRECISION3 *split_uvw = xalloc(sizeof(PRECISION3) * N);
    PRECISION2 *predicted = xalloc(sizeof(PRECISION2) * N);
    PRECISION2 *delta = xalloc(sizeof(PRECISION2) * N);
    Config *slice_configs = xalloc(sizeof(Config) * S);
    int *slice_counts = xalloc(sizeof(int) * S);
    int *slice_iter = xalloc(sizeof(int) * S);
    PRECISION3 *sources = xalloc(sizeof(PRECISION3) * c.max_sources);

    for (int cycle = 0; cycle < c.num_major; cycle++) {
        int maj_iter = cycle, num_corrected = 0, num_sources = 0;

        /* predict: the estimate's uv grid, one copy per slice */
        execute_inv_convolution_correction_actor(G, estimate, prolate, config, image);
        reciprocal_transform_actor(c.num_receivers, gains, config, gains_out);
        fft_shift_real_to_complex_actor(G, image, config, fourier);
        CUFFT_EXECUTE_FORWARD_C2C_actor(G, fourier, uv);
        fft_shift_complex_to_complex_actor(G, uv, config, input_grid);
        split_uvgrid(G, S, input_grid, slice_grids);

        /* calibrated visibilities on the fine grid, split into slices */
        apply_gains_only(c.num_receivers, c.num_baselines, N, measured, gains_out, pairs, config, corrected);
        correct_to_finegrid(N, G, c.oversampling, 0 /* PERFORM_SIMPLIFICATION */, uvw, corrected, config, fine_vis,
                            fine_uvw, &num_corrected);
        split_fftdegrid_visibilities(N, S, G, config, fine_uvw, fine_vis, &num_corrected, &maj_iter, split_vis,
                                     split_uvw, slice_configs, slice_counts, slice_iter);

        for (int s = 0; s < S; s++)
            std_degridding(G, per, c.num_kernels, c.total_kernel_samples, c.oversampling, dkernels, dsupports,
                           slice_grids + (size_t)s * cells, split_uvw + (size_t)s * per, &slice_counts[s],
                           &slice_configs[s], predicted + (size_t)s * per);
        for (int s = 0; s < S; s++)
            subtract_from_measurements_finegrid(per, split_vis + (size_t)s * per, predicted + (size_t)s * per,
                                                &slice_counts[s], delta + (size_t)s * per);
        for (int s = 0; s < S; s++)
            std_gridding(G, per, c.num_kernels, c.total_kernel_samples, c.oversampling, 0 /* BYPASS */,
                         &slice_iter[s], &slice_counts[s], kernels, supports, split_uvw + (size_t)s * per,
                         delta + (size_t)s * per, &slice_configs[s], prev_grid, out_grids + (size_t)s * cells);
        merge_gridded_visibilities(G, S, out_grids, merged);

        /* residual image and the minor cycle */
        fft_shift_complex_to_complex_actor(G, merged, config, uv);
        CUFFT_EXECUTE_INVERSE_C2C_actor(G, uv, fourier);
        fft_shift_complex_to_real_actor(G, fourier, config, dirty);
        execute_convolution_correction_actor(G, dirty, prolate, config, residual);
        hogbom_clean(G, c.num_minor, c.max_sources, residual, psf, &halfdims, config, estimate, &num_sources,
                     sources, model);

        write_file("residual", cycle, NULL, 0, residual, sizeof(PRECISION) * cells);
        write_file("model", cycle, NULL, 0, model, sizeof(PRECISION) * cells);
        write_file("sources", cycle, &num_sources, sizeof num_sources, sources, sizeof(PRECISION3) * num_sources);
        printf("cycle %d: %d corrected visibilities, slices", cycle, num_corrected);
        for (int s = 0; s < S; s++) printf(" %d", slice_counts[s]);
        printf(", %d components\n", num_sources);
        fflush(stdout);
        memcpy(estimate, model, sizeof(PRECISION) * cells);
    }
    return 0;
}
