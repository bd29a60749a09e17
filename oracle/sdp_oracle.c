/*
 * sdp_oracle.c -- CPU restatement of the reference actors.
 *
 * TEST INFRASTRUCTURE ONLY: used by tests/, by __graft_entry__.smoke() as
 * the checker and by bench.py as the timed CPU baseline ("port").  Nothing
 * in the product library links or calls it.
 *
 * Build: gcc -O2 -std=c99 -ffp-contract=off (the reference builds with
 * -O3 -std=c99, CMakeLists.txt:140,243, which also disables contraction).
 * All float arithmetic is IEEE binary32 on SSE (FLT_EVAL_METHOD == 0), so
 * the float/double promotions written below are the ones the reference's C
 * performs.  Reference paths are relative to
 * preesm_pipelines/simsdp_fft_imaging_pipeline/Code/.
 */
#include "sdp_oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

static inline float2 cmul(float2 a, float2 b)
{
    /* complex_mult, degridgrid.c:5-8 / complex_mult_CPU, gridding_run.c:148-151 */
    float2 r;
    r.x = a.x * b.x - a.y * b.y;
    r.y = a.y * b.x + a.x * b.y;
    return r;
}

/* w-plane index of std_gridding / correct_to_finegrid (degridgrid.c:27,84):
 * SQRT(ABS(z * w_scale)) + 0.5 with ABS = fabsf, SQRT = sqrtf, 0.5 double. */
static inline double plane_arg(float z, double w_scale)
{
    return (double)sqrtf(fabsf((float)((double)z * w_scale))) + 0.5;
}

/* The reference's finegrid hashmap stores the address of one stack variable
 * as every key (degridgrid.c:29,46 with map.c:220-238), so two entries match
 * when their 32-bit hashes match (map.c:201-209).  The hash of the 8-byte key
 * (map.c:131-170 with size 8) reduces to this expression. */
static inline uint32_t finegrid_key(uint64_t idx)
{
    uint64_t h = (UINT64_C(2166136261) ^ idx) * UINT64_C(0xbf58476d1ce4e5b9);
    return (uint32_t)(h ^ (h >> 32));
}

typedef struct { uint32_t key; int32_t slot; } key_slot;

static int find_or_insert(key_slot *table, uint64_t mask, uint32_t key, int32_t slot, int32_t *found)
{
    uint64_t i = (key * UINT64_C(0x9E3779B97F4A7C15)) >> 20 & mask;
    for (;;) {
        if (table[i].slot < 0) {
            table[i].key = key;
            table[i].slot = slot;
            return 0;
        }
        if (table[i].key == key) {
            *found = table[i].slot;
            return 1;
        }
        i = (i + 1) & mask;
    }
}

/* correct_to_finegrid, degridgrid.c:12-65.  The double -> unsigned long
 * conversion of a negative value (:24-25) is undefined in C; the oracle
 * defines it as signed truncation (int64), the same choice the GPU makes. */
void oracle_correct_to_finegrid(int NUM_VISIBILITIES, int GRID_SIZE, int OVERSAMPLING_FACTOR,
        int PERFORM_SIMPLIFICATION, const float3 *vis_uvw_coords, const float2 *input_visibilities,
        const Config *config, float2 *output_visibilities, float3 *output_finegrid_vis_coords,
        int *num_output_visibilities)
{
    memset(output_visibilities, 0, sizeof(float2) * (size_t)NUM_VISIBILITIES);
    *num_output_visibilities = 0;

    uint64_t cap = 64;
    while (cap < (uint64_t)NUM_VISIBILITIES * 2) cap <<= 1;
    key_slot *table = NULL;
    if (PERFORM_SIMPLIFICATION == 1) {
        table = (key_slot *)malloc(sizeof(key_slot) * cap);
        for (uint64_t i = 0; i < cap; ++i) table[i].slot = -1;
    }

    const double half = (double)(GRID_SIZE / 2);
    for (int i = 0; i < NUM_VISIBILITIES; ++i) {
        double fxd = ((double)vis_uvw_coords[i].x * config->uv_scale + half) * OVERSAMPLING_FACTOR + 0.5;
        double fyd = ((double)vis_uvw_coords[i].y * config->uv_scale + half) * OVERSAMPLING_FACTOR + 0.5;
        uint64_t fx = (uint64_t)(int64_t)fxd;
        uint64_t fy = (uint64_t)(int64_t)fyd;
        uint64_t w_idx = (uint64_t)(int64_t)plane_arg(vis_uvw_coords[i].z, config->w_scale);
        uint64_t idx = (fx + fy * (uint64_t)GRID_SIZE * (uint64_t)OVERSAMPLING_FACTOR)
                       + w_idx * ((uint64_t)GRID_SIZE * GRID_SIZE * OVERSAMPLING_FACTOR * OVERSAMPLING_FACTOR);
        int n = *num_output_visibilities;
        if (PERFORM_SIMPLIFICATION == 1) {
            int32_t found = -1;
            if (find_or_insert(table, cap - 1, finegrid_key(idx), n, &found)) {
                output_visibilities[found].x += input_visibilities[i].x;
                output_visibilities[found].y += input_visibilities[i].y;
                continue;
            }
        }
        output_visibilities[n] = input_visibilities[i];
        output_finegrid_vis_coords[n].x = (float)(((double)(int64_t)fx / OVERSAMPLING_FACTOR - half) / config->uv_scale);
        output_finegrid_vis_coords[n].y = (float)(((double)(int64_t)fy / OVERSAMPLING_FACTOR - half) / config->uv_scale);
        output_finegrid_vis_coords[n].z = vis_uvw_coords[i].z;
        *num_output_visibilities = n + 1;
    }
    free(table);
}

/* std_gridding, degridgrid.c:69-125 */
void oracle_std_gridding(int GRID_SIZE, int NUM_VISIBILITIES, int NUM_KERNELS, int TOTAL_KERNEL_SAMPLES,
        int OVERSAMPLING_FACTOR, int BYPASS, const int *maj_iter, const int *num_corrected_visibilities,
        const float2 *kernels, const int2 *kernel_supports, const float3 *corrected_vis_uvw_coords,
        const float2 *visibilities, const Config *config, const float2 *prev_grid, float2 *output_grid)
{
    (void)NUM_VISIBILITIES; (void)NUM_KERNELS; (void)TOTAL_KERNEL_SAMPLES;
    memset(output_grid, 0, sizeof(float2) * (size_t)GRID_SIZE * GRID_SIZE);
    if (BYPASS && *maj_iter > 0) {
        memcpy(output_grid, prev_grid, sizeof(float2) * (size_t)GRID_SIZE * GRID_SIZE);
        return;
    }
    const int gc = GRID_SIZE / 2;
    const float os = (float)OVERSAMPLING_FACTOR;
    for (int i = 0; i < *num_corrected_visibilities; ++i) {
        const float3 c = corrected_vis_uvw_coords[i];
        int w_idx = (int)plane_arg(c.z, config->w_scale);
        float gx = (float)((double)c.x * config->uv_scale);
        float gy = (float)((double)c.y * config->uv_scale);
        int h = kernel_supports[w_idx].x;
        float conj = (c.z < 0.0) ? -1.0f : 1.0f;
        int w_offset = kernel_supports[w_idx].y;
        int row = (h + 1) * OVERSAMPLING_FACTOR;
        for (int v = (int)ceilf(gy - (float)h); (float)v < ceilf(gy + (float)h); ++v) {
            if (v < -gc || v >= gc) continue;
            int ky = abs((int)roundf(((float)v - gy) * os));
            for (int u = (int)ceilf(gx - (float)h); (float)u < ceilf(gx + (float)h); ++u) {
                if (u < -gc || u >= gc) continue;
                int kx = abs((int)roundf(((float)u - gx) * os));
                float2 ks = kernels[w_offset + ky * row + kx];
                ks.y *= conj;
                float2 p = cmul(visibilities[i], ks);
                size_t g = (size_t)(v + gc) * GRID_SIZE + (u + gc);
                output_grid[g].x += p.x;
                output_grid[g].y += p.y;
            }
        }
    }
}

void oracle_std_gridding_f64(int GRID_SIZE, int num_vis, int OVERSAMPLING_FACTOR,
        const float2 *kernels, const int2 *kernel_supports, const float3 *uvw,
        const float2 *visibilities, double uv_scale, double w_scale, double *grid)
{
    memset(grid, 0, sizeof(double) * 2 * (size_t)GRID_SIZE * GRID_SIZE);
    const int gc = GRID_SIZE / 2;
    const float os = (float)OVERSAMPLING_FACTOR;
    for (int i = 0; i < num_vis; ++i) {
        const float3 c = uvw[i];
        int w_idx = (int)plane_arg(c.z, w_scale);
        float gx = (float)((double)c.x * uv_scale);
        float gy = (float)((double)c.y * uv_scale);
        int h = kernel_supports[w_idx].x;
        double conj = (c.z < 0.0) ? -1.0 : 1.0;
        int row = (h + 1) * OVERSAMPLING_FACTOR;
        for (int v = (int)ceilf(gy - (float)h); (float)v < ceilf(gy + (float)h); ++v) {
            if (v < -gc || v >= gc) continue;
            int ky = abs((int)roundf(((float)v - gy) * os));
            for (int u = (int)ceilf(gx - (float)h); (float)u < ceilf(gx + (float)h); ++u) {
                if (u < -gc || u >= gc) continue;
                int kx = abs((int)roundf(((float)u - gx) * os));
                float2 k = kernels[kernel_supports[w_idx].y + ky * row + kx];
                double kr = k.x, ki = (double)k.y * conj;
                double vr = visibilities[i].x, vi = visibilities[i].y;
                size_t g = (size_t)(v + gc) * GRID_SIZE + (u + gc);
                grid[2 * g] += vr * kr - vi * ki;
                grid[2 * g + 1] += vi * kr + vr * ki;
            }
        }
    }
}

/* std_degridding, degridgrid.c:127-202.  The kernel-weighted loop (:147-191)
 * is evaluated for its cost and then overwritten by the nearest grid cell
 * (:194-200), as in the reference.  Its mirrored index can reach one row
 * past the grid (v == G/2 gives corrected_v == G/2); such reads use 0. */
void oracle_std_degridding(int GRID_SIZE, int NUM_VISIBILITIES, int NUM_KERNELS, int TOTAL_KERNEL_SAMPLES,
        int OVERSAMPLING_FACTOR, const float2 *kernels, const int2 *kernel_supports,
        const float2 *input_grid, const float3 *corrected_vis_uvw_coords,
        const int *num_corrected_visibilities, const Config *config, float2 *output_visibilities)
{
    (void)NUM_KERNELS; (void)TOTAL_KERNEL_SAMPLES;
    memset(output_visibilities, 0, sizeof(float2) * (size_t)NUM_VISIBILITIES);
    const int gc = GRID_SIZE / 2;
    const float os = (float)OVERSAMPLING_FACTOR;
    const size_t ncell = (size_t)GRID_SIZE * GRID_SIZE;
    for (int i = 0; i < *num_corrected_visibilities; ++i) {
        const float3 c = corrected_vis_uvw_coords[i];
        int w_idx = (int)plane_arg(c.z, config->w_scale);
        int h = kernel_supports[w_idx].x;
        int w_offset = 0;
        float gx = (float)((double)c.x * config->uv_scale);
        float gy = (float)((double)c.y * config->uv_scale);
        float conj = (c.z < 0.0) ? -1.0f : 1.0f;
        float comm_norm = 0.f;
        float2 acc = {0.f, 0.f};
        for (int v = (int)ceilf(gy - (float)h); (float)v < ceilf(gy + (float)h); ++v) {
            int cv = v;
            if (v < -gc) cv = (-gc - v) - gc;
            else if (v >= gc) cv = (gc - v) + gc;
            int ky = abs((int)roundf(((float)cv - gy) * os));
            for (int u = (int)ceilf(gx - (float)h); (float)u < ceilf(gx + (float)h); ++u) {
                int cu = u;
                if (u < -gc) cu = (-gc - u) - gc;
                else if (u >= gc) cu = (gc - u) + gc;
                int kx = abs((int)roundf(((float)cu - gx) * os));
                float2 ks = kernels[w_offset + ky * (h + 1) * OVERSAMPLING_FACTOR + kx];
                ks.y *= conj;
                long g = (long)(cv + gc) * GRID_SIZE + (cu + gc);
                float2 gv = {0.f, 0.f};
                if (g >= 0 && (size_t)g < ncell) gv = input_grid[g];
                float2 p = cmul(gv, ks);
                comm_norm += ks.x * ks.x + ks.y * ks.y;
                acc.x += p.x;
                acc.y += p.y;
            }
        }
        comm_norm = (float)sqrt((double)comm_norm);
        output_visibilities[i].x = comm_norm < 1e-5f ? acc.x / 1e-5f : acc.x / comm_norm;
        output_visibilities[i].y = comm_norm < 1e-5f ? output_visibilities[i].x / 1e-5f : acc.y / comm_norm;
        /* nearest cell (:194-200): float + float, then + 0.5 in double */
        int x = (int)((double)(gx + (float)gc) + 0.5);
        int y = (int)((double)(gy + (float)gc) + 0.5);
        long idx = (long)x + (long)y * GRID_SIZE;
        float2 nv = {0.f, 0.f};
        if (idx >= 0 && (size_t)idx < ncell) nv = input_grid[idx];
        output_visibilities[i] = nv;
    }
}

void oracle_degrid_weighted_f64(int GRID_SIZE, int num_vis, int OVERSAMPLING_FACTOR,
        const float2 *kernels, const int2 *kernel_supports, const float2 *grid, const float3 *uvw,
        double uv_scale, double w_scale, double *out)
{
    const int gc = GRID_SIZE / 2;
    const float os = (float)OVERSAMPLING_FACTOR;
    for (int i = 0; i < num_vis; ++i) {
        const float3 c = uvw[i];
        int w_idx = (int)plane_arg(c.z, w_scale);
        int h = kernel_supports[w_idx].x;
        int off = kernel_supports[w_idx].y;
        float gx = (float)((double)c.x * uv_scale);
        float gy = (float)((double)c.y * uv_scale);
        double conj = (c.z < 0.0) ? -1.0 : 1.0;
        double sr = 0.0, si = 0.0;
        for (int v = (int)ceilf(gy - (float)h); (float)v < ceilf(gy + (float)h); ++v) {
            if (v < -gc || v >= gc) continue;
            int ky = abs((int)roundf(((float)v - gy) * os));
            for (int u = (int)ceilf(gx - (float)h); (float)u < ceilf(gx + (float)h); ++u) {
                if (u < -gc || u >= gc) continue;
                int kx = abs((int)roundf(((float)u - gx) * os));
                float2 k = kernels[off + ky * (h + 1) * OVERSAMPLING_FACTOR + kx];
                double kr = k.x, ki = (double)k.y * conj;
                float2 g = grid[(size_t)(v + gc) * GRID_SIZE + (u + gc)];
                sr += (double)g.x * kr - (double)g.y * ki;
                si += (double)g.y * kr + (double)g.x * ki;
            }
        }
        out[2 * i] = sr;
        out[2 * i + 1] = si;
    }
}

/* gridding_actor -> gridding_CPU, gridding_run.c:5-137 (OS fixed to 16 at :43) */
void oracle_gridding_actor(int GRID_SIZE, int NUM_VISIBILITIES, int NUM_KERNELS, int TOTAL_KERNEL_SAMPLES,
        const float2 *kernels, const int2 *kernel_supports, const float3 *vis_uvw_coords,
        const float2 *visibilities, const Config *config, float2 *uv_grid)
{
    (void)NUM_KERNELS; (void)TOTAL_KERNEL_SAMPLES;
    memset(uv_grid, 0, sizeof(float2) * (size_t)GRID_SIZE * GRID_SIZE);
    const int oversampling = 16;
    const float os = (float)oversampling;
    const double uv_scale = config->uv_scale, w_scale = config->w_scale;
    for (unsigned int vi = 0; vi < (unsigned int)NUM_VISIBILITIES; vi++) {
        const float3 c = vis_uvw_coords[vi];
        const int plane_index = (int)roundf(sqrtf(fabsf((float)((double)c.z * w_scale))));
        const float gx = (float)((double)c.x * uv_scale);
        const float gy = (float)((double)c.y * uv_scale);
        const int half_grid_size = GRID_SIZE / 2;
        const int half_support = kernel_supports[plane_index].x;
        float conj = (c.z < 0.0) ? -1.0f : 1.0f;
        const float sx = roundf(gx * os) / os;
        const float sy = roundf(gy * os) / os;
        const float minx = ceilf(sx - (float)half_support), miny = ceilf(sy - (float)half_support);
        const float maxx = floorf(sx + (float)half_support), maxy = floorf(sy + (float)half_support);
        int w_kernel_offset = kernel_supports[plane_index].y;
        for (int gv = (int)miny; (float)gv <= maxy; ++gv) {
            if (gv < -half_grid_size || gv >= half_grid_size) continue;
            int ky = abs((int)roundf(((float)gv - sy) * os));
            for (int gu = (int)minx; (float)gu <= maxx; ++gu) {
                if (gu < -half_grid_size || gu >= half_grid_size) continue;
                int kx = abs((int)roundf(((float)gu - sx) * os));
                int kidx = w_kernel_offset + ky * (half_support + 1) * oversampling + kx;
                float2 ks;
                ks.x = kernels[kidx].x;
                ks.y = kernels[kidx].y * conj;
                size_t g = (size_t)(gv + half_grid_size) * GRID_SIZE + (gu + half_grid_size);
                float2 p = cmul(visibilities[vi], ks);
                uv_grid[g].x += p.x;
                uv_grid[g].y += p.y;
            }
        }
    }
}

/* gridding_actor's index arithmetic (gridding_run.c:63-123: snap to 1/16,
 * inclusive range, roundf plane, OS fixed at 16) with the products and the
 * cell sums in double: the exact sum the float grid approximates */
void oracle_gridding_actor_f64(int GRID_SIZE, int NUM_VISIBILITIES, const float2 *kernels,
        const int2 *kernel_supports, const float3 *vis_uvw_coords, const float2 *visibilities,
        double uv_scale, double w_scale, double *grid)
{
    memset(grid, 0, sizeof(double) * 2 * (size_t)GRID_SIZE * GRID_SIZE);
    const int oversampling = 16;
    const float os = (float)oversampling;
    const int half_grid_size = GRID_SIZE / 2;
    for (int vi = 0; vi < NUM_VISIBILITIES; vi++) {
        const float3 c = vis_uvw_coords[vi];
        const int plane_index = (int)roundf(sqrtf(fabsf((float)((double)c.z * w_scale))));
        const float gx = (float)((double)c.x * uv_scale);
        const float gy = (float)((double)c.y * uv_scale);
        const int h = kernel_supports[plane_index].x;
        const double conj = (c.z < 0.0) ? -1.0 : 1.0;
        const float sx = roundf(gx * os) / os;
        const float sy = roundf(gy * os) / os;
        const float maxx = floorf(sx + (float)h), maxy = floorf(sy + (float)h);
        for (int gv = (int)ceilf(sy - (float)h); (float)gv <= maxy; ++gv) {
            if (gv < -half_grid_size || gv >= half_grid_size) continue;
            int ky = abs((int)roundf(((float)gv - sy) * os));
            for (int gu = (int)ceilf(sx - (float)h); (float)gu <= maxx; ++gu) {
                if (gu < -half_grid_size || gu >= half_grid_size) continue;
                int kx = abs((int)roundf(((float)gu - sx) * os));
                float2 k = kernels[kernel_supports[plane_index].y + ky * (h + 1) * oversampling + kx];
                double kr = k.x, ki = (double)k.y * conj;
                double vr = visibilities[vi].x, vim = visibilities[vi].y;
                size_t g = (size_t)(gv + half_grid_size) * GRID_SIZE + (gu + half_grid_size);
                grid[2 * g] += vr * kr - vim * ki;
                grid[2 * g + 1] += vim * kr + vr * ki;
            }
        }
    }
}

/* The "average" hogbom_clean computes in its peak scan
 * (deconvolution_run.c:223-227): m = |r[0]|, then m += (|r[i]| - m) / i in
 * float for i = 1 .. n-1.  In exact arithmetic this is the mean of
 * |r[1..n-1]|; in float its rounding drifts with n. */
float oracle_running_mean(const float *r, long n)
{
    float avg = 0.f;
    for (long idx = 0; idx < n; ++idx) {
        float val = fabsf(r[idx]);
        avg = idx == 0 ? val : avg + (val - avg) / (float)idx;
    }
    return avg;
}

/* dft_actor, dft_run.c:3-47 (approximation branch, :21-32) */
void oracle_dft_actor(int NUM_VISIBILITIES, int NUM_MAX_SOURCES, const float3 *sources,
        const float3 *vis_uvw_coords, const int *num_sources, const Config *config, float2 *visibilities)
{
    (void)NUM_MAX_SOURCES; (void)config;
    const float pi = (float)3.141592654;
    for (int j = 0; j < NUM_VISIBILITIES; j++) {
        const float two_pi = pi + pi;
        const float3 vis = vis_uvw_coords[j];
        float2 sum = {0.f, 0.f};
        for (int s = 0; s < *num_sources; ++s) {
            const float3 src = sources[s];
            float term = (float)(0.5 * (double)((src.x * src.x) + (src.y * src.y)));
            float w_correction = -term;
            float image_correction = (float)(1.0 - (double)term);
            float src_correction = src.z / image_correction;
            float theta = (vis.x * src.x + vis.y * src.y + vis.z * w_correction) * two_pi;
            float sn = sinf(theta), cs = cosf(theta);
            sum.x += cs * src_correction;
            sum.y += -sn * src_correction;
        }
        visibilities[j] = sum;
    }
}

void oracle_dft_f64(int num_vis, int num_sources, const float3 *sources, const float3 *uvw, double *out)
{
    for (int j = 0; j < num_vis; j++) {
        double sr = 0.0, si = 0.0;
        for (int s = 0; s < num_sources; ++s) {
            double l = sources[s].x, m = sources[s].y, f = sources[s].z;
            double term = 0.5 * (l * l + m * m);
            double phase = (double)uvw[j].x * l + (double)uvw[j].y * m - (double)uvw[j].z * term;
            phase -= nearbyint(phase);           /* cycles, exact reduction */
            double corr = f / (1.0 - term);
            sr += cos(2.0 * M_PI * phase) * corr;
            si -= sin(2.0 * M_PI * phase) * corr;
        }
        out[2 * j] = sr;
        out[2 * j + 1] = si;
    }
}

/* hogbom_clean, deconvolution_run.c:197-312 */
void oracle_hogbom_clean(int GRID_SIZE, int NUM_MINOR_CYCLES, int MAX_SOURCES, const float *residual,
        const float *psf, const Config *config, const float *current_model, int *num_sources_out,
        float3 *sources_out, float *output_model, float *residual_out)
{
    (void)MAX_SOURCES;
    const size_t n = (size_t)GRID_SIZE * GRID_SIZE;
    float *r = (float *)malloc(sizeof(float) * n);
    memcpy(r, residual, sizeof(float) * n);
    memcpy(output_model, current_model, sizeof(float) * n);
    *num_sources_out = 0;
    const double WEAK = config->weak_source_percent_img;
    const double GAIN = config->loop_gain;
    const double NOISE = config->noise_factor;

    for (int it = 0; it < NUM_MINOR_CYCLES; ++it) {
        float curr_max = 0.f;
        int mx = 0, my = 0;
        float avg = 0.f;
        for (int y = 0; y < GRID_SIZE; ++y) {
            for (int x = 0; x < GRID_SIZE; ++x) {
                int idx = y * GRID_SIZE + x;
                float val = fabsf(r[idx]);
                avg = idx == 0 ? val : avg + (val - avg) / (float)idx;
                if (val > curr_max) {
                    curr_max = (float)fabs((double)r[idx]);
                    mx = x; my = y;
                }
            }
        }
        int max_idx = my * GRID_SIZE + mx;
        bool extracting_noise = (double)fabsf(curr_max) < NOISE * (double)avg * GAIN;
        bool weak_source = *num_sources_out > 0 &&
                           (double)fabsf(curr_max) < ((double)fabsf(sources_out[0].z) * WEAK);
        if (extracting_noise || weak_source) break;

        const int pcx = GRID_SIZE / 2, pcy = GRID_SIZE / 2;
        float max_val = (float)((double)r[max_idx] / config->psf_max_value);
        int ns = *num_sources_out;
        sources_out[ns].x = (float)mx;
        sources_out[ns].y = (float)my;
        sources_out[ns].z = (float)((double)max_val * GAIN);
        *num_sources_out = ns + 1;
        output_model[max_idx] = (float)((double)output_model[max_idx] + (double)max_val * GAIN);

        int sx = mx < GRID_SIZE / 2 ? mx : GRID_SIZE / 2;
        int sy = my < GRID_SIZE / 2 ? my : GRID_SIZE / 2;
        int ex = (GRID_SIZE - mx) < GRID_SIZE / 2 ? (GRID_SIZE - mx) : GRID_SIZE / 2;
        int ey = (GRID_SIZE - my) < GRID_SIZE / 2 ? (GRID_SIZE - my) : GRID_SIZE / 2;
        for (int y = my - sy; y < my + ey; ++y) {
            for (int x = mx - sx; x < mx + ex; ++x) {
                int image_idx = y * GRID_SIZE + x;
                int psf_idx = (y - my + pcy) * GRID_SIZE + (GRID_SIZE - (x - mx + pcx));
                r[image_idx] = (float)((double)r[image_idx]
                                       - GAIN * (double)max_val * (double)psf[psf_idx] * config->psf_max_value);
            }
        }
    }
    if (*num_sources_out > 0) {
        for (int i = 0; i < *num_sources_out; i++) {
            sources_out[i].x = (float)((double)(sources_out[i].x - (float)(GRID_SIZE / 2)) * config->cell_size);
            sources_out[i].y = (float)((double)(sources_out[i].y - (float)(GRID_SIZE / 2)) * config->cell_size);
        }
    }
    if (residual_out) memcpy(residual_out, r, sizeof(float) * n);
    free(r);
}

/* fft_shift_complex_to_complex_actor, fft_run.c:5-20 */
void oracle_shift_c2c(int G, const float2 *in, float2 *out)
{
    for (int r = 0; r < G; r++)
        for (int c = 0; c < G; c++) {
            int a = 1 - 2 * ((r + c) & 1);
            size_t i = (size_t)r * G + c;
            out[i].x = in[i].x * (float)a;
            out[i].y = in[i].y * (float)a;
        }
}

/* fft_shift_complex_to_real_actor, fft_run.c:76-92 */
void oracle_shift_c2r(int G, const float2 *in, float *out)
{
    for (int r = 0; r < G; r++)
        for (int c = 0; c < G; c++) {
            int a = 1 - 2 * ((r + c) & 1);
            size_t i = (size_t)r * G + c;
            out[i] = in[i].x * (float)a;
        }
}

/* fft_shift_real_to_complex_actor, fft_run.c:94-111 */
void oracle_shift_r2c(int G, const float *in, float2 *out)
{
    for (int r = 0; r < G; r++)
        for (int c = 0; c < G; c++) {
            int a = 1 - 2 * ((r + c) & 1);
            size_t i = (size_t)r * G + c;
            out[i].x = in[i] * (float)a;
            out[i].y = 0;
        }
}

/* execute_convolution_correction_actor, convolution_correction_run.c:4-38 */
void oracle_convolution_correction(int G, const float *in, const float *prolate, float *out)
{
    for (int r = 0; r < G; r++)
        for (int c = 0; c < G; c++) {
            const int half = G / 2;
            size_t i = (size_t)r * G + c;
            float taper = prolate[abs(c - half)] * prolate[abs(r - half)];
            if (abs(c - half) == 1229 || abs(r - half) == 1229) taper = 0.0f;
            out[i] = ((double)fabsf(taper) > 1E-10) ? in[i] / taper : 0.0f;
        }
}

/* execute_inv_convolution_correction_actor, convolution_correction_run.c:41-71 (identity) */
void oracle_inv_convolution_correction(int G, const float *in, float *out)
{
    memmove(out, in, sizeof(float) * (size_t)G * G);
}

static void fft1d_f64(int n, double *re, double *im, int sign)
{
    for (int i = 1, j = 0; i < n; i++) {
        int bit = n >> 1;
        for (; j & bit; bit >>= 1) j ^= bit;
        j ^= bit;
        if (i < j) {
            double t = re[i]; re[i] = re[j]; re[j] = t;
            t = im[i]; im[i] = im[j]; im[j] = t;
        }
    }
    for (int len = 2; len <= n; len <<= 1) {
        for (int k = 0; k < len / 2; k++) {
            double a = sign * 2.0 * M_PI * k / len;
            double wr = cos(a), wi = sin(a);
            for (int i = k; i < n; i += len) {
                int j = i + len / 2;
                double xr = re[j] * wr - im[j] * wi;
                double xi = re[j] * wi + im[j] * wr;
                re[j] = re[i] - xr; im[j] = im[i] - xi;
                re[i] += xr; im[i] += xi;
            }
        }
    }
}

/* FFTW's FFTW_FORWARD (-1) / FFTW_BACKWARD (+1) unnormalised 2D transform
 * (fft_run.c:34,42,59,67), restated as row then column radix-2 passes. */
int oracle_fft2d_f64(int G, const double *in, double *out, int sign)
{
    if (G <= 0 || (G & (G - 1))) return -1;
    double *re = (double *)malloc(sizeof(double) * G), *im = (double *)malloc(sizeof(double) * G);
    size_t n = (size_t)G * G;
    if (out != in) memcpy(out, in, sizeof(double) * 2 * n);
    for (int r = 0; r < G; r++) {
        for (int c = 0; c < G; c++) { re[c] = out[2 * ((size_t)r * G + c)]; im[c] = out[2 * ((size_t)r * G + c) + 1]; }
        fft1d_f64(G, re, im, sign);
        for (int c = 0; c < G; c++) { out[2 * ((size_t)r * G + c)] = re[c]; out[2 * ((size_t)r * G + c) + 1] = im[c]; }
    }
    for (int c = 0; c < G; c++) {
        for (int r = 0; r < G; r++) { re[r] = out[2 * ((size_t)r * G + c)]; im[r] = out[2 * ((size_t)r * G + c) + 1]; }
        fft1d_f64(G, re, im, sign);
        for (int r = 0; r < G; r++) { out[2 * ((size_t)r * G + c)] = re[r]; out[2 * ((size_t)r * G + c) + 1] = im[r]; }
    }
    free(re); free(im);
    return 0;
}

void oracle_dft2d_direct_f64(int G, const double *in, double *out, int sign)
{
    for (int k1 = 0; k1 < G; k1++)
        for (int k2 = 0; k2 < G; k2++) {
            double sr = 0, si = 0;
            for (int j1 = 0; j1 < G; j1++)
                for (int j2 = 0; j2 < G; j2++) {
                    long ph = ((long)j1 * k1 + (long)j2 * k2) % G;
                    double a = sign * 2.0 * M_PI * (double)ph / G;
                    double xr = in[2 * ((size_t)j1 * G + j2)], xi = in[2 * ((size_t)j1 * G + j2) + 1];
                    sr += xr * cos(a) - xi * sin(a);
                    si += xr * sin(a) + xi * cos(a);
                }
            out[2 * ((size_t)k1 * G + k2)] = sr;
            out[2 * ((size_t)k1 * G + k2) + 1] = si;
        }
}

/* subtract_from_measurements_finegrid, gains_apply_run.c:120-144 */
void oracle_subtract_from_measurements_finegrid(int NUM_VISIBILITIES, const float2 *gridded_measured_vis,
        const float2 *input_vis, const int *num_finegrid_vis, float2 *delta_vis)
{
    (void)NUM_VISIBILITIES;
    for (int i = 0; i < *num_finegrid_vis; ++i) {
        delta_vis[i].x = gridded_measured_vis[i].x - input_vis[i].x;
        delta_vis[i].y = gridded_measured_vis[i].y - input_vis[i].y;
    }
}

static inline float2 cmul_apply(float2 a, float2 b)
{
    /* complex_multiply_apply, gains_apply_run.c:146-149 */
    float2 r;
    r.x = a.x * b.x - a.y * b.y;
    r.y = a.x * b.y + a.y * b.x;
    return r;
}

/* subtract_from_measurements, gains_apply_run.c:70-96 */
void oracle_subtract_from_measurements(int NUM_RECEIVERS, int NUM_BASELINES, int NUM_VISIBILITIES,
        const float2 *measured_vis, const float2 *visibilities_in, const float2 *gains,
        const int2 *receiver_pairs, float2 *visibilities_out)
{
    (void)NUM_RECEIVERS;
    for (int i = 0; i < NUM_VISIBILITIES; i++) {
        int b = i % NUM_BASELINES;
        float2 ga = gains[receiver_pairs[b].x];
        float2 gb = gains[receiver_pairs[b].y];
        gb.y = -gb.y;
        float2 m = cmul_apply(measured_vis[i], cmul_apply(ga, gb));
        visibilities_out[i].x = m.x - visibilities_in[i].x;
        visibilities_out[i].y = m.y - visibilities_in[i].y;
    }
}

/* reciprocal_transform_actor, gains_apply_run.c:4-17: each gain's
 * reciprocal as conj(z) / (x^2 + y^2), every operation in float */
void oracle_reciprocal_transform(int NUM_RECEIVERS, const float2 *gains_in, float2 *gains_out)
{
    for (int i = 0; i < NUM_RECEIVERS; i++) {
        const float2 z = gains_in[i];
        const float re = z.x / (z.x * z.x + z.y * z.y);
        const float im = z.y / (z.x * z.x + z.y * z.y);
        gains_out[i].x = re;
        gains_out[i].y = -im;
    }
}

/* apply_gains_only, gains_apply_run.c:98-118 */
void oracle_apply_gains_only(int NUM_RECEIVERS, int NUM_BASELINES, int NUM_VISIBILITIES,
        const float2 *measured_vis, const float2 *gains, const int2 *receiver_pairs,
        float2 *visibilities_out)
{
    (void)NUM_RECEIVERS;
    for (int i = 0; i < NUM_VISIBILITIES; i++) {
        int b = i % NUM_BASELINES;
        float2 ga = gains[receiver_pairs[b].x];
        float2 gb = gains[receiver_pairs[b].y];
        gb.y = -gb.y;
        visibilities_out[i] = cmul_apply(measured_vis[i], cmul_apply(ga, gb));
    }
}

/* add_visibilities, gridding_run.c:139-146 */
void oracle_add_visibilities(int NUM_VISIBILITIES, const float2 *v1, const float2 *v2, float2 *output)
{
    for (int i = 0; i < NUM_VISIBILITIES; ++i) {
        output[i].x = v1[i].x + v2[i].x;
        output[i].y = v1[i].y + v2[i].y;
    }
}

/* merge_gridded_visibilities, splitmerge.c:14-24 */
void oracle_merge_gridded_visibilities(int GRID_SIZE, int SLICES, const float2 *grids, float2 *output_grid)
{
    size_t n = (size_t)GRID_SIZE * GRID_SIZE;
    for (size_t i = 0; i < n; ++i) {
        float2 s = {0, 0};
        for (int j = 0; j < SLICES; ++j) {
            s.x += grids[(size_t)j * n + i].x;
            s.y += grids[(size_t)j * n + i].y;
        }
        output_grid[i] = s;
    }
}

/* slice counts of split_fftdegrid_visibilities, splitmerge.c:40-47 */
void oracle_split_fftdegrid_counts(int NUM_VISIBILITIES, int SLICES, int num_corrected_vis, int *num_vis)
{
    int left = num_corrected_vis, ideal = NUM_VISIBILITIES / SLICES;
    for (int i = 0; i < SLICES; ++i) {
        int cur = left > ideal ? ideal : left;
        left -= cur;
        num_vis[i] = cur;
    }
}

/* prolate_spheroidal, top.c:408-447: Schwab's rational approximation; the
 * coefficient table is the published one and must be the same values. */
double oracle_prolate_spheroidal(double nu)
{
    static const double p[] = {0.08203343, -0.3644705, 0.627866, -0.5335581, 0.2312756,
                               0.004028559, -0.03697768, 0.1021332, -0.1201436, 0.06412774};
    static const double q[] = {1.0, 0.8212018, 0.2078043, 1.0, 0.9599102, 0.2918724};
    int part;
    double nuend;
    if (nu >= 0.0 && nu < 0.75) { part = 0; nuend = 0.75; }
    else if (nu >= 0.75 && nu < 1.0) { part = 1; nuend = 1.0; }
    else return 0.0;
    double delta = nu * nu - nuend * nuend;
    double top = p[part * 5], bottom = q[part * 3];
    for (int i = 1; i < 5; i++) top += p[part * 5 + i] * pow(delta, i);
    for (int i = 1; i < 3; i++) bottom += q[part * 3 + i] * pow(delta, i);
    return (bottom == 0.0) ? 0.0 : top / bottom;
}

/* create_1D_half_prolate, top.c:395-404 */
void oracle_create_1D_half_prolate(float *prolate, int grid_size)
{
    int half = grid_size / 2;
    for (int i = 0; i < half; ++i) prolate[i] = (float)oracle_prolate_spheroidal((double)i / (double)half);
}

/* conv2D, major_loop_iter.c:8-46: scatter form, sources in raster order */
void oracle_conv2d(const float *image, const float *filter, float *result, int IMAGE_DIM, int hx, int hy)
{
    memset(result, 0, sizeof(float) * (size_t)IMAGE_DIM * IMAGE_DIM);
    for (int y = 0; y < IMAGE_DIM; ++y)
        for (int x = 0; x < IMAGE_DIM; ++x) {
            const int idx = y * IMAGE_DIM + x;
            if (image[idx] < 1e-5) continue;
            const int center = IMAGE_DIM / 2 + 1;
            for (int fy = -hy; fy < hy; ++fy)
                for (int fx = -hx; fx < hx; ++fx) {
                    const int ix = x + fx, iy = y + fy;
                    if (ix < 0 || ix >= IMAGE_DIM || iy < 0 || iy >= IMAGE_DIM) continue;
                    const int cfx = center + (2 * hx - (fx + hx) - hx);
                    const int cfy = center + (2 * hy - (fy + hy) - hy);
                    result[iy * IMAGE_DIM + ix] += image[idx] * filter[cfy * IMAGE_DIM + cfx];
                }
        }
}

/* gaussian2D, top.c:520-523 */
static float oracle_gaussian2d(float x, float y, float x_sigma, float y_sigma)
{
    const float pi = (float)3.141592654;
    return exp(-(x * x) / (2 * x_sigma * x_sigma) - (y * y) / (2 * y_sigma * y_sigma)) / (2 * pi * x_sigma * y_sigma);
}

/* clean_psf_host_set_up, top.c:527-592 (prev_sign is never updated there:
 * a walk ends at its first non-decreasing step) */
void oracle_clean_psf(int GRID_SIZE, double psf_max_value, const float *dirty_psf, float *clean_psf, int *halfdim)
{
    int center_box[4] = {GRID_SIZE / 2};
    const int dx[4] = {1, 0, -1, 0}, dy[4] = {0, 1, 0, -1};
    const int cx = GRID_SIZE / 2 + 1, cy = GRID_SIZE / 2 + 1;
    for (int i = 0; i < 4; ++i) {
        const int prev_sign = -1;
        for (int j = 1; j < GRID_SIZE / 2; ++j) {
            const int cur = (cx + dx[i] * j) + (cy + dy[i] * j) * GRID_SIZE;
            const int last = (cx + dx[i] * (j - 1)) + (cy + dy[i] * (j - 1)) * GRID_SIZE;
            const float deriv = dirty_psf[cur] - dirty_psf[last];
            const int sign = deriv >= 0.f ? 1 : -1;
            if (sign != prev_sign) {
                center_box[i] = j;
                break;
            }
        }
    }
    int h = center_box[0];
    for (int i = 1; i < 4; ++i)
        if (center_box[i] > h) h = center_box[i];
    if (h > GRID_SIZE / 2 - 2) h = GRID_SIZE / 2 - 2;   /* the build's cap: the box stays inside the image */
    *halfdim = h;
    memset(clean_psf, 0, sizeof(float) * (size_t)GRID_SIZE * GRID_SIZE);
    const float psf_max_val = (float)psf_max_value;
    const float x_sigma = h / 2.f, y_sigma = h / 2.f;
    float sum = 0.f;
    for (int y = -h; y < h; ++y)
        for (int x = -h; x < h; ++x) {
            const int idx = cx + x + (cy + y) * GRID_SIZE;
            clean_psf[idx] = oracle_gaussian2d(x, y, x_sigma, y_sigma);
            clean_psf[idx] *= clean_psf[idx];
            sum += clean_psf[idx];
        }
    for (int y = -h; y < h; ++y)
        for (int x = -h; x < h; ++x) {
            const int idx = cx + x + (cy + y) * GRID_SIZE;
            clean_psf[idx] = clean_psf[idx] / sum * psf_max_val;
        }
}

/* The reference's only intra-node parallelism (splitmerge.c:14-56): the
 * visibilities cut into SLICES equal runs, each gridded into its own grid
 * (or degridded) by its own thread, the grids then summed in slice order
 * (merge_gridded_visibilities).  For the CPU baseline's multi-core mode. */
typedef struct {
    int G, n, nk, tks, os;
    const float2 *kernels;
    const int2 *supports;
    const float3 *uvw;
    const float2 *vis;
    const Config *config;
    float2 *grid;          /* gridding: this slice's grid */
    const float2 *in_grid; /* degridding: the grid read */
    float2 *out_vis;       /* degridding: this slice's outputs */
    size_t c0, c1;         /* merge: cell range */
    const float2 *grids;
    int slices;
    float2 *merged;
} slice_job;

static void *grid_slice(void *p)
{
    slice_job *j = (slice_job *)p;
    const int zero = 0;
    oracle_std_gridding(j->G, j->n, j->nk, j->tks, j->os, 0, &zero, &j->n, j->kernels, j->supports, j->uvw, j->vis,
                        j->config, NULL, j->grid);
    return NULL;
}

static void *degrid_slice(void *p)
{
    slice_job *j = (slice_job *)p;
    oracle_std_degridding(j->G, j->n, j->nk, j->tks, j->os, j->kernels, j->supports, j->in_grid, j->uvw, &j->n,
                          j->config, j->out_vis);
    return NULL;
}

static void *merge_slice(void *p)
{
    slice_job *j = (slice_job *)p;
    const size_t n = (size_t)j->G * j->G;
    for (size_t i = j->c0; i < j->c1; ++i) {
        float2 s = {0, 0};
        for (int k = 0; k < j->slices; ++k) {
            s.x += j->grids[(size_t)k * n + i].x;
            s.y += j->grids[(size_t)k * n + i].y;
        }
        j->merged[i] = s;
    }
    return NULL;
}

int oracle_sliced_grid_degrid(int SLICES, int GRID_SIZE, int NUM_VISIBILITIES, int NUM_KERNELS,
        int TOTAL_KERNEL_SAMPLES, int OVERSAMPLING_FACTOR, const float2 *grid_kernels, const float2 *degrid_kernels,
        const int2 *kernel_supports, const float3 *uvw, const float2 *vis, const Config *config,
        float2 *output_grid, float2 *output_vis, double *seconds)
{
    if (SLICES < 1 || SLICES > 256) return -1;
    const size_t cells = (size_t)GRID_SIZE * GRID_SIZE;
    float2 *grids = (float2 *)calloc(cells * (size_t)SLICES, sizeof(float2));
    slice_job *jobs = (slice_job *)calloc((size_t)SLICES, sizeof(slice_job));
    pthread_t *th = (pthread_t *)calloc((size_t)SLICES, sizeof(pthread_t));
    if (!grids || !jobs || !th) {
        free(grids); free(jobs); free(th);
        return -1;
    }
    struct timespec t0, t1, t2;
    for (int k = 0; k < SLICES; ++k) {
        const int b = (int)((long long)NUM_VISIBILITIES * k / SLICES);
        const int e = (int)((long long)NUM_VISIBILITIES * (k + 1) / SLICES);
        slice_job j = {GRID_SIZE, e - b, NUM_KERNELS, TOTAL_KERNEL_SAMPLES, OVERSAMPLING_FACTOR, grid_kernels,
                       kernel_supports, uvw + b, vis + b, config, grids + cells * (size_t)k, output_grid,
                       output_vis + b, cells * (size_t)k / SLICES, cells * (size_t)(k + 1) / SLICES, grids, SLICES,
                       output_grid};
        jobs[k] = j;
    }
    clock_gettime(CLOCK_MONOTONIC, &t0);
    for (int k = 0; k < SLICES; ++k) pthread_create(&th[k], NULL, grid_slice, &jobs[k]);
    for (int k = 0; k < SLICES; ++k) pthread_join(th[k], NULL);
    for (int k = 0; k < SLICES; ++k) pthread_create(&th[k], NULL, merge_slice, &jobs[k]);
    for (int k = 0; k < SLICES; ++k) pthread_join(th[k], NULL);
    clock_gettime(CLOCK_MONOTONIC, &t1);
    for (int k = 0; k < SLICES; ++k) jobs[k].kernels = degrid_kernels;
    for (int k = 0; k < SLICES; ++k) pthread_create(&th[k], NULL, degrid_slice, &jobs[k]);
    for (int k = 0; k < SLICES; ++k) pthread_join(th[k], NULL);
    clock_gettime(CLOCK_MONOTONIC, &t2);
    seconds[0] = (double)(t1.tv_sec - t0.tv_sec) + 1e-9 * (double)(t1.tv_nsec - t0.tv_nsec);
    seconds[1] = (double)(t2.tv_sec - t1.tv_sec) + 1e-9 * (double)(t2.tv_nsec - t1.tv_nsec);
    free(grids);
    free(jobs);
    free(th);
    return 0;
}
