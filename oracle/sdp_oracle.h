/*
 * sdp_oracle.h -- CPU restatement of the reference actors (TEST
 * INFRASTRUCTURE ONLY).
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use
 * this library, and only as the checker or as the timed CPU baseline; the
 * product library never links or calls it.
 *
 * Each oracle_* function follows one reference function operation by
 * operation (same float/double promotions, same loop order, no FMA
 * contraction), cited at its definition.  Where the reference reads out of
 * bounds the oracle takes one padded element instead (documented per
 * function).
 */
#ifndef SDP_ORACLE_H
#define SDP_ORACLE_H

#include "sdp_types.h"

#ifdef __cplusplus
extern "C" {
#endif

void oracle_correct_to_finegrid(int NUM_VISIBILITIES, int GRID_SIZE, int OVERSAMPLING_FACTOR,
        int PERFORM_SIMPLIFICATION, const float3 *vis_uvw_coords, const float2 *input_visibilities,
        const Config *config, float2 *output_visibilities, float3 *output_finegrid_vis_coords,
        int *num_output_visibilities);

void oracle_std_gridding(int GRID_SIZE, int NUM_VISIBILITIES, int NUM_KERNELS, int TOTAL_KERNEL_SAMPLES,
        int OVERSAMPLING_FACTOR, int BYPASS, const int *maj_iter, const int *num_corrected_visibilities,
        const float2 *kernels, const int2 *kernel_supports, const float3 *corrected_vis_uvw_coords,
        const float2 *visibilities, const Config *config, const float2 *prev_grid, float2 *output_grid);

/* same index arithmetic as oracle_std_gridding, accumulation in double */
void oracle_std_gridding_f64(int GRID_SIZE, int num_vis, int OVERSAMPLING_FACTOR,
        const float2 *kernels, const int2 *kernel_supports, const float3 *uvw,
        const float2 *visibilities, double uv_scale, double w_scale, double *grid /* 2*G*G */);

void oracle_std_degridding(int GRID_SIZE, int NUM_VISIBILITIES, int NUM_KERNELS, int TOTAL_KERNEL_SAMPLES,
        int OVERSAMPLING_FACTOR, const float2 *kernels, const int2 *kernel_supports,
        const float2 *input_grid, const float3 *corrected_vis_uvw_coords,
        const int *num_corrected_visibilities, const Config *config, float2 *output_visibilities);

/* kernel-weighted degridding in double: the pre-override loop of
 * degridgrid.c:147-186 with plane offset supports[w].y, zero padding instead
 * of mirroring and no comm_norm division (the adjoint of std_gridding when
 * the table is conjugated). */
void oracle_degrid_weighted_f64(int GRID_SIZE, int num_vis, int OVERSAMPLING_FACTOR,
        const float2 *kernels, const int2 *kernel_supports, const float2 *grid, const float3 *uvw,
        double uv_scale, double w_scale, double *out /* 2*num_vis */);

void oracle_gridding_actor(int GRID_SIZE, int NUM_VISIBILITIES, int NUM_KERNELS, int TOTAL_KERNEL_SAMPLES,
        const float2 *kernels, const int2 *kernel_supports, const float3 *vis_uvw_coords,
        const float2 *visibilities, const Config *config, float2 *uv_grid);

/* gridding_actor's indices with double products and sums */
void oracle_gridding_actor_f64(int GRID_SIZE, int NUM_VISIBILITIES, const float2 *kernels,
        const int2 *kernel_supports, const float3 *vis_uvw_coords, const float2 *visibilities,
        double uv_scale, double w_scale, double *grid /* 2*G*G */);
/* the float recurrence hogbom_clean's peak scan calls its average */
float oracle_running_mean(const float *r, long n);

void oracle_dft_actor(int NUM_VISIBILITIES, int NUM_MAX_SOURCES, const float3 *sources,
        const float3 *vis_uvw_coords, const int *num_sources, const Config *config, float2 *visibilities);
/* same formula evaluated in double from the float inputs */
void oracle_dft_f64(int num_vis, int num_sources, const float3 *sources, const float3 *uvw, double *out);

/* psf: G*G+1 floats, psf[G*G] = 0 stands for the element the reference reads
 * one past the end (deconvolution_run.c:269).  residual_out (optional)
 * receives the residual after the last subtraction. */
void oracle_hogbom_clean(int GRID_SIZE, int NUM_MINOR_CYCLES, int MAX_SOURCES, const float *residual,
        const float *psf, const Config *config, const float *current_model, int *num_sources_out,
        float3 *sources_out, float *output_model, float *residual_out);

void oracle_shift_c2c(int GRID_SIZE, const float2 *in, float2 *out);
void oracle_shift_c2r(int GRID_SIZE, const float2 *in, float *out);
void oracle_shift_r2c(int GRID_SIZE, const float *in, float2 *out);
/* prolate: G/2+1 floats, prolate[G/2] = 0 (the element read one past the
 * end at convolution_correction_run.c:23) */
void oracle_convolution_correction(int GRID_SIZE, const float *in, const float *prolate, float *out);
void oracle_inv_convolution_correction(int GRID_SIZE, const float *in, float *out);

/* unnormalised 2D DFT of a G x G complex array in double, sign -1 or +1:
 * radix-2 FFT and the O(G^4) definition (tiny G only) */
int oracle_fft2d_f64(int G, const double *in, double *out, int sign);
void oracle_dft2d_direct_f64(int G, const double *in, double *out, int sign);

void oracle_subtract_from_measurements_finegrid(int NUM_VISIBILITIES, const float2 *gridded_measured_vis,
        const float2 *input_vis, const int *num_finegrid_vis, float2 *delta_vis);
void oracle_subtract_from_measurements(int NUM_RECEIVERS, int NUM_BASELINES, int NUM_VISIBILITIES,
        const float2 *measured_vis, const float2 *visibilities_in, const float2 *gains,
        const int2 *receiver_pairs, float2 *visibilities_out);
void oracle_reciprocal_transform(int NUM_RECEIVERS, const float2 *gains_in, float2 *gains_out);
void oracle_apply_gains_only(int NUM_RECEIVERS, int NUM_BASELINES, int NUM_VISIBILITIES,
        const float2 *measured_vis, const float2 *gains, const int2 *receiver_pairs,
        float2 *visibilities_out);
void oracle_add_visibilities(int NUM_VISIBILITIES, const float2 *v1, const float2 *v2, float2 *output);
void oracle_merge_gridded_visibilities(int GRID_SIZE, int SLICES, const float2 *grids, float2 *output_grid);
void oracle_split_fftdegrid_counts(int NUM_VISIBILITIES, int SLICES, int num_corrected_vis, int *num_vis);

double oracle_prolate_spheroidal(double nu);
void oracle_create_1D_half_prolate(float *prolate, int grid_size);

/* std_gridding then std_degridding over SLICES threads (splitmerge.c's
 * split, per-slice grids, slice-order merge); seconds[0] gridding with the
 * merge, seconds[1] degridding; 0 on success */
int oracle_sliced_grid_degrid(int SLICES, int GRID_SIZE, int NUM_VISIBILITIES, int NUM_KERNELS,
        int TOTAL_KERNEL_SAMPLES, int OVERSAMPLING_FACTOR, const float2 *grid_kernels, const float2 *degrid_kernels,
        const int2 *kernel_supports, const float3 *uvw, const float2 *vis, const Config *config,
        float2 *output_grid, float2 *output_vis, double *seconds);
void oracle_conv2d(const float *image, const float *filter, float *result, int IMAGE_DIM, int hx, int hy);
void oracle_clean_psf(int GRID_SIZE, double psf_max_value, const float *dirty_psf, float *clean_psf, int *halfdim);

#ifdef __cplusplus
}
#endif
#endif
