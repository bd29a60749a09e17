/*
 * sdp_actors.h -- drop-in actor C-ABI for the SDP imaging major cycle.
 *
 * Every function here keeps the exact name, argument order and argument
 * types of the reference actor it replaces, so that a PREESM-generated
 * schedule (e.g. param_code/code_fft_1core/generated/Core0.c) links against
 * libsdp_actors.so instead of the reference CPU/CUDA objects without any
 * change.  Paths in the citations are relative to
 * preesm_pipelines/simsdp_fft_imaging_pipeline/Code/ unless stated.
 *
 * Buffer contract (same as the reference): all pointers are host memory
 * owned by the caller; inputs are read only; outputs are written before the
 * call returns (the GPU work is synchronous from the caller's view).  The
 * FFT, shift and correction actors accept out == in.
 *
 * Errors: the reference aborts the process on fatal errors
 * (top.c:146, GPU/common.cu:200-208).  These actors do the same: on a HIP
 * error they print the failing call to stderr and abort().  sdp_last_error()
 * (sdp_device.h) reports the last message.
 */
#ifndef SDP_ACTORS_H
#define SDP_ACTORS_H

#include "sdp_types.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ---- gridding / degridding (include/degridgrid.h) ---------------------- */

/* replaces include/degridgrid.h:10 (src/CPU/degridgrid.c:12-65) */
void correct_to_finegrid(int NUM_VISIBILITIES, int GRID_SIZE, int OVERSAMPLING_FACTOR,
        int PERFORM_SIMPLIFICATION, PRECISION3* vis_uvw_coords, PRECISION2* input_visibilities,
        Config* config, PRECISION2* output_visibilities, PRECISION3* output_finegrid_vis_coords,
        int* num_output_visibilities);

/* replaces include/degridgrid.h:13 (src/CPU/degridgrid.c:69-125) */
void std_gridding(int GRID_SIZE, int NUM_VISIBILITIES, int NUM_KERNELS, int TOTAL_KERNEL_SAMPLES,
        int OVERSAMPLING_FACTOR, int BYPASS, int* maj_iter, int* num_corrected_visibilities,
        PRECISION2* kernels, int2* kernel_supports, PRECISION3* corrected_vis_uvw_coords,
        PRECISION2* visibilities, Config* config, PRECISION2* prev_grid, PRECISION2* output_grid);

/* replaces include/degridgrid.h:16 (src/CPU/degridgrid.c:127-202).
 * Default output is the reference's nearest-cell value (degridgrid.c:194-200);
 * sdp_set_degrid_mode(SDP_DEGRID_WEIGHTED) selects the kernel-weighted sum. */
void std_degridding(int GRID_SIZE, int NUM_VISIBILITIES, int NUM_KERNELS, int TOTAL_KERNEL_SAMPLES,
        int OVERSAMPLING_FACTOR, PRECISION2* kernels, int2* kernel_supports, PRECISION2* input_grid,
        PRECISION3* corrected_vis_uvw_coords, int* num_corrected_visibilities, Config* config,
        PRECISION2* output_visibilities);

/* ---- DFT-pipeline gridder (include/gridding_run.h) --------------------- */

/* replaces include/gridding_run.h:12 (src/CPU/gridding_run.c:5-137) */
void gridding_actor(int GRID_SIZE, int NUM_VISIBILITIES, int NUM_KERNELS, int TOTAL_KERNEL_SAMPLES,
        IN PRECISION2* kernels, IN int2* kernel_supports, IN PRECISION3* vis_uvw_coords,
        IN PRECISION2* visibilities, IN Config *config, OUT PRECISION2* uv_grid);

/* replaces include/gridding_run.h:21 (src/CPU/gridding_run.c:139-146) */
void add_visibilities(int NUM_VISIBILITIES, IN PRECISION2* v1, IN PRECISION2* v2,
        OUT PRECISION2* output);

/* ---- DFT predict (include/dft_run.h) ------------------------------------ */

/* replaces include/dft_run.h:15 (src/CPU/dft_run.c:3-47, src/GPU/dft_run.cu:80-116) */
void dft_actor(int NUM_VISIBILITIES, int NUM_MAX_SOURCES, IN PRECISION3 *sources,
        IN PRECISION3 *vis_uvw_coords, IN int *num_sources, IN Config *config,
        OUT PRECISION2 *visibilities);

/* ---- Hogbom minor cycle (include/deconvolution_run.h) ------------------- */

/* replaces include/deconvolution_run.h:14 (src/CPU/deconvolution_run.c:197-312).
 * The noise stop uses the reference's own float running average (sdp_hogbom
 * in sdp_device.h).  Unlike the reference, it stops after MAX_SOURCES
 * components instead of writing past sources_out. */
void hogbom_clean(int GRID_SIZE, int NUM_MINOR_CYCLES, int MAX_SOURCES, IN PRECISION* residual,
        IN PRECISION* partial_psf, int2* partial_psf_halfdims, IN Config* config,
        IN PRECISION* current_model, OUT int* num_sources_out, OUT PRECISION3* sources_out,
        OUT PRECISION* output_model);

/* ---- FFT and shifts (include/fft_run.h) --------------------------------- */

/* replaces include/fft_run.h:13 (src/CPU/fft_run.c:5-20) */
void fft_shift_complex_to_complex_actor(int GRID_SIZE, IN PRECISION2* uv_grid_in, IN Config *config,
        OUT PRECISION2* uv_grid_out);
/* replaces include/fft_run.h:15 (src/CPU/fft_run.c:50-73, FFTW_FORWARD, unnormalised) */
void CUFFT_EXECUTE_FORWARD_C2C_actor(int GRID_SIZE, IN PRECISION2* uv_grid_in, OUT PRECISION2* uv_grid_out);
/* replaces include/fft_run.h:17 (src/CPU/fft_run.c:23-48, FFTW_BACKWARD, unnormalised) */
void CUFFT_EXECUTE_INVERSE_C2C_actor(int GRID_SIZE, IN PRECISION2* uv_grid_in, OUT PRECISION2* uv_grid_out);
/* replaces include/fft_run.h:19 (src/CPU/fft_run.c:76-92) */
void fft_shift_complex_to_real_actor(int GRID_SIZE, IN PRECISION2* uv_grid, IN Config *config,
        OUT PRECISION* dirty_image);
/* replaces include/fft_run.h:21 (src/CPU/fft_run.c:94-111) */
void fft_shift_real_to_complex_actor(int GRID_SIZE, IN PRECISION *image, IN Config *config,
        OUT PRECISION2 *fourier);

/* ---- gridding correction (include/convolution_correction_run.h) --------- */

/* replaces include/convolution_correction_run.h:11 (src/CPU/convolution_correction_run.c:4-38) */
void execute_convolution_correction_actor(int GRID_SIZE, IN PRECISION* dirty_image_in,
        IN PRECISION* prolate, IN Config *config, OUT PRECISION* dirty_image_out);
/* replaces include/convolution_correction_run.h:12 (src/CPU/convolution_correction_run.c:41-71) */
void execute_inv_convolution_correction_actor(int GRID_SIZE, IN PRECISION* dirty_image_in,
        IN PRECISION* prolate, IN Config *config, OUT PRECISION* dirty_image_out);

/* ---- visibility arithmetic (include/gains_apply_run.h) ----------------- */

/* replaces include/gains_apply_run.h:14 (src/CPU/gains_apply_run.c:4-17): the
 * first actor of the FFT graph's correct_to_finegrid subgraph
 * (Algo/correct_to_finegrid.pi:11-24) */
void reciprocal_transform_actor(int NUM_RECEIVERS, IN PRECISION2* gains_in, IN Config *config,
        OUT PRECISION2* gains_out);
/* replaces include/gains_apply_run.h:23 (src/CPU/gains_apply_run.c:70-96) */
void subtract_from_measurements(int NUM_RECEIVERS, int NUM_BASELINES, int NUM_VISIBILITIES,
        IN PRECISION2* measured_vis, IN PRECISION2* visibilities_in, IN PRECISION2* gains,
        IN int2* receiver_pairs, IN Config *config, OUT PRECISION2* visibilities_out);
/* replaces include/gains_apply_run.h:28 (src/CPU/gains_apply_run.c:98-118) */
void apply_gains_only(int NUM_RECEIVERS, int NUM_BASELINES, int NUM_VISIBILITIES,
        IN PRECISION2* measured_vis, IN PRECISION2* gains, IN int2* receiver_pairs,
        IN Config *config, OUT PRECISION2* visibilities_out);
/* replaces include/gains_apply_run.h:31 (src/CPU/gains_apply_run.c:120-144) */
void subtract_from_measurements_finegrid(int NUM_VISIBILITIES, IN PRECISION2* gridded_measured_vis,
        IN PRECISION2* input_vis, int* num_finegrid_vis, OUT PRECISION2* delta_vis);

/* ---- slice split / merge (include/splitmerge.h) ------------------------- */

/* replaces include/splitmerge.h:15 (src/splitmerge.c:3-12) */
void split_visibilities(IN Config* config, IN int NUM_VISIBILITIES, IN int SLICES,
        IN PRECISION2* measured_visibilities, IN PRECISION3* vis_coords, OUT PRECISION2* visibilities,
        OUT PRECISION3* out_vis_coords, OUT Config* output_configs, OUT Config* output_config);
/* replaces include/splitmerge.h:17 (src/splitmerge.c:14-24) */
void merge_gridded_visibilities(IN int GRID_SIZE, IN int SLICES, IN PRECISION2* grids,
        OUT PRECISION2* output_grid);
/* replaces include/splitmerge.h:18 (src/splitmerge.c:26-28) */
void merge_configs(IN Config* in_configs, OUT Config* out_config);
/* replaces include/splitmerge.h:19 (src/splitmerge.c:30-48) */
void split_fftdegrid_visibilities(int NUM_VISIBILITIES, int SLICES, int GRID_SIZE, IN Config* config,
        IN PRECISION3* vis_coords, IN PRECISION2* measured_visibilities, IN int* num_corrected_vis,
        IN int* maj_iter, OUT PRECISION2* output_visibilities, OUT PRECISION3* out_vis_coords,
        OUT Config* output_configs, OUT int* num_vis, OUT int* maj_iter_out);
/* replaces include/splitmerge.h:21 (src/splitmerge.c:50-56) */
void split_uvgrid(int GRID_SIZE, int SLICES, IN PRECISION2* input_grid, OUT PRECISION2* output_grid);

/* ---- set-up actors (include/top.h) -------------------------------------- */

/* replaces include/top.h:17 (src/top.c:9-69) */
void config_struct_set_up(int GRID_SIZE, int NUM_KERNELS, OUT Config *config);
/* replaces include/top.h:27 (src/top.c:235-315): reads the support / real /
 * imaginary kernel CSV files named in config */
void kernel_host_set_up(int NUM_KERNELS, int TOTAL_KERNEL_SAMPLES, IN Config *config,
        OUT int2 *kernel_supports, OUT PRECISION2 *kernels);
/* replaces include/top.h:29 (src/top.c:318-379) */
void degridding_kernel_host_set_up(int NUM_DEGRIDDING_KERNELS, int TOTAL_DEGRIDDING_KERNEL_SAMPLES,
        IN Config *config, OUT int2 *degridding_kernel_supports, OUT PRECISION2 *degridding_kernels);
/* replaces include/top.h:32 (src/top.c:381-393) */
void correction_set_up(int GRID_SIZE, OUT PRECISION *prolate);
/* replaces include/top.h:34 (src/top.c:395-404) */
void create_1D_half_prolate(PRECISION *prolate, int grid_size);
/* replaces include/top.h:36 (src/top.c:408-447) */
double prolate_spheroidal(double nu);
/* replaces include/top.h:38 (src/top.c:449-512) */
void psf_host_set_up(int GRID_SIZE, int PSF_GRID_SIZE, IN Config *config, OUT PRECISION *psf,
        OUT double *psf_max_value);
/* replaces include/top.h:42 (src/top.c:527-592): the clean beam, a squared
 * Gaussian over the dirty PSF's central lobe; host-only (set-up) */
void clean_psf_host_set_up(int GRID_SIZE, int GAUSSIAN_CLEAN_PSF, IN Config *config, IN PRECISION *dirty_psf,
        OUT PRECISION *clean_psf, OUT int2 *partial_psf_halfdims);
/* replaces include/major_loop_iter.h:30 (src/major_loop_iter.c:253-278): the
 * output images as text files, the restored image (conv2D,
 * major_loop_iter.c:8-46) formed on the GPU by sdp_restore_conv2d */
void save_output(int GRID_SIZE, IN PRECISION *residual, IN PRECISION *model, IN PRECISION *clean_psf,
        int2 *clean_psf_halfdims, IN PRECISION *psf, IN Config *config, IN int *cycle);
/* replaces include/top.h:25 (src/top.c:140-232): the visibility text file,
 * a count then one "u v w re im weight" line per visibility (uvw in metres);
 * uvw scaled by frequency_hz / c, u and w negated when right_ascension, the
 * value weighted unless force_weight_to_one; host-only (file input) */
void visibility_host_set_up(int NUM_VISIBILITIES, IN Config *config, OUT PRECISION3 *vis_uvw_coords,
        OUT PRECISION2 *measured_vis);
/* replaces include/top.h:40 (src/top.c:514-518) */
void config_struct_set_up_sequel(IN Config *config_in, IN double *psf_max_value, OUT Config *config_out);

#ifdef __cplusplus
}
#endif

#endif /* SDP_ACTORS_H */
