/*
 * sdp_pipeline.h -- the imaging major cycle on one GPU, everything resident
 * in device memory.
 *
 * One major cycle follows the reference's FFT imaging graph
 * (preesm_pipelines/simsdp_fft_imaging_pipeline/Algo/fft_delta.pi:61-320:
 * fft_run, fft_degrid, subtract_from_measurements_finegrid, grid,
 * inverse_fft_run, convolution_correction_run, then the Hogbom actor):
 *
 *   model image --(taper divide, shifts, forward FFT)--> model grid
 *   model grid  --(kernel-weighted degrid, minus measured)--> residual vis
 *   residual vis --(grid)--> residual grid
 *   residual grid --(shifts, inverse FFT, real part, taper divide)--> residual image
 *   residual image --(Hogbom CLEAN, deconvolution_run.c:197-312)--> model image += components
 *
 * The PSF is made the same way from unit visibilities and normalised by its
 * peak, as psf_host_set_up does with the PSF it reads (top.c:449-512).
 *
 * Residual RMS is reported in Jy: sqrt(mean(residual^2)) / psf_max.
 *
 * Differences from the reference graph, each selectable:
 *  - the predict step divides the model by the taper before the forward FFT
 *    (the reference's execute_inv_convolution_correction_actor is an
 *    identity copy, convolution_correction_run.c:41-68);
 *    SDP_PIPE_REFERENCE_PREDICT restores the identity;
 *  - degridding is kernel-weighted (std_degridding's nearest-cell override,
 *    degridgrid.c:194-200, is available through the actor ABI);
 *  - Hogbom subtracts the PSF in its stored orientation;
 *    SDP_PIPE_PSF_MIRRORED restores the reference's x-mirrored read
 *    (deconvolution_run.c:269).
 */
#ifndef SDP_PIPELINE_H
#define SDP_PIPELINE_H

#include "sdp_device.h"

#ifdef __cplusplus
extern "C" {
#endif

enum {
    SDP_PIPE_GRAPH = 1,              /* replay each cycle from a captured graph */
    SDP_PIPE_REFERENCE_PREDICT = 2,  /* no taper division before the forward FFT */
    SDP_PIPE_PSF_MIRRORED = 4,       /* Hogbom reads the PSF x-mirrored */
    SDP_PIPE_PROFILE = 8             /* time each stage with events (cycles run eagerly) */
};

typedef struct sdp_pipeline_params {
    int grid_size;          /* G, power of two in [16, 8192] */
    int oversampling;       /* kernel oversampling OS */
    int n_kernels;          /* w-planes */
    int n_vis;
    double uv_scale;        /* G * cell_size */
    double w_scale;         /* (n_kernels - 1)^2 / max_w */
    double cell_size;       /* radians per pixel */
    int n_minor;            /* Hogbom iterations per cycle */
    int max_sources;        /* capacity of the per-cycle source list */
    double loop_gain;
    double noise_factor;
    double weak_source_percent;
    int flags;              /* SDP_PIPE_* */
    int tile, chunk;        /* gridder tiling, 0 for the defaults */
} sdp_pipeline_params;

typedef struct sdp_pipeline sdp_pipeline;

/* Builds the gridding plan, the PSF and all work buffers.  d_uvw (n_vis
 * float3) and d_vis (n_vis measured float2) are only read here.  The kernel
 * tables (supports: host int2 {h, offset} per plane as kernel_host_set_up
 * lays them out, top.c:235-379; d_degrid_kernels the conjugate table) and
 * d_taper (correction_set_up's prolate spheroidal, G/2 + 1 floats with the
 * last one 0) are used by every cycle and must outlive the pipeline. */
int sdp_pipeline_create(sdp_pipeline **out, const sdp_pipeline_params *params, const int *supports,
                        const float *d_uvw, const float *d_vis, const float *d_grid_kernels,
                        const float *d_degrid_kernels, const float *d_taper, sdp_stream s);
int sdp_pipeline_destroy(sdp_pipeline *p);

/* One major cycle.  rms_residual: RMS (Jy) of the residual image formed in
 * this cycle, before CLEAN; rms_after: RMS after this cycle's CLEAN
 * subtraction; ms: wall time of the cycle on the stream.  Any output may be
 * NULL.  Synchronous. */
int sdp_pipeline_cycle(sdp_pipeline *p, double *rms_residual, double *rms_after, float *ms);

/* With SDP_PIPE_PROFILE: milliseconds of each stage of the last cycle
 * (SDP_STAGE_DEGRID_SUBTRACT: the fused degrid + subtract + grid pass into
 * the work items' sub-grids; SDP_STAGE_GRID: their merge into the grid),
 * ms[SDP_STAGE_COUNT]. */
enum { SDP_STAGE_PREDICT_FFT = 0, SDP_STAGE_DEGRID_SUBTRACT, SDP_STAGE_GRID, SDP_STAGE_INVERSE_FFT,
       SDP_STAGE_CLEAN, SDP_STAGE_COUNT };
int sdp_pipeline_stage_ms(const sdp_pipeline *p, float *ms);

/* Learned minor cycle.  With a network set, each major cycle, after forming
 * the residual, proposes a model update with it instead of running Hogbom:
 * the residual is cut into patch x patch tiles (their centres, patch -
 * 2 margin cells wide, tile the inner 80% of the image); tiles whose peak
 * |residual| is at least `threshold` times the image's are fed to the
 * network with the central PSF patch, each divided by a scale: its own
 * peak (SDP_LEARNED_TILE_SCALE) or the image's peak |residual|
 * (SDP_LEARNED_GLOBAL_SCALE).  The network returns the sky in those units;
 * gain times its values above `floor` (at pixels where the tile's residual
 * / scale exceeds input_floor; a negative input_floor keeps every pixel),
 * rescaled by scale / psf_max, is added to the model over each tile's
 * centre.  With Hogbom iterations set (n_minor > 0) the residual is then
 * updated by the PSF-convolved proposal and cleaned by Hogbom (the
 * convention of the training pairs made by sdp_learned.py).  Hogbom is the
 * fallback: when a cycle's residual RMS comes out above the previous
 * cycle's, the last proposal is undone and the pipeline runs Hogbom cycles
 * from then on; it also does so after max_cycles learned cycles (0: no
 * limit).  The network (height = width = patch, in_ch 2, out_ch 1) must
 * outlive the pipeline; NULL switches back to Hogbom.  Learned cycles run
 * eagerly. */
enum { SDP_LEARNED_TILE_SCALE = 0, SDP_LEARNED_GLOBAL_SCALE = 1 };
typedef struct sdp_learned_params {
    int patch, margin;
    double gain, threshold, floor, input_floor;
    int max_cycles;
    int scale;
} sdp_learned_params;
int sdp_pipeline_set_learned(sdp_pipeline *p, sdp_cnn *net, const sdp_learned_params *params);
/* 1 while learned cycles are in use, 0 after a fallback to Hogbom or without a network */
int sdp_pipeline_learned_active(const sdp_pipeline *p);

/* Restarts from an empty model (PSF and plan kept). */
int sdp_pipeline_reset(sdp_pipeline *p);
/* Replaces the model image with G x G floats from device memory d_model
 * (a prior model, or a minor-cycle proposal made elsewhere). */
int sdp_pipeline_load_model(sdp_pipeline *p, const float *d_model);

enum { SDP_IMAGE_RESIDUAL = 0, SDP_IMAGE_MODEL = 1, SDP_IMAGE_PSF = 2 };
/* Device pointer to one of the pipeline's G x G float images. */
const float *sdp_pipeline_image(const sdp_pipeline *p, int which);
double sdp_pipeline_psf_max(const sdp_pipeline *p);
/* Components found by the last cycle: up to max_sources float3 (l, m,
 * flux) copied to the host, count in *n. */
int sdp_pipeline_sources(const sdp_pipeline *p, float *h_sources, int *n);
/* Gridding plan used by the pipeline (for its info). */
const sdp_grid_plan *sdp_pipeline_plan(const sdp_pipeline *p);

/* ---- batched multi-channel residual imaging -----------------------------
 * No reference counterpart: the reference images one channel per MPI rank
 * (PN/generated/main.c:44-51), each running the cycle above.  Here
 * n_channels channels share one GPU: each has its own visibilities (any
 * count, e.g. after per-channel flagging), model and residual images; one
 * gridding plan covers all channels (sdp_grid_plan_create_multi), and the
 * model images, complex grids and residual images are stacked.
 * sdp_batch_residuals runs the first four stages of the major cycle for
 * all channels at once (models -> batched predict FFT -> degrid and
 * subtract -> grid -> batched inverse FFT and correction -> residual
 * images), each stage in the launches one channel would take; with
 * SDP_BATCH_GRAPH the pass is captured once (after one eager pass) and
 * replayed as one graph per cycle.  tile and chunk (0: defaults) apply to
 * the plan (sdp_grid_autotune picks them).  Channel k's images match what a
 * plan of channel k alone gives (same tile and work-item size). */
enum { SDP_BATCH_GRAPH = 1, SDP_BATCH_PROFILE = 2 };
typedef struct sdp_batch_params {
    int n_channels, grid_size, oversampling, n_kernels;
    double uv_scale, w_scale;
    int tile, chunk, accumulate, flags;
} sdp_batch_params;
typedef struct sdp_batch sdp_batch;
/* supports: host (half_support, offset) pairs; n_vis: host counts per
 * channel; d_uvw, d_vis: host arrays of n_channels device pointers (uvw in
 * wavelengths for that channel's frequency, float3; measured float2). */
int sdp_batch_create(sdp_batch **batch, const sdp_batch_params *params, const int *supports, const int *n_vis,
                     const float *const *d_uvw, const float *const *d_vis, const float *d_grid_kernels,
                     const float *d_degrid_kernels, const float *d_taper, sdp_stream s);
int sdp_batch_destroy(sdp_batch *batch);
/* one multi-channel pass; *ms: its device time.  With SDP_BATCH_PROFILE
 * (eager, events between stages) stage_ms[SDP_STAGE_COUNT] receives the
 * stage times of the pass (SDP_STAGE_CLEAN stays 0). */
int sdp_batch_residuals(sdp_batch *batch, float *ms, float *stage_ms);
/* device G x G float images of one channel (the model may be written) */
float *sdp_batch_model(sdp_batch *batch, int channel);
const float *sdp_batch_residual(const sdp_batch *batch, int channel);
/* the plan over all channels (the same for every valid channel) */
const sdp_grid_plan *sdp_batch_plan(const sdp_batch *batch, int channel);

#ifdef __cplusplus
}
#endif
#endif
