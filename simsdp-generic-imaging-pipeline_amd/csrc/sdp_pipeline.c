/*
 * sdp_pipeline.c -- the imaging major cycle (include/sdp_pipeline.h) on top
 * of the device C-ABI (include/sdp_device.h).
 *
 * Everything stays in device memory between cycles: the measured
 * visibilities in the gridding plan's order, one complex grid (the model
 * grid during predict, then the residual grid), the residual, model and PSF
 * images.  A cycle is a fixed sequence of enqueued launches with no host
 * synchronisation inside, so after a first eager cycle its stages 1-5 can
 * be captured once into a graph and replayed (SDP_PIPE_GRAPH); Hogbom is
 * enqueued after the replay, where it can run as one cooperative launch.
 *
 * Stage order per cycle (preesm_pipelines/simsdp_fft_imaging_pipeline/
 * Algo/fft_delta.pi:61-320; actor sources under .../Code/src/CPU/):
 *   1. fft_run: fft_shift_real_to_complex, forward FFT,
 *      fft_shift_complex_to_complex (fft_run.c:5-20,50-73,94-111), with the
 *      taper division of the inverse correction fused on the way in;
 *   2. fft_degrid + subtract_from_measurements_finegrid (degridgrid.c:127-202,
 *      gains_apply_run.c:120-144), kernel-weighted, and the gridding of the
 *      residual into the work items' sub-grids (degridgrid.c:69-125): one
 *      fused pass (sdp_degrid_grid_accumulate);
 *   3. the sub-grids merged into the residual grid (sdp_grid_merge);
 *   4. inverse_fft_run + convolution_correction_run (fft_run.c:23-48,76-92,
 *      convolution_correction_run.c:4-37), one fused FFT;
 *   5. residual statistics;
 *   6. hogbom_clean (deconvolution_run.c:197-312) on the residual image,
 *      adding its components to the model in place;
 *   7. residual statistics after CLEAN.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "sdp_pipeline.h"

struct sdp_pipeline {
    sdp_pipeline_params p;
    sdp_stream s;
    sdp_grid_plan *plan;
    const float *gk, *dk, *taper;
    float *vis_meas;        /* n_vis float2, plan order */
    float *vis_resid;       /* n_vis float2, plan order */
    float *grid;            /* G*G float2 */
    float *residual;        /* G*G float */
    float *model;           /* G*G float */
    float *psf;             /* G*G + 1 float, peak-normalised, last 0 */
    float *sources;         /* max_sources float3 */
    int *num_sources;
    void *fft_work, *hog_work, *stats_work;
    double *stats;          /* {sum_sq, max, max_abs} before and after CLEAN */
    double psf_max;
    int cycles_done;
    sdp_graph graph;
    sdp_event ev0, ev1;
    sdp_event stage_ev[SDP_STAGE_COUNT + 1];
    float stage_ms[SDP_STAGE_COUNT];
    /* learned minor cycle */
    sdp_cnn *net;
    int patch, margin, batch_max, n_cand, learned, max_learned, learned_done;
    double gain, threshold, floor, in_floor, last_rms;
    int global_scale;
    int have_backup;
    int *origins, *h_origins;   /* device / host int pairs, n_cand of them */
    float *peaks, *h_peaks, *sel_peaks, *h_sel_peaks, *batch, *out, *model_backup;
    float *delta;     /* G*G: model update, then its dirty image */
    float *psf_ft;    /* G*G complex: psf_max / G^2 * FFT(PSF moved to the origin) */
};

#define TRY(x) do { int rc_ = (x); if (rc_ != SDP_OK) return rc_; } while (0)
#define FAIL(code, msg) sdp_set_last_error((code), "sdp_pipeline: " msg)

static size_t pixels(const sdp_pipeline *P) { return (size_t)P->p.grid_size * (size_t)P->p.grid_size; }

/* event between stages when profiling */
static int mark(sdp_pipeline *P, int stage)
{
    if (!(P->p.flags & SDP_PIPE_PROFILE)) return SDP_OK;
    return sdp_event_record(P->stage_ev[stage], P->s);
}

/* Stages 1-5 above: launches only, no synchronisation. */
static int enqueue_residual(sdp_pipeline *P)
{
    const sdp_pipeline_params *p = &P->p;
    const int G = p->grid_size;
    const int predict_flags = SDP_FFT_REAL_IN | SDP_FFT_SHIFT_IN | SDP_FFT_SHIFT_OUT |
                              ((p->flags & SDP_PIPE_REFERENCE_PREDICT) ? 0 : SDP_FFT_CORRECT_IN);
    TRY(mark(P, SDP_STAGE_PREDICT_FFT));
    TRY(sdp_fft2d(P->model, P->grid, G, SDP_FFT_FORWARD, predict_flags, P->taper, P->fft_work, P->s));
    TRY(mark(P, SDP_STAGE_DEGRID_SUBTRACT));
    /* the model grid is read by the fused pass before the merge overwrites it */
    TRY(sdp_degrid_grid_accumulate(P->plan, P->dk, P->gk, P->grid, P->vis_meas, SDP_ORDER_PLAN, P->vis_resid,
                                   P->s));
    TRY(mark(P, SDP_STAGE_GRID));
    TRY(sdp_grid_merge(P->plan, P->grid, P->s));
    TRY(mark(P, SDP_STAGE_INVERSE_FFT));
    TRY(sdp_fft2d(P->grid, P->residual, G, SDP_FFT_INVERSE,
                  SDP_FFT_SHIFT_IN | SDP_FFT_SHIFT_OUT | SDP_FFT_REAL_OUT | SDP_FFT_CORRECT_OUT, P->taper,
                  P->fft_work, P->s));
    TRY(mark(P, SDP_STAGE_CLEAN));
    TRY(sdp_image_stats_async(pixels(P), P->residual, P->stats, P->stats_work, P->s));
    return SDP_OK;
}

/* Stages 6-7: Hogbom on the residual, statistics after it. */
static int enqueue_clean(sdp_pipeline *P)
{
    const sdp_pipeline_params *p = &P->p;
    const int G = p->grid_size;
    TRY(sdp_hogbom_ex(G, p->n_minor, p->max_sources, P->residual, P->psf, P->psf_max, p->loop_gain,
                      p->noise_factor, p->weak_source_percent, p->cell_size, P->model, P->model, P->sources,
                      P->num_sources, P->hog_work,
                      (p->flags & SDP_PIPE_PSF_MIRRORED) ? 0 : SDP_HOGBOM_PSF_AS_STORED, P->s));
    TRY(sdp_image_stats_async(pixels(P), P->residual, P->stats + 3, P->stats_work, P->s));
    TRY(mark(P, SDP_STAGE_COUNT));
    return SDP_OK;
}

static int enqueue_cycle(sdp_pipeline *P)
{
    TRY(enqueue_residual(P));
    return enqueue_clean(P);
}

/* The network's proposal from the current residual (stats[0..2] already on
 * the host in h): tiles with a large enough peak, in batches. */
static int propose(sdp_pipeline *P, const double *h)
{
    const int G = P->p.grid_size;
    const double maxabs = h[2];
    if (!(maxabs > 0.0)) return SDP_OK;
    int n = 0;
    TRY(sdp_cnn_patch_peaks(G, P->patch, P->margin, P->residual, P->peaks, &n, P->s));
    TRY(sdp_memcpy_dtoh(P->h_peaks, P->peaks, sizeof(float) * (size_t)n, P->s));
    const int stride = P->patch - 2 * P->margin, npx = (G + stride - 1) / stride;
    int sel = 0;
    for (int q = 0; q < n; ++q) {
        if (!((double)P->h_peaks[q] >= P->threshold * maxabs) || !(P->h_peaks[q] > 0.f)) continue;
        /* only tiles whose centre lies in the inner 80% of the field, where
         * the gridding taper that the predict divides by stays large */
        const int cx = (q % npx) * stride + stride / 2 - G / 2, cy = (q / npx) * stride + stride / 2 - G / 2;
        if (10 * (cx < 0 ? -cx : cx) > 4 * G || 10 * (cy < 0 ? -cy : cy) > 4 * G) continue;
        P->h_origins[2 * sel] = (q % npx) * stride - P->margin;
        P->h_origins[2 * sel + 1] = (q / npx) * stride - P->margin;
        P->h_sel_peaks[sel] = P->global_scale ? (float)maxabs : P->h_peaks[q];
        ++sel;
    }
    if (sel == 0) return SDP_OK;
    TRY(sdp_memcpy_htod(P->origins, P->h_origins, sizeof(int) * 2 * (size_t)sel, P->s));
    TRY(sdp_memcpy_htod(P->sel_peaks, P->h_sel_peaks, sizeof(float) * (size_t)sel, P->s));
    /* each tile enters divided by its scale (its peak or the image's); the
     * network returns the sky in those units: model += gain * out * scale /
     * psf_max */
    const float out_scale = (float)(P->gain / P->psf_max);
    for (int b0 = 0; b0 < sel; b0 += P->batch_max) {
        const int nb = sel - b0 < P->batch_max ? sel - b0 : P->batch_max;
        TRY(sdp_cnn_gather_patches(G, P->patch, nb, P->origins + 2 * b0, P->sel_peaks + b0, P->residual, P->psf,
                                   P->batch, P->s));
        TRY(sdp_cnn_forward(P->net, P->batch, P->out, nb, P->s));
        TRY(sdp_cnn_scatter_add(G, P->patch, P->margin, nb, P->origins + 2 * b0, P->sel_peaks + b0, P->out,
                                out_scale, (float)P->floor, P->batch, (float)P->in_floor, P->model, P->s));
    }
    return SDP_OK;
}

/* A cycle that follows a proposal: form the residual; when the proposal
 * raised its RMS above the previous cycle's, undo it (the fallback) and
 * form the residual again.  Returns 1 when the proposal was undone. */
static int check_proposal(sdp_pipeline *P, double *h, int *undone)
{
    TRY(enqueue_residual(P));
    TRY(sdp_memcpy_dtoh(h, P->stats, sizeof(double) * 3, P->s));
    const double rms = sqrt(h[0] / (double)pixels(P));
    *undone = 0;
    if (P->have_backup && rms > P->last_rms) {
        P->learned = 0;
        *undone = 1;
        TRY(sdp_memcpy_dtod(P->model, P->model_backup, sizeof(float) * pixels(P), P->s));
        TRY(enqueue_residual(P));
        TRY(sdp_memcpy_dtoh(h, P->stats, sizeof(double) * 3, P->s));
    }
    P->have_backup = 0;
    P->last_rms = sqrt(h[0] / (double)pixels(P));
    return SDP_OK;
}

/* A Hogbom cycle after a proposal (the learned cycles are over or were
 * undone): checked residual, then Hogbom. */
static int checked_clean_cycle(sdp_pipeline *P, double *h)
{
    int undone = 0;
    TRY(check_proposal(P, h, &undone));
    TRY(enqueue_clean(P));
    TRY(sdp_memcpy_dtoh(h + 3, P->stats + 3, sizeof(double) * 3, P->s));
    return SDP_OK;
}

/* One learned cycle: checked residual, then the proposal (or, when the
 * previous proposal was undone, Hogbom). */
static int learned_cycle(sdp_pipeline *P, double *h)
{
    int undone = 0;
    TRY(check_proposal(P, h, &undone));
    if (undone) {
        TRY(enqueue_clean(P));
        TRY(sdp_memcpy_dtoh(h + 3, P->stats + 3, sizeof(double) * 3, P->s));
        return SDP_OK;
    }
    const double rms = P->last_rms;
    TRY(sdp_memcpy_dtod(P->model_backup, P->model, sizeof(float) * pixels(P), P->s));
    P->have_backup = 1;
    P->last_rms = rms;
    TRY(propose(P, h));
    if (P->p.n_minor > 0) {
        /* residual -= PSF (*) update (circular convolution, in the residual's
         * units), then Hogbom cleans what the proposal left */
        const int G = P->p.grid_size;
        const size_t n = pixels(P);
        TRY(sdp_axpby(n, 1.f, P->model, -1.f, P->model_backup, P->delta, P->s));
        TRY(sdp_fft2d(P->delta, P->grid, G, SDP_FFT_FORWARD, SDP_FFT_REAL_IN, NULL, P->fft_work, P->s));
        TRY(sdp_complex_multiply(n, P->grid, P->psf_ft, n, P->grid, P->s));
        TRY(sdp_fft2d(P->grid, P->delta, G, SDP_FFT_INVERSE, SDP_FFT_REAL_OUT, NULL, P->fft_work, P->s));
        TRY(sdp_axpby(n, 1.f, P->residual, -1.f, P->delta, P->residual, P->s));
        TRY(enqueue_clean(P));
        TRY(sdp_memcpy_dtoh(h + 3, P->stats + 3, sizeof(double) * 3, P->s));
    } else {
        h[3] = h[0];   /* no image-domain update: the next cycle measures the result */
        h[4] = h[1];
        h[5] = h[2];
    }
    return SDP_OK;
}

/* PSF: grid unit visibilities, inverse FFT with the same shifts and
 * correction as the residual, record the peak and divide by it
 * (psf_host_set_up, top.c:449-512). */
static int make_psf(sdp_pipeline *P)
{
    const int G = P->p.grid_size;
    TRY(sdp_fill_complex((size_t)P->p.n_vis, P->vis_resid, 1.0f, 0.0f, P->s));
    TRY(sdp_grid(P->plan, P->gk, P->vis_resid, SDP_ORDER_PLAN, P->grid, P->s));
    TRY(sdp_fft2d(P->grid, P->psf, G, SDP_FFT_INVERSE,
                  SDP_FFT_SHIFT_IN | SDP_FFT_SHIFT_OUT | SDP_FFT_REAL_OUT | SDP_FFT_CORRECT_OUT, P->taper,
                  P->fft_work, P->s));
    TRY(sdp_memset_zero(P->psf + pixels(P), sizeof(float), P->s));
    double h[3];
    TRY(sdp_image_stats_async(pixels(P), P->psf, P->stats, P->stats_work, P->s));
    TRY(sdp_memcpy_dtoh(h, P->stats, sizeof h, P->s));
    if (!(h[1] > 0.0)) return FAIL(SDP_E_ARG, "PSF peak is not positive (no visibility on the grid?)");
    P->psf_max = (double)(float)h[1];
    TRY(sdp_divide(pixels(P), P->psf, P->psf_max, P->s));
    return SDP_OK;
}

static void free_learned(sdp_pipeline *P);

static void release(sdp_pipeline *P)
{
    free_learned(P);
    if (P->graph) (void)sdp_graph_destroy(P->graph);
    if (P->ev0) (void)sdp_event_destroy(P->ev0);
    if (P->ev1) (void)sdp_event_destroy(P->ev1);
    for (int i = 0; i <= SDP_STAGE_COUNT; ++i)
        if (P->stage_ev[i]) (void)sdp_event_destroy(P->stage_ev[i]);
    if (P->plan) (void)sdp_grid_plan_destroy(P->plan);
    void *bufs[] = {P->vis_meas, P->vis_resid, P->grid, P->residual, P->model, P->psf, P->sources,
                    P->num_sources, P->fft_work, P->hog_work, P->stats_work, P->stats};
    for (size_t i = 0; i < sizeof bufs / sizeof bufs[0]; ++i) (void)sdp_device_free(bufs[i]);
    free(P);
}

int sdp_pipeline_create(sdp_pipeline **out, const sdp_pipeline_params *params, const int *supports,
                        const float *d_uvw, const float *d_vis, const float *d_grid_kernels,
                        const float *d_degrid_kernels, const float *d_taper, sdp_stream s)
{
    if (!out || !params || !supports || !d_uvw || !d_vis || !d_grid_kernels || !d_degrid_kernels || !d_taper)
        return FAIL(SDP_E_ARG, "null argument");
    *out = NULL;
    const sdp_pipeline_params *p = params;
    const int G = p->grid_size;
    if (G < 16 || G > 8192 || (G & (G - 1)) || p->n_vis < 1 || p->n_kernels < 1 || p->oversampling < 1 ||
        p->n_minor < 0 || p->max_sources < 1 || !(p->cell_size > 0.0))
        return FAIL(SDP_E_ARG, "bad parameters (grid size, counts or cell size)");
    TRY(sdp_init());

    sdp_pipeline *P = (sdp_pipeline *)calloc(1, sizeof *P);
    if (!P) return FAIL(SDP_E_NOMEM, "out of host memory");
    P->p = *p;
    P->s = s;
    P->gk = d_grid_kernels;
    P->dk = d_degrid_kernels;
    P->taper = d_taper;

    sdp_grid_params gp;
    memset(&gp, 0, sizeof gp);
    gp.grid_size = G;
    gp.oversampling = p->oversampling;
    gp.n_kernels = p->n_kernels;
    gp.snap_mode = SDP_SNAP_STD;
    gp.uv_scale = p->uv_scale;
    gp.w_scale = p->w_scale;
    gp.tile = p->tile;
    gp.chunk = p->chunk;

    const size_t npix = (size_t)G * G;
    const size_t vis_bytes = sizeof(float) * 2 * (size_t)p->n_vis;
    int rc = sdp_grid_plan_create(&P->plan, &gp, supports, d_uvw, p->n_vis, s);
    if (rc == SDP_OK) rc = sdp_device_malloc((void **)&P->vis_meas, vis_bytes);
    if (rc == SDP_OK) rc = sdp_device_malloc((void **)&P->vis_resid, vis_bytes);
    if (rc == SDP_OK) rc = sdp_device_malloc((void **)&P->grid, sizeof(float) * 2 * npix);
    if (rc == SDP_OK) rc = sdp_device_malloc((void **)&P->residual, sizeof(float) * npix);
    if (rc == SDP_OK) rc = sdp_device_malloc((void **)&P->model, sizeof(float) * npix);
    if (rc == SDP_OK) rc = sdp_device_malloc((void **)&P->psf, sizeof(float) * (npix + 4));
    if (rc == SDP_OK) rc = sdp_device_malloc((void **)&P->sources, sizeof(float) * 3 * (size_t)p->max_sources);
    if (rc == SDP_OK) rc = sdp_device_malloc((void **)&P->num_sources, 16);
    if (rc == SDP_OK) rc = sdp_device_malloc(&P->fft_work, sdp_fft_work_bytes(G));
    if (rc == SDP_OK) rc = sdp_device_malloc(&P->hog_work, sdp_hogbom_work_bytes(G));
    if (rc == SDP_OK) rc = sdp_device_malloc(&P->stats_work, sdp_image_stats_work_bytes());
    if (rc == SDP_OK) rc = sdp_device_malloc((void **)&P->stats, sizeof(double) * 8);
    if (rc == SDP_OK) rc = sdp_event_create(&P->ev0);
    if (rc == SDP_OK) rc = sdp_event_create(&P->ev1);
    for (int i = 0; rc == SDP_OK && i <= SDP_STAGE_COUNT && (p->flags & SDP_PIPE_PROFILE); ++i)
        rc = sdp_event_create(&P->stage_ev[i]);
    if (rc == SDP_OK) rc = sdp_grid_plan_to_plan_order(P->plan, d_vis, P->vis_meas, s);
    if (rc == SDP_OK) rc = make_psf(P);
    if (rc == SDP_OK) rc = sdp_pipeline_reset(P);
    if (rc != SDP_OK) {
        release(P);
        return rc;
    }
    *out = P;
    return SDP_OK;
}

int sdp_pipeline_destroy(sdp_pipeline *P)
{
    if (P) {
        (void)sdp_stream_sync(P->s);
        release(P);
    }
    return SDP_OK;
}

int sdp_pipeline_reset(sdp_pipeline *P)
{
    if (!P) return FAIL(SDP_E_ARG, "null pipeline");
    TRY(sdp_memset_zero(P->model, sizeof(float) * pixels(P), P->s));
    TRY(sdp_memset_zero(P->sources, sizeof(float) * 3 * (size_t)P->p.max_sources, P->s));
    TRY(sdp_memset_zero(P->num_sources, 16, P->s));
    P->cycles_done = 0;
    P->have_backup = 0;
    P->learned_done = 0;
    if (P->net) P->learned = 1;
    return sdp_stream_sync(P->s);
}

int sdp_pipeline_load_model(sdp_pipeline *P, const float *d_model)
{
    if (!P || !d_model) return FAIL(SDP_E_ARG, "null argument");
    TRY(sdp_memcpy_dtod(P->model, d_model, sizeof(float) * pixels(P), P->s));
    return sdp_stream_sync(P->s);
}

int sdp_pipeline_cycle(sdp_pipeline *P, double *rms_residual, double *rms_after, float *ms)
{
    if (!P) return FAIL(SDP_E_ARG, "null pipeline");
    TRY(sdp_event_record(P->ev0, P->s));
    double h[6];
    const int use_graph = (P->p.flags & SDP_PIPE_GRAPH) && !(P->p.flags & SDP_PIPE_PROFILE);
    if (P->net && P->learned && P->max_learned > 0 && P->learned_done >= P->max_learned) P->learned = 0;
    const int learned = P->net && P->learned;
    const int checked = !learned && P->net && P->have_backup;   /* first Hogbom cycle after a proposal */
    const int host_h = learned || checked;                        /* h filled on the host by the cycle */
    if (learned) P->learned_done++;
    if (learned) {
        TRY(learned_cycle(P, h));
    } else if (checked) {
        TRY(checked_clean_cycle(P, h));
    } else if (use_graph && P->cycles_done > 0) {
        if (!P->graph) {
            /* the first cycle ran eagerly, so every lazily created resource
             * (twiddle tables, kernel attributes) exists before capture */
            TRY(sdp_graph_begin_capture(P->s));
            int rc = enqueue_residual(P);
            sdp_graph g = NULL;
            int rc2 = sdp_graph_end_capture(P->s, &g);
            if (rc != SDP_OK) {
                if (g) (void)sdp_graph_destroy(g);
                return rc;
            }
            TRY(rc2);
            P->graph = g;
            TRY(sdp_event_record(P->ev0, P->s));
        }
        TRY(sdp_graph_launch(P->graph, P->s));
        TRY(enqueue_clean(P));
    } else {
        TRY(enqueue_cycle(P));
    }
    TRY(sdp_event_record(P->ev1, P->s));
    if (!host_h) TRY(sdp_memcpy_dtoh(h, P->stats, sizeof h, P->s));
    else TRY(sdp_event_sync(P->ev1));
    float t = 0.f;
    TRY(sdp_event_elapsed_ms(P->ev0, P->ev1, &t));
    if (P->p.flags & SDP_PIPE_PROFILE)
        for (int i = 0; i < SDP_STAGE_COUNT; ++i)
            TRY(sdp_event_elapsed_ms(P->stage_ev[i], P->stage_ev[i + 1], &P->stage_ms[i]));
    P->cycles_done++;
    const double n = (double)pixels(P);
    if (rms_residual) *rms_residual = sqrt(h[0] / n) / P->psf_max;
    if (rms_after) *rms_after = sqrt(h[3] / n) / P->psf_max;
    if (ms) *ms = t;
    return SDP_OK;
}

static void free_learned(sdp_pipeline *P)
{
    void *bufs[] = {P->origins, P->peaks, P->sel_peaks, P->batch, P->out, P->model_backup, P->delta, P->psf_ft};
    for (size_t i = 0; i < sizeof bufs / sizeof bufs[0]; ++i) (void)sdp_device_free(bufs[i]);
    free(P->h_origins);
    free(P->h_peaks);
    free(P->h_sel_peaks);
    P->origins = NULL;
    P->peaks = P->sel_peaks = P->batch = P->out = P->model_backup = P->delta = P->psf_ft = NULL;
    P->h_origins = NULL;
    P->h_peaks = P->h_sel_peaks = NULL;
    P->net = NULL;
    P->learned = 0;
    P->have_backup = 0;
}

int sdp_pipeline_set_learned(sdp_pipeline *P, sdp_cnn *net, const sdp_learned_params *lp)
{
    if (!P) return FAIL(SDP_E_ARG, "null pipeline");
    TRY(sdp_stream_sync(P->s));
    free_learned(P);
    if (!net) return SDP_OK;
    if (!lp) return FAIL(SDP_E_ARG, "null learned parameters");
    const int patch = lp->patch, margin = lp->margin;
    const double gain = lp->gain, threshold = lp->threshold;
    if (lp->scale != SDP_LEARNED_TILE_SCALE && lp->scale != SDP_LEARNED_GLOBAL_SCALE)
        return FAIL(SDP_E_ARG, "bad learned scale mode");
    if (lp->max_cycles < 0) return FAIL(SDP_E_ARG, "negative max_cycles");
    sdp_cnn_params cp;
    TRY(sdp_cnn_describe(net, &cp));
    if (cp.height != patch || cp.width != patch || cp.in_ch != 2 || cp.out_ch != 1)
        return FAIL(SDP_E_ARG, "network shape does not match the patch (height = width = patch, 2 in, 1 out)");
    if (margin < 0 || 2 * margin >= patch || !(gain > 0.0) || !(threshold >= 0.0))
        return FAIL(SDP_E_ARG, "bad margin, gain or threshold");
    const int G = P->p.grid_size;
    const int stride = patch - 2 * margin;
    const int npx = (G + stride - 1) / stride;
    P->n_cand = npx * npx;
    P->batch_max = cp.max_batch;
    const size_t tile = (size_t)patch * patch;
    int rc = sdp_device_malloc((void **)&P->origins, sizeof(int) * 2 * (size_t)P->n_cand);
    if (rc == SDP_OK) rc = sdp_device_malloc((void **)&P->peaks, sizeof(float) * (size_t)P->n_cand);
    if (rc == SDP_OK) rc = sdp_device_malloc((void **)&P->sel_peaks, sizeof(float) * (size_t)P->n_cand);
    if (rc == SDP_OK) rc = sdp_device_malloc((void **)&P->batch, sizeof(float) * 2 * tile * (size_t)P->batch_max);
    if (rc == SDP_OK) rc = sdp_device_malloc((void **)&P->out, sizeof(float) * tile * (size_t)P->batch_max);
    if (rc == SDP_OK) rc = sdp_device_malloc((void **)&P->model_backup, sizeof(float) * pixels(P));
    if (rc == SDP_OK) rc = sdp_device_malloc((void **)&P->delta, sizeof(float) * pixels(P));
    if (rc == SDP_OK) rc = sdp_device_malloc((void **)&P->psf_ft, sizeof(float) * 2 * pixels(P));
    if (rc == SDP_OK) {
        /* FFT of the PSF with its peak moved from (G/2, G/2) to the origin:
         * the output checkerboard is that shift, for even G */
        rc = sdp_fft2d(P->psf, P->psf_ft, G, SDP_FFT_FORWARD, SDP_FFT_REAL_IN | SDP_FFT_SHIFT_OUT, NULL, P->fft_work,
                       P->s);
        if (rc == SDP_OK) rc = sdp_divide(2 * pixels(P), P->psf_ft, (double)pixels(P) / P->psf_max, P->s);
        if (rc == SDP_OK) rc = sdp_stream_sync(P->s);
    }
    P->h_origins = (int *)malloc(sizeof(int) * 2 * (size_t)P->n_cand);
    P->h_peaks = (float *)malloc(sizeof(float) * (size_t)P->n_cand);
    P->h_sel_peaks = (float *)malloc(sizeof(float) * (size_t)P->n_cand);
    if (rc == SDP_OK && (!P->h_origins || !P->h_peaks || !P->h_sel_peaks)) rc = FAIL(SDP_E_NOMEM, "out of host memory");
    if (rc != SDP_OK) {
        free_learned(P);
        return rc;
    }
    P->net = net;
    P->patch = patch;
    P->margin = margin;
    P->gain = gain;
    P->threshold = threshold;
    P->floor = lp->floor;
    P->in_floor = lp->input_floor;
    P->max_learned = lp->max_cycles;
    P->global_scale = lp->scale == SDP_LEARNED_GLOBAL_SCALE;
    P->learned_done = 0;
    P->learned = 1;
    return SDP_OK;
}

int sdp_pipeline_learned_active(const sdp_pipeline *P) { return P && P->net && P->learned; }

const float *sdp_pipeline_image(const sdp_pipeline *P, int which)
{
    if (!P) return NULL;
    switch (which) {
    case SDP_IMAGE_RESIDUAL: return P->residual;
    case SDP_IMAGE_MODEL: return P->model;
    case SDP_IMAGE_PSF: return P->psf;
    default: return NULL;
    }
}

int sdp_pipeline_stage_ms(const sdp_pipeline *P, float *ms)
{
    if (!P || !ms) return FAIL(SDP_E_ARG, "null argument");
    if (!(P->p.flags & SDP_PIPE_PROFILE)) return FAIL(SDP_E_ARG, "pipeline created without SDP_PIPE_PROFILE");
    memcpy(ms, P->stage_ms, sizeof P->stage_ms);
    return SDP_OK;
}

double sdp_pipeline_psf_max(const sdp_pipeline *P) { return P ? P->psf_max : 0.0; }

int sdp_pipeline_sources(const sdp_pipeline *P, float *h_sources, int *n)
{
    if (!P || !n) return FAIL(SDP_E_ARG, "null argument");
    int cnt = 0;
    TRY(sdp_memcpy_dtoh(&cnt, P->num_sources, sizeof cnt, P->s));
    if (cnt < 0) cnt = 0;
    if (cnt > P->p.max_sources) cnt = P->p.max_sources;
    if (h_sources && cnt) TRY(sdp_memcpy_dtoh(h_sources, P->sources, sizeof(float) * 3 * (size_t)cnt, P->s));
    *n = cnt;
    return SDP_OK;
}

const sdp_grid_plan *sdp_pipeline_plan(const sdp_pipeline *P) { return P ? P->plan : NULL; }
