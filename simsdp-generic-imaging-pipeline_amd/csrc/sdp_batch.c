/*
 * sdp_batch.c -- batched multi-channel residual imaging
 * (include/sdp_pipeline.h, sdp_batch_*).
 *
 * The reference images each frequency channel on its own MPI rank
 * (PN/generated/main.c:44-51) with the FFT graph's per-channel stages
 * (Algo/fft_delta.pi:61-320): fft_run, fft_degrid +
 * subtract_from_measurements_finegrid, grid, inverse_fft_run +
 * convolution_correction_run.  Here all channels live on one GPU as stacked
 * arrays: one gridding plan over every channel's visibilities (the channel
 * is part of the plan's (class, tile) keys, so one launch per kernel shape
 * covers all channels), measured and residual visibilities of all channels
 * in plan order, and n_channels consecutive model images, complex grids and
 * residual images.  A pass is one batched forward FFT (three launches for
 * all grids), the fused degrid + subtract + grid launches of the plan, its
 * merge, and one batched inverse FFT: a fixed launch sequence with no host
 * synchronisation, captured into one graph after an eager first pass.
 */
#include <stdlib.h>
#include <string.h>

#include "sdp_pipeline.h"

struct sdp_batch {
    sdp_batch_params p;
    sdp_stream s;
    const float *gk, *dk, *taper;
    sdp_grid_plan *plan;          /* all channels */
    float *vis_meas, *vis_resid;  /* all channels' visibilities, plan order, float2 */
    float *model, *residual;      /* n_channels x G*G floats */
    float *grid;                  /* n_channels x G*G float2 */
    void *fft_work;               /* n_channels x sdp_fft_work_bytes(G) */
    sdp_graph graph;
    int passes;
    sdp_event ev0, ev1;
    sdp_event stage_ev[SDP_STAGE_COUNT + 1];
};

#define TRY(x) do { int rc_ = (x); if (rc_ != SDP_OK) return rc_; } while (0)
#define FAIL(code, msg) sdp_set_last_error((code), "sdp_batch: " msg)

static size_t pixels(const sdp_batch *B) { return (size_t)B->p.grid_size * (size_t)B->p.grid_size; }

/* the residual images of every channel: launches only */
static int enqueue_pass(sdp_batch *B, int profile)
{
    const int G = B->p.grid_size, nc = B->p.n_channels;
    if (profile) TRY(sdp_event_record(B->stage_ev[0], B->s));
    TRY(sdp_fft2d_batch(B->model, B->grid, G, nc, SDP_FFT_FORWARD,
                        SDP_FFT_REAL_IN | SDP_FFT_SHIFT_IN | SDP_FFT_SHIFT_OUT | SDP_FFT_CORRECT_IN, B->taper,
                        B->fft_work, B->s));
    if (profile) TRY(sdp_event_record(B->stage_ev[1], B->s));
    TRY(sdp_degrid_grid_accumulate(B->plan, B->dk, B->gk, B->grid, B->vis_meas, SDP_ORDER_PLAN, B->vis_resid,
                                   B->s));
    if (profile) TRY(sdp_event_record(B->stage_ev[2], B->s));
    TRY(sdp_grid_merge(B->plan, B->grid, B->s));
    if (profile) TRY(sdp_event_record(B->stage_ev[3], B->s));
    TRY(sdp_fft2d_batch(B->grid, B->residual, G, nc, SDP_FFT_INVERSE,
                        SDP_FFT_SHIFT_IN | SDP_FFT_SHIFT_OUT | SDP_FFT_REAL_OUT | SDP_FFT_CORRECT_OUT, B->taper,
                        B->fft_work, B->s));
    if (profile) TRY(sdp_event_record(B->stage_ev[4], B->s));
    return SDP_OK;
}

static void release(sdp_batch *B)
{
    if (B->graph) (void)sdp_graph_destroy(B->graph);
    if (B->ev0) (void)sdp_event_destroy(B->ev0);
    if (B->ev1) (void)sdp_event_destroy(B->ev1);
    for (int i = 0; i <= SDP_STAGE_COUNT; ++i)
        if (B->stage_ev[i]) (void)sdp_event_destroy(B->stage_ev[i]);
    if (B->plan) (void)sdp_grid_plan_destroy(B->plan);
    (void)sdp_device_free(B->vis_meas);
    (void)sdp_device_free(B->vis_resid);
    (void)sdp_device_free(B->model);
    (void)sdp_device_free(B->residual);
    (void)sdp_device_free(B->grid);
    (void)sdp_device_free(B->fft_work);
    free(B);
}

int sdp_batch_create(sdp_batch **out, const sdp_batch_params *params, const int *supports, const int *n_vis,
                     const float *const *d_uvw, const float *const *d_vis, const float *d_grid_kernels,
                     const float *d_degrid_kernels, const float *d_taper, sdp_stream s)
{
    if (!out || !params || !supports || !n_vis || !d_uvw || !d_vis || !d_grid_kernels || !d_degrid_kernels ||
        !d_taper)
        return FAIL(SDP_E_ARG, "null argument");
    *out = NULL;
    const sdp_batch_params *p = params;
    const int G = p->grid_size;
    if (p->n_channels < 1 || p->n_channels > 65535 || G < 16 || G > 8192 || (G & (G - 1)) || p->n_kernels < 1 ||
        p->oversampling < 1 || !(p->uv_scale > 0.0) || (p->flags & ~(SDP_BATCH_GRAPH | SDP_BATCH_PROFILE)))
        return FAIL(SDP_E_ARG, "bad parameters (channels, grid size, counts, scale or flags)");
    long long total = 0;
    for (int k = 0; k < p->n_channels; ++k) {
        if (n_vis[k] < 1 || !d_uvw[k] || !d_vis[k]) return FAIL(SDP_E_ARG, "a channel without visibilities");
        total += n_vis[k];
    }
    if (total >= (1LL << 31)) return FAIL(SDP_E_UNSUPPORTED, "more than 2^31 - 1 visibilities over all channels");
    TRY(sdp_init());
    sdp_batch *B = (sdp_batch *)calloc(1, sizeof *B);
    if (!B) return FAIL(SDP_E_NOMEM, "out of host memory");
    B->p = *p;
    B->s = s;
    B->gk = d_grid_kernels;
    B->dk = d_degrid_kernels;
    B->taper = d_taper;
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
    gp.accumulate = p->accumulate;
    const size_t nc = (size_t)p->n_channels, npix = pixels(B), vb = sizeof(float) * 2 * (size_t)total;
    float *caller = NULL;   /* measured visibilities of all channels in channel order, then freed */
    int rc = sdp_grid_plan_create_multi(&B->plan, &gp, supports, p->n_channels, d_uvw, n_vis, s);
    if (rc == SDP_OK) rc = sdp_device_malloc((void **)&B->grid, sizeof(float) * 2 * npix * nc);
    if (rc == SDP_OK) rc = sdp_device_malloc(&B->fft_work, sdp_fft_work_bytes(G) * nc);
    if (rc == SDP_OK) rc = sdp_device_malloc((void **)&B->model, sizeof(float) * npix * nc);
    if (rc == SDP_OK) rc = sdp_device_malloc((void **)&B->residual, sizeof(float) * npix * nc);
    if (rc == SDP_OK) rc = sdp_device_malloc((void **)&B->vis_meas, vb);
    if (rc == SDP_OK) rc = sdp_device_malloc((void **)&B->vis_resid, vb);
    if (rc == SDP_OK) rc = sdp_device_malloc((void **)&caller, vb);
    for (int k = 0, first = 0; rc == SDP_OK && k < p->n_channels; first += n_vis[k], ++k)
        rc = sdp_memcpy_dtod(caller + 2 * (size_t)first, d_vis[k], sizeof(float) * 2 * (size_t)n_vis[k], s);
    if (rc == SDP_OK) rc = sdp_grid_plan_to_plan_order(B->plan, caller, B->vis_meas, s);
    if (rc == SDP_OK) rc = sdp_memset_zero(B->model, sizeof(float) * npix * nc, s);
    if (rc == SDP_OK) rc = sdp_event_create(&B->ev0);
    if (rc == SDP_OK) rc = sdp_event_create(&B->ev1);
    for (int i = 0; rc == SDP_OK && i <= SDP_STAGE_COUNT; ++i) rc = sdp_event_create(&B->stage_ev[i]);
    const int rs = sdp_stream_sync(s);   /* also on the error path: nothing may still use `caller` */
    if (rc == SDP_OK) rc = rs;
    (void)sdp_device_free(caller);
    if (rc != SDP_OK) {
        release(B);
        return rc;
    }
    *out = B;
    return SDP_OK;
}

int sdp_batch_destroy(sdp_batch *B)
{
    if (B) {
        (void)sdp_stream_sync(B->s);
        release(B);
    }
    return SDP_OK;
}

int sdp_batch_residuals(sdp_batch *B, float *ms, float *stage_ms)
{
    if (!B) return FAIL(SDP_E_ARG, "null batch");
    const int profile = (B->p.flags & SDP_BATCH_PROFILE) != 0;
    float acc[SDP_STAGE_COUNT] = {0};
    TRY(sdp_event_record(B->ev0, B->s));
    if (profile) {
        TRY(enqueue_pass(B, 1));
        TRY(sdp_event_sync(B->stage_ev[4]));
        for (int i = 0; i < 4; ++i) TRY(sdp_event_elapsed_ms(B->stage_ev[i], B->stage_ev[i + 1], &acc[i]));
    } else if ((B->p.flags & SDP_BATCH_GRAPH) && B->passes > 0) {
        if (!B->graph) {
            /* the first pass ran eagerly, so every lazily created resource
             * exists before capture */
            TRY(sdp_graph_begin_capture(B->s));
            int rc = enqueue_pass(B, 0);
            sdp_graph g = NULL;
            int rc2 = sdp_graph_end_capture(B->s, &g);
            if (rc != SDP_OK) {
                if (g) (void)sdp_graph_destroy(g);
                return rc;
            }
            TRY(rc2);
            B->graph = g;
            TRY(sdp_event_record(B->ev0, B->s));
        }
        TRY(sdp_graph_launch(B->graph, B->s));
    } else {
        TRY(enqueue_pass(B, 0));
    }
    TRY(sdp_event_record(B->ev1, B->s));
    TRY(sdp_event_sync(B->ev1));
    float t = 0.f;
    TRY(sdp_event_elapsed_ms(B->ev0, B->ev1, &t));
    if (ms) *ms = t;
    if (stage_ms)
        for (int i = 0; i < SDP_STAGE_COUNT; ++i) stage_ms[i] = acc[i];
    B->passes++;
    return SDP_OK;
}

float *sdp_batch_model(sdp_batch *B, int k)
{
    return (B && k >= 0 && k < B->p.n_channels) ? B->model + (size_t)k * pixels(B) : NULL;
}

const float *sdp_batch_residual(const sdp_batch *B, int k)
{
    return (B && k >= 0 && k < B->p.n_channels) ? B->residual + (size_t)k * pixels(B) : NULL;
}

const sdp_grid_plan *sdp_batch_plan(const sdp_batch *B, int k)
{
    return (B && k >= 0 && k < B->p.n_channels) ? B->plan : NULL;
}
