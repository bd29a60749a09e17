// sdp_hogbom.hip -- Hogbom CLEAN minor cycle, hogbom_clean
// (src/CPU/deconvolution_run.c:197-312).
//
// One launch per minor iteration.  Launch i first applies the PSF
// subtraction chosen by launch i-1 to its pixel range, then scans the same
// range for the largest |residual| (first index on ties, as the reference's
// strict '>' row-major scan, :222-233) and for the sum of |residual| over
// pixels 1..N-1 (the reference's running mean, :226, equals that sum over
// N-1).  The last workgroup to finish (ticket counter) reduces the
// per-workgroup partials in a fixed order, applies the stop tests (:238-243),
// records the component (:247-256) and leaves the next subtraction pending.
// Once a stop test fires every later launch returns at once, so a fixed
// number of launches can be queued (or captured in a graph).
//
// The subtraction (:266-273) is evaluated in double exactly as the
// reference writes it, ((gain * max_val) * psf) * psf_max, then subtracted
// from the float residual in double and rounded to float, so the residual,
// model and component list match the reference bit for bit whenever the
// stop decision does.
#include "sdp_internal.h"

namespace {

constexpr int kThreads = 256;
constexpr int kPixelsPerBlock = 16384;

struct Partial {
    float absmax;   // largest |r| in the block that is > 0, else 0
    int idx;        // its first linear index, or INT_MAX
    float value;    // r at idx
    int pad;
    double sum;     // sum of |r| over the block's pixels with index >= 1
};

struct State {
    int stop, pending, mx, my;
    float max_val;
    int ticket, nsrc, pad;
    float src0;
    int pad2[3];
};

struct HogbomArgs {
    int G, n_minor, max_sources, nblocks, rows_per_block, mirror;
    double psf_max, gain, noise, weak;
};

__device__ __forceinline__ bool better(float av, int ai, float bv, int bi)
{
    return av > bv || (av == bv && ai < bi);
}

__global__ void hogbom_init_kernel(State *st, const float *model_in, float *model_out, size_t n)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) {
        st->stop = 0;
        st->pending = 0;
        st->ticket = 0;
        st->nsrc = 0;
        st->src0 = 0.f;
    }
    for (size_t k = i; k < n; k += (size_t)gridDim.x * blockDim.x) model_out[k] = model_in[k];
}

__global__ void __launch_bounds__(kThreads)
hogbom_step_kernel(int iter, HogbomArgs a, float *__restrict__ r, const float *__restrict__ psf,
                   float *__restrict__ model, float *__restrict__ sources, int *__restrict__ num_sources,
                   State *__restrict__ st, Partial *__restrict__ parts)
{
    __shared__ int s_pending, s_mx, s_my;
    __shared__ float s_maxval;
    __shared__ float red_v[kThreads], red_val[kThreads];
    __shared__ int red_i[kThreads];
    __shared__ double red_s[kThreads];
    __shared__ bool s_last;

    if (threadIdx.x == 0) {
        s_pending = st->stop ? -1 : st->pending;
        s_mx = st->mx;
        s_my = st->my;
        s_maxval = st->max_val;
    }
    __syncthreads();
    if (s_pending < 0) return;

    const int G = a.G;
    const size_t npix = (size_t)G * G;
    const int y0 = blockIdx.x * a.rows_per_block;
    const int y1 = min(G, y0 + a.rows_per_block);

    if (s_pending) {
        const int mx = s_mx, my = s_my, half = G / 2;
        const int sx = min(mx, half), sy = min(my, half);
        const int ex = min(G - mx, half), ey = min(G - my, half);
        const int ylo = max(y0, my - sy), yhi = min(y1, my + ey);
        const double scale = a.gain * (double)s_maxval;
        for (int y = ylo; y < yhi; ++y) {
            float *row = r + (size_t)y * G;
            if (a.mirror) {
                // reference index (y - my + half) * G + (G - (x - mx + half)) = base - x
                const long long base = (long long)(y - my + half) * G + (G - half + mx);
                for (int x = mx - sx + (int)threadIdx.x; x < mx + ex; x += kThreads)
                    row[x] = (float)((double)row[x] - scale * (double)psf[base - x] * a.psf_max);
            } else {
                // PSF in its own orientation: (y - my + half) * G + (x - mx + half) = base + x
                const long long base = (long long)(y - my + half) * G + (half - mx);
                for (int x = mx - sx + (int)threadIdx.x; x < mx + ex; x += kThreads)
                    row[x] = (float)((double)row[x] - scale * (double)psf[base + x] * a.psf_max);
            }
        }
    }
    if (iter >= a.n_minor) return;   // final launch: subtraction only

    float bv = 0.f, bval = 0.f;
    int bi = 0x7fffffff;
    double bs = 0.0;
    for (int y = y0; y < y1; ++y) {
        const float *row = r + (size_t)y * G;
        for (int x = threadIdx.x; x < G; x += kThreads) {
            const float v = row[x];
            const float av = fabsf(v);
            const int p = y * G + x;
            if (p >= 1) bs += (double)av;
            if (av > 0.f && better(av, p, bv, bi)) {
                bv = av;
                bi = p;
                bval = v;
            }
        }
    }
    red_v[threadIdx.x] = bv;
    red_i[threadIdx.x] = bi;
    red_val[threadIdx.x] = bval;
    red_s[threadIdx.x] = bs;
    __syncthreads();
    for (int off = kThreads / 2; off > 0; off >>= 1) {
        if (threadIdx.x < off) {
            const int o = threadIdx.x + off;
            if (better(red_v[o], red_i[o], red_v[threadIdx.x], red_i[threadIdx.x])) {
                red_v[threadIdx.x] = red_v[o];
                red_i[threadIdx.x] = red_i[o];
                red_val[threadIdx.x] = red_val[o];
            }
            red_s[threadIdx.x] += red_s[o];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        Partial q;
        q.absmax = red_v[0];
        q.idx = red_i[0];
        q.value = red_val[0];
        q.pad = 0;
        q.sum = red_s[0];
        parts[blockIdx.x] = q;
        __threadfence();
        const int t = atomicAdd(&st->ticket, 1);
        s_last = (t == a.nblocks - 1);
    }
    __syncthreads();
    if (!s_last) return;
    __threadfence();   // acquire: the other workgroups' partials are visible

    // last workgroup: reduce the partials in block order
    bv = 0.f; bi = 0x7fffffff; bval = 0.f; bs = 0.0;
    for (int b = threadIdx.x; b < a.nblocks; b += kThreads) {
        const Partial q = parts[b];
        if (q.absmax > 0.f && better(q.absmax, q.idx, bv, bi)) {
            bv = q.absmax;
            bi = q.idx;
            bval = q.value;
        }
    }
    red_v[threadIdx.x] = bv;
    red_i[threadIdx.x] = bi;
    red_val[threadIdx.x] = bval;
    __syncthreads();
    for (int off = kThreads / 2; off > 0; off >>= 1) {
        if (threadIdx.x < off) {
            const int o = threadIdx.x + off;
            if (better(red_v[o], red_i[o], red_v[threadIdx.x], red_i[threadIdx.x])) {
                red_v[threadIdx.x] = red_v[o];
                red_i[threadIdx.x] = red_i[o];
                red_val[threadIdx.x] = red_val[o];
            }
        }
        __syncthreads();
    }
    if (threadIdx.x != 0) return;
    double sum = 0.0;
    for (int b = 0; b < a.nblocks; ++b) sum += parts[b].sum;

    float curr_max = red_v[0];
    int max_idx = red_i[0];
    float peak = red_val[0];
    if (!(curr_max > 0.f)) {   // nothing beat the initial 0: the reference keeps pixel (0, 0)
        curr_max = 0.f;
        max_idx = 0;
        peak = __hip_atomic_load(&r[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    const float mean = (float)(sum / (double)(npix - 1));
    const int nsrc = st->nsrc;
    const bool extracting_noise = (double)curr_max < a.noise * (double)mean * a.gain;
    const bool weak_source = nsrc > 0 && (double)curr_max < (double)fabsf(st->src0) * a.weak;
    st->ticket = 0;
    if (extracting_noise || weak_source || nsrc >= a.max_sources) {
        st->stop = 1;
        st->pending = 0;
        return;
    }
    const float max_val = (float)((double)peak / a.psf_max);
    const int mx = max_idx % G, my = max_idx / G;
    const float flux = (float)((double)max_val * a.gain);
    sources[3 * nsrc] = (float)mx;
    sources[3 * nsrc + 1] = (float)my;
    sources[3 * nsrc + 2] = flux;
    if (nsrc == 0) st->src0 = flux;
    model[max_idx] = (float)((double)model[max_idx] + (double)max_val * a.gain);
    st->nsrc = nsrc + 1;
    *num_sources = nsrc + 1;
    st->mx = mx;
    st->my = my;
    st->max_val = max_val;
    st->pending = 1;
}

// deconvolution_run.c:297-306: pixel coordinates to radians
__global__ void hogbom_finish_kernel(const State *st, int G, double cell_size, float *sources, int *num_sources)
{
    const int n = st->nsrc;
    if (threadIdx.x == 0 && blockIdx.x == 0) *num_sources = n;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        sources[3 * i] = (float)((double)(sources[3 * i] - (float)(G / 2)) * cell_size);
        sources[3 * i + 1] = (float)((double)(sources[3 * i + 1] - (float)(G / 2)) * cell_size);
    }
}

size_t partial_offset() { return 256; }

}  // namespace

extern "C" {

size_t sdp_hogbom_work_bytes(int G)
{
    if (G < 2) return 0;
    const size_t npix = (size_t)G * G;
    const int rpb = G >= kPixelsPerBlock ? 1 : kPixelsPerBlock / G;
    const size_t nb = ((size_t)G + rpb - 1) / rpb;
    (void)npix;
    return partial_offset() + nb * sizeof(Partial);
}

int sdp_hogbom(int G, int n_minor, int max_sources, float *d_residual, const float *d_psf, double psf_max,
               double loop_gain, double noise_factor, double weak_source_percent, double cell_size,
               const float *d_model_in, float *d_model_out, float *d_sources, int *d_num_sources, void *d_work,
               sdp_stream s)
{
    return sdp_hogbom_ex(G, n_minor, max_sources, d_residual, d_psf, psf_max, loop_gain, noise_factor,
                         weak_source_percent, cell_size, d_model_in, d_model_out, d_sources, d_num_sources, d_work,
                         0, s);
}

int sdp_hogbom_ex(int G, int n_minor, int max_sources, float *d_residual, const float *d_psf, double psf_max,
                  double loop_gain, double noise_factor, double weak_source_percent, double cell_size,
                  const float *d_model_in, float *d_model_out, float *d_sources, int *d_num_sources, void *d_work,
                  int flags, sdp_stream s)
{
    SDP_REQUIRE(G >= 2 && G % 2 == 0 && G <= 32768, "bad grid size %d", G);
    SDP_REQUIRE(n_minor >= 0 && max_sources >= 0, "bad iteration / source counts");
    SDP_REQUIRE(d_residual && d_psf && d_model_in && d_model_out && d_sources && d_num_sources && d_work,
                "sdp_hogbom: null pointer");
    SDP_REQUIRE(((uintptr_t)d_work & 15) == 0, "work buffer must be 16-byte aligned");
    hipStream_t st = sdp::stream_of(s);
    const size_t npix = (size_t)G * G;
    HogbomArgs a;
    a.G = G;
    a.n_minor = n_minor;
    a.max_sources = max_sources;
    a.rows_per_block = G >= kPixelsPerBlock ? 1 : kPixelsPerBlock / G;
    a.mirror = (flags & SDP_HOGBOM_PSF_AS_STORED) ? 0 : 1;
    a.nblocks = (G + a.rows_per_block - 1) / a.rows_per_block;
    a.psf_max = psf_max;
    a.gain = loop_gain;
    a.noise = noise_factor;
    a.weak = weak_source_percent;
    State *state = (State *)d_work;
    Partial *parts = (Partial *)((char *)d_work + partial_offset());

    hogbom_init_kernel<<<sdp_div_up(npix, 256) > 4096 ? 4096 : sdp_div_up(npix, 256), 256, 0, st>>>(
        state, d_model_in, d_model_out, d_model_in == d_model_out ? 0 : npix);
    SDP_LAUNCH_CHECK();
    for (int it = 0; it <= n_minor; ++it) {
        hogbom_step_kernel<<<a.nblocks, kThreads, 0, st>>>(it, a, d_residual, d_psf, d_model_out, d_sources,
                                                           d_num_sources, state, parts);
        SDP_LAUNCH_CHECK();
    }
    hogbom_finish_kernel<<<4, 256, 0, st>>>(state, G, cell_size, d_sources, d_num_sources);
    SDP_LAUNCH_CHECK();
    return SDP_OK;
}

}  // extern "C"
