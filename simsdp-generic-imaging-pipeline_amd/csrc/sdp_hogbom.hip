// sdp_hogbom.hip -- Hogbom CLEAN minor cycle, hogbom_clean
// (src/CPU/deconvolution_run.c:197-312).
//
// One launch per minor iteration, every workgroup doing the same three
// things:
//   1. reduce the per-workgroup partials the previous launch wrote (largest
//      |residual|, first index on ties as the reference's strict '>'
//      row-major scan, :222-233; sum of |residual| over pixels 1..N-1, which
//      the reference's running mean, :226, equals over N-1), in a fixed
//      order, so every workgroup reaches the same component and the same stop
//      decision (:238-243); workgroup 0 records the component (:247-256);
//   2. subtract that component's PSF from its pixels (:266-273);
//   3. scan its pixels and write its partial for the next launch.
// State and partials are double-buffered by launch parity, so no workgroup
// waits on another inside a launch.  Once a stop test fires every later
// launch returns at once, so a fixed number of launches can be queued (or
// captured in a graph).
//
// Each workgroup owns chunks of consecutive pixels, loaded as 16-byte groups
// into registers, updated there by the subtraction and scanned from there:
// per iteration one read of the residual, one read of the PSF window and one
// write of the window (the window is nearly the whole image, :259-264).
//
// The subtraction is evaluated in double exactly as the reference writes it,
// ((gain * max_val) * psf) * psf_max, then subtracted from the float residual
// in double and rounded to float, so the residual, model and component list
// match the reference bit for bit whenever the stop decision does.
#include <algorithm>
#include <cstdlib>

#include "sdp_internal.h"

namespace {

constexpr int kThreads = 256;
constexpr int kMinPixelsPerChunk = kThreads * 4 * 4;   // smallest chunk (4 groups per thread)

struct Partial {
    float absmax;   // largest |r| in the workgroup that is > 0, else 0
    int idx;        // its first linear index, or INT_MAX
    float value;    // r at idx
    float v0;       // r at pixel 0 (workgroup 0 only)
    double sum;     // sum of |r| over the workgroup's pixels with index >= 1
};

// one slot per launch parity: launch `it` reads slot it & 1, writes the other
struct State {
    int stop, nsrc;
    float src0;
    int pad;
};

struct HogbomArgs {
    int G, n_minor, max_sources, nblocks, mirror, vec;
    double psf_max, gain, noise, weak;
};

__device__ __forceinline__ bool better(float av, int ai, float bv, int bi)
{
    return av > bv || (av == bv && ai < bi);
}

__global__ void hogbom_init_kernel(State *st, const float *model_in, float *model_out, size_t n)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < 2) st[i] = State{0, 0, 0.f, 0};
    for (size_t k = i; k < n; k += (size_t)gridDim.x * blockDim.x) model_out[k] = model_in[k];
}

// residual - ((gain * max_val) * psf) * psf_max, in double, rounded to float
// (deconvolution_run.c:266-273)
__device__ __forceinline__ float subtract_one(float r, double scale, float psf, double psf_max)
{
    return (float)((double)r - scale * (double)psf * psf_max);
}

// the 4 pixels p0 .. p0+3 one by one (window edges, or rows of G % 4 == 2)
__device__ __forceinline__ float4 subtract_lanes(float4 v, int p0, int G, int mx, int my, int half, int xlo, int xhi,
                                                 int ylo, int yhi, int mirror, const float *__restrict__ psf,
                                                 double scale, double pm)
{
    float e[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const int p = p0 + j;
        const int y = p / G, x = p - y * G;
        if (y >= ylo && y < yhi && x >= xlo && x < xhi) {
            const int prow = (y - my + half) * G;
            const int idx = mirror ? prow + (G - (x - mx + half)) : prow + (x - mx + half);
            e[j] = subtract_one(e[j], scale, psf[idx], pm);
        }
    }
    return make_float4(e[0], e[1], e[2], e[3]);
}

// Every workgroup of launch `it` reduces the partials of launch it - 1 the
// same way; the result lands in shared memory.
__device__ void reduce_partials(const Partial *__restrict__ parts, int nparts, float *red_v, int *red_i,
                                float *red_val, double *red_s)
{
    float bv = 0.f, bval = 0.f;
    int bi = 0x7fffffff;
    double bs = 0.0;
    for (int b = threadIdx.x; b < nparts; b += kThreads) {
        const Partial q = parts[b];
        if (q.absmax > 0.f && better(q.absmax, q.idx, bv, bi)) {
            bv = q.absmax;
            bi = q.idx;
            bval = q.value;
        }
        bs += q.sum;
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
}

template <int kVecPerThread>
__global__ void __launch_bounds__(kThreads)
hogbom_step_kernel(int iter, HogbomArgs a, float *__restrict__ r, const float *__restrict__ psf,
                   float *__restrict__ model, float *__restrict__ sources, int *__restrict__ num_sources,
                   State *__restrict__ st, Partial *__restrict__ parts)
{
    __shared__ float red_v[kThreads], red_val[kThreads];
    __shared__ int red_i[kThreads];
    __shared__ double red_s[kThreads];
    __shared__ int s_go, s_mx, s_my;
    __shared__ float s_maxval;

    const State in = st[iter & 1];
    State *out = &st[(iter + 1) & 1];
    const bool leader = blockIdx.x == 0 && threadIdx.x == 0;
    if (in.stop) {
        if (leader) *out = in;
        return;
    }
    const int G = a.G;
    const int npix = G * G;                  // G <= 32768
    Partial *prev = parts + (size_t)((iter + 1) & 1) * a.nblocks;   // written by launch it - 1
    Partial *next = parts + (size_t)(iter & 1) * a.nblocks;         // read by launch it + 1

    bool subtract = false;
    if (iter > 0) {
        // component `iter - 1`: the largest |r| the previous launch found
        reduce_partials(prev, a.nblocks, red_v, red_i, red_val, red_s);
        if (threadIdx.x == 0) {
            float curr_max = red_v[0];
            int max_idx = red_i[0];
            float peak = red_val[0];
            if (!(curr_max > 0.f)) {   // nothing beat the initial 0: the reference keeps pixel (0, 0)
                curr_max = 0.f;
                max_idx = 0;
                peak = prev[0].v0;
            }
            const float mean = (float)(red_s[0] / (double)(npix - 1));
            const int nsrc = in.nsrc;
            const bool extracting_noise = (double)curr_max < a.noise * (double)mean * a.gain;
            const bool weak_source = nsrc > 0 && (double)curr_max < (double)fabsf(in.src0) * a.weak;
            if (extracting_noise || weak_source || nsrc >= a.max_sources) {
                s_go = 0;
                if (blockIdx.x == 0) *out = State{1, nsrc, in.src0, 0};
            } else {
                const float max_val = (float)((double)peak / a.psf_max);
                s_go = 1;
                s_mx = max_idx % G;
                s_my = max_idx / G;
                s_maxval = max_val;
                if (blockIdx.x == 0) {
                    const float flux = (float)((double)max_val * a.gain);
                    sources[3 * nsrc] = (float)s_mx;
                    sources[3 * nsrc + 1] = (float)s_my;
                    sources[3 * nsrc + 2] = flux;
                    model[max_idx] = (float)((double)model[max_idx] + (double)max_val * a.gain);
                    *num_sources = nsrc + 1;
                    *out = State{0, nsrc + 1, nsrc == 0 ? flux : in.src0, 0};
                }
            }
        }
        __syncthreads();
        if (!s_go) return;
        subtract = true;
    } else if (leader) {
        *out = in;
    }

    const bool scan = iter < a.n_minor;      // the final launch only subtracts
    const int n4 = npix / 4;
    const int gpr = G / 4;                   // 16-byte groups per row when G % 4 == 0
    float4 *r4 = reinterpret_cast<float4 *>(r);

    // subtraction window (deconvolution_run.c:259-264): rows [ylo, yhi),
    // columns [xlo, xhi), PSF centre at (G/2, G/2)
    const int mx = s_mx, my = s_my, half = G / 2;
    const int xlo = mx - min(mx, half), xhi = mx + min(G - mx, half);
    const int ylo = my - min(my, half), yhi = my + min(G - my, half);
    const double scale = a.gain * (double)s_maxval;
    const double pm = a.psf_max;

    float bv = 0.f, bval = 0.f, v0 = 0.f;
    int bi = 0x7fffffff;
    double bs = 0.0;

    // chunks of kThreads * kVecPerThread groups, grid-stride over the image
    for (int c = blockIdx.x; c * (kThreads * kVecPerThread) < n4; c += gridDim.x) {
        const int q0 = c * (kThreads * kVecPerThread);
        float4 v[kVecPerThread];
#pragma unroll
        for (int k = 0; k < kVecPerThread; ++k) {
            const int q = q0 + k * kThreads + threadIdx.x;
            v[k] = q < n4 ? r4[q] : make_float4(0.f, 0.f, 0.f, 0.f);
        }

        if (subtract) {
            if (a.vec) {
                // G % 4 == 0: a group never crosses a row.  Groups wholly
                // inside the window read their 4 PSF values with aligned
                // 16-byte loads, all issued before any is used; groups on
                // the window's left or right edge go lane by lane.
                float4 pa[kVecPerThread], pb[kVecPerThread];
                unsigned full = 0, part = 0;
#pragma unroll
                for (int k = 0; k < kVecPerThread; ++k) {
                    const int q = q0 + k * kThreads + threadIdx.x;
                    pa[k] = pb[k] = make_float4(0.f, 0.f, 0.f, 0.f);
                    if (q >= n4) continue;
                    const int y = q / gpr, x0 = (q - y * gpr) * 4;
                    if (y < ylo || y >= yhi || x0 + 3 < xlo || x0 >= xhi) continue;
                    const int prow = (y - my + half) * G;
                    const int lo = a.mirror ? prow + (G - (x0 + 3 - mx + half)) : prow + (x0 - mx + half);
                    const int d = lo & 3;
                    if (x0 >= xlo && x0 + 3 < xhi && (lo - d) + (d ? 8 : 4) <= npix + 1) {
                        const float4 *g = reinterpret_cast<const float4 *>(psf + (lo - d));
                        pa[k] = g[0];
                        if (d) pb[k] = g[1];
                        full |= 1u << k;
                    } else {
                        part |= 1u << k;
                    }
                }
#pragma unroll
                for (int k = 0; k < kVecPerThread; ++k) {
                    const int q = q0 + k * kThreads + threadIdx.x;
                    if (full & (1u << k)) {
                        const int y = q / gpr, x0 = (q - y * gpr) * 4;
                        const int prow = (y - my + half) * G;
                        const int lo = a.mirror ? prow + (G - (x0 + 3 - mx + half)) : prow + (x0 - mx + half);
                        const int d = lo & 3;
                        const float4 A = pa[k], B = pb[k];
                        const float4 w = d == 0 ? A
                                       : d == 1 ? make_float4(A.y, A.z, A.w, B.x)
                                       : d == 2 ? make_float4(A.z, A.w, B.x, B.y)
                                                : make_float4(A.w, B.x, B.y, B.z);    // psf[lo .. lo+3]
                        float4 e = v[k];
                        if (a.mirror) {   // index decreases with x: lane j reads psf[lo + 3 - j]
                            e.x = subtract_one(e.x, scale, w.w, pm);
                            e.y = subtract_one(e.y, scale, w.z, pm);
                            e.z = subtract_one(e.z, scale, w.y, pm);
                            e.w = subtract_one(e.w, scale, w.x, pm);
                        } else {
                            e.x = subtract_one(e.x, scale, w.x, pm);
                            e.y = subtract_one(e.y, scale, w.y, pm);
                            e.z = subtract_one(e.z, scale, w.z, pm);
                            e.w = subtract_one(e.w, scale, w.w, pm);
                        }
                        v[k] = e;
                        r4[q] = e;
                    } else if (part & (1u << k)) {
                        v[k] = subtract_lanes(v[k], q * 4, G, mx, my, half, xlo, xhi, ylo, yhi, a.mirror, psf,
                                              scale, pm);
                        r4[q] = v[k];
                    }
                }
            } else {
                // G % 4 == 2: groups may cross a row end; every lane on its own
#pragma unroll
                for (int k = 0; k < kVecPerThread; ++k) {
                    const int q = q0 + k * kThreads + threadIdx.x;
                    if (q >= n4) continue;
                    const float4 e = subtract_lanes(v[k], q * 4, G, mx, my, half, xlo, xhi, ylo, yhi, a.mirror,
                                                    psf, scale, pm);
                    r4[q] = e;
                    v[k] = e;
                }
            }
        }
        if (!scan) continue;

        // scan the updated values: largest |r| (first index on ties) and the
        // sum of |r| over pixels 1..N-1
#pragma unroll
        for (int k = 0; k < kVecPerThread; ++k) {
            const int q = q0 + k * kThreads + threadIdx.x;
            if (q >= n4) continue;
            const int p0 = q * 4;
            const float e[4] = {v[k].x, v[k].y, v[k].z, v[k].w};
            if (p0 == 0) v0 = e[0];
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const float av = fabsf(e[j]);
                if (p0 + j >= 1) bs += (double)av;
                if (av > 0.f && better(av, p0 + j, bv, bi)) {
                    bv = av;
                    bi = p0 + j;
                    bval = e[j];
                }
            }
        }
    }
    if (!scan) return;

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
        q.v0 = v0;   // pixel 0 belongs to thread 0 of workgroup 0
        q.sum = red_s[0];
        next[blockIdx.x] = q;
    }
}

// deconvolution_run.c:297-306: pixel coordinates to radians
__global__ void hogbom_finish_kernel(const State *st, int G, double cell_size, float *sources, int *num_sources)
{
    const int n = st->nsrc;   // the slot the last launch wrote
    if (threadIdx.x == 0 && blockIdx.x == 0) *num_sources = n;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        sources[3 * i] = (float)((double)(sources[3 * i] - (float)(G / 2)) * cell_size);
        sources[3 * i + 1] = (float)((double)(sources[3 * i + 1] - (float)(G / 2)) * cell_size);
    }
}

size_t partial_offset() { return 256; }

// groups per thread (4 or 8); SDP_HOGBOM_GROUPS overrides the default for tuning
int hogbom_groups()
{
    static int g = 0;
    if (g == 0) {
        const char *e = getenv("SDP_HOGBOM_GROUPS");
        g = (e && atoi(e) == 4) ? 4 : 8;
    }
    return g;
}

int compute_units()
{
    static int n = 0;
    if (n == 0) {
        int dev = 0, cu = 0;
        if (hipGetDevice(&dev) == hipSuccess &&
            hipDeviceGetAttribute(&cu, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && cu > 0)
            n = cu;
        else
            n = 256;
    }
    return n;
}

}  // namespace

extern "C" {

size_t sdp_hogbom_work_bytes(int G)
{
    if (G < 2) return 0;
    const size_t npix = (size_t)G * G;
    const size_t nb = (npix + kMinPixelsPerChunk - 1) / kMinPixelsPerChunk;
    return partial_offset() + 2 * nb * sizeof(Partial);   // partials of two consecutive launches
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
    SDP_REQUIRE(((uintptr_t)d_work & 15) == 0 && ((uintptr_t)d_residual & 15) == 0,
                "residual and work buffers must be 16-byte aligned");
    hipStream_t st = sdp::stream_of(s);
    const size_t npix = (size_t)G * G;
    HogbomArgs a;
    a.G = G;
    a.n_minor = n_minor;
    a.max_sources = max_sources;
    a.mirror = (flags & SDP_HOGBOM_PSF_AS_STORED) ? 0 : 1;
    a.vec = (G % 4) == 0;
    // 16-byte groups per thread per chunk: 8 keeps more loads in flight, 4
    // more workgroups resident; the workgroup count is what the CUs hold at
    // once (no more than there are chunks), so one resident wave covers the
    // image and no second, partial wave trails it
    const int groups = hogbom_groups();
    auto kern = groups == 4 ? hogbom_step_kernel<4> : hogbom_step_kernel<8>;
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, kThreads, 0) != hipSuccess || per_cu < 1)
        per_cu = 1;
    const size_t chunks = (npix + (size_t)kThreads * groups * 4 - 1) / ((size_t)kThreads * groups * 4);
    a.nblocks = (int)std::min<size_t>(chunks, (size_t)per_cu * compute_units());
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
        kern<<<a.nblocks, kThreads, 0, st>>>(it, a, d_residual, d_psf, d_model_out, d_sources, d_num_sources,
                                             state, parts);
        SDP_LAUNCH_CHECK();
    }
    hogbom_finish_kernel<<<4, 256, 0, st>>>(state + ((n_minor + 1) & 1), G, cell_size, d_sources, d_num_sources);
    SDP_LAUNCH_CHECK();
    return SDP_OK;
}

}  // extern "C"
