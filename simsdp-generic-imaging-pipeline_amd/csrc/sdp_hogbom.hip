// sdp_hogbom.hip -- Hogbom CLEAN minor cycle, hogbom_clean
// (src/CPU/deconvolution_run.c:197-312).
//
// One launch per minor iteration, every workgroup doing the same three
// things:
//   1. reduce the per-workgroup partials the previous launch wrote (largest
//      |residual|, first index on ties as the reference's strict '>'
//      row-major scan, :222-233; sum of |residual| over pixels 1..N-1) in a
//      fixed order, so every workgroup reaches the same component and the
//      same stop decision (:238-243); workgroup 0 records the component
//      (:247-256);
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
// The PSF is read from a copy in the work buffer with 4 zero floats on each
// side, through a buffer resource: a group's 4 PSF values are one 16-byte
// load at any dword offset, issued for every group of a chunk before any is
// used (no per-group branches); lanes outside the window are masked.
//
// The subtraction is evaluated in double exactly as the reference writes it,
// ((gain * max_val) * psf) * psf_max, then subtracted from the float residual
// in double and rounded to float, so the residual, model and component list
// match the reference bit for bit whenever the stop decision does.
//
// The noise stop compares the peak with noise_factor * gain * "average",
// where the reference's average (:226) is the float recurrence m = |r0|,
// m += (|r_i| - m) / i.  In exact arithmetic that is the mean of
// |r_1 .. r_{N-1}|, but in float it drifts as N grows (about 2e-6 of the
// mean at 128^2, 1e-4 at 2048^2 and 0.7% at 4096^2 on noise-like images),
// so no parallel sum reproduces it.  The stop can only fire when
// noise_factor * gain exceeds 1/2 (the recurrence never exceeds the largest
// |r| by more than a few ulps); then, by default, one wavefront runs the
// reference's recurrence itself before each iteration
// (hogbom_running_mean_kernel: sequential, about 0.1 s per iteration at
// 2048^2), and the decision is the reference's bit for bit.
// SDP_HOGBOM_DOUBLE_MEAN takes the parallel double sum over N - 1 instead,
// which agrees with the reference except when the peak lies within that
// drift of the threshold.
#include <algorithm>
#include <cstdlib>

#include "sdp_internal.h"

namespace {

constexpr int kThreads = 256;
constexpr int kMinPixelsPerChunk = kThreads * 4 * 2;   // smallest chunk (2 groups per thread)

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
    float mean;   // the reference's running "average" of the current residual (exact-mean mode)
};

struct HogbomArgs {
    int G, n_minor, max_sources, nblocks, mirror, vec, exact_mean;
    int psf_pad_bytes;
    double psf_max, gain, noise, weak;
};

typedef float v4f_t __attribute__((ext_vector_type(4)));

__device__ __forceinline__ bool better(float av, int ai, float bv, int bi)
{
    return av > bv || (av == bv && ai < bi);
}

constexpr int kPsfPad = 4;   // zero floats before and after the PSF copy

// state slots, the model copy (n_model floats, 0 when in place) and the
// padded PSF copy: kPsfPad zeros, the npix + 1 PSF floats, kPsfPad - 1 zeros
__global__ void hogbom_init_kernel(State *st, const float *model_in, float *model_out, size_t n_model,
                                   const float *psf, float *psf_pad, size_t npix)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    if (i < 2) st[i] = State{0, 0, 0.f, 0.f};
    for (size_t k = i; k < n_model; k += stride) model_out[k] = model_in[k];
    const size_t total = npix + 2 * kPsfPad;
    for (size_t k = i; k < total; k += stride)
        psf_pad[k] = (k >= (size_t)kPsfPad && k <= npix + kPsfPad) ? psf[k - kPsfPad] : 0.f;
}

// deconvolution_run.c:223-227: avg = |r0|, then avg += (|r_i| - avg) / i in
// float, i converted to float, for i = 1 .. npix-1 in row-major order.  One
// wavefront: each step loads 64 consecutive pixels, one per lane, and every
// lane runs the same sequential recurrence over them, taking pixel j from
// lane j by readlane, while the next 64 pixels are already loading.
// Skipped once a stop has fired.
__global__ void __launch_bounds__(64) hogbom_running_mean_kernel(const float *__restrict__ r, int npix, State *slot)
{
    if (slot->stop) return;
    const int lane = threadIdx.x;
    float avg = 0.f;
    float next = lane < npix ? r[lane] : 0.f;
    for (int base = 0; base < npix; base += 64) {
        const float cur = next;
        const int q = base + 64 + lane;
        next = q < npix ? r[q] : 0.f;
        const int n = min(64, npix - base);
        const int bits = __float_as_int(cur);
        if (n == 64) {
#pragma unroll
            for (int j = 0; j < 64; ++j) {
                const float val = fabsf(__int_as_float(__builtin_amdgcn_readlane(bits, j)));
                const int idx = base + j;
                avg = idx == 0 ? val : avg + (val - avg) / (float)idx;
            }
        } else {
            for (int j = 0; j < n; ++j) {
                const float val = fabsf(__int_as_float(__builtin_amdgcn_readlane(bits, j)));
                const int idx = base + j;
                avg = idx == 0 ? val : avg + (val - avg) / (float)idx;
            }
        }
    }
    if (lane == 0) slot->mean = avg;
}

constexpr unsigned kOob = 0x80000000u;   // a buffer offset past any PSF copy: reads as zero

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
                   const float *__restrict__ psf_pad, float *__restrict__ model, float *__restrict__ sources,
                   int *__restrict__ num_sources, State *__restrict__ st, Partial *__restrict__ parts)
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
    const int npix = G * G;                  // G <= 23168
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
            const float mean = a.exact_mean ? in.mean : (float)(red_s[0] / (double)(npix - 1));
            const int nsrc = in.nsrc;
            const bool extracting_noise = (double)curr_max < a.noise * (double)mean * a.gain;
            const bool weak_source = nsrc > 0 && (double)curr_max < (double)fabsf(in.src0) * a.weak;
            if (extracting_noise || weak_source || nsrc >= a.max_sources) {
                s_go = 0;
                if (blockIdx.x == 0) *out = State{1, nsrc, in.src0, 0.f};
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
                    *out = State{0, nsrc + 1, nsrc == 0 ? flux : in.src0, 0.f};
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
                // G % 4 == 0: a group never crosses a row.  Its 4 PSF values
                // psf[lo .. lo+3] are one 16-byte load from the padded copy
                // (lo >= -4 and lo + 3 <= npix + 3 whenever a lane of the
                // group is in the window); groups with no lane in the
                // window load from past the copy and read zeros.
                const __amdgpu_buffer_rsrc_t rs =
                    __builtin_amdgcn_make_buffer_rsrc((void *)psf_pad, (short)0, a.psf_pad_bytes, 0x00020000);
                v4f_t w[kVecPerThread];
                unsigned lanes[kVecPerThread];
#pragma unroll
                for (int k = 0; k < kVecPerThread; ++k) {
                    const int q = q0 + k * kThreads + threadIdx.x;
                    const int y = q / gpr, x0 = (q - y * gpr) * 4;
                    const bool row = q < n4 && y >= ylo && y < yhi;
                    unsigned m = 0;
#pragma unroll
                    for (int j = 0; j < 4; ++j) m |= (row && x0 + j >= xlo && x0 + j < xhi) ? 1u << j : 0u;
                    lanes[k] = m;
                    const int prow = (y - my + half) * G;
                    const int lo = a.mirror ? prow + (G - (x0 + 3 - mx + half)) : prow + (x0 - mx + half);
                    const unsigned off = m ? (unsigned)(lo + kPsfPad) * 4u : kOob;
                    w[k] = __builtin_bit_cast(v4f_t, __builtin_amdgcn_raw_buffer_load_b128(rs, off, 0, 0));
                }
#pragma unroll
                for (int k = 0; k < kVecPerThread; ++k) {
                    const unsigned m = lanes[k];
                    if (!m) continue;
                    const int q = q0 + k * kThreads + threadIdx.x;
                    // lane j reads psf[lo + j], or psf[lo + 3 - j] mirrored
                    const float p0 = a.mirror ? w[k].w : w[k].x, p1 = a.mirror ? w[k].z : w[k].y;
                    const float p2 = a.mirror ? w[k].y : w[k].z, p3 = a.mirror ? w[k].x : w[k].w;
                    float4 e = v[k];
                    if (m & 1u) e.x = subtract_one(e.x, scale, p0, pm);
                    if (m & 2u) e.y = subtract_one(e.y, scale, p1, pm);
                    if (m & 4u) e.z = subtract_one(e.z, scale, p2, pm);
                    if (m & 8u) e.w = subtract_one(e.w, scale, p3, pm);
                    v[k] = e;
                    r4[q] = e;
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

size_t partial_count(size_t npix) { return (npix + kMinPixelsPerChunk - 1) / kMinPixelsPerChunk; }

// the padded PSF copy follows the two launches' partials, 256-byte aligned
size_t psf_pad_offset(size_t npix)
{
    return (partial_offset() + 2 * partial_count(npix) * sizeof(Partial) + 255) & ~(size_t)255;
}

// 16-byte groups per thread per chunk (2, 4 or 8): the most that still
// gives every CU two chunks, so small images spread over the whole GPU;
// SDP_HOGBOM_GROUPS overrides the choice for tuning
int hogbom_groups(size_t npix, int cus)
{
    static int env = -1;
    if (env < 0) {
        const char *e = getenv("SDP_HOGBOM_GROUPS");
        const int v = e ? atoi(e) : 0;
        env = (v == 2 || v == 4 || v == 8) ? v : 0;
    }
    if (env > 0) return env;
    for (int g = 8; g > 2; g >>= 1)
        if (npix / ((size_t)kThreads * 4 * g) >= 2 * (size_t)cus) return g;
    return 2;
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
    // partials of two consecutive launches, then the padded PSF copy
    return psf_pad_offset(npix) + (npix + 2 * kPsfPad) * sizeof(float);
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
    SDP_REQUIRE(G >= 2 && G % 2 == 0 && G <= 23168, "bad grid size %d", G);   // padded PSF bytes fit an int
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
    // the noise stop can fire only when noise_factor * gain > 1/2
    a.exact_mean = !(flags & SDP_HOGBOM_DOUBLE_MEAN) && noise_factor * loop_gain > 0.5;
    // 16-byte groups per thread per chunk: 8 keeps more loads in flight, 4
    // more workgroups resident; the workgroup count is what the CUs hold at
    // once (no more than there are chunks), so one resident wave covers the
    // image and no second, partial wave trails it
    const int groups = hogbom_groups(npix, compute_units());
    auto kern = groups == 2 ? hogbom_step_kernel<2> : groups == 4 ? hogbom_step_kernel<4> : hogbom_step_kernel<8>;
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
    float *psf_pad = (float *)((char *)d_work + psf_pad_offset(npix));
    a.psf_pad_bytes = (int)((npix + 2 * kPsfPad) * sizeof(float));

    hogbom_init_kernel<<<sdp_div_up(npix, 256) > 4096 ? 4096 : sdp_div_up(npix, 256), 256, 0, st>>>(
        state, d_model_in, d_model_out, d_model_in == d_model_out ? 0 : npix, d_psf, psf_pad, npix);
    SDP_LAUNCH_CHECK();
    for (int it = 0; it <= n_minor; ++it) {
        if (a.exact_mean && it > 0) {
            // the residual the previous launch left, for launch it's decision
            hogbom_running_mean_kernel<<<1, 64, 0, st>>>(d_residual, (int)npix, state + (it & 1));
            SDP_LAUNCH_CHECK();
        }
        kern<<<a.nblocks, kThreads, 0, st>>>(it, a, d_residual, d_psf, psf_pad, d_model_out, d_sources,
                                             d_num_sources, state, parts);
        SDP_LAUNCH_CHECK();
    }
    hogbom_finish_kernel<<<4, 256, 0, st>>>(state + ((n_minor + 1) & 1), G, cell_size, d_sources, d_num_sources);
    SDP_LAUNCH_CHECK();
    return SDP_OK;
}

}  // extern "C"
