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
#include <cstdint>
#include <cstdlib>
#include <type_traits>

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

// ---- one launch for all iterations: the residual held in registers ------
//
// A cooperative launch (every workgroup resident at once, one per CU at
// most) runs all minor iterations.  Workgroup b owns the 16-byte groups
// q = b * TPB NG + k * TPB + t (thread t, k < NG) and keeps them in
// registers from the first iteration to the last, so an iteration reads
// only the PSF window: the residual is read once and written once.
//
// Per iteration each workgroup finds its best pixel (largest |r|, first
// index on ties: the largest |r| in one pass, then the first index holding
// it) and publishes it as one 64-bit word (|r| bits, index, sign, iteration
// tag); thread t of every workgroup reads workgroup t's word until it
// carries this iteration's tag, and each workgroup reduces the words the
// same way, so all reach the same component and stop decision; workgroup 0
// records the component and every workgroup subtracts it from its own
// pixels.  Words are double-buffered by iteration parity: a workgroup
// publishes iteration i + 2 into a buffer only after it has read every
// workgroup's word of iteration i + 1, each published after its workgroup
// had read all of iteration i.
//
// Waits are bounded: a word that does not arrive within 20 ms (it cannot,
// short of a fault elsewhere) sets an abort flag, which thread 0 of every
// workgroup reads at each exchange; every workgroup leaves and the run
// reports -1 sources.
//
// The arithmetic, tie order and stop tests are those of hogbom_step_kernel
// (bit for bit); the noise stop is left out, so the form is taken only when
// it cannot fire (noise_factor * gain <= 1/2).
constexpr unsigned long long kWaitTicks = 2000000ull;   // 20 ms of the 100 MHz counter
// Every word is stored in kWordReplicas copies, kWordStride words apart
// (8.5 KB: different memory channels), and workgroup b polls copy
// b % kWordReplicas: each copy's lines are polled by an eighth of the
// workgroups instead of all of them.
constexpr int kWordReplicas = 8;
constexpr int kWordStride = 1088;   // >= the largest workgroup count (one per CU)

struct PersistArgs {
    int G, n_minor, max_sources, nblocks, mirror, psf_pad_bytes, row_shift;   // row_shift: log2(G / 4)
    double psf_max, gain, weak;
};

// lane moves for wavefront reductions: DPP within 16-lane rows, then the
// gfx950 row swaps across them (no lane-address registers)
template <int CTRL> __device__ __forceinline__ int dpp_mov(int x)
{
    return __builtin_amdgcn_update_dpp(0, x, CTRL, 0xf, 0xf, false);
}
__device__ __forceinline__ int swap_rows16(int x)
{
    return (int)__builtin_amdgcn_permlane16_swap((unsigned)x, (unsigned)x, false, false)[1];
}
__device__ __forceinline__ int swap_rows32(int x)
{
    return (int)__builtin_amdgcn_permlane32_swap((unsigned)x, (unsigned)x, false, false)[1];
}
// a 64-lane wavefront's largest 64-bit key in every lane
template <int CTRL> __device__ __forceinline__ uint64_t dpp_mov64(uint64_t x)
{
    const unsigned lo = (unsigned)dpp_mov<CTRL>((int)(unsigned)x);
    const unsigned hi = (unsigned)dpp_mov<CTRL>((int)(unsigned)(x >> 32));
    return (uint64_t)hi << 32 | lo;
}
__device__ __forceinline__ uint64_t swap64(uint64_t x, bool rows32)
{
    const unsigned lo = (unsigned)(rows32 ? swap_rows32((int)(unsigned)x) : swap_rows16((int)(unsigned)x));
    const unsigned hi = (unsigned)(rows32 ? swap_rows32((int)(unsigned)(x >> 32)) : swap_rows16((int)(unsigned)(x >> 32)));
    return (uint64_t)hi << 32 | lo;
}
__device__ __forceinline__ uint64_t max64(uint64_t a, uint64_t b) { return a > b ? a : b; }
__device__ __forceinline__ uint64_t wave_max64(uint64_t x)
{
    x = max64(x, dpp_mov64<0xb1>(x));
    x = max64(x, dpp_mov64<0x4e>(x));
    x = max64(x, dpp_mov64<0x141>(x));
    x = max64(x, dpp_mov64<0x140>(x));
    x = max64(x, swap64(x, false));
    return max64(x, swap64(x, true));
}
// the workgroup's largest key, in every thread: one barrier.  lds holds
// TPB / 64 slots; consecutive calls must alternate between two such arrays
// (a call's slots are rewritten two calls later, after the barrier of the
// call in between has seen every thread finish reading them).
template <int TPB> __device__ uint64_t block_max64(uint64_t x, uint64_t *lds)
{
    x = wave_max64(x);
    if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = x;
    __syncthreads();
    uint64_t r = lds[0];
#pragma unroll
    for (int w = 1; w < TPB / 64; ++w) r = max64(r, lds[w]);
    return r;
}

// a pixel's reduction key: |r| bits above, the complement of index * 2 +
// sign below, so the largest key is the largest |r| at its first index;
// 0 for |r| = 0 (never a component unless every pixel is 0)
__device__ __forceinline__ uint64_t pixel_key(float absr, unsigned idx2)
{
    return absr > 0.f ? (uint64_t)__float_as_uint(absr) << 32 | (uint32_t)~idx2 : 0ull;
}

// word: |r| bits << 32 | (index * 2 + sign) << 7 | iteration tag (7 bits;
// index < 2^24); a workgroup with no |r| > 0 publishes |r| bits 0 and, for
// workgroup 0, the sign of pixel 0 (the reference's peak when the whole
// image is zero)
__device__ __forceinline__ uint64_t pack_word(float absmax, int idx, unsigned sign, unsigned tag)
{
    return (uint64_t)__float_as_uint(absmax) << 32 | ((uint64_t)idx * 2 + sign) << 7 | tag;
}

// take ? a : b, as bit operations
__device__ __forceinline__ float blend(unsigned take, float a, float b)
{
    const unsigned m = take ? ~0u : 0u;
    return __uint_as_float((__float_as_uint(a) & m) | (__float_as_uint(b) & ~m));
}

// NG 16-byte groups per thread: the first NG - NL in registers, the last NL
// in LDS (each thread reads and writes only its own LDS slots, so they need
// no barriers); the scan for the next component runs in the same pass as
// the subtraction, so a group is read and written once per iteration.
template <int NG, int NL, int TPB, int B>
__global__ void __launch_bounds__(TPB)
hogbom_resident_kernel(PersistArgs a, float *__restrict__ r, const float *__restrict__ psf_pad,
                       float *__restrict__ model, float *__restrict__ sources, int *__restrict__ num_sources,
                       State *__restrict__ st_final, uint64_t *__restrict__ words, unsigned *__restrict__ sync)
{
    constexpr int NR = NG - NL;
    static_assert(NR >= 1 && NG % B == 0 && NR % B == 0, "group split");
    __shared__ uint64_t lds_k[2][TPB / 64];   // block_max64's slots, used alternately
    __shared__ int s_ok;
    __shared__ float4 lds_v[NL > 0 ? NL * TPB : 1];
    unsigned *abort_flag = sync + 64;
    if (threadIdx.x == 0) s_ok = 1;   // read first after block_max's barriers
    const int G = a.G, half = G / 2;
    const int n4 = G * G / 4;
    const int gmask = (1 << a.row_shift) - 1;
    float4 *r4 = reinterpret_cast<float4 *>(r);
    const int qb = blockIdx.x * (TPB * NG) + threadIdx.x;
    float4 v[NR];
    auto lds_slot = [&](int k) -> float4 & { return lds_v[(k - NR) * TPB + threadIdx.x]; };
    // the scan state: the thread's largest |r| (positive pixels only count:
    // fmaxf over |r| >= 0 leaves 0 when there are none), its first group
    // holding it, and that group's values
    float m = 0.f;
    int kmax = 0;
    float4 e4 = make_float4(0.f, 0.f, 0.f, 0.f);
    auto scan_group = [&](int k, const float4 &g) {
        const float gm = fmaxf(fmaxf(fabsf(g.x), fabsf(g.y)), fmaxf(fabsf(g.z), fabsf(g.w)));
        const bool up = gm > m;
        kmax = up ? k : kmax;
        e4.x = up ? g.x : e4.x;
        e4.y = up ? g.y : e4.y;
        e4.z = up ? g.z : e4.z;
        e4.w = up ? g.w : e4.w;
        m = fmaxf(m, gm);
    };
#pragma unroll
    for (int k = 0; k < NG; ++k) {
        const int q = qb + k * TPB;
        const float4 g = q < n4 ? r4[q] : make_float4(0.f, 0.f, 0.f, 0.f);
        if (k < NR) v[k < NR ? k : 0] = g;
        else lds_slot(k) = g;
        scan_group(k, g);
    }
    const __amdgpu_buffer_rsrc_t rs =
        __builtin_amdgcn_make_buffer_rsrc((void *)psf_pad, (short)0, a.psf_pad_bytes, 0x00020000);
    int nsrc = 0, stop = 0;
    float src0 = 0.f;
    for (int it = 0; it < a.n_minor; ++it) {
        // opaque to the optimiser: the groups' coordinates are formed again
        // per iteration, not hoisted into registers held all along
        int qi = qb;
        asm volatile("" : "+v"(qi));
        // the workgroup's best pixel: the largest key over its threads, each
        // the thread's largest |r| (m) at the first index holding it
        uint64_t key = 0;
        if (m > 0.f) {
            const int j = fabsf(e4.x) == m ? 0 : fabsf(e4.y) == m ? 1 : fabsf(e4.z) == m ? 2 : 3;
            const float ej = j == 0 ? e4.x : j == 1 ? e4.y : j == 2 ? e4.z : e4.w;
            key = pixel_key(m, (unsigned)(((qi + kmax * TPB) * 4 + j) * 2) + (__float_as_uint(ej) >> 31));
        }
        key = block_max64<TPB>(key, lds_k[0]);
        // copy c of parity p at words + (c * 2 + p) * kWordStride
        uint64_t *par = words + (size_t)(it & 1) * kWordStride;
        uint64_t *buf = par + (size_t)(blockIdx.x % kWordReplicas) * 2 * kWordStride;   // the copy this workgroup reads
        const unsigned tag = (unsigned)it & 0x3fu;   // 0x7f never: the reset words carry it
        if (threadIdx.x < 64) {
            // pixel 0 lives in thread 0 of workgroup 0 (group 0: registers)
            const unsigned sign0 = (unsigned)__builtin_amdgcn_readfirstlane((int)__float_as_uint(v[0].x)) >> 31;
            const uint64_t w = key ? (key & 0xffffffff00000000ull) | (uint64_t)(~(uint32_t)key) << 7 | tag
                                   : pack_word(0.f, 0, sign0, tag);
            if (threadIdx.x < kWordReplicas)
                __hip_atomic_store(par + (size_t)threadIdx.x * 2 * kWordStride + blockIdx.x, w, __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_AGENT);
        }
        // all words: thread t < nblocks reads workgroup t's until it carries
        // this iteration's tag; thread 0 reads the abort flag with every read
        // of its word (an abort raised anywhere ends every workgroup at its
        // next exchange) and passes it on through s_ok, which the other
        // waiting threads read each round
        uint64_t wkey = 0;
        if ((int)threadIdx.x < a.nblocks) {
            const unsigned long long t0 = __builtin_amdgcn_s_memrealtime();
            uint64_t w;
            for (;;) {
                w = __hip_atomic_load(buf + threadIdx.x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                if (threadIdx.x == 0) {
                    if (__hip_atomic_load(abort_flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
                        s_ok = 0;
                        break;
                    }
                } else if (!*(volatile int *)&s_ok) {
                    break;
                }
                if ((w & 0x7fu) == tag) break;
                if (__builtin_amdgcn_s_memrealtime() - t0 > kWaitTicks) {
                    __hip_atomic_store(abort_flag, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    s_ok = 0;
                    break;
                }
                __builtin_amdgcn_s_sleep(3);
            }
            wkey = pixel_key(__uint_as_float((uint32_t)(w >> 32)), (unsigned)(w >> 7) & 0x1ffffffu);
        }
        // the same reduction over the words in every workgroup (its barrier
        // also publishes s_ok)
        const uint64_t gkey = block_max64<TPB>(wkey, lds_k[1]);
        if (!s_ok) {
            stop = 2;
            break;   // every workgroup leaves: the abort flag is set
        }
        float curr_max, peak;
        int max_idx;
        if (gkey) {
            const unsigned idx2 = ~(uint32_t)gkey;
            curr_max = __uint_as_float((uint32_t)(gkey >> 32));
            max_idx = (int)(idx2 >> 1);
            peak = (idx2 & 1u) ? -curr_max : curr_max;
        } else {
            // nothing beat the initial 0: the reference keeps pixel (0, 0)
            // with its value, a zero of pixel 0's sign (workgroup 0's word,
            // this iteration's: workgroup 0 publishes no later one before
            // this workgroup's next word)
            const uint64_t w0 = __hip_atomic_load(buf, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            curr_max = 0.f;
            max_idx = 0;
            peak = ((w0 >> 7) & 1u) ? -0.f : 0.f;
        }
        const bool weak_source = nsrc > 0 && (double)curr_max < (double)fabsf(src0) * a.weak;
        if (weak_source || nsrc >= a.max_sources) {
            stop = 1;
            break;   // the same decision in every workgroup
        }
        const float max_val = (float)((double)peak / a.psf_max);
        const int mx = max_idx % G, my = max_idx / G;
        const float flux = (float)((double)max_val * a.gain);
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            sources[3 * nsrc] = (float)mx;
            sources[3 * nsrc + 1] = (float)my;
            sources[3 * nsrc + 2] = flux;
            model[max_idx] = (float)((double)model[max_idx] + (double)max_val * a.gain);
            *num_sources = nsrc + 1;
        }
        if (nsrc == 0) src0 = flux;
        ++nsrc;
        // subtraction window (deconvolution_run.c:259-264), and in the same
        // pass the scan for the next component
        const int xlo = mx - min(mx, half), xhi = mx + min(G - mx, half);
        const int ylo = my - min(my, half), yhi = my + min(G - my, half);
        const double scale = a.gain * (double)max_val;
        const double pm = a.psf_max;
        m = 0.f;
        kmax = 0;
#pragma unroll
        for (int k0 = 0; k0 < NG; k0 += B) {
            v4f_t w[B];
            unsigned lanes[B];
            float4 e[B];
#pragma unroll
            for (int k = 0; k < B; ++k) {
                const int q = qi + (k0 + k) * TPB;
                const int y = q >> a.row_shift, x0 = (q & gmask) * 4;
                // bitwise tests (no short-circuit branches)
                const unsigned row = (unsigned)(q < n4) & (unsigned)(y >= ylo) & (unsigned)(y < yhi);
                unsigned mk = 0;
#pragma unroll
                for (int j = 0; j < 4; ++j) mk |= (row & (unsigned)(x0 + j >= xlo) & (unsigned)(x0 + j < xhi)) << j;
                lanes[k] = mk;
                const int prow = (y - my + half) * G;
                const int lo = a.mirror ? prow + (G - (x0 + 3 - mx + half)) : prow + (x0 - mx + half);
                const unsigned off = mk ? (unsigned)(lo + kPsfPad) * 4u : kOob;
                w[k] = __builtin_bit_cast(v4f_t, __builtin_amdgcn_raw_buffer_load_b128(rs, off, 0, 0));
                e[k] = k0 + k < NR ? v[k0 + k < NR ? k0 + k : 0] : lds_slot(k0 + k);
            }
#pragma unroll
            for (int k = 0; k < B; ++k) {
                const unsigned mk = lanes[k];
                const float p0 = a.mirror ? w[k].w : w[k].x, p1 = a.mirror ? w[k].z : w[k].y;
                const float p2 = a.mirror ? w[k].y : w[k].z, p3 = a.mirror ? w[k].x : w[k].w;
                // every lane computed, the window's kept by a bitwise blend
                // (no per-pixel branches)
                float4 g = e[k];
                g.x = blend(mk & 1u, subtract_one(g.x, scale, p0, pm), g.x);
                g.y = blend(mk & 2u, subtract_one(g.y, scale, p1, pm), g.y);
                g.z = blend(mk & 4u, subtract_one(g.z, scale, p2, pm), g.z);
                g.w = blend(mk & 8u, subtract_one(g.w, scale, p3, pm), g.w);
                if (k0 + k < NR) v[k0 + k < NR ? k0 + k : 0] = g;
                else lds_slot(k0 + k) = g;
                scan_group(k0 + k, g);
            }
        }
    }
#pragma unroll
    for (int k = 0; k < NG; ++k) {
        const int q = qb + k * TPB;
        if (q < n4) r4[q] = k < NR ? v[k < NR ? k : 0] : lds_slot(k);
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) *st_final = State{stop ? 1 : 0, stop == 2 ? -1 : nsrc, src0, 0.f};
}

// the abort flag to `abort` (1 only to exercise the exit path: every
// workgroup then leaves at its first exchange), the words of both parities
// in every copy to a tag no iteration carries
__global__ void hogbom_resident_init_kernel(unsigned *sync, uint64_t *words, int n, unsigned abort)
{
    if (threadIdx.x == 64) sync[64] = abort;
    for (int c = 0; c < 2 * kWordReplicas; ++c)
        for (int i = threadIdx.x; i < n; i += blockDim.x) words[(size_t)c * kWordStride + i] = 0x7full;
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

// then the one-launch form's abort flag (at byte 256 of a 512-byte block)
size_t sync_offset(size_t npix)
{
    return (psf_pad_offset(npix) + (npix + 2 * kPsfPad) * sizeof(float) + 255) & ~(size_t)255;
}

// then its words: both parities in every copy
size_t words_offset(size_t npix) { return sync_offset(npix) + 512; }

size_t words_bytes() { return sizeof(uint64_t) * 2 * kWordReplicas * kWordStride; }

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

// All iterations in one cooperative launch (hogbom_resident_kernel) when it
// applies: a power-of-two side whose pixels fit the registers of one
// workgroup per CU, a noise stop that cannot fire, a stream that is not
// being captured (a graph keeps per-iteration launches) and a device that
// takes the cooperative launch.  Returns 1 when launched, 0 when the
// caller should queue per-iteration launches.  SDP_HOGBOM_RESIDENT=0 in the
// environment turns it off for the process.
int launch_resident(int G, int n_minor, int max_sources, const HogbomArgs &h, double noise_gain, float *d_residual,
                    const float *psf_pad, float *d_model, float *d_sources, int *d_num_sources, State *st_final,
                    void *d_work, bool preset_abort, hipStream_t st)
{
    static int env = -1;
    if (env < 0) {
        const char *e = getenv("SDP_HOGBOM_RESIDENT");
        env = (e && atoi(e) == 0) ? 0 : 1;
    }
    if (!env || h.exact_mean || noise_gain > 0.5 || n_minor < 1 || G < 16 || G > 4096 || (G & (G - 1)) != 0)
        return 0;
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(st, &cs) != hipSuccess || cs != hipStreamCaptureStatusNone) return 0;
    static int coop = -1;
    if (coop < 0) {
        int dev = 0, v = 0;
        coop = (hipGetDevice(&dev) == hipSuccess &&
                hipDeviceGetAttribute(&v, hipDeviceAttributeCooperativeLaunch, dev) == hipSuccess && v)
                   ? 1
                   : 0;
    }
    if (!coop) return 0;
    const size_t npix = (size_t)G * G, n4 = npix / 4;
    // (groups per thread, threads per workgroup): 1024 threads up to 8
    // groups (128 registers a thread); beyond, 512 threads and the
    // registers for more PSF loads in flight
    struct Shape {
        int ng, tpb;
        const void *k;
    };
    const Shape shapes[5] = {{1, 1024, (const void *)hogbom_resident_kernel<1, 0, 1024, 1>},
                             {2, 1024, (const void *)hogbom_resident_kernel<2, 0, 1024, 2>},
                             {4, 1024, (const void *)hogbom_resident_kernel<4, 0, 1024, 4>},
                             {8, 1024, (const void *)hogbom_resident_kernel<8, 0, 1024, 4>},
                             {16, 1024, (const void *)hogbom_resident_kernel<16, 8, 1024, 4>}};
    for (const Shape &sh : shapes) {
        const size_t nb = (n4 + (size_t)sh.tpb * sh.ng - 1) / ((size_t)sh.tpb * sh.ng);
        int per_cu = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, sh.k, sh.tpb, 0) != hipSuccess || per_cu < 1)
            continue;
        // one workgroup per CU at most; one waiting thread per word
        if (nb > (size_t)compute_units() || nb > (size_t)sh.tpb || nb > (size_t)kWordStride) continue;
        PersistArgs a;
        a.G = G;
        a.n_minor = n_minor;
        a.max_sources = max_sources;
        a.nblocks = (int)nb;
        a.mirror = h.mirror;
        a.psf_pad_bytes = h.psf_pad_bytes;
        a.row_shift = 0;
        while ((1 << a.row_shift) < G / 4) ++a.row_shift;
        a.psf_max = h.psf_max;
        a.gain = h.gain;
        a.weak = h.weak;
        uint64_t *words = (uint64_t *)((char *)d_work + words_offset(npix));
        unsigned *sync = (unsigned *)((char *)d_work + sync_offset(npix));
        hogbom_resident_init_kernel<<<1, 256, 0, st>>>(sync, words, (int)nb, preset_abort ? 1u : 0u);
        SDP_LAUNCH_CHECK();
        void *args[] = {&a, &d_residual, (void *)&psf_pad, &d_model, &d_sources, &d_num_sources, &st_final,
                        &words, &sync};
        // A cooperative launch can be refused (another process may hold the
        // device's cooperative resources); nothing is enqueued then, and the
        // iterations go as separate launches.
        if (hipLaunchCooperativeKernel(sh.k, dim3((unsigned)nb), dim3(sh.tpb), args, 0, st) != hipSuccess) {
            (void)hipGetLastError();
            return 0;
        }
        return 1;
    }
    return 0;
}

}  // namespace

extern "C" {

size_t sdp_hogbom_work_bytes(int G)
{
    if (G < 2) return 0;
    const size_t npix = (size_t)G * G;
    // partials of two consecutive launches, the padded PSF copy, the abort
    // flag, the one-launch form's words
    return words_offset(npix) + words_bytes();
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
    if (!(flags & SDP_HOGBOM_PER_ITERATION) &&
        launch_resident(G, n_minor, max_sources, a, noise_factor * loop_gain, d_residual, psf_pad, d_model_out,
                        d_sources, d_num_sources, state + ((n_minor + 1) & 1), d_work,
                        (flags & SDP_HOGBOM_TEST_ABORT) != 0, st)) {
        hogbom_finish_kernel<<<4, 256, 0, st>>>(state + ((n_minor + 1) & 1), G, cell_size, d_sources,
                                                d_num_sources);
        SDP_LAUNCH_CHECK();
        return SDP_OK;
    }
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
