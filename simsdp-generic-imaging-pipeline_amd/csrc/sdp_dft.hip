// sdp_dft.hip -- direct Fourier transform predict of visibilities from a
// point-source list, dft_actor (src/CPU/dft_run.c:3-47, CUDA twin
// src/GPU/dft_run.cu:80-116):
//   V_j = sum_s f_s / (1 - r_s^2/2) * exp(-i 2 pi (u_j l_s + v_j m_s - w_j r_s^2/2)),
// with r^2 = l^2 + m^2 (the reference's small-angle form of n - 1, :21-23).
//
// One thread per visibility; sources are staged through LDS in batches with
// their per-source terms precomputed once per workgroup.
//  SDP_DFT_REFERENCE_FP32: the reference's float arithmetic (phase in float,
//      sinf/cosf of the full phase), for parity with the CPU actor.
//  SDP_DFT_FP64_PHASE: phase (in turns) formed in double and reduced to
//      [-1/2, 1/2] exactly, then sin/cos of the reduced float argument;
//      float accumulation.  Accurate to float rounding independent of |uvw|.
//
// sdp_dft_predict_array: the same sum for a whole array, factorised per
// antenna.  A baseline's phase is the difference of its two antennas'
// phases, so per instant t the visibilities are the complex matrix product
//   V(t)[a][b] = sum_s conj(E_as) g_s E_bs
// of the antenna-by-source phasor matrix with itself: n_ant * n_src
// phasors per instant instead of n_ant^2 / 2 * n_src, and the contraction
// runs on the matrix cores (v_mfma_f32_16x16x32_{bf16,f16}).  Real form,
// k running over (re, im) of each source:
//   Re V = sum_k A_r[a][k] B[b][k],  A_r[a] = (Re E_a, Im E_a)
//   Im V = sum_k A_i[a][k] B[b][k],  A_i[a] = (-Im E_a, Re E_a)
//   B[b] = g (Re E_b, Im E_b) * scale
// One workgroup per (instant, 128 x 128 antenna tile with a-tile <= b-tile);
// per chunk of 16 sources its threads form the tile's antennas' phasors
// (FP64 phase and range reduction; a diagonal tile's two sides are the same
// antennas and are formed once), split them into hi and lo parts and stage
// them in LDS; each of the 8 waves then runs 4 x 2 x 2 x 3 MFMAs for its
// 64 x 32 block.  scale is a power of two that keeps FP16 operands in range.
#include <cstdlib>
#include <mutex>

#include "sdp_internal.h"

namespace {

constexpr int kThreads = 256;
constexpr int kBatch = 256;

struct SrcRef {   // reference float terms (dft_run.c:21-25)
    float l, m, wcorr, corr;
};

struct SrcF64 {
    double l, m, wterm;   // wterm = -r^2/2
    float corr, pad;
};

__global__ void __launch_bounds__(kThreads)
dft_ref_kernel(int n_vis, int n_src, const float *__restrict__ src, const float *__restrict__ uvw, float2 *__restrict__ out)
{
    __shared__ SrcRef s[kBatch];
    const int j = blockIdx.x * kThreads + threadIdx.x;
    float u = 0.f, v = 0.f, w = 0.f;
    if (j < n_vis) {
        u = uvw[3 * (size_t)j];
        v = uvw[3 * (size_t)j + 1];
        w = uvw[3 * (size_t)j + 2];
    }
    const float pi = 3.141592654f;
    const float two_pi = pi + pi;
    float sr = 0.f, si = 0.f;
    for (int b = 0; b < n_src; b += kBatch) {
        __syncthreads();
        const int k = b + threadIdx.x;
        if (threadIdx.x < kBatch && k < n_src) {
            const float l = src[3 * k], m = src[3 * k + 1], f = src[3 * k + 2];
            const float term = (float)(0.5 * (double)(l * l + m * m));
            const float image_correction = (float)(1.0 - (double)term);
            s[threadIdx.x] = SrcRef{l, m, -term, f / image_correction};
        }
        __syncthreads();
        const int nb = min(kBatch, n_src - b);
        for (int q = 0; q < nb; ++q) {
            const SrcRef t = s[q];
            const float theta = (u * t.l + v * t.m + w * t.wcorr) * two_pi;
            const float sn = sinf(theta), cs = cosf(theta);
            sr += cs * t.corr;
            si += -sn * t.corr;
        }
    }
    if (j < n_vis) out[j] = make_float2(sr, si);
}

__global__ void __launch_bounds__(kThreads)
dft_f64_kernel(int n_vis, int n_src, const float *__restrict__ src, const float *__restrict__ uvw, float2 *__restrict__ out)
{
    __shared__ SrcF64 s[kBatch];
    const int j = blockIdx.x * kThreads + threadIdx.x;
    double u = 0.0, v = 0.0, w = 0.0;
    if (j < n_vis) {
        u = uvw[3 * (size_t)j];
        v = uvw[3 * (size_t)j + 1];
        w = uvw[3 * (size_t)j + 2];
    }
    float sr = 0.f, si = 0.f;
    for (int b = 0; b < n_src; b += kBatch) {
        __syncthreads();
        const int k = b + threadIdx.x;
        if (k < n_src) {
            const double l = src[3 * k], m = src[3 * k + 1];
            const double term = 0.5 * (l * l + m * m);
            s[threadIdx.x] = SrcF64{l, m, -term, (float)((double)src[3 * k + 2] / (1.0 - term)), 0.f};
        }
        __syncthreads();
        const int nb = min(kBatch, n_src - b);
        for (int q = 0; q < nb; ++q) {
            const SrcF64 t = s[q];
            double turns = __builtin_fma(w, t.wterm, __builtin_fma(v, t.m, u * t.l));
            turns -= rint(turns);                       // exact: |turns - rint| <= 1/2
            float sn, cs;
            sincospif(2.0f * (float)turns, &sn, &cs);   // sin, cos of 2 pi * turns
            sr += cs * t.corr;
            si -= sn * t.corr;
        }
    }
    if (j < n_vis) out[j] = make_float2(sr, si);
}

// ---------------------------------------------------------------- array form

constexpr int kTile = 128;     // antennas per tile side
constexpr int kArrayThreads = 512;   // 8 waves: 2 x 4 blocks of 64 x 32 outputs
constexpr int kChunk = 16;     // sources per MFMA k-step (k = 32: re, im)
constexpr int kRow = 40;       // LDS row stride in 16-bit elements (32 + 8 pad)

struct SrcTerm {   // per source: l, m, -r^2/2 (FP64), g * scale
    double l, m, n;
    float g, pad;
};

typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

// source terms and the power-of-two scale: one workgroup
__global__ void __launch_bounds__(kThreads)
dft_array_prep_kernel(int n_src, int n_pad, const float *__restrict__ src, SrcTerm *__restrict__ terms,
                      float *__restrict__ scale_out, int fp16)
{
    __shared__ float red[kThreads];
    float mx = 0.f;
    for (int k = threadIdx.x; k < n_src; k += kThreads) {
        const double l = src[3 * k], m = src[3 * k + 1];
        const double term = 0.5 * (l * l + m * m);
        mx = fmaxf(mx, fabsf((float)((double)src[3 * k + 2] / (1.0 - term))));
    }
    red[threadIdx.x] = mx;
    __syncthreads();
    for (int off = kThreads / 2; off > 0; off >>= 1) {
        if (threadIdx.x < off) red[threadIdx.x] = fmaxf(red[threadIdx.x], red[threadIdx.x + off]);
        __syncthreads();
    }
    // FP16: the largest |g| scaled into [2^13, 2^14), so hi and lo parts
    // stay normal numbers for all but the faintest sources
    float scale = 1.f;
    if (fp16 && red[0] > 0.f) {
        int e;
        frexpf(red[0], &e);               // red[0] = f 2^e, f in [0.5, 1)
        scale = ldexpf(1.f, 14 - e);
    }
    if (threadIdx.x == 0) *scale_out = scale;
    for (int k = threadIdx.x; k < n_pad; k += kThreads) {
        SrcTerm t{0.0, 0.0, 0.0, 0.f, 0.f};
        if (k < n_src) {
            const double l = src[3 * k], m = src[3 * k + 1];
            const double term = 0.5 * (l * l + m * m);
            t = SrcTerm{l, m, -term, (float)((double)src[3 * k + 2] / (1.0 - term)) * scale, 0.f};
        }
        terms[k] = t;
    }
}

template <typename T> struct Mfma;
template <> struct Mfma<__bf16> {
    typedef __bf16 v8 __attribute__((ext_vector_type(8)));
    static __device__ __forceinline__ f32x4 run(v8 a, v8 b, f32x4 c)
    {
        return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
    }
};
template <> struct Mfma<_Float16> {
    typedef _Float16 v8 __attribute__((ext_vector_type(8)));
    static __device__ __forceinline__ f32x4 run(v8 a, v8 b, f32x4 c)
    {
        return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0);
    }
};

// (re, im) pairs -> (-im, re): A_i from A_r, on the packed 16-bit halves
template <typename V> __device__ __forceinline__ V rotate_pairs(V a)
{
    u32x4 x = __builtin_bit_cast(u32x4, a);
#pragma unroll
    for (int i = 0; i < 4; ++i) x[i] = ((x[i] >> 16) | (x[i] << 16)) ^ 0x00008000u;
    return __builtin_bit_cast(V, x);
}

// sin and cos of 2 pi x for x in [-1/2, 1/2]: the library routine, or the
// hardware v_sin_f32 / v_cos_f32, whose argument is in revolutions
template <bool kHw> __device__ __forceinline__ void sincos_turns(float x, float *sn, float *cs)
{
    if (kHw) {
        *sn = __builtin_amdgcn_sinf(x);
        *cs = __builtin_amdgcn_cosf(x);
    } else {
        sincospif(2.0f * x, sn, cs);
    }
}

// kPer sources' phasors for one antenna, split into hi and lo parts (the
// weighted ones when `weighted`), and with kTwin also the weighted twins
template <typename T, bool kHw, int kPer, bool kTwin>
__device__ __forceinline__ void form_phasors(const SrcTerm *__restrict__ terms, double u, double v, double w,
                                             bool live, bool weighted, T *hi, T *lo, T *whi, T *wlo)
{
#pragma unroll
    for (int q = 0; q < kPer; ++q) {
        const SrcTerm st = terms[q];
        double turns = __builtin_fma(w, st.n, __builtin_fma(v, st.m, u * st.l));
        turns -= rint(turns);
        float sn, cs;
        sincos_turns<kHw>((float)turns, &sn, &cs);
        const float re = live ? cs : 0.f, im = live ? -sn : 0.f;
        const float wgt = weighted ? st.g : 1.f;
        float r1 = re * wgt, i1 = im * wgt;
        // the products are rounded to float before the split: without this
        // register-only barrier the compiler may fold a product into the
        // 16-bit conversion that forms hi for lo (v_fma_mix, one rounding)
        // but not into the one that stores hi (two roundings), and the two
        // hi values then differ by an ulp where the roundings disagree
        asm volatile("" : "+v"(r1), "+v"(i1));
        hi[2 * q] = (T)r1;
        hi[2 * q + 1] = (T)i1;
        lo[2 * q] = (T)(r1 - (float)hi[2 * q]);
        lo[2 * q + 1] = (T)(i1 - (float)hi[2 * q + 1]);
        if (kTwin) {
            float r2 = re * st.g, i2 = im * st.g;
            asm volatile("" : "+v"(r2), "+v"(i2));
            whi[2 * q] = (T)r2;
            whi[2 * q + 1] = (T)i2;
            wlo[2 * q] = (T)(r2 - (float)whi[2 * q]);
            wlo[2 * q + 1] = (T)(i2 - (float)whi[2 * q + 1]);
        }
    }
}

template <typename T, bool kHw>
__global__ void __launch_bounds__(kArrayThreads)
dft_array_kernel(int n_ant, int n_chunks, const double *__restrict__ ant_uvw, const SrcTerm *__restrict__ terms,
                 const float *__restrict__ scale_in, float2 *__restrict__ out, int split)
{
    typedef typename Mfma<T>::v8 v8;
    // [hi/lo][antenna slot][k]: slots 0..kTile-1 the a-tile (conjugated
    // side), kTile..2 kTile-1 the b-tile (weighted side)
    __shared__ __attribute__((aligned(16))) T stage[2][2][2 * kTile][kRow];   // [buffer][hi/lo][slot][k]

    // triangle tile index -> (ta, tb), ta <= tb
    int tb = 0, rem = blockIdx.x;
    while (rem > tb) {
        rem -= tb + 1;
        ++tb;
    }
    const int ta = rem;
    const bool diag = ta == tb;
    const int t = blockIdx.y;

    // phasor duty: off the diagonal, thread -> (slot, 8 sources) over all
    // 2 kTile slots; on it the two sides are the same antennas, so thread
    // -> (a-side slot, 4 sources) and each phasor is written to both sides
    // (the source part is uniform over a wavefront, so its terms are scalar
    // loads shared by the wavefront's lanes)
    const int per = diag ? 4 : 8;
    const int slot = diag ? (int)(threadIdx.x & (kTile - 1)) : (int)(threadIdx.x & (2 * kTile - 1));
    const int part = __builtin_amdgcn_readfirstlane(diag ? (int)(threadIdx.x >> 7) : (int)(threadIdx.x >> 8));
    const int ant = slot < kTile ? ta * kTile + slot : tb * kTile + (slot - kTile);
    const bool weighted = slot >= kTile;
    double u = 0.0, v = 0.0, w = 0.0;
    const bool live = ant < n_ant;
    if (live) {
        const double *p = ant_uvw + ((size_t)t * n_ant + ant) * 3;
        u = p[0];
        v = p[1];
        w = p[2];
    }

    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int wr = (wave >> 2) * 64, wc = (wave & 3) * 32;   // this wave's 64 x 32 block (rows a, cols b)
    const int kq = (lane >> 4) * 8;
    f32x4 acc_r[4][2], acc_i[4][2];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc_r[i][j] = acc_i[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};

    // Software pipeline over chunks of 16 sources, two LDS buffers: while
    // the matrix cores consume chunk c from buffer c & 1, the vector ALUs
    // form chunk c + 1's phasors, which then go to the other buffer; one
    // barrier per chunk.  The work buffer holds one chunk of zero terms
    // past the last, so the final iteration's look-ahead reads valid data.
    T hi[16], lo[16], whi[8], wlo[8];
    auto form = [&](int c) {
        if (diag)
            form_phasors<T, kHw, 4, true>(terms + c * kChunk + part * per, u, v, w, live, weighted, hi, lo, whi, wlo);
        else
            form_phasors<T, kHw, 8, false>(terms + c * kChunk + part * per, u, v, w, live, weighted, hi, lo, whi,
                                           wlo);
    };
    auto store = [&](int buf) {
        if (diag) {
            const int k0 = part * 8;
            *reinterpret_cast<v8 *>(&stage[buf][0][slot][k0]) =
                v8{hi[0], hi[1], hi[2], hi[3], hi[4], hi[5], hi[6], hi[7]};
            *reinterpret_cast<v8 *>(&stage[buf][1][slot][k0]) =
                v8{lo[0], lo[1], lo[2], lo[3], lo[4], lo[5], lo[6], lo[7]};
            *reinterpret_cast<v8 *>(&stage[buf][0][kTile + slot][k0]) =
                v8{whi[0], whi[1], whi[2], whi[3], whi[4], whi[5], whi[6], whi[7]};
            *reinterpret_cast<v8 *>(&stage[buf][1][kTile + slot][k0]) =
                v8{wlo[0], wlo[1], wlo[2], wlo[3], wlo[4], wlo[5], wlo[6], wlo[7]};
        } else {
#pragma unroll
            for (int q = 0; q < 16; q += 8) {
                *reinterpret_cast<v8 *>(&stage[buf][0][slot][part * 16 + q]) =
                    v8{hi[q], hi[q + 1], hi[q + 2], hi[q + 3], hi[q + 4], hi[q + 5], hi[q + 6], hi[q + 7]};
                *reinterpret_cast<v8 *>(&stage[buf][1][slot][part * 16 + q]) =
                    v8{lo[q], lo[q + 1], lo[q + 2], lo[q + 3], lo[q + 4], lo[q + 5], lo[q + 6], lo[q + 7]};
            }
        }
    };
    form(0);
    store(0);
    __syncthreads();
    for (int c = 0; c < n_chunks; ++c) {
        const int buf = c & 1;
        form(c + 1);   // look-ahead (chunk n_chunks is the zero pad)
        v8 bb[2][2];   // [hi/lo][column group]
#pragma unroll
        for (int h = 0; h < 2; ++h)
#pragma unroll
            for (int j = 0; j < 2; ++j)
                bb[h][j] = *reinterpret_cast<const v8 *>(&stage[buf][h][kTile + wc + j * 16 + (lane & 15)][kq]);
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const v8 ar0 = *reinterpret_cast<const v8 *>(&stage[buf][0][wr + i * 16 + (lane & 15)][kq]);
            const v8 ar1 = *reinterpret_cast<const v8 *>(&stage[buf][1][wr + i * 16 + (lane & 15)][kq]);
            const v8 ai0 = rotate_pairs(ar0), ai1 = rotate_pairs(ar1);
#pragma unroll
            for (int j = 0; j < 2; ++j) {
                acc_r[i][j] = Mfma<T>::run(ar0, bb[0][j], acc_r[i][j]);
                acc_i[i][j] = Mfma<T>::run(ai0, bb[0][j], acc_i[i][j]);
                if (split) {
                    acc_r[i][j] = Mfma<T>::run(ar0, bb[1][j], acc_r[i][j]);
                    acc_r[i][j] = Mfma<T>::run(ar1, bb[0][j], acc_r[i][j]);
                    acc_i[i][j] = Mfma<T>::run(ai0, bb[1][j], acc_i[i][j]);
                    acc_i[i][j] = Mfma<T>::run(ai1, bb[0][j], acc_i[i][j]);
                }
            }
        }
        // buffer buf ^ 1 was last read in iteration c - 1, before its barrier
        store(buf ^ 1);
        __syncthreads();
    }

    // C[a][b], a = row (4 (lane / 16) + r), b = column (lane % 16): for a
    // fixed a, 16 consecutive b are 16 consecutive baselines
    const float inv = 1.f / *scale_in;
    const size_t n_pairs = (size_t)n_ant * (n_ant - 1) / 2;
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int a = ta * kTile + wr + i * 16 + (lane >> 4) * 4 + r;
                const int b = tb * kTile + wc + j * 16 + (lane & 15);
                if (a < b && b < n_ant) {
                    const size_t k = (size_t)a * n_ant - (size_t)a * (a + 1) / 2 + (size_t)(b - a - 1);
                    out[(size_t)t * n_pairs + k] = make_float2(acc_r[i][j][r] * inv, acc_i[i][j][r] * inv);
                }
            }
}

size_t array_terms_offset() { return 256; }

// The allocation-free entry point's work buffer: the scale and the source
// terms of up to SDP_DFT_NOALLOC_MAX_SOURCES sources (plus the zero chunk),
// 2.0 MiB held by each GPU the library runs on until the process ends.  Host
// code never uses this name as an address: each device's copy is reached
// through the pointer hipGetSymbolAddress returns with that device current.
constexpr size_t kScratchBytes = 256 + (size_t)(SDP_DFT_NOALLOC_MAX_SOURCES + kChunk) * sizeof(SrcTerm);
static __device__ __align__(256) char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[kMaxDevices] = {};

// the current device's scratch address, looked up on its first use
int scratch_address(void **out)
{
    int dev = 0;
    SDP_HIP_TRY(hipGetDevice(&dev));
    SDP_REQUIRE(dev >= 0 && dev < kMaxDevices, "device %d beyond the scratch table", dev);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return sdp::set_error(SDP_E_HIP, "hipGetSymbolAddress(scratch) failed: %s", hipGetErrorString(e));
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return SDP_OK;
}

}  // namespace

extern "C" {

size_t sdp_dft_array_work_bytes(int n_sources)
{
    const size_t n_pad = (size_t)((n_sources > 0 ? n_sources : 0) + kChunk - 1) / kChunk * kChunk;
    return array_terms_offset() + (n_pad + kChunk) * sizeof(SrcTerm);   // one zero chunk past the last
}

int sdp_dft_predict_array(int n_time, int n_ant, const double *d_ant_uvw, int n_sources, const float *d_sources,
                          float *d_vis_out, int flags, void *d_work, sdp_stream s)
{
    SDP_REQUIRE(n_time >= 0 && n_ant >= 0 && n_sources >= 0, "sdp_dft_predict_array: negative counts");
    SDP_REQUIRE(n_ant <= 46341, "sdp_dft_predict_array: too many antennas");
    SDP_REQUIRE((flags & ~(SDP_DFT_ARRAY_FP16 | SDP_DFT_ARRAY_NO_SPLIT)) == 0, "sdp_dft_predict_array: bad flags");
    SDP_REQUIRE(n_time <= 65535, "sdp_dft_predict_array: too many instants for one launch");
    if (n_time == 0 || n_ant < 2) return SDP_OK;
    SDP_REQUIRE(d_ant_uvw && d_vis_out && d_work && (n_sources == 0 || d_sources),
                "sdp_dft_predict_array: null pointer");
    SDP_REQUIRE(((uintptr_t)d_work & 15) == 0, "sdp_dft_predict_array: work buffer must be 16-byte aligned");
    hipStream_t st = sdp::stream_of(s);
    const int n_pad = (n_sources + kChunk - 1) / kChunk * kChunk;
    float *scale = (float *)d_work;
    SrcTerm *terms = (SrcTerm *)((char *)d_work + array_terms_offset());
    const int fp16 = (flags & SDP_DFT_ARRAY_FP16) != 0;
    dft_array_prep_kernel<<<1, kThreads, 0, st>>>(n_sources, n_pad + kChunk, d_sources, terms, scale, fp16);
    SDP_LAUNCH_CHECK();
    const int nt = (n_ant + kTile - 1) / kTile;
    const dim3 grid(nt * (nt + 1) / 2, n_time);
    const int split = (flags & SDP_DFT_ARRAY_NO_SPLIT) ? 0 : 1;
    static int hw = -1;   // SDP_DFT_SINCOS=lib selects the library sin/cos (for comparison)
    if (hw < 0) {
        const char *e = getenv("SDP_DFT_SINCOS");
        hw = (e && e[0] == 'l') ? 0 : 1;
    }
    auto kern = fp16 ? (hw ? dft_array_kernel<_Float16, true> : dft_array_kernel<_Float16, false>)
                     : (hw ? dft_array_kernel<__bf16, true> : dft_array_kernel<__bf16, false>);
    kern<<<grid, kArrayThreads, 0, st>>>(n_ant, n_pad / kChunk, d_ant_uvw, terms, scale, (float2 *)d_vis_out, split);
    SDP_LAUNCH_CHECK();
    return SDP_OK;
}

int sdp_dft_predict_array_noalloc(int n_time, int n_ant, const double *d_ant_uvw, int n_sources,
                                  const float *d_sources, float *d_vis_out, int flags, sdp_stream s)
{
    SDP_REQUIRE(n_sources >= 0 && n_sources <= SDP_DFT_NOALLOC_MAX_SOURCES,
                "sdp_dft_predict_array_noalloc: %d sources, at most %d", n_sources, SDP_DFT_NOALLOC_MAX_SOURCES);
    static_assert(256 + (size_t)(SDP_DFT_NOALLOC_MAX_SOURCES + kChunk) * sizeof(SrcTerm) == kScratchBytes,
                  "scratch holds the work buffer of the largest call");
    void *work = nullptr;
    if (int e = scratch_address(&work)) return e;
    // every byte the kernels read is written by this call's prep kernel
    return sdp_dft_predict_array(n_time, n_ant, d_ant_uvw, n_sources, d_sources, d_vis_out, flags, work, s);
}

int sdp_dft_predict(int n_vis, int n_sources, const float *d_sources, const float *d_uvw, float *d_vis_out, int mode,
                    sdp_stream s)
{
    SDP_REQUIRE(n_vis >= 0 && n_sources >= 0, "negative counts");
    SDP_REQUIRE(mode == SDP_DFT_REFERENCE_FP32 || mode == SDP_DFT_FP64_PHASE, "bad DFT mode %d", mode);
    if (n_vis == 0) return SDP_OK;
    hipStream_t st = sdp::stream_of(s);
    const unsigned int blocks = sdp_div_up(n_vis, kThreads);
    if (mode == SDP_DFT_REFERENCE_FP32)
        dft_ref_kernel<<<blocks, kThreads, 0, st>>>(n_vis, n_sources, d_sources, d_uvw, (float2 *)d_vis_out);
    else
        dft_f64_kernel<<<blocks, kThreads, 0, st>>>(n_vis, n_sources, d_sources, d_uvw, (float2 *)d_vis_out);
    SDP_LAUNCH_CHECK();
    return SDP_OK;
}

}  // extern "C"
