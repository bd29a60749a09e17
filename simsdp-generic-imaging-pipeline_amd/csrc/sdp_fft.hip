// sdp_fft.hip -- unnormalised 2D complex FFT of a G x G grid (G a power of
// two, 16 <= G <= 8192), the arithmetic of the FFTW calls in
// CUFFT_EXECUTE_FORWARD/INVERSE_C2C_actor (src/CPU/fft_run.c:23-73:
// FFTW_FORWARD = exp(-2 pi i jk/G), FFTW_BACKWARD = exp(+2 pi i jk/G), no
// scaling), with the neighbouring element-wise actors fused in.
//
// Three HBM passes, each an LDS-resident Stockham autosort FFT (radix 4,
// plus one radix-2 stage when log2 of the length is odd):
//   R : every row, length G                     in   -> work
//   C1: column sub-FFTs of length G1 over rows j1*G2 + j2, times the
//       four-step twiddle W_G^(j2 k1), in place on work
//   C2: column sub-FFTs of length G2 over rows k1*G2 + j2 (a contiguous
//       block), written to rows k1 + G1*k2      work -> out
// where G = G1 * G2.  Column passes move 16-32 adjacent columns per
// workgroup so every global access is a contiguous row segment.
// Fused on the way in (pass R): fft_shift_real_to_complex_actor /
// fft_shift_complex_to_complex_actor (fft_run.c:5-20,94-111); on the way out
// (pass C2): the checkerboard shift, fft_shift_complex_to_real_actor
// (fft_run.c:76-92) and execute_convolution_correction_actor
// (convolution_correction_run.c:17-31).
//
// Twiddles come from one table per G of exp(2 pi i m / G) computed in double
// precision on the host and rounded once to float.
//
// A batch of grids (consecutive G x G arrays) takes the same three launches,
// the grid index in blockIdx.z.
#include <algorithm>
#include <cmath>
#include <map>
#include <mutex>
#include <vector>

#include "sdp_internal.h"

namespace {

constexpr int kFftThreads = 256;

struct Line {
    int N, nlines, pstride, lstride;
};

__device__ __forceinline__ float2 cmulf(float2 a, float2 b)
{
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

__device__ __forceinline__ float2 tw_at(const float2 *tw, int m, float sign)
{
    const float2 t = tw[m];
    return make_float2(t.x, sign * t.y);
}

// In-LDS Stockham FFT of `nlines` lines of N points: point p of line l at
// buf[l * lstride + p * pstride].  tw has Gtw = N << tw_shift entries.
// Consecutive threads take consecutive points of one line when points are
// contiguous (rows) and the same point of consecutive lines otherwise
// (columns), so LDS accesses stay on distinct banks.
template <int MAXB>
__device__ void lds_fft(float2 *buf, Line L, float sign, const float2 *__restrict__ tw, int tw_shift)
{
    const int tid = threadIdx.x;
    const int N = L.N;
    int Ns = 1;
    if (__builtin_ctz(N) & 1) {   // one radix-2 stage
        const int nb = (N >> 1) * L.nlines;
        float2 v[MAXB][2];
        int base[MAXB], jj[MAXB];
#pragma unroll
        for (int q = 0; q < MAXB; ++q) {
            const int b = tid + q * kFftThreads;
            if (b < nb) {
                const int nbl = N >> 1;
                const int line = L.pstride == 1 ? b / nbl : b % L.nlines;
                const int j = L.pstride == 1 ? b - line * nbl : b / L.nlines;
                base[q] = line * L.lstride;
                jj[q] = j;
                v[q][0] = buf[base[q] + j * L.pstride];
                v[q][1] = buf[base[q] + (j + (N >> 1)) * L.pstride];
            }
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < MAXB; ++q) {
            const int b = tid + q * kFftThreads;
            if (b < nb) {
                const float2 a = v[q][0], c = v[q][1];
                buf[base[q] + (2 * jj[q]) * L.pstride] = make_float2(a.x + c.x, a.y + c.y);
                buf[base[q] + (2 * jj[q] + 1) * L.pstride] = make_float2(a.x - c.x, a.y - c.y);
            }
        }
        __syncthreads();
        Ns = 2;
    }
    for (; Ns < N; Ns <<= 2) {
        const int nb = (N >> 2) * L.nlines;
        const int quarter = N >> 2;
        float2 v[MAXB][4];
        int base[MAXB], dst[MAXB];
#pragma unroll
        for (int q = 0; q < MAXB; ++q) {
            const int b = tid + q * kFftThreads;
            if (b < nb) {
                const int line = L.pstride == 1 ? b / quarter : b % L.nlines;
                const int j = L.pstride == 1 ? b - line * quarter : b / L.nlines;
                base[q] = line * L.lstride;
#pragma unroll
                for (int r = 0; r < 4; ++r) v[q][r] = buf[base[q] + (j + r * quarter) * L.pstride];
                const int k = j & (Ns - 1);
                const int step = (k * (N / (4 * Ns))) << tw_shift;
                if (k) {
                    v[q][1] = cmulf(v[q][1], tw_at(tw, step, sign));
                    v[q][2] = cmulf(v[q][2], tw_at(tw, 2 * step, sign));
                    v[q][3] = cmulf(v[q][3], tw_at(tw, 3 * step, sign));
                }
                // 4-point DFT with kernel exp(sign * 2 pi i / 4)
                const float2 t0 = make_float2(v[q][0].x + v[q][2].x, v[q][0].y + v[q][2].y);
                const float2 t1 = make_float2(v[q][0].x - v[q][2].x, v[q][0].y - v[q][2].y);
                const float2 t2 = make_float2(v[q][1].x + v[q][3].x, v[q][1].y + v[q][3].y);
                const float2 d = make_float2(v[q][1].x - v[q][3].x, v[q][1].y - v[q][3].y);
                const float2 t3 = make_float2(-sign * d.y, sign * d.x);
                v[q][0] = make_float2(t0.x + t2.x, t0.y + t2.y);
                v[q][2] = make_float2(t0.x - t2.x, t0.y - t2.y);
                v[q][1] = make_float2(t1.x + t3.x, t1.y + t3.y);
                v[q][3] = make_float2(t1.x - t3.x, t1.y - t3.y);
                dst[q] = (j / Ns) * Ns * 4 + k;
            }
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < MAXB; ++q) {
            const int b = tid + q * kFftThreads;
            if (b < nb) {
#pragma unroll
                for (int r = 0; r < 4; ++r) buf[base[q] + (dst[q] + r * Ns) * L.pstride] = v[q][r];
            }
        }
        __syncthreads();
    }
}

__device__ __forceinline__ float checker(int r, int c) { return ((r + c) & 1) ? -1.0f : 1.0f; }

__device__ __forceinline__ float taper_divide(float x, const float *taper, int r, int c, int half)
{
    const int dc = abs(c - half), dr = abs(r - half);
    float t = taper[dc] * taper[dr];
    if (dc == 1229 || dr == 1229) t = 0.0f;
    return ((double)fabsf(t) > 1e-10) ? x / t : 0.0f;
}


// Pass R: rows_per_block rows of length G.
template <int MAXB>
__global__ void __launch_bounds__(kFftThreads)
fft_rows_kernel(const void *__restrict__ in, float2 *__restrict__ out, int G, int rows_per_block, int flags,
                float sign, const float2 *__restrict__ tw, const float *__restrict__ taper)
{
    extern __shared__ float2 lds[];
    const size_t gofs = (size_t)blockIdx.z * G * G;   // this grid's first element
    if (flags & SDP_FFT_REAL_IN)
        in = (const float *)in + gofs;
    else
        in = (const float2 *)in + gofs;
    out += gofs;
    const int row0 = blockIdx.x * rows_per_block;
    const int n = rows_per_block * G;
    for (int i = threadIdx.x; i < n; i += kFftThreads) {
        const int r = row0 + i / G, c = i % G;
        float2 v;
        if (flags & SDP_FFT_REAL_IN)
            v = make_float2(((const float *)in)[(size_t)r * G + c], 0.0f);
        else
            v = ((const float2 *)in)[(size_t)r * G + c];
        if (flags & SDP_FFT_CORRECT_IN) v.x = taper_divide(v.x, taper, r, c, G / 2);
        if (flags & SDP_FFT_SHIFT_IN) {
            const float a = checker(r, c);
            v.x *= a;
            if (!(flags & SDP_FFT_REAL_IN)) v.y *= a;
        }
        lds[i] = v;
    }
    __syncthreads();
    lds_fft<MAXB>(lds, Line{G, rows_per_block, 1, G}, sign, tw, 0);
    for (int i = threadIdx.x; i < n; i += kFftThreads) out[(size_t)row0 * G + i] = lds[i];
}

// Pass C1: columns [cb*CB, cb*CB + CB), rows j1*G2 + j2, in place, then
// times W_G^(sign j2 k1).
template <int MAXB>
__global__ void __launch_bounds__(kFftThreads)
fft_cols1_kernel(float2 *__restrict__ data, int G, int G1, int G2, int CB, int tw_shift, float sign,
                 const float2 *__restrict__ tw)
{
    extern __shared__ float2 lds[];
    data += (size_t)blockIdx.z * G * G;
    const int c0 = blockIdx.x * CB, j2 = blockIdx.y;
    const int n = G1 * CB;
    for (int i = threadIdx.x; i < n; i += kFftThreads) {
        const int j1 = i / CB, c = i % CB;
        lds[i] = data[(size_t)(j1 * G2 + j2) * G + c0 + c];
    }
    __syncthreads();
    lds_fft<MAXB>(lds, Line{G1, CB, CB, 1}, sign, tw, tw_shift);
    for (int i = threadIdx.x; i < n; i += kFftThreads) {
        const int k1 = i / CB, c = i % CB;
        float2 v = lds[i];
        const int m = (int)(((long long)j2 * k1) % G);
        if (m) v = cmulf(v, tw_at(tw, m, sign));
        data[(size_t)(k1 * G2 + j2) * G + c0 + c] = v;
    }
}

struct OutArgs {
    int flags;
    const float *taper;   // G/2+1 entries, last one 0
};

// Pass C2: rows k1*G2 + j2 (j2 = 0..G2-1) -> output rows k1 + G1*k2.
template <int MAXB>
__global__ void __launch_bounds__(kFftThreads)
fft_cols2_kernel(const float2 *__restrict__ data, void *__restrict__ out, int G, int G1, int G2, int CB, int tw_shift,
                 float sign, const float2 *__restrict__ tw, OutArgs o)
{
    extern __shared__ float2 lds[];
    const size_t gofs = (size_t)blockIdx.z * G * G;
    data += gofs;
    if (o.flags & SDP_FFT_REAL_OUT)
        out = (float *)out + gofs;
    else
        out = (float2 *)out + gofs;
    const int c0 = blockIdx.x * CB, k1 = blockIdx.y;
    const int n = G2 * CB;
    for (int i = threadIdx.x; i < n; i += kFftThreads) {
        const int j2 = i / CB, c = i % CB;
        lds[i] = data[(size_t)(k1 * G2 + j2) * G + c0 + c];
    }
    __syncthreads();
    lds_fft<MAXB>(lds, Line{G2, CB, CB, 1}, sign, tw, tw_shift);
    const int half = G / 2;
    for (int i = threadIdx.x; i < n; i += kFftThreads) {
        const int k2 = i / CB, c = c0 + i % CB;
        const int r = k1 + G1 * k2;
        float2 v = lds[i];
        if (o.flags & SDP_FFT_SHIFT_OUT) {
            const float a = checker(r, c);
            v.x *= a;
            v.y *= a;
        }
        if (o.flags & SDP_FFT_REAL_OUT) {
            float x = v.x;
            if (o.flags & SDP_FFT_CORRECT_OUT) x = taper_divide(x, o.taper, r, c, half);
            ((float *)out)[(size_t)r * G + c] = x;
        } else {
            ((float2 *)out)[(size_t)r * G + c] = v;
        }
    }
}

std::mutex g_tw_mutex;
std::map<int, float2 *> g_twiddles;   // per G, lives until process exit

int twiddles_for(int G, const float2 **out)
{
    std::lock_guard<std::mutex> lock(g_tw_mutex);
    auto it = g_twiddles.find(G);
    if (it != g_twiddles.end()) {
        *out = it->second;
        return SDP_OK;
    }
    std::vector<float2> h(G);
    for (int m = 0; m < G; ++m) {
        const double a = 2.0 * M_PI * (double)m / (double)G;
        h[m] = make_float2((float)cos(a), (float)sin(a));
    }
    float2 *d = nullptr;
    SDP_HIP_TRY(hipMalloc(&d, sizeof(float2) * G));
    SDP_HIP_TRY(hipMemcpy(d, h.data(), sizeof(float2) * G, hipMemcpyHostToDevice));
    g_twiddles[G] = d;
    *out = d;
    return SDP_OK;
}

int ilog2(int x) { return 31 - __builtin_clz((unsigned)x); }

// largest number of butterflies one thread holds in a stage of an N-point
// FFT over `lines` lines (the radix-2 stage, when present, is the widest)
int butterflies_per_thread(int N, int lines)
{
    const int nb = ((ilog2(N) & 1) ? N / 2 : N / 4) * lines;
    return (nb + kFftThreads - 1) / kFftThreads;
}

template <typename K>
void allow_lds(K kernel, size_t bytes)
{
    if (bytes > 65536) (void)hipFuncSetAttribute((const void *)kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
}

}  // namespace

extern "C" {

size_t sdp_fft_work_bytes(int G) { return (G > 0) ? sizeof(float2) * (size_t)G * G : 0; }

int sdp_fft2d(const void *d_in, void *d_out, int G, int direction, int flags, const float *d_taper, void *d_work,
              sdp_stream s)
{
    return sdp_fft2d_batch(d_in, d_out, G, 1, direction, flags, d_taper, d_work, s);
}

int sdp_fft2d_batch(const void *d_in, void *d_out, int G, int n_batch, int direction, int flags,
                    const float *d_taper, void *d_work, sdp_stream s)
{
    SDP_REQUIRE(G >= 16 && G <= 8192 && (G & (G - 1)) == 0, "FFT size %d unsupported (power of two in [16, 8192])", G);
    SDP_REQUIRE(n_batch >= 1 && n_batch <= 65535, "FFT batch of %d grids unsupported (1 to 65535)", n_batch);
    SDP_REQUIRE(direction == SDP_FFT_FORWARD || direction == SDP_FFT_INVERSE, "bad FFT direction");
    SDP_REQUIRE(d_in && d_out && d_work, "sdp_fft2d: null pointer");
    SDP_REQUIRE(!(flags & SDP_FFT_CORRECT_OUT) || ((flags & SDP_FFT_REAL_OUT) && d_taper),
                "CORRECT_OUT needs REAL_OUT and a taper");
    SDP_REQUIRE(!(flags & SDP_FFT_CORRECT_IN) || ((flags & SDP_FFT_REAL_IN) && d_taper),
                "CORRECT_IN needs REAL_IN and a taper");
    SDP_REQUIRE(d_work != d_in && d_work != d_out, "work buffer must not alias input or output");
    const float2 *tw = nullptr;
    if (int e = twiddles_for(G, &tw)) return e;
    hipStream_t st = sdp::stream_of(s);
    const float sign = (float)direction;
    float2 *work = (float2 *)d_work;

    const int lg = ilog2(G);
    // pass R
    {
        const int rpb = G >= 2048 ? 1 : std::min(G, 2048 / G);   // rows per workgroup
        const size_t lds = sizeof(float2) * (size_t)rpb * G;
        const int maxb = butterflies_per_thread(G, rpb);
        dim3 grid(G / rpb, 1, n_batch);
        if (maxb <= 4) {
            fft_rows_kernel<4><<<grid, kFftThreads, lds, st>>>(d_in, work, G, rpb, flags, sign, tw, d_taper);
        } else if (maxb <= 8) {
            allow_lds(fft_rows_kernel<8>, lds);
            fft_rows_kernel<8><<<grid, kFftThreads, lds, st>>>(d_in, work, G, rpb, flags, sign, tw, d_taper);
        } else {
            allow_lds(fft_rows_kernel<16>, lds);
            fft_rows_kernel<16><<<grid, kFftThreads, lds, st>>>(d_in, work, G, rpb, flags, sign, tw, d_taper);
        }
        SDP_LAUNCH_CHECK();
    }
    // column passes
    const int G1 = 1 << ((lg + 1) / 2), G2 = G / G1;
    const int CB = G >= 32 ? 32 : G;
    {
        const size_t lds = sizeof(float2) * (size_t)G1 * CB;
        dim3 grid(G / CB, G2, n_batch);
        if (butterflies_per_thread(G1, CB) <= 4)
            fft_cols1_kernel<4><<<grid, kFftThreads, lds, st>>>(work, G, G1, G2, CB, ilog2(G / G1), sign, tw);
        else
            fft_cols1_kernel<8><<<grid, kFftThreads, lds, st>>>(work, G, G1, G2, CB, ilog2(G / G1), sign, tw);
        SDP_LAUNCH_CHECK();
    }
    {
        const size_t lds = sizeof(float2) * (size_t)G2 * CB;
        dim3 grid(G / CB, G1, n_batch);
        OutArgs o{flags, d_taper};
        if (butterflies_per_thread(G2, CB) <= 4)
            fft_cols2_kernel<4><<<grid, kFftThreads, lds, st>>>(work, d_out, G, G1, G2, CB, ilog2(G / G2), sign, tw, o);
        else
            fft_cols2_kernel<8><<<grid, kFftThreads, lds, st>>>(work, d_out, G, G1, G2, CB, ilog2(G / G2), sign, tw, o);
        SDP_LAUNCH_CHECK();
    }
    return SDP_OK;
}

}  // extern "C"
