// sdp_fft.hip -- unnormalised 2D complex FFT of a G x G grid (G a power of
// two, 16 <= G <= 8192), the arithmetic of the FFTW calls in
// CUFFT_EXECUTE_FORWARD/INVERSE_C2C_actor (src/CPU/fft_run.c:23-73:
// FFTW_FORWARD = exp(-2 pi i jk/G), FFTW_BACKWARD = exp(+2 pi i jk/G), no
// scaling), with the neighbouring element-wise actors fused in.
//
// 256 <= G <= 4096: two HBM passes (rows, then columns), each line held in
// registers and transformed in radix-16 stages with LDS exchanges between
// them (the two-pass form below).
//
// Other sizes (and SDP_FFT_THREE_PASS=1): three HBM passes, each an
// LDS-resident Stockham autosort FFT (radix 4, plus one radix-2 stage when
// log2 of the length is odd):
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
// A batch of grids (consecutive G x G arrays) takes the same launches, the
// grid index in blockIdx.z.
#include <algorithm>
#include <cmath>
#include <cstdlib>
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

// ---------------------------------------------------------------------------
// Two-pass form, 256 <= G <= 4096: pass A transforms every row (in -> work),
// pass B every column (work -> out), each reading and writing the grid once.
//
// A line (row or column) of N points is held in registers by T = N/16
// threads, thread j holding points j + m*T (m = 0..15).  Three Stockham
// stages: radix 16 (Ns = 1), radix 16 (Ns = 16), radix N/256 (Ns = 256;
// none for N = 256).  A stage of radix r takes inputs x[b + m*N/r] of
// butterfly b, multiplies input m by W_(r*Ns)^(m*(b mod Ns)), and writes
// output m to (b / Ns)*r*Ns + (b mod Ns) + m*Ns.  Between stages the points
// pass through LDS (point i of a line at i + i/16: the 16-point groups that
// one thread writes start on different banks).  After the last stage thread j
// holds outputs j + m*T, the natural order, so loads and stores are direct.
// Column blocks take 8 (G >= 1024) or 16 adjacent columns; each thread owns
// two adjacent columns (16-byte loads and stores), transformed one after the
// other through the same LDS lines.

constexpr float kCos1 = 0.92387953251128674f;   // cos(pi/8)
constexpr float kSin1 = 0.38268343236508977f;   // sin(pi/8)
constexpr float kHalfSqrt2 = 0.70710678118654752f;

__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }

// S*i*z for S = +1 or -1
template <int S>
__device__ __forceinline__ float2 mul_si(float2 z)
{
    return S > 0 ? make_float2(-z.y, z.x) : make_float2(z.y, -z.x);
}

// z * exp(S * 2 pi i P / 16)
template <int S, int P>
__device__ __forceinline__ float2 rot16(float2 z)
{
    if constexpr (P == 0) {
        return z;
    } else if constexpr (P == 4) {
        return mul_si<S>(z);
    } else if constexpr (P == 2) {   // (1 + S i) / sqrt 2
        return make_float2(kHalfSqrt2 * (z.x - S * z.y), kHalfSqrt2 * (z.y + S * z.x));
    } else if constexpr (P == 6) {   // (-1 + S i) / sqrt 2
        return make_float2(-kHalfSqrt2 * (z.x + S * z.y), kHalfSqrt2 * (S * z.x - z.y));
    } else {
        static_assert(P == 1 || P == 3 || P == 9, "rot16: unused angle");
        constexpr float c = P == 1 ? kCos1 : P == 3 ? kSin1 : -kCos1;
        constexpr float s = S * (P == 1 ? kSin1 : P == 3 ? kCos1 : -kSin1);
        return make_float2(z.x * c - z.y * s, z.x * s + z.y * c);
    }
}

template <int S>
__device__ __forceinline__ void dft2(float2 &a, float2 &b)
{
    const float2 t = a;
    a = cadd(t, b);
    b = csub(t, b);
}

// y_k = sum_n x_n exp(S 2 pi i n k / 4), in place
template <int S>
__device__ __forceinline__ void dft4(float2 &a, float2 &b, float2 &c, float2 &d)
{
    const float2 t0 = cadd(a, c), t1 = csub(a, c), t2 = cadd(b, d), t3 = mul_si<S>(csub(b, d));
    a = cadd(t0, t2);
    c = csub(t0, t2);
    b = cadd(t1, t3);
    d = csub(t1, t3);
}

// 8 points: n = n1 + 2 n2, k = k1 + 4 k2
template <int S>
__device__ __forceinline__ void dft8(float2 (&v)[8])
{
    dft4<S>(v[0], v[2], v[4], v[6]);
    dft4<S>(v[1], v[3], v[5], v[7]);
    v[3] = rot16<S, 2>(v[3]);
    v[5] = rot16<S, 4>(v[5]);
    v[7] = rot16<S, 6>(v[7]);
    float2 y[8];
#pragma unroll
    for (int k1 = 0; k1 < 4; ++k1) {
        y[k1] = cadd(v[2 * k1], v[2 * k1 + 1]);
        y[k1 + 4] = csub(v[2 * k1], v[2 * k1 + 1]);
    }
#pragma unroll
    for (int k = 0; k < 8; ++k) v[k] = y[k];
}

// 16 points: n = n1 + 4 n2, k = k1 + 4 k2
template <int S>
__device__ __forceinline__ void dft16(float2 (&e)[16])
{
#pragma unroll
    for (int n1 = 0; n1 < 4; ++n1) dft4<S>(e[n1], e[n1 + 4], e[n1 + 8], e[n1 + 12]);
    // e[n1 + 4 k1] times W16^(n1 k1)
    e[5] = rot16<S, 1>(e[5]);
    e[9] = rot16<S, 2>(e[9]);
    e[13] = rot16<S, 3>(e[13]);
    e[6] = rot16<S, 2>(e[6]);
    e[10] = rot16<S, 4>(e[10]);
    e[14] = rot16<S, 6>(e[14]);
    e[7] = rot16<S, 3>(e[7]);
    e[11] = rot16<S, 6>(e[11]);
    e[15] = rot16<S, 9>(e[15]);
#pragma unroll
    for (int k1 = 0; k1 < 4; ++k1) dft4<S>(e[4 * k1], e[4 * k1 + 1], e[4 * k1 + 2], e[4 * k1 + 3]);
    float2 y[16];
#pragma unroll
    for (int k1 = 0; k1 < 4; ++k1)
#pragma unroll
        for (int k2 = 0; k2 < 4; ++k2) y[k1 + 4 * k2] = e[4 * k1 + k2];
#pragma unroll
    for (int k = 0; k < 16; ++k) e[k] = y[k];
}

// W_N^m with the transform's sign (the table holds exp(+2 pi i m / N))
template <int S>
__device__ __forceinline__ float2 twiddle(const float2 *__restrict__ tw, int m)
{
    const float2 t = tw[m];
    return S > 0 ? t : make_float2(t.x, -t.y);
}

// e[m] *= W_N^(m * step), m = 1..15, the table read four points at a time:
// the scheduling barriers keep the compiler from hoisting all fifteen table
// reads (30 registers) ahead of the multiplications
// x, with the compiler unable to see that it equals an earlier value: stops
// it from keeping offsets alive from one use to the next (in registers the
// kernel needs for points) instead of recomputing them
__device__ __forceinline__ unsigned opaque(unsigned x)
{
    asm volatile("" : "+v"(x));
    return x;
}

template <int S>
__device__ __forceinline__ void twiddle_points(float2 (&e)[16], const float2 *__restrict__ tw, int step)
{
    step = (int)opaque((unsigned)step);
#pragma unroll
    for (int m0 = 1; m0 < 16; m0 += 4) {
        float2 w[4];
#pragma unroll
        for (int i = 0; i < 4; ++i)
            if (m0 + i < 16) w[i] = twiddle<S>(tw, (m0 + i) * step);
#pragma unroll
        for (int i = 0; i < 4; ++i)
            if (m0 + i < 16) e[m0 + i] = cmulf(e[m0 + i], w[i]);
        __builtin_amdgcn_sched_barrier(0);
    }
}

__device__ __forceinline__ int lds_at(int i) { return i + (i >> 4); }

template <int N>
struct Fft2Shape {
    static constexpr int T = N / 16;                 // threads per line
    static constexpr int R3 = N / 256;               // radix of the last stage (1: none)
    static constexpr int LS = N + N / 16 + 8;        // LDS line stride (float2)
};

// The line transform of the thread's 16 points; `line` is its LDS line.
// Every thread of the workgroup calls it the same number of times.  On
// return the thread's LDS reads are complete only after a barrier.
template <int N, int S>
__device__ __forceinline__ void line_fft(float2 (&e)[16], float2 *line, int j, const float2 *__restrict__ tw)
{
    constexpr int T = Fft2Shape<N>::T, R3 = Fft2Shape<N>::R3;
    dft16<S>(e);
#pragma unroll
    for (int m = 0; m < 16; ++m) line[lds_at(16 * j + m)] = e[m];
    __syncthreads();
#pragma unroll
    for (int m = 0; m < 16; ++m) e[m] = line[lds_at(j + m * T)];
    const int k = j & 15;
    twiddle_points<S>(e, tw, k * (N / 256));
    dft16<S>(e);
    if constexpr (R3 > 1) {
        __syncthreads();
#pragma unroll
        for (int m = 0; m < 16; ++m) line[lds_at((j >> 4) * 256 + k + 16 * m)] = e[m];
        __syncthreads();
#pragma unroll
        for (int m = 0; m < 16; ++m) e[m] = line[lds_at(j + m * T)];
        if constexpr (R3 == 16) {
            twiddle_points<S>(e, tw, j);
            dft16<S>(e);
        } else if constexpr (R3 == 8) {
#pragma unroll
            for (int q = 0; q < 2; ++q) {
                const int b = j + q * T;
                float2 v[8];
#pragma unroll
                for (int m = 0; m < 8; ++m) v[m] = m ? cmulf(e[q + 2 * m], twiddle<S>(tw, m * b)) : e[q];
                dft8<S>(v);
#pragma unroll
                for (int m = 0; m < 8; ++m) e[q + 2 * m] = v[m];
            }
        } else if constexpr (R3 == 4) {
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int b = j + q * T;
#pragma unroll
                for (int m = 1; m < 4; ++m) e[q + 4 * m] = cmulf(e[q + 4 * m], twiddle<S>(tw, m * b));
                dft4<S>(e[q], e[q + 4], e[q + 8], e[q + 12]);
            }
        } else {
            static_assert(R3 == 2, "line_fft: N must be 256 * (1, 2, 4, 8 or 16)");
#pragma unroll
            for (int q = 0; q < 8; ++q) {
                e[q + 8] = cmulf(e[q + 8], twiddle<S>(tw, j + q * T));
                dft2<S>(e[q], e[q + 8]);
            }
        }
    }
}

// Pass A: 4096 / N rows per workgroup of 256 threads.
template <int N, int S>
__global__ void __launch_bounds__(256)
fft2_rows_kernel(const void *__restrict__ in, float2 *__restrict__ out, int flags, const float2 *__restrict__ tw,
                 const float *__restrict__ taper)
{
    constexpr int T = Fft2Shape<N>::T, LS = Fft2Shape<N>::LS, ROWS = 256 / T;
    __shared__ float2 lds[ROWS * LS];
    const int line = threadIdx.x / T, j = threadIdx.x % T;
    const int r = blockIdx.x * ROWS + line;
    const size_t row_ofs = (size_t)blockIdx.z * N * N + (size_t)r * N;
    float2 e[16];
    if (flags & SDP_FFT_REAL_IN) {
        const float *p = (const float *)in + row_ofs;
#pragma unroll
        for (int m = 0; m < 16; ++m) e[m] = make_float2(p[j + m * T], 0.0f);
        if (flags & SDP_FFT_CORRECT_IN) {
#pragma unroll
            for (int m = 0; m < 16; ++m) e[m].x = taper_divide(e[m].x, taper, r, j + m * T, N / 2);
        }
        if ((flags & SDP_FFT_SHIFT_IN) && ((r + j) & 1)) {   // T is even: column parity is j's
#pragma unroll
            for (int m = 0; m < 16; ++m) e[m].x = -e[m].x;
        }
    } else {
        const float2 *p = (const float2 *)in + row_ofs;
#pragma unroll
        for (int m = 0; m < 16; ++m) e[m] = p[j + m * T];
        if ((flags & SDP_FFT_SHIFT_IN) && ((r + j) & 1)) {
#pragma unroll
            for (int m = 0; m < 16; ++m) e[m] = make_float2(-e[m].x, -e[m].y);
        }
    }
    line_fft<N, S>(e, lds + line * LS, j, tw);
    float2 *o = out + row_ofs;
#pragma unroll
    for (int m = 0; m < 16; ++m) o[j + m * T] = e[m];
}

// Pass B: CB adjacent columns per workgroup, two per thread.
template <int N, int CB, int S>
__global__ void __launch_bounds__(CB / 2 * (N / 16))
fft2_cols_kernel(const float2 *__restrict__ in, void *__restrict__ out, int flags, const float2 *__restrict__ tw,
                 const float *__restrict__ taper)
{
    constexpr int T = Fft2Shape<N>::T, LS = Fft2Shape<N>::LS, NL = CB / 2;
    extern __shared__ float2 lds[];
    const int cq = threadIdx.x % NL, j = threadIdx.x / NL;
    // Column blocks 2i and 2i + 1 share 128-byte lines when CB = 8: place
    // them on one XCD (workgroups go to XCDs round robin) in consecutive slots.
    int cb = blockIdx.x;
    if (CB == 8 && (gridDim.x & 15) == 0) {
        const int xcd = cb & 7, slot = cb >> 3;
        cb = ((slot >> 1) << 4) | (xcd << 1) | (slot & 1);
    }
    const int c = cb * CB + 2 * cq;
    const size_t gofs = (size_t)blockIdx.z * N * N;
    // point m of the thread's columns: element (j + m T) * N + c of the grid
    const char *src = (const char *)(in + gofs);
    const unsigned off0 = (unsigned)(j * N + c) * 8u;
    constexpr unsigned kStep = (unsigned)T * N * 8u;
    float2 e[2][16];
#pragma unroll
    for (int m = 0; m < 16; ++m) {
        const float4 v = *(const float4 *)(src + (off0 + m * kStep));
        e[0][m] = make_float2(v.x, v.y);
        e[1][m] = make_float2(v.z, v.w);
    }
    // one column at a time (scheduling barriers: the second column's
    // arithmetic is not interleaved with the first's, which would double the
    // live registers)
    __builtin_amdgcn_sched_barrier(0);
    line_fft<N, S>(e[0], lds + cq * LS, j, tw);
    __syncthreads();
    __builtin_amdgcn_sched_barrier(0);
    line_fft<N, S>(e[1], lds + cq * LS, j, tw);
    __builtin_amdgcn_sched_barrier(0);
    // output row r = j + m T (T even: row parity is j's), columns c, c + 1
    const bool flip = (flags & SDP_FFT_SHIFT_OUT) && ((j + c) & 1);   // c is even: column c + 1 flips the other way
    const float a0 = flip ? -1.0f : 1.0f;
    const float a1 = (flags & SDP_FFT_SHIFT_OUT) ? -a0 : 1.0f;
    const unsigned idx0 = opaque((unsigned)(j * N + c));
    if (flags & SDP_FFT_REAL_OUT) {
        char *dst = (char *)((float *)out + gofs);
#pragma unroll
        for (int m = 0; m < 16; ++m) {
            float x0 = e[0][m].x * a0, x1 = e[1][m].x * a1;
            if (flags & SDP_FFT_CORRECT_OUT) {
                const int r = j + m * T;
                x0 = taper_divide(x0, taper, r, c, N / 2);
                x1 = taper_divide(x1, taper, r, c + 1, N / 2);
            }
            *(float2 *)(dst + (idx0 * 4u + m * (kStep / 2))) = make_float2(x0, x1);
        }
    } else {
        char *dst = (char *)((float2 *)out + gofs);
#pragma unroll
        for (int m = 0; m < 16; ++m)
            *(float4 *)(dst + (idx0 * 8u + m * kStep)) =
                make_float4(e[0][m].x * a0, e[0][m].y * a0, e[1][m].x * a1, e[1][m].y * a1);
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

template <int N, int S>
int launch_two_pass(const void *in, void *out, float2 *work, int nb, int flags, const float2 *tw, const float *taper,
                    hipStream_t st)
{
    constexpr int ROWS = 4096 / N;   // rows per 256-thread workgroup
    fft2_rows_kernel<N, S><<<dim3(N / ROWS, 1, nb), 256, 0, st>>>(in, work, flags, tw, taper);
    SDP_LAUNCH_CHECK();
    // 16 columns (whole 128-byte lines) once there are at least 512 column
    // blocks of 16 to spread over the CUs; 8 otherwise (measured: 64 grids
    // of 2048^2 1.16 ms against 1.26 ms per column pass, one grid 19 us
    // against 31 us).  SDP_FFT_COL_BLOCK=8 or 16 overrides.  The per-column
    // arithmetic is the same either way.
    static const int cb_env = [] {
        const char *e = getenv("SDP_FFT_COL_BLOCK");
        return e ? atoi(e) : 0;
    }();
    if constexpr (N <= 2048) {
        if (cb_env == 16 || (cb_env == 0 && (N < 1024 || (long)(N / 16) * nb >= 512))) {
            constexpr size_t lds = sizeof(float2) * 8 * Fft2Shape<N>::LS;
            allow_lds(fft2_cols_kernel<N, 16, S>, lds);
            fft2_cols_kernel<N, 16, S><<<dim3(N / 16, 1, nb), 8 * (N / 16), lds, st>>>(work, out, flags, tw, taper);
            SDP_LAUNCH_CHECK();
            return SDP_OK;
        }
    }
    constexpr size_t lds = sizeof(float2) * 4 * Fft2Shape<N>::LS;
    allow_lds(fft2_cols_kernel<N, 8, S>, lds);
    fft2_cols_kernel<N, 8, S><<<dim3(N / 8, 1, nb), 4 * (N / 16), lds, st>>>(work, out, flags, tw, taper);
    SDP_LAUNCH_CHECK();
    return SDP_OK;
}

template <int S>
int two_pass(int G, const void *in, void *out, float2 *work, int nb, int flags, const float2 *tw, const float *taper,
             hipStream_t st)
{
    switch (G) {
    case 256: return launch_two_pass<256, S>(in, out, work, nb, flags, tw, taper, st);
    case 512: return launch_two_pass<512, S>(in, out, work, nb, flags, tw, taper, st);
    case 1024: return launch_two_pass<1024, S>(in, out, work, nb, flags, tw, taper, st);
    case 2048: return launch_two_pass<2048, S>(in, out, work, nb, flags, tw, taper, st);
    default: return launch_two_pass<4096, S>(in, out, work, nb, flags, tw, taper, st);
    }
}

// SDP_FFT_THREE_PASS=1 selects the three-pass form for every size (comparison)
bool three_pass_forced()
{
    static const bool forced = [] {
        const char *e = getenv("SDP_FFT_THREE_PASS");
        return e && atoi(e) != 0;
    }();
    return forced;
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

    if (G >= 256 && G <= 4096 && !three_pass_forced())
        return direction > 0 ? two_pass<1>(G, d_in, d_out, work, n_batch, flags, tw, d_taper, st)
                             : two_pass<-1>(G, d_in, d_out, work, n_batch, flags, tw, d_taper, st);
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
