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

}  // namespace

extern "C" {

int sdp_dft_predict(int n_vis, int n_sources, const float *d_sources, const float *d_uvw, float *d_vis_out, int mode,
                    sdp_stream s)
{
    SDP_REQUIRE(n_vis >= 0 && n_sources >= 0, "negative counts");
    SDP_REQUIRE(mode == SDP_DFT_REFERENCE_FP32 || mode == SDP_DFT_FP64_PHASE || mode == SDP_DFT_MFMA,
                "bad DFT mode %d", mode);
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
