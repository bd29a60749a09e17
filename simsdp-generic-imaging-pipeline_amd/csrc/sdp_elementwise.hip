// sdp_elementwise.hip -- HBM-bound element-wise actors: checkerboard
// shifts, gridding correction, visibility arithmetic and the slice merge.
//
// All kernels use 16-byte accesses and grid-stride loops.  The library is
// built with -ffp-contract=off, so every expression rounds exactly as the
// reference's C does (bit-identical outputs).
#include <algorithm>
#include <cmath>
#include <vector>

#include "sdp_internal.h"

namespace {

constexpr int kBlock = 256;

unsigned int grid_for(size_t n)
{
    size_t b = (n + kBlock - 1) / kBlock;
    if (b > 8192) b = 8192;
    return b ? (unsigned int)b : 1u;
}

// (-1)^(r+c) for an even grid: r*G + c has the parity of r + c only if G is
// even, which every supported G is.
__device__ __forceinline__ float checker(size_t row, size_t col) { return ((row + col) & 1) ? -1.0f : 1.0f; }

// fft_shift_complex_to_complex_actor (fft_run.c:5-20): 2 cells per thread.
__global__ void shift_c2c_kernel(const float4 *in, float4 *out, int G)
{
    const size_t n2 = (size_t)G * G / 2;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n2; i += (size_t)gridDim.x * blockDim.x) {
        size_t cell = 2 * i, row = cell / G, col = cell % G;
        float a = checker(row, col);
        float4 v = in[i];
        v.x *= a; v.y *= a; v.z *= -a; v.w *= -a;
        out[i] = v;
    }
}

// fft_shift_complex_to_real_actor (fft_run.c:76-92)
__global__ void shift_c2r_kernel(const float4 *in, float2 *out, int G)
{
    const size_t n2 = (size_t)G * G / 2;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n2; i += (size_t)gridDim.x * blockDim.x) {
        size_t cell = 2 * i, row = cell / G, col = cell % G;
        float a = checker(row, col);
        float4 v = in[i];
        out[i] = make_float2(v.x * a, v.z * -a);
    }
}

// fft_shift_real_to_complex_actor (fft_run.c:94-111)
__global__ void shift_r2c_kernel(const float2 *in, float4 *out, int G)
{
    const size_t n2 = (size_t)G * G / 2;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n2; i += (size_t)gridDim.x * blockDim.x) {
        size_t cell = 2 * i, row = cell / G, col = cell % G;
        float a = checker(row, col);
        float2 v = in[i];
        out[i] = make_float4(v.x * a, 0.0f, v.y * -a, 0.0f);
    }
}

// execute_convolution_correction_actor (convolution_correction_run.c:17-31).
// taper has G/2+1 entries, taper[G/2] == 0 stands for the element the
// reference reads past the end for row or column 0.
__device__ __forceinline__ float correct_one(float x, const float *taper, int r, int c, int half)
{
    int dc = abs(c - half), dr = abs(r - half);
    float t = taper[dc] * taper[dr];
    if (dc == 1229 || dr == 1229) t = 0.0f;
    return ((double)fabsf(t) > 1e-10) ? x / t : 0.0f;
}

__global__ void correct_kernel(const float2 *in, const float *taper, float2 *out, int G)
{
    const size_t n2 = (size_t)G * G / 2;
    const int half = G / 2;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n2; i += (size_t)gridDim.x * blockDim.x) {
        size_t cell = 2 * i;
        int r = (int)(cell / G), c = (int)(cell % G);
        float2 v = in[i];
        out[i] = make_float2(correct_one(v.x, taper, r, c, half), correct_one(v.y, taper, r, c + 1, half));
    }
}

__global__ void vis_sub_kernel(const float2 *a, const float2 *b, float2 *out, int n)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        float2 x = a[i], y = b[i];
        out[i] = make_float2(x.x - y.x, x.y - y.y);
    }
}

__global__ void vis_add_kernel(const float2 *a, const float2 *b, float2 *out, int n)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        float2 x = a[i], y = b[i];
        out[i] = make_float2(x.x + y.x, x.y + y.y);
    }
}

// complex_multiply_apply (gains_apply_run.c:146-149)
__device__ __forceinline__ float2 cmul_apply(float2 a, float2 b)
{
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// reciprocal_transform_actor (gains_apply_run.c:4-17): out = conj(z) / |z|^2
// = 1 / z, in the reference's float order (|z|^2 formed for each part)
__global__ void reciprocal_kernel(const float2 *in, int n, float2 *out)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float2 z = in[i];
    const float re = z.x / (z.x * z.x + z.y * z.y);
    const float im = z.y / (z.x * z.x + z.y * z.y);
    out[i] = make_float2(re, -im);
}

// conv2D (major_loop_iter.c:8-46), the restored image: every pixel of the
// model at or above 1e-5 spreads model * filter over the [-h, h) box around
// it, the filter read flipped about (D/2 + 1, D/2 + 1).  Gathered here: an
// output pixel adds its sources in their raster order, which is the order
// in which the reference's scatter adds them, so each sum rounds the same.
__global__ void restore_conv2d_kernel(const float *image, const float *filter, int D, int hx, int hy, float *out)
{
    const int tx = blockIdx.x * blockDim.x + threadIdx.x, ty = blockIdx.y;
    if (tx >= D) return;
    const int center = D / 2 + 1;
    float acc = 0.f;
    for (int y = max(ty - hy + 1, 0); y <= min(ty + hy, D - 1); ++y) {
        const int fy = ty - y;                               // in [-hy, hy)
        const float *frow = filter + (size_t)(center - fy) * D + center;
        const float *irow = image + (size_t)y * D;
        for (int x = max(tx - hx + 1, 0); x <= min(tx + hx, D - 1); ++x) {
            const float v = irow[x];
            // the reference's float < 1e-5 (a double): v <= float(1e-5), which
            // lies below 1e-5, i.e. v < the next float 0x1.4f8b5ap-17
            if (v < 0x1.4f8b5ap-17f) continue;
            acc += v * frow[-(tx - x)];
        }
    }
    out[(size_t)ty * D + tx] = acc;
}

// subtract_from_measurements / apply_gains_only (gains_apply_run.c:79-92,105-115)
__global__ void gains_kernel(const float2 *measured, const float2 *pred, const float2 *gains, const int2 *pairs,
                             int n, int n_baselines, int apply_only, float2 *out)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        int b = i % n_baselines;
        int2 p = pairs[b];
        float2 ga = gains[p.x], gb = gains[p.y];
        gb.y = -gb.y;
        float2 m = cmul_apply(measured[i], cmul_apply(ga, gb));
        if (!apply_only) {
            float2 q = pred[i];
            m = make_float2(m.x - q.x, m.y - q.y);
        }
        out[i] = m;
    }
}

// merge_gridded_visibilities (splitmerge.c:14-24): slice-ordered float sum.
__global__ void merge_kernel(const float4 *grids, float4 *out, size_t n2, int slices)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n2; i += (size_t)gridDim.x * blockDim.x) {
        float4 s = make_float4(0.f, 0.f, 0.f, 0.f);
        for (int j = 0; j < slices; ++j) {
            float4 g = grids[(size_t)j * n2 + i];
            s.x += g.x; s.y += g.y; s.z += g.z; s.w += g.w;
        }
        out[i] = s;
    }
}

// sum of squares and largest value / |value| of an image, one partial per block
__global__ void image_stats_kernel(const float *img, size_t n, double *part)
{
    __shared__ double s_sq[256];
    __shared__ float s_max[256], s_amax[256];
    double sq = 0.0;
    float mx = -INFINITY, amx = 0.f;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const float v = img[i];
        sq += (double)v * (double)v;
        mx = fmaxf(mx, v);
        amx = fmaxf(amx, fabsf(v));
    }
    s_sq[threadIdx.x] = sq;
    s_max[threadIdx.x] = mx;
    s_amax[threadIdx.x] = amx;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if (threadIdx.x < off) {
            s_sq[threadIdx.x] += s_sq[threadIdx.x + off];
            s_max[threadIdx.x] = fmaxf(s_max[threadIdx.x], s_max[threadIdx.x + off]);
            s_amax[threadIdx.x] = fmaxf(s_amax[threadIdx.x], s_amax[threadIdx.x + off]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part[3 * blockIdx.x] = s_sq[0];
        part[3 * blockIdx.x + 1] = s_max[0];
        part[3 * blockIdx.x + 2] = s_amax[0];
    }
}

// combines the per-block partials: out = {sum_sq, max, max_abs}.  One
// workgroup of 256 threads: thread t sums partials t, t + 256, ... in that
// order, then a fixed tree over the threads (the same order every call)
__global__ void __launch_bounds__(256) image_stats_final_kernel(const double *part, int nparts, double *out)
{
    __shared__ double s_sq[256], s_mx[256], s_amx[256];
    double sq = 0.0, mx = -INFINITY, amx = 0.0;
    for (int b = threadIdx.x; b < nparts; b += 256) {
        sq += part[3 * b];
        mx = fmax(mx, part[3 * b + 1]);
        amx = fmax(amx, part[3 * b + 2]);
    }
    s_sq[threadIdx.x] = sq;
    s_mx[threadIdx.x] = mx;
    s_amx[threadIdx.x] = amx;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) {
            s_sq[threadIdx.x] += s_sq[threadIdx.x + off];
            s_mx[threadIdx.x] = fmax(s_mx[threadIdx.x], s_mx[threadIdx.x + off]);
            s_amx[threadIdx.x] = fmax(s_amx[threadIdx.x], s_amx[threadIdx.x + off]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out[0] = s_sq[0];
        out[1] = s_mx[0];
        out[2] = s_amx[0];
    }
}

constexpr unsigned int kStatsBlocks = 1024;

__global__ void fill_complex_kernel(float2 *d, size_t n, float re, float im)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        d[i] = make_float2(re, im);
}

// psf_host_set_up's normalisation (top.c:505-507): float / double, rounded to float
__global__ void divide_kernel(float *d, size_t n, double by)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        d[i] = (float)((double)d[i] / by);
}

// out[i] = a[i] * b[i % nb] (complex)
__global__ void complex_multiply_kernel(const float2 *a, const float2 *b, size_t n, size_t nb, float2 *out)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const float2 x = a[i], y = b[i % nb];
        out[i] = make_float2(x.x * y.x - x.y * y.y, x.x * y.y + x.y * y.x);
    }
}

// out[i] = a x[i] + b y[i]
__global__ void axpby_kernel(size_t n, float a, const float *x, float b, const float *y, float *out)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        out[i] = a * x[i] + b * y[i];
}

int check_grid(int G)
{
    if (G < 2 || (G & 1)) return sdp::set_error(SDP_E_ARG, "grid size %d must be even and >= 2", G);
    return SDP_OK;
}

}  // namespace

extern "C" {

int sdp_shift_c2c(int G, const float *d_in, float *d_out, sdp_stream s)
{
    if (int e = check_grid(G)) return e;
    size_t n2 = (size_t)G * G / 2;
    shift_c2c_kernel<<<grid_for(n2), kBlock, 0, sdp::stream_of(s)>>>((const float4 *)d_in, (float4 *)d_out, G);
    SDP_LAUNCH_CHECK();
    return SDP_OK;
}

int sdp_shift_c2r(int G, const float *d_in, float *d_out, sdp_stream s)
{
    if (int e = check_grid(G)) return e;
    SDP_REQUIRE((const void *)d_in != (const void *)d_out, "sdp_shift_c2r: in and out must differ");
    size_t n2 = (size_t)G * G / 2;
    shift_c2r_kernel<<<grid_for(n2), kBlock, 0, sdp::stream_of(s)>>>((const float4 *)d_in, (float2 *)d_out, G);
    SDP_LAUNCH_CHECK();
    return SDP_OK;
}

int sdp_shift_r2c(int G, const float *d_in, float *d_out, sdp_stream s)
{
    if (int e = check_grid(G)) return e;
    SDP_REQUIRE((const void *)d_in != (const void *)d_out, "sdp_shift_r2c: in and out must differ");
    size_t n2 = (size_t)G * G / 2;
    shift_r2c_kernel<<<grid_for(n2), kBlock, 0, sdp::stream_of(s)>>>((const float2 *)d_in, (float4 *)d_out, G);
    SDP_LAUNCH_CHECK();
    return SDP_OK;
}

int sdp_convolution_correct(int G, const float *d_in, const float *d_taper, float *d_out, sdp_stream s)
{
    if (int e = check_grid(G)) return e;
    size_t n2 = (size_t)G * G / 2;
    correct_kernel<<<grid_for(n2), kBlock, 0, sdp::stream_of(s)>>>((const float2 *)d_in, d_taper, (float2 *)d_out, G);
    SDP_LAUNCH_CHECK();
    return SDP_OK;
}

int sdp_vis_subtract(int n, const float *d_a, const float *d_b, float *d_out, sdp_stream s)
{
    if (n <= 0) return SDP_OK;
    vis_sub_kernel<<<grid_for(n), kBlock, 0, sdp::stream_of(s)>>>((const float2 *)d_a, (const float2 *)d_b,
                                                                  (float2 *)d_out, n);
    SDP_LAUNCH_CHECK();
    return SDP_OK;
}

int sdp_vis_add(int n, const float *d_a, const float *d_b, float *d_out, sdp_stream s)
{
    if (n <= 0) return SDP_OK;
    vis_add_kernel<<<grid_for(n), kBlock, 0, sdp::stream_of(s)>>>((const float2 *)d_a, (const float2 *)d_b,
                                                                  (float2 *)d_out, n);
    SDP_LAUNCH_CHECK();
    return SDP_OK;
}

int sdp_vis_subtract_gains(int n, int n_baselines, const float *d_measured, const float *d_pred,
                           const float *d_gains, const int *d_pairs, int apply_only, float *d_out, sdp_stream s)
{
    if (n <= 0) return SDP_OK;
    SDP_REQUIRE(n_baselines > 0, "n_baselines must be positive");
    gains_kernel<<<grid_for(n), kBlock, 0, sdp::stream_of(s)>>>((const float2 *)d_measured, (const float2 *)d_pred,
                                                                (const float2 *)d_gains, (const int2 *)d_pairs, n,
                                                                n_baselines, apply_only, (float2 *)d_out);
    SDP_LAUNCH_CHECK();
    return SDP_OK;
}

int sdp_restore_conv2d(int D, const float *d_image, const float *d_filter, int hx, int hy, float *d_out,
                       sdp_stream s)
{
    SDP_REQUIRE(D >= 4 && d_image && d_filter && d_out, "sdp_restore_conv2d: bad arguments");
    // the flipped filter rows and columns D/2 + 1 - f, f in [-h, h), stay inside the image
    SDP_REQUIRE(hx >= 0 && hy >= 0 && D / 2 + 1 + hx < D && D / 2 + 1 + hy < D,
                "sdp_restore_conv2d: filter half sizes (%d, %d) too large for %d^2", hx, hy, D);
    dim3 grid((D + kBlock - 1) / kBlock, D);
    restore_conv2d_kernel<<<grid, kBlock, 0, sdp::stream_of(s)>>>(d_image, d_filter, D, hx, hy, d_out);
    SDP_LAUNCH_CHECK();
    return SDP_OK;
}

int sdp_gains_reciprocal(int n, const float *d_in, float *d_out, sdp_stream s)
{
    if (n <= 0) return SDP_OK;
    reciprocal_kernel<<<grid_for(n), kBlock, 0, sdp::stream_of(s)>>>((const float2 *)d_in, n, (float2 *)d_out);
    SDP_LAUNCH_CHECK();
    return SDP_OK;
}

int sdp_merge_grids(int G, int slices, const float *d_grids, float *d_out, sdp_stream s)
{
    if (int e = check_grid(G)) return e;
    SDP_REQUIRE(slices >= 1, "slices must be >= 1");
    size_t n2 = (size_t)G * G / 2;
    merge_kernel<<<grid_for(n2), kBlock, 0, sdp::stream_of(s)>>>((const float4 *)d_grids, (float4 *)d_out, n2, slices);
    SDP_LAUNCH_CHECK();
    return SDP_OK;
}

size_t sdp_image_stats_work_bytes(void) { return sizeof(double) * 3 * kStatsBlocks; }

int sdp_image_stats_async(size_t n, const float *d_img, double *d_stats, void *d_work, sdp_stream s)
{
    SDP_REQUIRE(d_img && n > 0, "sdp_image_stats: empty image");
    SDP_REQUIRE(d_stats && d_work, "sdp_image_stats: null output or work buffer");
    hipStream_t st = sdp::stream_of(s);
    const unsigned int blocks = std::min<unsigned int>(kStatsBlocks, sdp_div_up(n, 256));
    image_stats_kernel<<<blocks, 256, 0, st>>>(d_img, n, (double *)d_work);
    SDP_LAUNCH_CHECK();
    image_stats_final_kernel<<<1, 256, 0, st>>>((const double *)d_work, (int)blocks, d_stats);
    SDP_LAUNCH_CHECK();
    return SDP_OK;
}

int sdp_image_stats(size_t n, const float *d_img, double *sum_sq, float *max_value, float *max_abs, sdp_stream s)
{
    SDP_REQUIRE(d_img && n > 0, "sdp_image_stats: empty image");
    hipStream_t st = sdp::stream_of(s);
    double *buf = nullptr;
    SDP_HIP_TRY(hipMalloc(&buf, sdp_image_stats_work_bytes() + 4 * sizeof(double)));
    double h[3] = {0.0, 0.0, 0.0};
    int rc = sdp_image_stats_async(n, d_img, buf, buf + 4, s);
    hipError_t e = hipSuccess;
    if (rc == SDP_OK) {
        e = hipMemcpyAsync(h, buf, sizeof h, hipMemcpyDeviceToHost, st);
        if (e == hipSuccess) e = hipStreamSynchronize(st);
    }
    (void)hipFree(buf);
    if (rc != SDP_OK) return rc;
    if (e != hipSuccess) return sdp::set_error(SDP_E_HIP, "sdp_image_stats: %s", hipGetErrorString(e));
    if (sum_sq) *sum_sq = h[0];
    if (max_value) *max_value = (float)h[1];
    if (max_abs) *max_abs = (float)h[2];
    return SDP_OK;
}

int sdp_complex_multiply(size_t n, const float *d_a, const float *d_b, size_t nb, float *d_out, sdp_stream s)
{
    if (!n) return SDP_OK;
    SDP_REQUIRE(d_a && d_b && d_out && nb > 0, "sdp_complex_multiply: bad argument");
    complex_multiply_kernel<<<grid_for(n), kBlock, 0, sdp::stream_of(s)>>>((const float2 *)d_a, (const float2 *)d_b,
                                                                           n, nb, (float2 *)d_out);
    SDP_LAUNCH_CHECK();
    return SDP_OK;
}

int sdp_axpby(size_t n, float a, const float *d_x, float b, const float *d_y, float *d_out, sdp_stream s)
{
    if (!n) return SDP_OK;
    SDP_REQUIRE(d_x && d_y && d_out, "sdp_axpby: null argument");
    axpby_kernel<<<grid_for(n), kBlock, 0, sdp::stream_of(s)>>>(n, a, d_x, b, d_y, d_out);
    SDP_LAUNCH_CHECK();
    return SDP_OK;
}

int sdp_fill_complex(size_t n, float *d, float re, float im, sdp_stream s)
{
    if (!n) return SDP_OK;
    fill_complex_kernel<<<grid_for(n), kBlock, 0, sdp::stream_of(s)>>>((float2 *)d, n, re, im);
    SDP_LAUNCH_CHECK();
    return SDP_OK;
}

int sdp_divide(size_t n, float *d, double by, sdp_stream s)
{
    if (!n) return SDP_OK;
    divide_kernel<<<grid_for(n), kBlock, 0, sdp::stream_of(s)>>>(d, n, by);
    SDP_LAUNCH_CHECK();
    return SDP_OK;
}

}  // extern "C"
