// sdp_finegrid.hip -- fine-grid snapping (correct_to_finegrid,
// src/CPU/degridgrid.c:12-65) and the nearest-cell value std_degridding
// returns (src/CPU/degridgrid.c:194-200): index arithmetic that must match
// the reference bit for bit, one thread per visibility.
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

#include <algorithm>

#include "sdp_internal.h"

namespace {

// std_degridding's returned value: the nearest grid cell (degridgrid.c:194-200)
__global__ void degrid_nearest_kernel(int n, int n_out, int G, double uv_scale, const float2 *grid,
                                      const float *uvw, float2 *out)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_out) return;
    float2 v = make_float2(0.f, 0.f);
    if (i < n) {
        const int gc = G / 2;
        float gx = (float)((double)uvw[3 * (size_t)i] * uv_scale);
        float gy = (float)((double)uvw[3 * (size_t)i + 1] * uv_scale);
        double xd = (double)(gx + (float)gc) + 0.5, yd = (double)(gy + (float)gc) + 0.5;
        if (xd > -2.0e9 && xd < 2.0e9 && yd > -2.0e9 && yd < 2.0e9) {
            long long idx = (long long)(int)xd + (long long)(int)yd * G;
            if (idx >= 0 && idx < (long long)G * G) v = grid[idx];
        }
    }
    out[i] = v;
}

// correct_to_finegrid coordinates (degridgrid.c:24-25,52-54); negative
// fine-grid positions use signed truncation (the reference's conversion to
// unsigned long is undefined there).
__global__ void finegrid_kernel(int n, int G, int OS, double uv_scale, const float *in, float *out)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double half = (double)(G / 2);
    const float x = in[3 * (size_t)i], y = in[3 * (size_t)i + 1], z = in[3 * (size_t)i + 2];
    long long fx = (long long)(((double)x * uv_scale + half) * (double)OS + 0.5);
    long long fy = (long long)(((double)y * uv_scale + half) * (double)OS + 0.5);
    out[3 * (size_t)i] = (float)(((double)fx / (double)OS - half) / uv_scale);
    out[3 * (size_t)i + 1] = (float)(((double)fy / (double)OS - half) / uv_scale);
    out[3 * (size_t)i + 2] = z;
}

}  // namespace

extern "C" {

int sdp_degrid_nearest(int n_vis, int n_out, int G, double uv_scale, const float *d_grid, const float *d_uvw,
                       float *d_vis_out, sdp_stream s)
{
    SDP_REQUIRE(G >= 2 && G % 2 == 0, "bad grid size");
    SDP_REQUIRE(n_vis >= 0 && n_out >= n_vis, "bad visibility counts");
    if (n_out == 0) return SDP_OK;
    degrid_nearest_kernel<<<sdp_div_up(n_out, 256), 256, 0, sdp::stream_of(s)>>>(
        n_vis, n_out, G, uv_scale, (const float2 *)d_grid, d_uvw, (float2 *)d_vis_out);
    SDP_LAUNCH_CHECK();
    return SDP_OK;
}

int sdp_finegrid_correct(int n_vis, int G, int OS, double uv_scale, const float *d_uvw_in, float *d_uvw_out,
                         sdp_stream s)
{
    SDP_REQUIRE(G >= 2 && OS >= 1 && n_vis >= 0, "bad arguments");
    if (n_vis == 0) return SDP_OK;
    finegrid_kernel<<<sdp_div_up(n_vis, 256), 256, 0, sdp::stream_of(s)>>>(n_vis, G, OS, uv_scale, d_uvw_in, d_uvw_out);
    SDP_LAUNCH_CHECK();
    return SDP_OK;
}

}  // extern "C"

// ---- correct_to_finegrid with PERFORM_SIMPLIFICATION = 1 ------------------
//
// degridgrid.c:33-49 merges visibilities whose fine-grid key matches a
// previous one.  Its hashmap keeps the address of one stack variable as every
// key (degridgrid.c:29,46; map.c:220-238), so an entry matches when the
// 32-bit key hash matches (map.c:201-209); the hash of the 8-byte key
// (map.c:131-170) is the expression in finegrid_key.  On the GPU: key per
// visibility, stable radix sort by key, one thread per run of equal keys sums
// the run in index order (the reference's order of float additions), and the
// runs are compacted in order of their first index.
namespace {

__device__ __forceinline__ uint32_t finegrid_key(uint64_t idx)
{
    uint64_t h = (2166136261ull ^ idx) * 0xbf58476d1ce4e5b9ull;
    return (uint32_t)(h ^ (h >> 32));
}

__global__ void simplify_keys_kernel(int n, int G, int OS, double uv_scale, double w_scale, const float *uvw,
                                     uint32_t *keys, uint32_t *vals, float *uvw_fine)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double half = (double)(G / 2);
    const float x = uvw[3 * (size_t)i], y = uvw[3 * (size_t)i + 1], z = uvw[3 * (size_t)i + 2];
    long long fx = (long long)(((double)x * uv_scale + half) * (double)OS + 0.5);
    long long fy = (long long)(((double)y * uv_scale + half) * (double)OS + 0.5);
    long long wi = (long long)((double)__fsqrt_rn(fabsf((float)((double)z * w_scale))) + 0.5);
    uint64_t idx = ((uint64_t)fx + (uint64_t)fy * (uint64_t)G * (uint64_t)OS) +
                   (uint64_t)wi * ((uint64_t)G * G * OS * OS);
    keys[i] = finegrid_key(idx);
    vals[i] = (uint32_t)i;
    uvw_fine[3 * (size_t)i] = (float)(((double)fx / (double)OS - half) / uv_scale);
    uvw_fine[3 * (size_t)i + 1] = (float)(((double)fy / (double)OS - half) / uv_scale);
    uvw_fine[3 * (size_t)i + 2] = z;
}

// head flags in caller order: first[i] = 1 if visibility i starts a run
__global__ void simplify_heads_kernel(int n, const uint32_t *skeys, const uint32_t *svals, int *first)
{
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    first[svals[j]] = (j == 0 || skeys[j] != skeys[j - 1]) ? 1 : 0;
}

__global__ void simplify_sum_kernel(int n, const uint32_t *skeys, const uint32_t *svals, const int *pos,
                                    const float2 *vis, const float *uvw_fine, float2 *vis_out, float *uvw_out)
{
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n || (j > 0 && skeys[j] == skeys[j - 1])) return;
    const uint32_t rep = svals[j];
    float2 s = vis[rep];
    for (int k = j + 1; k < n && skeys[k] == skeys[j]; ++k) {
        const float2 d = vis[svals[k]];
        s.x += d.x;
        s.y += d.y;
    }
    const int o = pos[rep];
    vis_out[o] = s;
    uvw_out[3 * (size_t)o] = uvw_fine[3 * (size_t)rep];
    uvw_out[3 * (size_t)o + 1] = uvw_fine[3 * (size_t)rep + 1];
    uvw_out[3 * (size_t)o + 2] = uvw_fine[3 * (size_t)rep + 2];
}

}  // namespace

extern "C" int sdp_finegrid_simplify(int n_vis, int G, int OS, double uv_scale, double w_scale, const float *d_uvw_in,
                                     const float *d_vis_in, float *d_uvw_out, float *d_vis_out, int *n_out,
                                     sdp_stream s)
{
    SDP_REQUIRE(G >= 2 && OS >= 1 && n_vis >= 0 && n_out, "bad arguments");
    *n_out = 0;
    if (n_vis == 0) return SDP_OK;
    hipStream_t st = sdp::stream_of(s);
    uint32_t *keys = nullptr, *skeys = nullptr, *vals = nullptr, *svals = nullptr;
    int *first = nullptr, *pos = nullptr;
    float *fine = nullptr;
    void *temp = nullptr;
    int rc = SDP_OK;
    size_t tb1 = 0, tb2 = 0;
    int last_first = 0, last_pos = 0;
    const size_t n = (size_t)n_vis;
#define SIMP_TRY(expr)                                                                              \
    do {                                                                                            \
        hipError_t e_ = (expr);                                                                     \
        if (e_ != hipSuccess) {                                                                     \
            rc = sdp::set_error(SDP_E_HIP, "%s failed: %s", #expr, hipGetErrorString(e_));          \
            goto done;                                                                              \
        }                                                                                           \
    } while (0)
    SIMP_TRY(hipMalloc(&keys, 4 * n));
    SIMP_TRY(hipMalloc(&skeys, 4 * n));
    SIMP_TRY(hipMalloc(&vals, 4 * n));
    SIMP_TRY(hipMalloc(&svals, 4 * n));
    SIMP_TRY(hipMalloc(&first, 4 * n));
    SIMP_TRY(hipMalloc(&pos, 4 * n));
    SIMP_TRY(hipMalloc(&fine, 12 * n));
    simplify_keys_kernel<<<sdp_div_up(n, 256), 256, 0, st>>>(n_vis, G, OS, uv_scale, w_scale, d_uvw_in, keys, vals, fine);
    SIMP_TRY(hipGetLastError());
    SIMP_TRY(rocprim::radix_sort_pairs(nullptr, tb1, keys, skeys, vals, svals, (size_t)n_vis, 0, 32, st));
    SIMP_TRY(rocprim::exclusive_scan(nullptr, tb2, first, pos, 0, (size_t)n_vis, rocprim::plus<int>(), st));
    SIMP_TRY(hipMalloc(&temp, std::max(tb1, tb2)));
    SIMP_TRY(rocprim::radix_sort_pairs(temp, tb1, keys, skeys, vals, svals, (size_t)n_vis, 0, 32, st));
    simplify_heads_kernel<<<sdp_div_up(n, 256), 256, 0, st>>>(n_vis, skeys, svals, first);
    SIMP_TRY(hipGetLastError());
    SIMP_TRY(rocprim::exclusive_scan(temp, tb2, first, pos, 0, (size_t)n_vis, rocprim::plus<int>(), st));
    simplify_sum_kernel<<<sdp_div_up(n, 256), 256, 0, st>>>(n_vis, skeys, svals, pos, (const float2 *)d_vis_in, fine,
                                                             (float2 *)d_vis_out, d_uvw_out);
    SIMP_TRY(hipGetLastError());
    SIMP_TRY(hipMemcpyAsync(&last_first, first + n - 1, 4, hipMemcpyDeviceToHost, st));
    SIMP_TRY(hipMemcpyAsync(&last_pos, pos + n - 1, 4, hipMemcpyDeviceToHost, st));
    SIMP_TRY(hipStreamSynchronize(st));
    *n_out = last_pos + last_first;
done:
#undef SIMP_TRY
    (void)hipStreamSynchronize(st);
    (void)hipFree(keys); (void)hipFree(skeys); (void)hipFree(vals); (void)hipFree(svals);
    (void)hipFree(first); (void)hipFree(pos); (void)hipFree(fine); (void)hipFree(temp);
    return rc;
}
