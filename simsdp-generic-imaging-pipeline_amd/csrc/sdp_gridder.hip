// sdp_gridder.hip -- convolutional gridding and degridding on gfx950.
//
// Reference arithmetic: std_gridding (degridgrid.c:69-125), std_degridding
// (degridgrid.c:127-202), gridding_CPU (gridding_run.c:55-137) and
// correct_to_finegrid (degridgrid.c:12-65), paths relative to
// preesm_pipelines/simsdp_fft_imaging_pipeline/Code/src/CPU/.
//
// Design
//  * Plan (once per set of coordinates): one thread per visibility repeats
//    the reference's float/double index arithmetic exactly (w-plane, grid
//    position, footprint bounds), clips the footprint to the grid and writes
//    a 16-byte record.  Visibilities are stable-sorted by the uv tile that
//    holds their footprint's first cell, and each tile's list is cut into
//    work items ("chunks") of bounded tap count.
//  * Gridding: one 256-thread workgroup per chunk accumulates into an LDS
//    sub-grid of S x S cells (S = tile + 2 * max half-support) that covers
//    every footprint of the tile.  Each thread owns the cells whose (row,
//    column) are congruent to its (ty, tx) modulo 16, so no two threads ever
//    touch the same cell: no atomics, a fixed visibility order per cell
//    (deterministic) and a per-cell TwoSum compensation term.  The sub-grid
//    is written to a per-chunk slab; a merge pass sums, for every grid cell,
//    the slabs that cover it in a fixed order (compensated), writing the
//    whole grid once with coalesced stores.
//  * Degridding: one workgroup per chunk stages the same S x S window of the
//    grid in LDS; each wavefront takes one visibility at a time, its lanes
//    sweep the footprint taps and a shuffle tree reduces the 64 partial sums.
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cstring>
#include <vector>

#include "sdp_internal.h"

namespace {

constexpr int kGridThreads = 256;   // 16 x 16 cell-ownership classes
constexpr int kDegridThreads = 256; // 4 wavefronts
constexpr int kStage = 256;         // visibilities staged per LDS batch
// dynamic LDS of the gridding workgroup (accumulators); the rest of the
// 160 KiB holds the staged visibilities
constexpr size_t kMaxGridLds = 152 * 1024;

struct VisRec {
    float gx, gy;     // grid position in cells (reference grid_pos / snapped coordinate)
    uint32_t start;   // first footprint cell after clipping: u | v << 16 (0-based)
    uint32_t meta;    // nu | nv << 8 | plane << 16 (15 bits) | conj << 31
};
static_assert(sizeof(VisRec) == 16, "VisRec is 16 bytes");

__device__ __forceinline__ float2 cmul(float2 a, float2 b)
{
    // complex_mult (degridgrid.c:5-8): separate roundings, no contraction
    return make_float2(a.x * b.x - a.y * b.y, a.y * b.x + a.x * b.y);
}

// Knuth TwoSum: s + e == a + b exactly.
__device__ __forceinline__ void two_sum_acc(float &s, float &c, float b)
{
    float t = s + b;
    float bb = t - s;
    float err = (s - (t - bb)) + (b - bb);
    s = t;
    c += err;
}

__device__ __forceinline__ int kernel_tap(int cell_abs, float pos, float os)
{
    // abs((int)ROUND((v - grid_pos) * OVERSAMPLING_FACTOR)), degridgrid.c:101,110
    return abs((int)roundf(((float)cell_abs - pos) * os));
}

// ---- plan construction ---------------------------------------------------

struct PlanArgs {
    int n, G, OS, n_kernels, snap_mode, tile, ntx, nty;
    double uv_scale, w_scale;
};

__global__ void plan_records_kernel(const float *uvw, const int2 *supports, PlanArgs a, VisRec *recs,
                                    uint32_t *keys, uint32_t *vals, unsigned int *tile_count,
                                    unsigned long long *tile_taps)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.n) return;
    const float x = uvw[3 * (size_t)i], y = uvw[3 * (size_t)i + 1], z = uvw[3 * (size_t)i + 2];
    const int gc = a.G / 2;
    const uint32_t dropped = (uint32_t)(a.ntx * a.nty);
    keys[i] = dropped;
    vals[i] = (uint32_t)i;

    float wabs = fabsf((float)((double)z * a.w_scale));
    float root = __fsqrt_rn(wabs);
    int plane;
    if (a.snap_mode == SDP_SNAP_STD)
        plane = (int)((double)root + 0.5);          // degridgrid.c:84
    else
        plane = (int)roundf(root);                   // gridding_run.c:67
    if (!(root == root) || plane < 0 || plane >= a.n_kernels) return;
    const int h = supports[plane].x;
    if (h < 0 || 2 * h + 1 > 255) return;

    float gx = (float)((double)x * a.uv_scale);
    float gy = (float)((double)y * a.uv_scale);
    if (!(fabsf(gx) < 1.0e9f) || !(fabsf(gy) < 1.0e9f)) return;
    float lo_x, hi_x, lo_y, hi_y;   // half-open [lo, hi) in absolute cells
    if (a.snap_mode == SDP_SNAP_STD) {
        lo_x = ceilf(gx - (float)h); hi_x = ceilf(gx + (float)h);       // degridgrid.c:93,103
        lo_y = ceilf(gy - (float)h); hi_y = ceilf(gy + (float)h);
    } else {
        const float os = (float)a.OS;
        gx = roundf(gx * os) / os;                                       // gridding_run.c:79-82
        gy = roundf(gy * os) / os;
        lo_x = ceilf(gx - (float)h); hi_x = floorf(gx + (float)h) + 1.0f; // inclusive upper bound :104,112
        lo_y = ceilf(gy - (float)h); hi_y = floorf(gy + (float)h) + 1.0f;
    }
    int ua = max((int)lo_x + gc, 0), ub = min((int)hi_x + gc, a.G);
    int va = max((int)lo_y + gc, 0), vb = min((int)hi_y + gc, a.G);
    if (ua >= ub || va >= vb) return;
    int nu = ub - ua, nv = vb - va;
    VisRec r;
    r.gx = gx;
    r.gy = gy;
    r.start = (uint32_t)ua | ((uint32_t)va << 16);
    r.meta = (uint32_t)nu | ((uint32_t)nv << 8) | ((uint32_t)plane << 16) | ((z < 0.0f) ? 0x80000000u : 0u);
    recs[i] = r;
    uint32_t tile = (uint32_t)((va / a.tile) * a.ntx + (ua / a.tile));
    keys[i] = tile;
    atomicAdd(&tile_count[tile], 1u);
    atomicAdd(&tile_taps[tile], (unsigned long long)(nu * nv));
}

__global__ void gather_records_kernel(const VisRec *recs, const uint32_t *order, int n, VisRec *out)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = recs[order[i]];
}

__global__ void permute_kernel(const float2 *in, const uint32_t *perm, int n, float2 *out, int to_plan)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (to_plan)
        out[i] = in[perm[i]];
    else
        out[perm[i]] = in[i];
}

// ---- gridding ------------------------------------------------------------

struct GridArgs {
    int G, OS, tile, ntx, S, stride;   // stride: LDS row pitch in cells
};

__global__ void __launch_bounds__(kGridThreads)
grid_chunk_kernel(const VisRec *__restrict__ recs, const float2 *__restrict__ vis, const uint32_t *__restrict__ perm,
                  const int2 *__restrict__ plane_tab, const float2 *__restrict__ kernels,
                  const int4 *__restrict__ chunks, GridArgs a, float2 *__restrict__ slabs)
{
    extern __shared__ float4 acc[];   // S rows x stride cells: (sum.re, sum.im, comp.re, comp.im)
    __shared__ VisRec srec[kStage];
    __shared__ float2 svis[kStage];

    const int4 ch = chunks[blockIdx.x];   // tile, first record, count, slab index
    const int tid = threadIdx.x;
    const int ty = tid >> 4, tx = tid & 15;
    const int tile_u0 = (ch.x % a.ntx) * a.tile, tile_v0 = (ch.x / a.ntx) * a.tile;
    const int gc = a.G / 2;
    const float os = (float)a.OS;

    for (int i = tid; i < a.S * a.stride; i += kGridThreads) acc[i] = make_float4(0.f, 0.f, 0.f, 0.f);

    for (int b = 0; b < ch.z; b += kStage) {
        const int nb = min(kStage, ch.z - b);
        __syncthreads();
        if (tid < nb) {
            int j = ch.y + b + tid;
            srec[tid] = recs[j];
            svis[tid] = vis[perm ? perm[j] : (uint32_t)j];
        }
        __syncthreads();
        for (int k = 0; k < nb; ++k) {
            const VisRec r = srec[k];
            const float2 v = svis[k];
            const int lu0 = (int)(r.start & 0xffffu) - tile_u0, lv0 = (int)(r.start >> 16) - tile_v0;
            const int nu = r.meta & 0xffu, nv = (r.meta >> 8) & 0xffu;
            const int2 pt = plane_tab[(r.meta >> 16) & 0x7fffu];
            const float conj = (r.meta & 0x80000000u) ? -1.0f : 1.0f;
            for (int row = lv0 + ((ty - lv0) & 15); row < lv0 + nv; row += 16) {
                const int ky = kernel_tap(tile_v0 + row - gc, r.gy, os);
                const float2 *krow = kernels + pt.x + ky * pt.y;
                float4 *arow = acc + row * a.stride;
                for (int col = lu0 + ((tx - lu0) & 15); col < lu0 + nu; col += 16) {
                    const int kx = kernel_tap(tile_u0 + col - gc, r.gx, os);
                    float2 ks = krow[kx];
                    ks.y *= conj;
                    const float2 p = cmul(v, ks);
                    float4 c = arow[col];
                    two_sum_acc(c.x, c.z, p.x);
                    two_sum_acc(c.y, c.w, p.y);
                    arow[col] = c;
                }
            }
        }
    }
    __syncthreads();
    float2 *slab = slabs + (size_t)ch.w * a.S * a.S;
    for (int i = tid; i < a.S * a.S; i += kGridThreads) {
        const int row = i / a.S, col = i - row * a.S;
        const float4 c = acc[row * a.stride + col];
        slab[i] = make_float2(c.x + c.z, c.y + c.w);
    }
}

struct MergeArgs {
    int G, tile, ntx, nty, S;
};

// grid[y][x] = compensated sum over every chunk slab covering (x, y), in
// (tile row, tile column, chunk) order.  Writes every cell.
__global__ void merge_slabs_kernel(const float2 *__restrict__ slabs, const int *__restrict__ tile_chunk_begin,
                                   MergeArgs a, float2 *__restrict__ grid)
{
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.y;
    if (x >= a.G) return;
    float sr = 0.f, si = 0.f, cr = 0.f, ci = 0.f;
    const int ty_lo = max(0, (y - a.S + a.tile) / a.tile), ty_hi = min(a.nty - 1, y / a.tile);
    const int tx_lo = max(0, (x - a.S + a.tile) / a.tile), tx_hi = min(a.ntx - 1, x / a.tile);
    for (int ty = ty_lo; ty <= ty_hi; ++ty) {
        const int ly = y - ty * a.tile;
        if (ly >= a.S) continue;
        for (int tx = tx_lo; tx <= tx_hi; ++tx) {
            const int lx = x - tx * a.tile;
            if (lx >= a.S) continue;
            const int t = ty * a.ntx + tx;
            const int c0 = tile_chunk_begin[t], c1 = tile_chunk_begin[t + 1];
            for (int c = c0; c < c1; ++c) {
                const float2 v = slabs[(size_t)c * a.S * a.S + ly * a.S + lx];
                two_sum_acc(sr, cr, v.x);
                two_sum_acc(si, ci, v.y);
            }
        }
    }
    grid[(size_t)y * a.G + x] = make_float2(sr + cr, si + ci);
}

// ---- degridding ------------------------------------------------------------

__device__ __forceinline__ float wave_sum(float v)
{
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

__global__ void __launch_bounds__(kDegridThreads)
degrid_chunk_kernel(const VisRec *__restrict__ recs, const uint32_t *__restrict__ perm,
                    const int2 *__restrict__ plane_tab, const float2 *__restrict__ kernels,
                    const int4 *__restrict__ chunks, GridArgs a, const float2 *__restrict__ grid,
                    const float2 *__restrict__ measured, float2 *__restrict__ out)
{
    extern __shared__ float2 sub[];   // S rows x stride cells of the grid window
    const int4 ch = chunks[blockIdx.x];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int tile_u0 = (ch.x % a.ntx) * a.tile, tile_v0 = (ch.x / a.ntx) * a.tile;
    const int gc = a.G / 2;
    const float os = (float)a.OS;

    for (int i = tid; i < a.S * a.S; i += kDegridThreads) {
        const int row = i / a.S, col = i - row * a.S;
        const int gy = tile_v0 + row, gx = tile_u0 + col;
        float2 g = make_float2(0.f, 0.f);
        if (gy < a.G && gx < a.G) g = grid[(size_t)gy * a.G + gx];
        sub[row * a.stride + col] = g;
    }
    __syncthreads();

    for (int k = wave; k < ch.z; k += kDegridThreads / 64) {
        const int j = ch.y + k;
        const VisRec r = recs[j];
        const int lu0 = (int)(r.start & 0xffffu) - tile_u0, lv0 = (int)(r.start >> 16) - tile_v0;
        const int nu = r.meta & 0xffu, nv = (r.meta >> 8) & 0xffu;
        const int2 pt = plane_tab[(r.meta >> 16) & 0x7fffu];
        const float conj = (r.meta & 0x80000000u) ? -1.0f : 1.0f;
        const int dr = 64 / nu, dc = 64 - dr * nu;
        int row = lane / nu, col = lane - row * nu;
        float sr = 0.f, si = 0.f;
        while (row < nv) {
            const int ky = kernel_tap(tile_v0 + lv0 + row - gc, r.gy, os);
            const int kx = kernel_tap(tile_u0 + lu0 + col - gc, r.gx, os);
            float2 ks = kernels[pt.x + ky * pt.y + kx];
            ks.y *= conj;
            const float2 g = sub[(lv0 + row) * a.stride + lu0 + col];
            const float2 p = cmul(g, ks);
            sr += p.x;
            si += p.y;
            col += dc;
            row += dr;
            if (col >= nu) { col -= nu; row += 1; }
        }
        sr = wave_sum(sr);
        si = wave_sum(si);
        if (lane == 0) {
            const uint32_t o = perm ? perm[j] : (uint32_t)j;
            float2 res = make_float2(sr, si);
            if (measured) {
                const float2 m = measured[o];
                res = make_float2(m.x - sr, m.y - si);
            }
            out[o] = res;
        }
    }
}

// visibilities the plan dropped (outside the grid): prediction 0
__global__ void degrid_dropped_kernel(const uint32_t *perm, int first, int n, const float2 *measured, float2 *out)
{
    int j = first + blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    const uint32_t o = perm ? perm[j] : (uint32_t)j;
    out[o] = measured ? measured[o] : make_float2(0.f, 0.f);
}

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

// ---- plan object -----------------------------------------------------------

struct sdp_grid_plan {
    sdp_grid_params p;
    int n = 0, n_kept = 0, tile = 0, chunk_taps = 0, ntx = 0, nty = 0, S = 0, stride = 0, H = 0;
    int n_chunks = 0, n_tiles_used = 0;
    VisRec *recs = nullptr;        // plan order
    uint32_t *perm = nullptr;      // plan position -> caller index
    int2 *plane_tab = nullptr;     // (sample offset, (h+1)*OS) per plane
    int4 *chunks = nullptr;        // (tile, first, count, slab)
    int *tile_chunk_begin = nullptr;
    float2 *slabs = nullptr;
    size_t slab_bytes = 0;
};

namespace {

void plan_free(sdp_grid_plan *p)
{
    if (!p) return;
    (void)hipFree(p->recs);
    (void)hipFree(p->perm);
    (void)hipFree(p->plane_tab);
    (void)hipFree(p->chunks);
    (void)hipFree(p->tile_chunk_begin);
    (void)hipFree(p->slabs);
    delete p;
}

// Tile edge and tap budget per work item.  The tile keeps the LDS sub-grid
// (tile + 2H cells square, 16 B per cell) at or below 64 x 64 cells so two
// gridding workgroups fit in one CU's 160 KiB of LDS.
void choose_tiling(int G, int H, int req_tile, int req_chunk, int *tile, int *chunk_taps)
{
    int t = req_tile > 0 ? req_tile : std::max(16, 64 - 2 * H);
    t = std::min(t, G);
    *tile = t;
    *chunk_taps = req_chunk > 0 ? req_chunk : (1 << 19);
}

}  // namespace

extern "C" {

int sdp_grid_plan_create(sdp_grid_plan **out, const sdp_grid_params *pp, const int *supports,
                         const float *d_uvw, int n_vis, sdp_stream s)
{
    SDP_REQUIRE(out && pp && supports, "sdp_grid_plan_create: null argument");
    *out = nullptr;
    const sdp_grid_params p = *pp;
    SDP_REQUIRE(p.grid_size >= 2 && p.grid_size % 2 == 0 && p.grid_size <= 32768,
                "grid size %d must be even and in [2, 32768]", p.grid_size);
    SDP_REQUIRE(p.oversampling >= 1 && p.n_kernels >= 1 && p.n_kernels < 32768, "bad oversampling / n_kernels");
    SDP_REQUIRE(n_vis >= 0, "negative visibility count");
    SDP_REQUIRE(p.snap_mode == SDP_SNAP_STD || p.snap_mode == SDP_SNAP_GRIDDING_ACTOR, "bad snap mode");
    hipStream_t st = sdp::stream_of(s);

    int H = 0;
    std::vector<int2> tab(p.n_kernels);
    for (int k = 0; k < p.n_kernels; ++k) {
        const int h = supports[2 * k], off = supports[2 * k + 1];
        SDP_REQUIRE(h >= 0 && h <= 127, "half support %d of plane %d outside [0, 127]", h, k);
        H = std::max(H, h);
        tab[k] = make_int2(off, (h + 1) * p.oversampling);
    }

    sdp_grid_plan *plan = new sdp_grid_plan();
    plan->p = p;
    plan->n = n_vis;
    plan->H = H;
    choose_tiling(p.grid_size, H, p.tile, p.chunk, &plan->tile, &plan->chunk_taps);
    plan->ntx = (p.grid_size + plan->tile - 1) / plan->tile;
    plan->nty = plan->ntx;
    plan->S = plan->tile + 2 * H;
    plan->stride = plan->S + 1;
    const int ntiles = plan->ntx * plan->nty;
    if (sizeof(float4) * (size_t)plan->S * plan->stride > kMaxGridLds) {
        const int S = plan->S;
        delete plan;
        return sdp::set_error(SDP_E_UNSUPPORTED,
                              "sub-grid of %d x %d cells (tile %d + 2 x half support %d) exceeds the LDS budget", S,
                              S, p.tile > 0 ? p.tile : std::max(16, 64 - 2 * H), H);
    }

    auto fail = [&](int code) { plan_free(plan); return code; };
#define PLAN_TRY(expr)                                                                              \
    do {                                                                                            \
        hipError_t e_ = (expr);                                                                     \
        if (e_ != hipSuccess)                                                                       \
            return fail(sdp::set_error(SDP_E_HIP, "%s failed: %s", #expr, hipGetErrorString(e_)));  \
    } while (0)

    PLAN_TRY(hipMalloc(&plan->plane_tab, sizeof(int2) * p.n_kernels));
    PLAN_TRY(hipMemcpyAsync(plan->plane_tab, tab.data(), sizeof(int2) * p.n_kernels, hipMemcpyHostToDevice, st));
    int2 *d_supports = nullptr;
    std::vector<int2> sup(p.n_kernels);
    for (int k = 0; k < p.n_kernels; ++k) sup[k] = make_int2(supports[2 * k], supports[2 * k + 1]);
    PLAN_TRY(hipMalloc(&d_supports, sizeof(int2) * p.n_kernels));
    PLAN_TRY(hipMemcpyAsync(d_supports, sup.data(), sizeof(int2) * p.n_kernels, hipMemcpyHostToDevice, st));

    const size_t n = (size_t)std::max(n_vis, 1);
    VisRec *recs_tmp = nullptr;
    uint32_t *keys = nullptr, *keys_sorted = nullptr, *vals = nullptr;
    unsigned int *tile_count = nullptr;
    unsigned long long *tile_taps = nullptr;
    PLAN_TRY(hipMalloc(&recs_tmp, sizeof(VisRec) * n));
    PLAN_TRY(hipMalloc(&plan->recs, sizeof(VisRec) * n));
    PLAN_TRY(hipMalloc(&plan->perm, sizeof(uint32_t) * n));
    PLAN_TRY(hipMalloc(&keys, sizeof(uint32_t) * n));
    PLAN_TRY(hipMalloc(&keys_sorted, sizeof(uint32_t) * n));
    PLAN_TRY(hipMalloc(&vals, sizeof(uint32_t) * n));
    PLAN_TRY(hipMalloc(&tile_count, sizeof(unsigned int) * ntiles));
    PLAN_TRY(hipMalloc(&tile_taps, sizeof(unsigned long long) * ntiles));
    PLAN_TRY(hipMemsetAsync(tile_count, 0, sizeof(unsigned int) * ntiles, st));
    PLAN_TRY(hipMemsetAsync(tile_taps, 0, sizeof(unsigned long long) * ntiles, st));

    if (n_vis > 0) {
        PlanArgs a{n_vis, p.grid_size, p.oversampling, p.n_kernels, p.snap_mode, plan->tile, plan->ntx, plan->nty,
                   p.uv_scale, p.w_scale};
        plan_records_kernel<<<sdp_div_up(n_vis, 256), 256, 0, st>>>(d_uvw, d_supports, a, recs_tmp, keys, vals,
                                                                    tile_count, tile_taps);
        PLAN_TRY(hipGetLastError());
        int end_bit = 1;
        while ((1u << end_bit) <= (unsigned)ntiles) ++end_bit;
        size_t temp_bytes = 0;
        PLAN_TRY(hipcub::DeviceRadixSort::SortPairs(nullptr, temp_bytes, keys, keys_sorted, vals, plan->perm, n_vis,
                                                    0, end_bit, st));
        void *temp = nullptr;
        PLAN_TRY(hipMalloc(&temp, temp_bytes));
        hipError_t se = hipcub::DeviceRadixSort::SortPairs(temp, temp_bytes, keys, keys_sorted, vals, plan->perm,
                                                           n_vis, 0, end_bit, st);
        (void)hipStreamSynchronize(st);
        (void)hipFree(temp);
        PLAN_TRY(se);
        gather_records_kernel<<<sdp_div_up(n_vis, 256), 256, 0, st>>>(recs_tmp, plan->perm, n_vis, plan->recs);
        PLAN_TRY(hipGetLastError());
    }

    std::vector<unsigned int> counts(ntiles);
    std::vector<unsigned long long> taps(ntiles);
    PLAN_TRY(hipMemcpyAsync(counts.data(), tile_count, sizeof(unsigned int) * ntiles, hipMemcpyDeviceToHost, st));
    PLAN_TRY(hipMemcpyAsync(taps.data(), tile_taps, sizeof(unsigned long long) * ntiles, hipMemcpyDeviceToHost, st));
    PLAN_TRY(hipStreamSynchronize(st));
    (void)hipFree(recs_tmp);
    (void)hipFree(keys);
    (void)hipFree(keys_sorted);
    (void)hipFree(vals);
    (void)hipFree(tile_count);
    (void)hipFree(tile_taps);
    (void)hipFree(d_supports);

    // work items: tile-major, each tile cut into pieces of about chunk_taps taps
    std::vector<int4> chunks;
    std::vector<int> tile_chunk_begin(ntiles + 1, 0);
    int first = 0;
    for (int t = 0; t < ntiles; ++t) {
        tile_chunk_begin[t] = (int)chunks.size();
        const int cnt = (int)counts[t];
        if (cnt > 0) {
            ++plan->n_tiles_used;
            long long pieces = (long long)((taps[t] + plan->chunk_taps - 1) / plan->chunk_taps);
            pieces = std::max(1LL, std::min(pieces, (long long)cnt));
            for (long long q = 0; q < pieces; ++q) {
                const int b = first + (int)(cnt * q / pieces), e = first + (int)(cnt * (q + 1) / pieces);
                chunks.push_back(make_int4(t, b, e - b, (int)chunks.size()));
            }
        }
        first += cnt;
    }
    tile_chunk_begin[ntiles] = (int)chunks.size();
    plan->n_kept = first;
    plan->n_chunks = (int)chunks.size();

    PLAN_TRY(hipMalloc(&plan->tile_chunk_begin, sizeof(int) * (ntiles + 1)));
    PLAN_TRY(hipMemcpyAsync(plan->tile_chunk_begin, tile_chunk_begin.data(), sizeof(int) * (ntiles + 1),
                            hipMemcpyHostToDevice, st));
    if (plan->n_chunks > 0) {
        PLAN_TRY(hipMalloc(&plan->chunks, sizeof(int4) * plan->n_chunks));
        PLAN_TRY(hipMemcpyAsync(plan->chunks, chunks.data(), sizeof(int4) * plan->n_chunks, hipMemcpyHostToDevice, st));
        plan->slab_bytes = sizeof(float2) * (size_t)plan->n_chunks * plan->S * plan->S;
        PLAN_TRY(hipMalloc(&plan->slabs, plan->slab_bytes));
    }
    PLAN_TRY(hipStreamSynchronize(st));
#undef PLAN_TRY
    *out = plan;
    return SDP_OK;
}

int sdp_grid_plan_destroy(sdp_grid_plan *plan)
{
    plan_free(plan);
    return SDP_OK;
}

int sdp_grid_plan_info(const sdp_grid_plan *plan, int64_t *info)
{
    SDP_REQUIRE(plan && info, "null argument");
    info[0] = plan->n_kept;
    info[1] = plan->n - plan->n_kept;
    info[2] = plan->n_tiles_used;
    info[3] = plan->n_chunks;
    info[4] = plan->tile;
    info[5] = plan->chunk_taps;
    info[6] = plan->H;
    info[7] = (int64_t)plan->slab_bytes;
    return SDP_OK;
}

int sdp_grid_plan_to_plan_order(const sdp_grid_plan *plan, const float *d_in, float *d_out, sdp_stream s)
{
    SDP_REQUIRE(plan && d_in != d_out, "sdp_grid_plan_to_plan_order: bad arguments");
    if (plan->n == 0) return SDP_OK;
    permute_kernel<<<sdp_div_up(plan->n, 256), 256, 0, sdp::stream_of(s)>>>((const float2 *)d_in, plan->perm, plan->n,
                                                                             (float2 *)d_out, 1);
    SDP_LAUNCH_CHECK();
    return SDP_OK;
}

int sdp_grid_plan_to_original_order(const sdp_grid_plan *plan, const float *d_in, float *d_out, sdp_stream s)
{
    SDP_REQUIRE(plan && d_in != d_out, "sdp_grid_plan_to_original_order: bad arguments");
    if (plan->n == 0) return SDP_OK;
    permute_kernel<<<sdp_div_up(plan->n, 256), 256, 0, sdp::stream_of(s)>>>((const float2 *)d_in, plan->perm, plan->n,
                                                                             (float2 *)d_out, 0);
    SDP_LAUNCH_CHECK();
    return SDP_OK;
}

int sdp_grid_accumulate(const sdp_grid_plan *plan, const float *d_kernels, const float *d_vis, int order,
                        sdp_stream s)
{
    SDP_REQUIRE(plan, "sdp_grid_accumulate: null plan");
    SDP_REQUIRE(order == SDP_ORDER_ORIGINAL || order == SDP_ORDER_PLAN, "sdp_grid: bad order");
    if (plan->n_chunks == 0) return SDP_OK;
    SDP_REQUIRE(d_kernels && d_vis, "sdp_grid_accumulate: null argument");
    GridArgs a{plan->p.grid_size, plan->p.oversampling, plan->tile, plan->ntx, plan->S, plan->stride};
    const size_t lds = sizeof(float4) * (size_t)plan->S * plan->stride;
    static bool lds_set = false;
    if (!lds_set) {
        (void)hipFuncSetAttribute((const void *)grid_chunk_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                                  160 * 1024 - (int)(sizeof(VisRec) + sizeof(float2)) * kStage);
        lds_set = true;
    }
    grid_chunk_kernel<<<plan->n_chunks, kGridThreads, lds, sdp::stream_of(s)>>>(
        plan->recs, (const float2 *)d_vis, order == SDP_ORDER_ORIGINAL ? plan->perm : nullptr, plan->plane_tab,
        (const float2 *)d_kernels, plan->chunks, a, plan->slabs);
    SDP_LAUNCH_CHECK();
    return SDP_OK;
}

int sdp_grid_merge(const sdp_grid_plan *plan, float *d_grid, sdp_stream s)
{
    SDP_REQUIRE(plan && d_grid, "sdp_grid_merge: null argument");
    const int G = plan->p.grid_size;
    MergeArgs m{G, plan->tile, plan->ntx, plan->nty, plan->S};
    dim3 mg(sdp_div_up(G, 256), G);
    merge_slabs_kernel<<<mg, 256, 0, sdp::stream_of(s)>>>(plan->slabs, plan->tile_chunk_begin, m, (float2 *)d_grid);
    SDP_LAUNCH_CHECK();
    return SDP_OK;
}

int sdp_grid(const sdp_grid_plan *plan, const float *d_kernels, const float *d_vis, int order, float *d_grid,
             sdp_stream s)
{
    if (int e = sdp_grid_accumulate(plan, d_kernels, d_vis, order, s)) return e;
    return sdp_grid_merge(plan, d_grid, s);
}

int sdp_degrid(const sdp_grid_plan *plan, const float *d_kernels, const float *d_grid, const float *d_vis_measured,
               int order, float *d_vis_out, sdp_stream s)
{
    SDP_REQUIRE(plan && d_grid && d_vis_out, "sdp_degrid: null argument");
    SDP_REQUIRE(order == SDP_ORDER_ORIGINAL || order == SDP_ORDER_PLAN, "sdp_degrid: bad order");
    hipStream_t st = sdp::stream_of(s);
    const uint32_t *perm = order == SDP_ORDER_ORIGINAL ? plan->perm : nullptr;
    GridArgs a{plan->p.grid_size, plan->p.oversampling, plan->tile, plan->ntx, plan->S, plan->stride};
    if (plan->n_chunks > 0) {
        const size_t lds = sizeof(float2) * (size_t)plan->S * plan->stride;
        degrid_chunk_kernel<<<plan->n_chunks, kDegridThreads, lds, st>>>(
            plan->recs, perm, plan->plane_tab, (const float2 *)d_kernels, plan->chunks, a, (const float2 *)d_grid,
            (const float2 *)d_vis_measured, (float2 *)d_vis_out);
        SDP_LAUNCH_CHECK();
    }
    if (plan->n > plan->n_kept) {
        degrid_dropped_kernel<<<sdp_div_up(plan->n - plan->n_kept, 256), 256, 0, st>>>(
            perm, plan->n_kept, plan->n, (const float2 *)d_vis_measured, (float2 *)d_vis_out);
        SDP_LAUNCH_CHECK();
    }
    return SDP_OK;
}

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
    SIMP_TRY(hipcub::DeviceRadixSort::SortPairs(nullptr, tb1, keys, skeys, vals, svals, n_vis, 0, 32, st));
    SIMP_TRY(hipcub::DeviceScan::ExclusiveSum(nullptr, tb2, first, pos, n_vis, st));
    SIMP_TRY(hipMalloc(&temp, std::max(tb1, tb2)));
    SIMP_TRY(hipcub::DeviceRadixSort::SortPairs(temp, tb1, keys, skeys, vals, svals, n_vis, 0, 32, st));
    simplify_heads_kernel<<<sdp_div_up(n, 256), 256, 0, st>>>(n_vis, skeys, svals, first);
    SIMP_TRY(hipGetLastError());
    SIMP_TRY(hipcub::DeviceScan::ExclusiveSum(temp, tb2, first, pos, n_vis, st));
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
