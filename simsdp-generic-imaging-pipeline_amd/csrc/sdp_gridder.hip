// sdp_gridder.hip -- convolutional gridding and degridding on gfx950.
//
// Reference arithmetic: std_gridding (degridgrid.c:69-125), the kernel-
// weighted sum of std_degridding (degridgrid.c:147-186), gridding_CPU
// (gridding_run.c:55-137) and subtract_from_measurements_finegrid
// (gains_apply_run.c:133-141); paths relative to
// preesm_pipelines/simsdp_fft_imaging_pipeline/Code/src/CPU/.
//
// Plan (once per set of coordinates)
//   One thread per visibility repeats the reference's float/double index
//   arithmetic (w-plane, grid position, footprint bounds), clips the
//   footprint to the grid and writes a 16-byte record.  Visibilities are
//   stable-sorted by (support class, uv tile of the footprint's first cell).
//   A support class groups w-planes of similar half support so that its
//   work items can use an LDS window sized for that class:
//   S = tile + 2 * (largest half support in the class).  Each (class, tile)
//   list is cut into work items ("chunks") of about `chunk_taps` taps.
//
// Kernels: one wavefront per chunk.
//   grid    The wavefront owns a private S x S LDS sub-grid of compensated
//           accumulators (sum and TwoSum error per component).  It takes the
//           chunk's visibilities in order; for each one its 64 lanes cover
//           the footprint's taps (lane = one column, consecutive lanes
//           consecutive columns, rows strided), so no two lanes touch one
//           cell within a visibility and LDS order serialises visibilities:
//           no atomics, deterministic, every lane busy.  The sub-grid is
//           written to the chunk's slab; a merge pass sums, per grid cell,
//           the slabs covering it in a fixed order (compensated) and writes
//           the whole grid once.
//   degrid  The wavefront stages the grid's S x S window in LDS; lanes sweep
//           the taps with the same mapping and a shuffle tree reduces them.
//   degrid+subtract+grid (major cycle): both of the above per visibility
//           in one pass: predict, subtract from the measured value, grid the
//           residual into the sub-grid.
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cstring>
#include <vector>

#include "sdp_internal.h"

namespace {

constexpr int kWave = 64;
constexpr int kBatch = 8;   // taps per lane whose kernel samples are loaded together
// half-support limits of the support classes; a class's window is
// tile + 2 * (largest half support present in it)
constexpr int kNumClasses = 6;
constexpr int kClassMaxH[kNumClasses] = {8, 12, 16, 20, 24, 32};
// LDS of one workgroup (one wavefront): accumulators, or accumulators plus
// the model-grid window for the fused kernel
constexpr size_t kMaxLds = 152 * 1024;

enum { MODE_GRID = 0, MODE_DEGRID = 1, MODE_DEGRID_GRID = 2 };

struct VisRec {
    float gx, gy;     // grid position in cells (reference grid_pos / snapped coordinate)
    uint32_t start;   // first footprint cell after clipping: u | v << 16 (0-based)
    uint32_t meta;    // nu | nv << 8 | plane << 16 (15 bits) | conj << 31
};
static_assert(sizeof(VisRec) == 16, "VisRec is 16 bytes");

int class_of(int h)
{
    for (int c = 0; c < kNumClasses; ++c)
        if (h <= kClassMaxH[c]) return c;
    return -1;
}

// Knuth TwoSum: s + e == a + b exactly.
__device__ __forceinline__ void two_sum_acc(float &s, float &c, float b)
{
    const float t = s + b;
    const float bb = t - s;
    const float err = (s - (t - bb)) + (b - bb);
    s = t;
    c += err;
}

__device__ __forceinline__ int kernel_tap(int cell_abs, float pos, float os)
{
    // abs((int)ROUND((v - grid_pos) * OVERSAMPLING_FACTOR)), degridgrid.c:101,110
    return abs((int)roundf(((float)cell_abs - pos) * os));
}

__device__ __forceinline__ float wave_sum(float v)
{
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
    return v;
}

// ---- plan construction ---------------------------------------------------

struct PlanArgs {
    int n, G, OS, n_kernels, snap_mode, tile, ntx, nty, ntiles, nkeys;
    double uv_scale, w_scale;
};

__global__ void plan_records_kernel(const float *uvw, const int2 *supports, const int *plane_class, PlanArgs a,
                                    VisRec *recs, uint32_t *keys, uint32_t *vals, unsigned int *key_count,
                                    unsigned long long *key_taps)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.n) return;
    const float x = uvw[3 * (size_t)i], y = uvw[3 * (size_t)i + 1], z = uvw[3 * (size_t)i + 2];
    const int gc = a.G / 2;
    keys[i] = (uint32_t)a.nkeys;   // dropped: sorts last
    vals[i] = (uint32_t)i;

    const float wabs = fabsf((float)((double)z * a.w_scale));
    const float root = __fsqrt_rn(wabs);
    int plane;
    if (a.snap_mode == SDP_SNAP_STD)
        plane = (int)((double)root + 0.5);          // degridgrid.c:84
    else
        plane = (int)roundf(root);                   // gridding_run.c:67
    if (!(root == root) || plane < 0 || plane >= a.n_kernels) return;
    const int h = supports[plane].x;
    const int cls = plane_class[plane];
    if (cls < 0) return;

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
    const int ua = max((int)lo_x + gc, 0), ub = min((int)hi_x + gc, a.G);
    const int va = max((int)lo_y + gc, 0), vb = min((int)hi_y + gc, a.G);
    if (ua >= ub || va >= vb) return;
    const int nu = ub - ua, nv = vb - va;
    VisRec r;
    r.gx = gx;
    r.gy = gy;
    r.start = (uint32_t)ua | ((uint32_t)va << 16);
    r.meta = (uint32_t)nu | ((uint32_t)nv << 8) | ((uint32_t)plane << 16) | ((z < 0.0f) ? 0x80000000u : 0u);
    recs[i] = r;
    const uint32_t key = (uint32_t)(cls * a.ntiles + (va / a.tile) * a.ntx + (ua / a.tile));
    keys[i] = key;
    atomicAdd(&key_count[key], 1u);
    atomicAdd(&key_taps[key], (unsigned long long)(nu * nv));
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

// ---- chunk kernels -------------------------------------------------------

struct ChunkArgs {
    int G, OS, tile, ntx, chunk_base;
};

struct ChunkIO {
    const float2 *kg;          // gridding kernel table
    const float2 *kd;          // degridding kernel table
    const float2 *vis;         // MODE_GRID: visibilities to grid
    const float2 *measured;    // MODE_DEGRID (optional) / MODE_DEGRID_GRID: measured visibilities
    const float2 *grid;        // MODE_DEGRID / MODE_DEGRID_GRID: model grid
    float2 *vis_out;           // MODE_DEGRID: prediction or residual; MODE_DEGRID_GRID: residual (optional)
    float2 *slabs;             // MODE_GRID / MODE_DEGRID_GRID
};

// Per visibility: lanes cover the footprint's taps.  With nu <= 64 columns,
// lane l takes column l % nu and rows l / nu, l / nu + 64 / nu, ... (lanes
// past (64 / nu) * nu idle); wider footprints loop over column blocks.
struct TapMap {
    int col0, row0, row_step, ncol_blocks;
    bool active;
};

__device__ __forceinline__ TapMap tap_map(int lane, int nu)
{
    TapMap m;
    if (nu <= kWave) {
        // (k + 0.5) / nu is at least 0.5 / nu away from an integer, far more
        // than the error of the hardware reciprocal, so these floors are exact
        const float rnu = __builtin_amdgcn_rcpf((float)nu);
        const int rps = (int)(((float)kWave + 0.5f) * rnu);     // floor(64 / nu)
        const int lr = (int)(((float)lane + 0.5f) * rnu);       // floor(lane / nu)
        m.col0 = lane - lr * nu;
        m.row0 = lr;
        m.row_step = rps;
        m.active = lr < rps;
        m.ncol_blocks = 1;
    } else {
        m.col0 = lane;
        m.row0 = 0;
        m.row_step = 1;
        m.active = true;
        m.ncol_blocks = (nu + kWave - 1) / kWave;
    }
    return m;
}

template <int MODE>
__global__ void __launch_bounds__(kWave)
chunk_kernel(const VisRec *__restrict__ recs, const uint32_t *__restrict__ perm, const int2 *__restrict__ plane_tab,
             const int4 *__restrict__ chunks, const long long *__restrict__ slab_off, ChunkArgs a, ChunkIO io)
{
    extern __shared__ float4 lds4[];
    const int4 ch = chunks[a.chunk_base + blockIdx.x];   // tile, first record, count, window S
    const int S = ch.w, stride = S + 1;
    const int lane = threadIdx.x;
    const int tile_u0 = (ch.x % a.ntx) * a.tile, tile_v0 = (ch.x / a.ntx) * a.tile;
    const int gc = a.G / 2;
    const float os = (float)a.OS;
    constexpr bool kGrid = MODE != MODE_DEGRID;
    constexpr bool kDegrid = MODE != MODE_GRID;

    float4 *acc = lds4;                                                             // S x stride accumulators
    float2 *win = (float2 *)(kGrid ? (void *)(lds4 + S * stride) : (void *)lds4);  // S x stride model window
    if (kGrid)
        for (int i = lane; i < S * stride; i += kWave) acc[i] = make_float4(0.f, 0.f, 0.f, 0.f);
    if (kDegrid) {
        for (int i = lane; i < S * S; i += kWave) {
            const int row = i / S, col = i - row * S;
            const int gy = tile_v0 + row, gx = tile_u0 + col;
            float2 g = make_float2(0.f, 0.f);
            if (gy < a.G && gx < a.G) g = io.grid[(size_t)gy * a.G + gx];
            win[row * stride + col] = g;
        }
    }
    __syncthreads();

    for (int k = 0; k < ch.z; ++k) {
        const int j = ch.y + k;
        const VisRec r = recs[j];
        const uint32_t o = perm ? perm[j] : (uint32_t)j;
        const int lu0 = (int)(r.start & 0xffffu) - tile_u0, lv0 = (int)(r.start >> 16) - tile_v0;
        const int nu = r.meta & 0xffu, nv = (r.meta >> 8) & 0xffu;
        const int2 pt = plane_tab[(r.meta >> 16) & 0x7fffu];
        const float conj = (r.meta & 0x80000000u) ? -1.0f : 1.0f;
        const TapMap m = tap_map(lane, nu);
        const int cell_u0 = tile_u0 + lu0 - gc, cell_v0 = tile_v0 + lv0 - gc;

        float2 v = make_float2(0.f, 0.f);
        if (MODE == MODE_GRID) v = io.vis[o];
        if (kDegrid) {
            float sr = 0.f, si = 0.f;
            if (m.active) {
                for (int cb = 0; cb < m.ncol_blocks; ++cb) {
                    const int col = m.col0 + cb * kWave;
                    if (col >= nu) break;
                    const int kx = kernel_tap(cell_u0 + col, r.gx, os);
                    const float2 *kcol = io.kd + pt.x + kx;
                    const float2 *wcol = win + lv0 * stride + lu0 + col;
                    for (int row0 = m.row0; row0 < nv; row0 += kBatch * m.row_step) {
                        float2 kk[kBatch], g[kBatch];
#pragma unroll
                        for (int q = 0; q < kBatch; ++q) {
                            const int row = row0 + q * m.row_step;
                            if (row < nv) {
                                kk[q] = kcol[kernel_tap(cell_v0 + row, r.gy, os) * pt.y];
                                g[q] = wcol[row * stride];
                            }
                        }
#pragma unroll
                        for (int q = 0; q < kBatch; ++q) {
                            if (row0 + q * m.row_step < nv) {
                                const float kyc = conj * kk[q].y;
                                sr = fmaf(g[q].x, kk[q].x, fmaf(-g[q].y, kyc, sr));
                                si = fmaf(g[q].y, kk[q].x, fmaf(g[q].x, kyc, si));
                            }
                        }
                    }
                }
            }
            sr = wave_sum(sr);
            si = wave_sum(si);
            if (MODE == MODE_DEGRID) {
                if (lane == 0) {
                    float2 res = make_float2(sr, si);
                    if (io.measured) {
                        const float2 mv = io.measured[o];
                        res = make_float2(mv.x - sr, mv.y - si);
                    }
                    io.vis_out[o] = res;
                }
                continue;
            }
            const float2 mv = io.measured[o];
            v = make_float2(mv.x - sr, mv.y - si);
            if (io.vis_out && lane == 0) io.vis_out[o] = v;
        }
        if (kGrid && m.active) {
            const float vr = v.x, vi = v.y, vrc = conj * v.x, vic = conj * v.y;
            for (int cb = 0; cb < m.ncol_blocks; ++cb) {
                const int col = m.col0 + cb * kWave;
                if (col >= nu) break;
                const int kx = kernel_tap(cell_u0 + col, r.gx, os);
                const float2 *kcol = io.kg + pt.x + kx;
                float4 *acol = acc + lv0 * stride + lu0 + col;
                for (int row0 = m.row0; row0 < nv; row0 += kBatch * m.row_step) {
                    float2 kk[kBatch];
#pragma unroll
                    for (int q = 0; q < kBatch; ++q) {
                        const int row = row0 + q * m.row_step;
                        if (row < nv) kk[q] = kcol[kernel_tap(cell_v0 + row, r.gy, os) * pt.y];
                    }
                    // distinct rows: the cells of one batch never alias, so all
                    // reads are issued before the writes
                    float4 cell[kBatch];
#pragma unroll
                    for (int q = 0; q < kBatch; ++q)
                        if (row0 + q * m.row_step < nv) cell[q] = acol[(row0 + q * m.row_step) * stride];
#pragma unroll
                    for (int q = 0; q < kBatch; ++q) {
                        if (row0 + q * m.row_step < nv) {
                            // vis * (k.x, conj * k.y)
                            const float px = fmaf(vr, kk[q].x, -(vic * kk[q].y));
                            const float py = fmaf(vi, kk[q].x, vrc * kk[q].y);
                            two_sum_acc(cell[q].x, cell[q].z, px);
                            two_sum_acc(cell[q].y, cell[q].w, py);
                        }
                    }
#pragma unroll
                    for (int q = 0; q < kBatch; ++q)
                        if (row0 + q * m.row_step < nv) acol[(row0 + q * m.row_step) * stride] = cell[q];
                }
            }
        }
    }
    if (kGrid) {
        __syncthreads();
        float2 *slab = io.slabs + slab_off[a.chunk_base + blockIdx.x];
        for (int i = lane; i < S * S; i += kWave) {
            const int row = i / S, col = i - row * S;
            const float4 c = acc[row * stride + col];
            slab[i] = make_float2(c.x + c.z, c.y + c.w);
        }
    }
}

struct MergeArgs {
    int G, tile, ntx, nty, smax;
};

constexpr int kMergeThreads = 256;
constexpr int kMergeList = 512;    // key-slab descriptors staged per pass
constexpr int kMergeTiles = 256;   // covering tiles per output tile (checked by the plan)

// Sum the slabs of a (class, tile) key's pieces, in piece order, into the
// key's slab.  One workgroup per key with more than one piece; each thread
// owns cells and issues the pieces' loads in groups of 8.
__global__ void __launch_bounds__(kMergeThreads)
reduce_pieces_kernel(const float2 *__restrict__ slabs, const long long *__restrict__ slab_off,
                     const int4 *__restrict__ multi, float2 *__restrict__ key_slabs)
{
    const int4 k = multi[blockIdx.x];   // first chunk, piece count, S*S, key slab start (cells, / 1)
    const long long out0 = ((const long long *)(multi + gridDim.x))[blockIdx.x];
    for (int i = threadIdx.x; i < k.z; i += kMergeThreads) {
        float sr = 0.f, si = 0.f, cr = 0.f, ci = 0.f;
        for (int p0 = 0; p0 < k.y; p0 += 8) {
            float2 v[8];
#pragma unroll
            for (int q = 0; q < 8; ++q)
                if (p0 + q < k.y) v[q] = slabs[slab_off[k.x + p0 + q] + i];
#pragma unroll
            for (int q = 0; q < 8; ++q)
                if (p0 + q < k.y) {
                    two_sum_acc(sr, cr, v[q].x);
                    two_sum_acc(si, ci, v[q].y);
                }
        }
        key_slabs[out0 + i] = make_float2(sr + cr, si + ci);
    }
}

// One workgroup per output tile: grid cell = compensated sum of the key
// slabs covering it, in (tile row, tile column, class) order.  The
// workgroup loads the covering keys' descriptors into LDS in parallel; with
// tiles of at most 128 cells the threads split the descriptors into
// interleaved groups whose partial sums are combined in group order.
// Writes every cell.
__global__ void __launch_bounds__(kMergeThreads)
merge_slabs_kernel(const float2 *__restrict__ kslabs, const int *__restrict__ key_S,
                   const long long *__restrict__ key_off, const int *__restrict__ tile_key_begin,
                   MergeArgs a, float2 *__restrict__ grid)
{
    constexpr int kCells = 4;   // output cells per thread when a tile has more than 256 cells
    __shared__ long long s_off[kMergeList];
    __shared__ int s_geo[kMergeList];          // S | by << 10 | bx << 20
    __shared__ int s_base[kMergeTiles + 1];    // first list position of each covering tile
    __shared__ int s_k0[kMergeTiles];          // its first key
    __shared__ int s_b[kMergeTiles];           // by | bx << 10 for that tile
    __shared__ float4 s_part[kMergeThreads];
    const int oty = blockIdx.y, otx = blockIdx.x;
    const int oy0 = oty * a.tile, ox0 = otx * a.tile;
    const int ncell = a.tile * a.tile;
    const int groups = ncell <= 128 ? kMergeThreads / ncell : 1;   // ncell divides 256 when <= 128? not needed
    const int grp = ncell <= 128 ? threadIdx.x / ncell : 0;
    const bool in_group = ncell <= 128 ? grp < groups && threadIdx.x < groups * ncell : true;
    float sr[kCells], si[kCells], cr[kCells], ci[kCells];
    int dy[kCells], dx[kCells];
    bool live[kCells];
#pragma unroll
    for (int q = 0; q < kCells; ++q) {
        const int i = ncell <= 128 ? (q == 0 ? threadIdx.x - grp * ncell : ncell) : threadIdx.x + q * kMergeThreads;
        dy[q] = i / a.tile;
        dx[q] = i - dy[q] * a.tile;
        live[q] = in_group && i < ncell && oy0 + dy[q] < a.G && ox0 + dx[q] < a.G;
        sr[q] = si[q] = cr[q] = ci[q] = 0.f;
    }
    const int ty_lo = max(0, (oy0 - a.smax + a.tile) / a.tile);
    const int tx_lo = max(0, (ox0 - a.smax + a.tile) / a.tile);
    const int ntx_c = otx - tx_lo + 1;
    const int nt = (oty - ty_lo + 1) * ntx_c;
    for (int k = threadIdx.x; k < nt; k += kMergeThreads) {
        const int ty = ty_lo + k / ntx_c, tx = tx_lo + k % ntx_c;
        const int t = ty * a.ntx + tx;
        s_k0[k] = tile_key_begin[t];
        s_base[k + 1] = tile_key_begin[t + 1] - tile_key_begin[t];
        s_b[k] = (oy0 - ty * a.tile) | ((ox0 - tx * a.tile) << 10);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        s_base[0] = 0;
        for (int k = 0; k < nt; ++k) s_base[k + 1] += s_base[k];
    }
    __syncthreads();
    const int total = s_base[nt];
    for (int base = 0; base < total; base += kMergeList) {
        const int n = min(kMergeList, total - base);
        for (int i = threadIdx.x; i < n; i += kMergeThreads) {
            const int g = base + i;
            int lo = 0, hi = nt - 1;   // covering tile k with s_base[k] <= g < s_base[k + 1]
            while (lo < hi) {
                const int mid = (lo + hi + 1) >> 1;
                if (s_base[mid] <= g) lo = mid; else hi = mid - 1;
            }
            const int key = s_k0[lo] + g - s_base[lo];
            s_off[i] = key_off[key];
            s_geo[i] = key_S[key] | (s_b[lo] << 10);
        }
        __syncthreads();
        for (int e = grp; e < n; e += groups) {
            const int geo = s_geo[e];
            const int S = geo & 1023, by = (geo >> 10) & 1023, bx = (geo >> 20) & 1023;
            const float2 *slab = kslabs + s_off[e];
#pragma unroll
            for (int q = 0; q < kCells; ++q) {
                const int ly = by + dy[q], lx = bx + dx[q];
                if (live[q] && ly < S && lx < S) {
                    const float2 v = slab[ly * S + lx];
                    two_sum_acc(sr[q], cr[q], v.x);
                    two_sum_acc(si[q], ci[q], v.y);
                }
            }
        }
        __syncthreads();
    }
    if (ncell <= 128) {
        // combine the groups' partial sums in group order
        s_part[threadIdx.x] = make_float4(sr[0], si[0], cr[0], ci[0]);
        __syncthreads();
        if (grp == 0 && live[0]) {
            float tr = 0.f, ti = 0.f, er = 0.f, ei = 0.f;
            for (int g = 0; g < groups; ++g) {
                const float4 p = s_part[g * ncell + threadIdx.x];
                two_sum_acc(tr, er, p.x);
                two_sum_acc(ti, ei, p.y);
                two_sum_acc(tr, er, p.z);
                two_sum_acc(ti, ei, p.w);
            }
            grid[(size_t)(oy0 + dy[0]) * a.G + ox0 + dx[0]] = make_float2(tr + er, ti + ei);
        }
        return;
    }
#pragma unroll
    for (int q = 0; q < kCells; ++q)
        if (live[q]) grid[(size_t)(oy0 + dy[q]) * a.G + ox0 + dx[q]] = make_float2(sr[q] + cr[q], si[q] + ci[q]);
}

// visibilities the plan dropped (outside the grid): prediction 0
__global__ void degrid_dropped_kernel(const uint32_t *perm, int first, int n, const float2 *measured, float2 *out)
{
    int j = first + blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    const uint32_t o = perm ? perm[j] : (uint32_t)j;
    out[o] = measured ? measured[o] : make_float2(0.f, 0.f);
}

}  // namespace

// ---- plan object -----------------------------------------------------------

struct sdp_grid_plan {
    sdp_grid_params p;
    int n = 0, n_kept = 0, tile = 0, chunk_taps = 0, ntx = 0, nty = 0, H = 0, smax = 0;
    int n_chunks = 0, n_tiles_used = 0, n_class_launches = 0;
    long long total_taps = 0;
    int class_S[kNumClasses] = {};
    int class_chunk_begin[kNumClasses + 1] = {};
    VisRec *recs = nullptr;        // plan order
    uint32_t *perm = nullptr;      // plan position -> caller index
    int2 *plane_tab = nullptr;     // (sample offset, (h+1)*OS) per plane
    int4 *chunks = nullptr;        // (tile, first, count, S), class-major
    long long *slab_off = nullptr; // first slab cell of each chunk
    float2 *slabs = nullptr;       // one slab per chunk
    size_t slab_bytes = 0;
    // per (class, tile) key that holds visibilities: window S and slab
    // (a chunk slab when the key has one piece, else a key_slabs entry),
    // grouped by tile in class order
    int n_keys_used = 0, n_multi = 0;
    int *key_S = nullptr;
    long long *key_off = nullptr;  // into key_base (see key_base_of)
    int *tile_key_begin = nullptr;
    int4 *multi = nullptr;         // multi-piece keys: first chunk, pieces, S*S, -, then their output starts
    float2 *key_slabs = nullptr;   // contiguous with slabs: slabs followed by the reduced key slabs
};

namespace {

void plan_free(sdp_grid_plan *p)
{
    if (!p) return;
    (void)hipFree(p->recs);
    (void)hipFree(p->perm);
    (void)hipFree(p->plane_tab);
    (void)hipFree(p->chunks);
    (void)hipFree(p->slab_off);
    (void)hipFree(p->slabs);
    (void)hipFree(p->key_S);
    (void)hipFree(p->key_off);
    (void)hipFree(p->tile_key_begin);
    (void)hipFree(p->multi);
    delete p;
}

size_t lds_bytes(int mode, int S)
{
    const size_t cells = (size_t)S * (S + 1);
    if (mode == MODE_GRID) return cells * sizeof(float4);
    if (mode == MODE_DEGRID) return cells * sizeof(float2);
    return cells * (sizeof(float4) + sizeof(float2));
}

template <int MODE>
int launch_chunks(const sdp_grid_plan *plan, const ChunkIO &io, const uint32_t *perm, hipStream_t st)
{
    static bool attr_set = false;
    if (!attr_set) {
        (void)hipFuncSetAttribute((const void *)chunk_kernel<MODE>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                  (int)kMaxLds);
        attr_set = true;
    }
    for (int c = 0; c < kNumClasses; ++c) {
        const int b = plan->class_chunk_begin[c], e = plan->class_chunk_begin[c + 1];
        if (e <= b) continue;
        ChunkArgs a{plan->p.grid_size, plan->p.oversampling, plan->tile, plan->ntx, b};
        chunk_kernel<MODE><<<e - b, kWave, lds_bytes(MODE, plan->class_S[c]), st>>>(
            plan->recs, perm, plan->plane_tab, plan->chunks, plan->slab_off, a, io);
        SDP_LAUNCH_CHECK();
    }
    return SDP_OK;
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
    SDP_REQUIRE(p.tile >= 0 && p.tile <= 32 && p.chunk >= 0, "tile must be in [1, 32] (0 = default)");
    hipStream_t st = sdp::stream_of(s);

    std::vector<int2> tab(p.n_kernels), sup(p.n_kernels);
    std::vector<int> pclass(p.n_kernels);
    int class_h[kNumClasses] = {-1, -1, -1, -1, -1, -1};
    int H = 0;
    for (int k = 0; k < p.n_kernels; ++k) {
        const int h = supports[2 * k], off = supports[2 * k + 1];
        SDP_REQUIRE(h >= 0, "negative half support for plane %d", k);
        if (h > kClassMaxH[kNumClasses - 1])
            return sdp::set_error(SDP_E_UNSUPPORTED, "half support %d of plane %d exceeds %d", h, k,
                                  kClassMaxH[kNumClasses - 1]);
        H = std::max(H, h);
        pclass[k] = class_of(h);
        class_h[pclass[k]] = std::max(class_h[pclass[k]], h);
        tab[k] = make_int2(off, (h + 1) * p.oversampling);
        sup[k] = make_int2(h, off);
    }

    sdp_grid_plan *plan = new sdp_grid_plan();
    plan->p = p;
    plan->n = n_vis;
    plan->H = H;
    plan->tile = std::min(p.tile > 0 ? p.tile : 8, p.grid_size);
    plan->chunk_taps = p.chunk > 0 ? p.chunk : (1 << 17);
    plan->ntx = (p.grid_size + plan->tile - 1) / plan->tile;
    plan->nty = plan->ntx;
    for (int c = 0; c < kNumClasses; ++c) {
        plan->class_S[c] = class_h[c] >= 0 ? plan->tile + 2 * class_h[c] : 0;
        plan->smax = std::max(plan->smax, plan->class_S[c]);
        if (class_h[c] >= 0 && lds_bytes(MODE_DEGRID_GRID, plan->class_S[c]) > kMaxLds) {
            const int S = plan->class_S[c], tile = plan->tile;
            delete plan;
            return sdp::set_error(SDP_E_UNSUPPORTED, "window of %d cells (tile %d, half support %d) exceeds LDS", S,
                                  tile, class_h[c]);
        }
    }
    const int ntiles = plan->ntx * plan->nty;
    const int nkeys = kNumClasses * ntiles;
    {
        const int span = (plan->smax + 2 * plan->tile - 1) / plan->tile;   // covering tiles per axis, upper bound
        if (span * span > kMergeTiles) {
            const int tile = plan->tile;
            delete plan;
            return sdp::set_error(SDP_E_UNSUPPORTED, "tile %d too small for half support %d", tile, H);
        }
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
    int *d_pclass = nullptr;
    PLAN_TRY(hipMalloc(&d_supports, sizeof(int2) * p.n_kernels));
    PLAN_TRY(hipMemcpyAsync(d_supports, sup.data(), sizeof(int2) * p.n_kernels, hipMemcpyHostToDevice, st));
    PLAN_TRY(hipMalloc(&d_pclass, sizeof(int) * p.n_kernels));
    PLAN_TRY(hipMemcpyAsync(d_pclass, pclass.data(), sizeof(int) * p.n_kernels, hipMemcpyHostToDevice, st));

    const size_t n = (size_t)std::max(n_vis, 1);
    VisRec *recs_tmp = nullptr;
    uint32_t *keys = nullptr, *keys_sorted = nullptr, *vals = nullptr;
    unsigned int *key_count = nullptr;
    unsigned long long *key_taps = nullptr;
    PLAN_TRY(hipMalloc(&recs_tmp, sizeof(VisRec) * n));
    PLAN_TRY(hipMalloc(&plan->recs, sizeof(VisRec) * n));
    PLAN_TRY(hipMalloc(&plan->perm, sizeof(uint32_t) * n));
    PLAN_TRY(hipMalloc(&keys, sizeof(uint32_t) * n));
    PLAN_TRY(hipMalloc(&keys_sorted, sizeof(uint32_t) * n));
    PLAN_TRY(hipMalloc(&vals, sizeof(uint32_t) * n));
    PLAN_TRY(hipMalloc(&key_count, sizeof(unsigned int) * nkeys));
    PLAN_TRY(hipMalloc(&key_taps, sizeof(unsigned long long) * nkeys));
    PLAN_TRY(hipMemsetAsync(key_count, 0, sizeof(unsigned int) * nkeys, st));
    PLAN_TRY(hipMemsetAsync(key_taps, 0, sizeof(unsigned long long) * nkeys, st));

    if (n_vis > 0) {
        PlanArgs a{n_vis, p.grid_size, p.oversampling, p.n_kernels, p.snap_mode, plan->tile, plan->ntx, plan->nty,
                   ntiles, nkeys, p.uv_scale, p.w_scale};
        plan_records_kernel<<<sdp_div_up(n_vis, 256), 256, 0, st>>>(d_uvw, d_supports, d_pclass, a, recs_tmp, keys,
                                                                    vals, key_count, key_taps);
        PLAN_TRY(hipGetLastError());
        int end_bit = 1;
        while ((1u << end_bit) <= (unsigned)nkeys) ++end_bit;
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

    std::vector<unsigned int> counts(nkeys);
    std::vector<unsigned long long> taps(nkeys);
    PLAN_TRY(hipMemcpyAsync(counts.data(), key_count, sizeof(unsigned int) * nkeys, hipMemcpyDeviceToHost, st));
    PLAN_TRY(hipMemcpyAsync(taps.data(), key_taps, sizeof(unsigned long long) * nkeys, hipMemcpyDeviceToHost, st));
    PLAN_TRY(hipStreamSynchronize(st));
    (void)hipFree(recs_tmp);
    (void)hipFree(keys);
    (void)hipFree(keys_sorted);
    (void)hipFree(vals);
    (void)hipFree(key_count);
    (void)hipFree(key_taps);
    (void)hipFree(d_supports);
    (void)hipFree(d_pclass);

    // work items, class-major then tile-major, each (class, tile) list cut
    // into pieces of about chunk_taps taps
    std::vector<int4> chunks;
    std::vector<long long> slab_off;
    std::vector<std::vector<int>> tile_keys(ntiles);    // per tile: key indices, class order
    std::vector<int> key_S;
    std::vector<long long> key_first_chunk, key_pieces;
    std::vector<char> tile_used(ntiles, 0);
    long long cells = 0;
    int first = 0;
    for (int c = 0; c < kNumClasses; ++c) {
        plan->class_chunk_begin[c] = (int)chunks.size();
        for (int t = 0; t < ntiles; ++t) {
            const int key = c * ntiles + t;
            const int cnt = (int)counts[key];
            if (cnt == 0) continue;
            tile_used[t] = 1;
            long long pieces = (long long)((taps[key] + plan->chunk_taps - 1) / plan->chunk_taps);
            pieces = std::max(1LL, std::min(pieces, (long long)cnt));
            tile_keys[t].push_back((int)key_S.size());
            key_S.push_back(plan->class_S[c]);
            key_first_chunk.push_back((long long)chunks.size());
            key_pieces.push_back(pieces);
            for (long long q = 0; q < pieces; ++q) {
                const int b = first + (int)(cnt * q / pieces), e = first + (int)(cnt * (q + 1) / pieces);
                chunks.push_back(make_int4(t, b, e - b, plan->class_S[c]));
                slab_off.push_back(cells);
                cells += (long long)plan->class_S[c] * plan->class_S[c];
            }
            first += cnt;
        }
    }
    plan->class_chunk_begin[kNumClasses] = (int)chunks.size();
    for (int c = 0; c < kNumClasses; ++c)
        plan->n_class_launches += plan->class_chunk_begin[c + 1] > plan->class_chunk_begin[c];
    for (int k = 0; k < nkeys; ++k) plan->total_taps += (long long)taps[k];
    plan->n_kept = first;
    plan->n_chunks = (int)chunks.size();
    for (int t = 0; t < ntiles; ++t) plan->n_tiles_used += tile_used[t];

    // key slabs: single-piece keys use their chunk's slab; multi-piece keys
    // get a reduced slab placed after the chunk slabs
    const int nk = (int)key_S.size();
    std::vector<long long> key_off(nk);
    std::vector<int4> multi;
    std::vector<long long> multi_out;
    long long extra = 0;
    for (int k = 0; k < nk; ++k) {
        const long long S2 = (long long)key_S[k] * key_S[k];
        if (key_pieces[k] == 1) {
            key_off[k] = slab_off[key_first_chunk[k]];
        } else {
            key_off[k] = cells + extra;
            multi.push_back(make_int4((int)key_first_chunk[k], (int)key_pieces[k], (int)S2, 0));
            multi_out.push_back(cells + extra);
            extra += S2;
        }
    }
    std::vector<int> tile_key_begin(ntiles + 1, 0), tile_key_list;
    tile_key_list.reserve(nk);
    for (int t = 0; t < ntiles; ++t) {
        tile_key_begin[t] = (int)tile_key_list.size();
        tile_key_list.insert(tile_key_list.end(), tile_keys[t].begin(), tile_keys[t].end());
    }
    tile_key_begin[ntiles] = (int)tile_key_list.size();
    // key tables in tile order so that a tile's keys are contiguous
    std::vector<int> kS_t(nk);
    std::vector<long long> kOff_t(nk);
    for (int i = 0; i < nk; ++i) {
        kS_t[i] = key_S[tile_key_list[i]];
        kOff_t[i] = key_off[tile_key_list[i]];
    }
    plan->n_keys_used = nk;
    plan->n_multi = (int)multi.size();

    PLAN_TRY(hipMalloc(&plan->tile_key_begin, sizeof(int) * (ntiles + 1)));
    PLAN_TRY(hipMemcpyAsync(plan->tile_key_begin, tile_key_begin.data(), sizeof(int) * (ntiles + 1),
                            hipMemcpyHostToDevice, st));
    if (plan->n_chunks > 0) {
        PLAN_TRY(hipMalloc(&plan->chunks, sizeof(int4) * plan->n_chunks));
        PLAN_TRY(hipMemcpyAsync(plan->chunks, chunks.data(), sizeof(int4) * plan->n_chunks, hipMemcpyHostToDevice, st));
        PLAN_TRY(hipMalloc(&plan->slab_off, sizeof(long long) * plan->n_chunks));
        PLAN_TRY(hipMemcpyAsync(plan->slab_off, slab_off.data(), sizeof(long long) * plan->n_chunks,
                                hipMemcpyHostToDevice, st));
        PLAN_TRY(hipMalloc(&plan->key_S, sizeof(int) * nk));
        PLAN_TRY(hipMemcpyAsync(plan->key_S, kS_t.data(), sizeof(int) * nk, hipMemcpyHostToDevice, st));
        PLAN_TRY(hipMalloc(&plan->key_off, sizeof(long long) * nk));
        PLAN_TRY(hipMemcpyAsync(plan->key_off, kOff_t.data(), sizeof(long long) * nk, hipMemcpyHostToDevice, st));
        if (plan->n_multi > 0) {
            // int4 descriptors followed by their long long output starts
            const size_t bytes = sizeof(int4) * plan->n_multi + sizeof(long long) * plan->n_multi;
            PLAN_TRY(hipMalloc(&plan->multi, bytes));
            PLAN_TRY(hipMemcpyAsync(plan->multi, multi.data(), sizeof(int4) * plan->n_multi, hipMemcpyHostToDevice,
                                    st));
            PLAN_TRY(hipMemcpyAsync((char *)plan->multi + sizeof(int4) * plan->n_multi, multi_out.data(),
                                    sizeof(long long) * plan->n_multi, hipMemcpyHostToDevice, st));
        }
        plan->slab_bytes = sizeof(float2) * (size_t)(cells + extra);
        PLAN_TRY(hipMalloc(&plan->slabs, plan->slab_bytes));
        plan->key_slabs = plan->slabs;
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
    info[8] = plan->n_class_launches;
    info[9] = plan->total_taps;
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
    ChunkIO io{};
    io.kg = (const float2 *)d_kernels;
    io.vis = (const float2 *)d_vis;
    io.slabs = plan->slabs;
    return launch_chunks<MODE_GRID>(plan, io, order == SDP_ORDER_ORIGINAL ? plan->perm : nullptr,
                                    sdp::stream_of(s));
}

int sdp_grid_merge(const sdp_grid_plan *plan, float *d_grid, sdp_stream s)
{
    SDP_REQUIRE(plan && d_grid, "sdp_grid_merge: null argument");
    hipStream_t st = sdp::stream_of(s);
    const int G = plan->p.grid_size;
    if (plan->n_multi > 0) {
        reduce_pieces_kernel<<<plan->n_multi, kMergeThreads, 0, st>>>(plan->slabs, plan->slab_off, plan->multi,
                                                                       plan->key_slabs);
        SDP_LAUNCH_CHECK();
    }
    MergeArgs m{G, plan->tile, plan->ntx, plan->nty, std::max(plan->smax, 1)};
    dim3 mg(plan->ntx, plan->nty);
    merge_slabs_kernel<<<mg, kMergeThreads, 0, st>>>(plan->key_slabs, plan->key_S, plan->key_off,
                                                     plan->tile_key_begin, m, (float2 *)d_grid);
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
    if (plan->n_chunks > 0) {
        SDP_REQUIRE(d_kernels, "sdp_degrid: null kernels");
        ChunkIO io{};
        io.kd = (const float2 *)d_kernels;
        io.grid = (const float2 *)d_grid;
        io.measured = (const float2 *)d_vis_measured;
        io.vis_out = (float2 *)d_vis_out;
        if (int e = launch_chunks<MODE_DEGRID>(plan, io, perm, st)) return e;
    }
    if (plan->n > plan->n_kept) {
        degrid_dropped_kernel<<<sdp_div_up(plan->n - plan->n_kept, 256), 256, 0, st>>>(
            perm, plan->n_kept, plan->n, (const float2 *)d_vis_measured, (float2 *)d_vis_out);
        SDP_LAUNCH_CHECK();
    }
    return SDP_OK;
}

int sdp_degrid_grid_accumulate(const sdp_grid_plan *plan, const float *d_degrid_kernels,
                               const float *d_grid_kernels, const float *d_model_grid, const float *d_vis_measured,
                               int order, float *d_residual_vis, sdp_stream s)
{
    SDP_REQUIRE(plan && d_model_grid && d_vis_measured, "sdp_degrid_grid: null argument");
    SDP_REQUIRE(order == SDP_ORDER_ORIGINAL || order == SDP_ORDER_PLAN, "sdp_degrid_grid: bad order");
    hipStream_t st = sdp::stream_of(s);
    const uint32_t *perm = order == SDP_ORDER_ORIGINAL ? plan->perm : nullptr;
    if (plan->n_chunks > 0) {
        SDP_REQUIRE(d_degrid_kernels && d_grid_kernels, "sdp_degrid_grid: null kernels");
        ChunkIO io{};
        io.kd = (const float2 *)d_degrid_kernels;
        io.kg = (const float2 *)d_grid_kernels;
        io.grid = (const float2 *)d_model_grid;
        io.measured = (const float2 *)d_vis_measured;
        io.vis_out = (float2 *)d_residual_vis;
        io.slabs = plan->slabs;
        if (int e = launch_chunks<MODE_DEGRID_GRID>(plan, io, perm, st)) return e;
    }
    if (d_residual_vis && plan->n > plan->n_kept) {
        degrid_dropped_kernel<<<sdp_div_up(plan->n - plan->n_kept, 256), 256, 0, st>>>(
            perm, plan->n_kept, plan->n, (const float2 *)d_vis_measured, (float2 *)d_residual_vis);
        SDP_LAUNCH_CHECK();
    }
    return SDP_OK;
}

int sdp_degrid_grid(const sdp_grid_plan *plan, const float *d_degrid_kernels, const float *d_grid_kernels,
                    const float *d_model_grid, const float *d_vis_measured, int order, float *d_residual_vis,
                    float *d_residual_grid, sdp_stream s)
{
    SDP_REQUIRE(plan && d_residual_grid, "sdp_degrid_grid: null argument");
    SDP_REQUIRE(d_model_grid != d_residual_grid, "sdp_degrid_grid: model and residual grids must differ");
    if (int e = sdp_degrid_grid_accumulate(plan, d_degrid_kernels, d_grid_kernels, d_model_grid, d_vis_measured,
                                           order, d_residual_vis, s))
        return e;
    return sdp_grid_merge(plan, d_residual_grid, s);
}

int sdp_grid_autotune(const sdp_grid_params *pp, const int *supports, const float *d_uvw, int n_vis,
                      const float *d_grid_kernels, const float *d_degrid_kernels, const int *tiles, int n_tiles,
                      const int *chunks, int n_chunk, int *best_tile, int *best_chunk, float *best_ms, sdp_stream s)
{
    SDP_REQUIRE(pp && supports && tiles && chunks && n_tiles > 0 && n_chunk > 0 && best_tile && best_chunk,
                "sdp_grid_autotune: bad arguments");
    hipStream_t st = sdp::stream_of(s);
    const size_t nvis = (size_t)std::max(n_vis, 1), ncell = (size_t)pp->grid_size * pp->grid_size;
    float2 *vis = nullptr, *grid = nullptr;
    hipEvent_t e0, e1;
    SDP_HIP_TRY(hipEventCreate(&e0));
    SDP_HIP_TRY(hipEventCreate(&e1));
    int rc = SDP_OK;
    float best = -1.f;
    if (hipMalloc(&vis, sizeof(float2) * nvis) != hipSuccess || hipMalloc(&grid, sizeof(float2) * ncell) != hipSuccess) {
        rc = sdp::set_error(SDP_E_NOMEM, "sdp_grid_autotune: allocation failed");
    } else {
        (void)hipMemsetAsync(vis, 0, sizeof(float2) * nvis, st);
        (void)hipMemsetAsync(grid, 0, sizeof(float2) * ncell, st);
        for (int a = 0; a < n_tiles && rc == SDP_OK; ++a) {
            for (int b = 0; b < n_chunk && rc == SDP_OK; ++b) {
                sdp_grid_params p = *pp;
                p.tile = tiles[a];
                p.chunk = chunks[b];
                sdp_grid_plan *plan = nullptr;
                if (sdp_grid_plan_create(&plan, &p, supports, d_uvw, n_vis, s) != SDP_OK) continue;   // shape not supported
                float ms = 0.f;
                // one untimed pass, then degrid + grid + merge timed together
                rc = sdp_degrid(plan, d_degrid_kernels, (const float *)grid, (const float *)vis, SDP_ORDER_PLAN,
                                (float *)vis, s);
                if (rc == SDP_OK) rc = sdp_grid(plan, d_grid_kernels, (const float *)vis, SDP_ORDER_PLAN, (float *)grid, s);
                if (rc == SDP_OK) {
                    (void)hipEventRecord(e0, st);
                    rc = sdp_degrid(plan, d_degrid_kernels, (const float *)grid, (const float *)vis, SDP_ORDER_PLAN,
                                    (float *)vis, s);
                    if (rc == SDP_OK)
                        rc = sdp_grid(plan, d_grid_kernels, (const float *)vis, SDP_ORDER_PLAN, (float *)grid, s);
                    (void)hipEventRecord(e1, st);
                    (void)hipEventSynchronize(e1);
                    (void)hipEventElapsedTime(&ms, e0, e1);
                }
                sdp_grid_plan_destroy(plan);
                if (rc == SDP_OK && (best < 0.f || ms < best)) {
                    best = ms;
                    *best_tile = tiles[a];
                    *best_chunk = chunks[b];
                }
            }
        }
        if (rc == SDP_OK && best < 0.f) rc = sdp::set_error(SDP_E_UNSUPPORTED, "sdp_grid_autotune: no candidate fits");
    }
    (void)hipFree(vis);
    (void)hipFree(grid);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (best_ms) *best_ms = best;
    return rc;
}

}  // extern "C"
