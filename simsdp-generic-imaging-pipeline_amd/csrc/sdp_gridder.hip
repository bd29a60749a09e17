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
// Exact-footprint classes and the register gridder
//   A visibility whose footprint is exactly 2h x 2h cells (not clipped at the
//   grid edge, not widened to 2h + 1 by float rounding) with 1 <= h <= 16,
//   and whose per-tap kernel offsets |round((u - p) OS)| are |d OS - c| for
//   the footprint's d = 0 .. 2h-1 (checked tap by tap against the reference
//   expression when the plan is built; half-way cases fail it), joins the
//   class of its own h.  Its sort key carries its first cell's position
//   (t_u, t_v) inside the uv tile.  Gridding these classes needs no index
//   arithmetic per tap: an "unfolded" copy of each plane's table,
//   [fy][fx][jy][jx] with jy, jx in [-h, h] holding the sample at
//   (|jy OS + fy|, |jx OS + fx|), puts tap (d_u, d_v) of a visibility at
//   base + d_v (2h + 1) + d_u, one base per visibility formed by the plan.
//   One wavefront covers 64 of the W^2 = (2h)^2 footprint positions (a, b)
//   of a work item; position (a, b) takes footprint tap
//   ((a - t_u) mod W, (b - t_v) mod W) of every visibility, which lands on
//   window column a or a + W and row b or b + W: each window cell belongs to
//   exactly one lane, whatever the visibility.  While (t_u, t_v) stays the
//   same (visibilities are sorted by it) a lane's tap offset and target cell
//   stay fixed, so a tap is one buffer load (the per-visibility base as the
//   scalar offset) and four FMAs into a register; a change of (t_u, t_v)
//   adds the register into the lane's cell (LDS private to the lane).  The
//   lanes then write the work item's window to its slab.  Degridding treats
//   these classes like any other (their window is tile + 2h - 1).
//
// Kernels: one wavefront per chunk.
//   grid    The wavefront is split into lane groups, each with a private
//           S x S LDS sub-grid copy.  A group takes its share of the chunk's
//           visibilities in order; its lanes cover a visibility's taps (lane
//           = one column of a row slot, consecutive lanes consecutive
//           columns), so no two lanes touch one cell within a visibility and
//           LDS order serialises visibilities: no atomics, deterministic.
//           The copies are summed in group order (TwoSum) into the chunk's
//           slab.  The merge then sums a (class, tile)'s pieces into one
//           slab (reduce_pieces_kernel) and, per 16 x 16 block of grid cells,
//           the slabs covering each cell in a fixed order (compensated),
//           writing the whole grid once (merge_region_kernel).
//   degrid  The wavefront stages the grid's S x S window in LDS; lanes sweep
//           the taps with the same mapping and a shuffle tree reduces them.
//   degrid+subtract+grid (major cycle): both of the above per visibility
//           in one pass: predict, subtract from the measured value, grid the
//           residual into the sub-grid.
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

#include <algorithm>
#include <type_traits>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <vector>

#include "sdp_internal.h"

namespace {

constexpr int kWave = 64;
// half-support limits of the support classes; a class's window is
// tile + 2 * (largest half support present in it).  Footprints span up to
// 2h + 1 cells, so h <= 7 fits 16-lane groups and h <= 15 32-lane groups.
constexpr int kNumClasses = 6;
constexpr int kClassMaxH[kNumClasses] = {7, 11, 15, 20, 24, 31};
// exact-footprint classes follow: class kNumClasses + h - 1 for h = 1 .. kFastMaxH
constexpr int kFastMaxH = 16;
constexpr int kAllClasses = kNumClasses + kFastMaxH;
__host__ __device__ constexpr int fast_class(int h) { return kNumClasses + h - 1; }
constexpr bool is_fast(int c) { return c >= kNumClasses; }
// LDS of one workgroup (one wavefront): accumulators, or accumulators plus
// the model-grid window for the fused kernel
constexpr size_t kMaxLds = 152 * 1024;
constexpr int kLdsPlanes = 256;   // plane tables up to this many planes are read from LDS

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
    // abs((int)ROUND((v - grid_pos) * OVERSAMPLING_FACTOR)), degridgrid.c:101,110.
    // ROUND takes halves away from zero, so this is roundf(|x|) =
    // trunc(|x|) + (|x| - trunc(|x|) >= 0.5); the difference is exact.
    const float y = fabsf(((float)cell_abs - pos) * os);
    const float t = truncf(y);
    return (int)t + (y - t >= 0.5f ? 1 : 0);
}

// ---- plan construction ---------------------------------------------------

struct PlanArgs {
    int n, G, OS, n_kernels, snap_mode, tile, ntx, nty, ntiles, nkeys;
    int plane_bits;   // sort key = (class, tile) key << plane_bits | low field (plane, or t_u + t_v tile)
    double uv_scale, w_scale;
    int grid, first;  // this launch's grid and the caller index of its first visibility
    int ntiles_all;   // tiles over all grids: a key is rank * ntiles_all + grid * ntiles + tile
    int class_rank[kAllClasses];   // key rank of each support class present (classes absent: -1)
};

// The exact-footprint test of one axis: first cell c0 (signed grid
// coordinate) of a 2h-cell footprint at position p.  With c the reference's
// tap offset of the first cell, every tap d must read |d OS - c| (the
// reference expression, kernel_tap, evaluated for each d), and the signed
// offsets d OS - c = j OS + f (0 <= f < OS) must keep j in [-h, h]; then
// *j0 = j at d = 0 and *f the phase.
__device__ bool exact_axis(int c0, float p, int h, int OS, float os, int *j0, int *f)
{
    const int c = kernel_tap(c0, p, os);
    for (int d = 1; d < 2 * h; ++d)
        if (kernel_tap(c0 + d, p, os) != abs(d * OS - c)) return false;
    const int j = -((c + OS - 1) / OS);   // floor(-c / OS), c >= 0
    *j0 = j;
    *f = -c - j * OS;
    return j >= -h && j + 2 * h - 1 <= h;
}

// plane_utab[p]: the plane's offset in the unfolded table, or -1 when its
// visibilities never join an exact-footprint class
__global__ void plan_records_kernel(const float *uvw, const int2 *supports, const int *plane_class,
                                    const int *plane_utab, PlanArgs a, VisRec *recs, uint2 *frecs, uint32_t *keys,
                                    uint32_t *vals)
{
    const int li = blockIdx.x * blockDim.x + threadIdx.x;   // within this grid's visibilities
    if (li >= a.n) return;
    const float x = uvw[3 * (size_t)li], y = uvw[3 * (size_t)li + 1], z = uvw[3 * (size_t)li + 2];
    const int i = a.first + li;                             // caller index over all grids
    const int gc = a.G / 2;
    keys[i] = (uint32_t)a.nkeys << a.plane_bits;   // dropped: sorts last
    vals[i] = (uint32_t)i;
    frecs[i] = make_uint2(0u, 0u);

    const float wabs = fabsf((float)((double)z * a.w_scale));
    const float root = __fsqrt_rn(wabs);
    int plane;
    if (a.snap_mode == SDP_SNAP_STD)
        plane = (int)((double)root + 0.5);          // degridgrid.c:84
    else
        plane = (int)roundf(root);                   // gridding_run.c:67
    if (!(root == root) || plane < 0 || plane >= a.n_kernels) return;
    const int h = supports[plane].x;
    int cls = plane_class[plane];
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
    // within a (class, tile) list, records in plane order: neighbours share
    // a kernel table and a footprint size
    uint32_t low = (uint32_t)plane;
    uint2 fr = make_uint2(0u, 0u);
    const int ut = plane_utab[plane];
    if (ut >= 0 && nu == 2 * h && nv == 2 * h && ua == (int)lo_x + gc && va == (int)lo_y + gc) {
        int jx, fx, jy, fy;
        const float os = (float)a.OS;
        if (exact_axis((int)lo_x, gx, h, a.OS, os, &jx, &fx) && exact_axis((int)lo_y, gy, h, a.OS, os, &jy, &fy)) {
            const int n = 2 * h + 1;
            const int tu = ua % a.tile, tv = va % a.tile;
            cls = fast_class(h);
            // exact-footprint lists in (t_u, t_v, conj) order: a lane's tap
            // stays put, and the conjugation is one per run
            const uint32_t conj = r.meta >> 31;
            low = (uint32_t)(tu + tv * a.tile) << 1 | conj;
            fr.x = 8u * (uint32_t)(ut + ((fy * a.OS + fx) * n + jy + h) * n + jx + h);   // bytes
            fr.y = (uint32_t)tu | (uint32_t)tv << 5 | conj << 10;
        }
    }
    frecs[i] = fr;
    const uint32_t key = (uint32_t)(a.class_rank[cls] * a.ntiles_all + a.grid * a.ntiles + (va / a.tile) * a.ntx +
                                    (ua / a.tile));
    keys[i] = key << a.plane_bits | (low & ((1u << a.plane_bits) - 1u));
}

// The (class, tile) lists of the sorted keys as runs: a run of equal keys
// is one list.  The first kernel flags each run's first record and writes
// every record's tap count; after an exclusive scan of the flags (run
// numbers) and an inclusive scan of the tap counts, the second records each
// run's key and first position at its start, and the third its length and
// tap total at its end.  Dropped records (key == nkeys) sort last and are
// skipped.  The sort keys carry a low field of plane_bits bits.
struct KeyRun {
    uint32_t key, first, count, pad;
    unsigned long long taps;
};

__global__ void run_flags_kernel(const uint32_t *keys, const VisRec *recs, int n, int nkeys, int plane_bits,
                                 uint32_t *flags, unsigned long long *taps)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t k = keys[i] >> plane_bits;
    if (k >= (uint32_t)nkeys) {
        taps[i] = 0;
        flags[i] = 0;
        return;
    }
    const uint32_t m = recs[i].meta;
    taps[i] = (unsigned long long)((m & 0xffu) * ((m >> 8) & 0xffu));
    flags[i] = (i == 0 || (keys[i - 1] >> plane_bits) != k) ? 1u : 0u;
}

__global__ void run_start_kernel(const uint32_t *keys, const uint32_t *flags, const uint32_t *run_no, int n,
                                 int plane_bits, KeyRun *runs)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || !flags[i]) return;
    KeyRun r;
    r.key = keys[i] >> plane_bits;
    r.first = (uint32_t)i;
    r.count = 0;
    r.pad = 0;
    r.taps = 0;
    runs[run_no[i]] = r;
}

__global__ void run_end_kernel(const uint32_t *keys, const uint32_t *flags, const uint32_t *run_no, int n,
                               int nkeys, int plane_bits, const unsigned long long *tap_scan, KeyRun *runs)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t k = keys[i] >> plane_bits;
    if (k >= (uint32_t)nkeys || (i + 1 < n && (keys[i + 1] >> plane_bits) == k)) return;
    KeyRun *r = &runs[run_no[i] + flags[i] - 1];
    const uint32_t f = r->first;
    r->count = (uint32_t)(i + 1) - f;
    r->taps = tap_scan[i] - (f > 0 ? tap_scan[f - 1] : 0ull);
}


// Runs of records (sorted order) with one (class, tile) key and one
// (t_u, t_v, conj) field: each record's bits 11-31 receive the records left
// in its run from it on (capped at 2^21 - 1), so the register gridder takes
// a run in one stretch.  Run starts are flagged, numbered by an inclusive
// scan of the flags, each run's end is recorded at its last record, and
// every record then reads its run's end.
constexpr uint32_t kRunField = 0x7ffu;   // t_u | t_v << 5 | conj << 10
constexpr uint32_t kRunCap = (1u << 21) - 1u;
__global__ void fast_run_flags_kernel(const uint32_t *keys, const uint2 *frecs, int n, int plane_bits,
                                      uint32_t *flags)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    flags[i] = (i == 0 || (keys[i - 1] >> plane_bits) != (keys[i] >> plane_bits) ||
                ((frecs[i - 1].y ^ frecs[i].y) & kRunField) != 0u) ? 1u : 0u;
}

__global__ void fast_run_ends_kernel(const uint32_t *flags, const uint32_t *run_id, int n, uint32_t *run_end)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (i + 1 == n || flags[i + 1]) run_end[run_id[i] - 1] = (uint32_t)(i + 1);
}

__global__ void fast_run_rem_kernel(const uint32_t *run_id, const uint32_t *run_end, int n, uint2 *frecs)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t rem = min(run_end[run_id[i] - 1] - (uint32_t)i, kRunCap);
    frecs[i].y = (frecs[i].y & kRunField) | rem << 11;
}

// the records and the exact-footprint records in sorted order
__global__ void gather_records_kernel(const VisRec *recs, const uint2 *frecs, const uint32_t *order, int n,
                                      VisRec *out, uint2 *fout)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t o = order[i];
    out[i] = recs[o];
    fout[i] = frecs[o];
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
    int G, OS, tile, ntx, ntiles, chunk_base;
    int lane_shift;   // log2 of the lanes per visibility (64 / copies)
    int copies;       // lane groups (and sub-grid copies) per wavefront
    int os_shift;     // log2 OS when OS is a power of two, else -1
    int tab_bytes;    // size of one phase-major kernel table
    int lds_planes;   // planes whose (offset, row length) entries are copied to LDS (0: read from memory)
    int pt_lds_off;   // byte offset of that copy in LDS
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

// Lane groups.  A wavefront splits into C groups of L lanes (L a power of
// two no smaller than the class's largest footprint extent, C = 64 / L or
// fewer when LDS runs short).  Group g takes the chunk's visibilities
// g, g + C, g + 2C, ... in order; within a visibility the group's lanes
// cover floor(L / nu) footprint rows at a time, one lane per tap, so a
// group's lanes never touch one cell at the same time.  Each group accumulates into its own copy of the sub-grid; the copies are
// summed in group order when the slab is written, so the result does not
// depend on timing.  The kernel-table offsets of footprint row r and column
// u are computed once per visibility, by lanes r and u of the group, and
// fetched with lane shuffles.  Taps of idle lanes (past the footprint, or
// past the chunk's last visibility) are predicated: they read zeros and
// accumulate into a per-lane spare cell.
// Phase-major kernel tables.  The reference lays a plane out as
// [ky][kx] with ky, kx = |offset| * OS in [0, (h+1) OS) (kernel_host_set_up,
// top.c:252-257), so the taps of one visibility, OS samples apart, fall on
// different cache lines.  Each call first copies the tables, value for
// value, into the order [fy][fx][iy][ix] with ky = iy OS + fy, kx = ix OS + fx:
// a visibility's taps along a row are then consecutive.  Same samples, same
// products: results are those of the reference layout bit for bit.
// os_shift >= 0: OS = 1 << os_shift < 4096 (shifts and 24-bit products:
// fy * OS < 2^24, and the sum stays below the plane's (hp1 * OS)^2 < 2^26
// samples); otherwise divide by OS
__device__ __forceinline__ int row_part(int ky, int OS, int os_shift, int hp1)
{
    if (os_shift >= 0) {
        const unsigned iy = (unsigned)ky >> os_shift, fy = (unsigned)ky & (unsigned)(OS - 1);
        return (int)(__umul24(__umul24(fy, OS), __umul24(hp1, hp1)) + __umul24(iy, hp1));
    }
    const int iy = ky / OS;
    const int fy = ky - iy * OS;
    return fy * OS * hp1 * hp1 + iy * hp1;
}
__device__ __forceinline__ int col_part(int kx, int OS, int os_shift, int hp1)
{
    if (os_shift >= 0) {
        const unsigned ix = (unsigned)kx >> os_shift, fx = (unsigned)kx & (unsigned)(OS - 1);
        return (int)(__umul24(fx, __umul24(hp1, hp1)) + ix);
    }
    const int ix = kx / OS;
    const int fx = kx - ix * OS;
    return fx * hp1 * hp1 + ix;
}

// the phase-major copy of one plane; t = (offset, h + 1); block bx of nbx
__device__ void relayout_plane(const float2 *__restrict__ in, int2 t, int OS, float2 *__restrict__ out, int bx,
                               int nbx)
{
    const int n = t.y * OS;
    const int total = n * n;
    for (int i = bx * blockDim.x + threadIdx.x; i < total; i += nbx * blockDim.x) {
        const int ky = i / n, kx = i - ky * n;
        out[t.x + row_part(ky, OS, -1, t.y) + col_part(kx, OS, -1, t.y)] = in[t.x + i];
    }
}

// kernel-table samples through a buffer resource: an offset past the table
// reads as zero, which is how idle taps are made to add nothing
constexpr int kOob = 0x40000000;
typedef float v2f_t __attribute__((ext_vector_type(2)));

__device__ __forceinline__ float2 tab_load(__amdgpu_buffer_rsrc_t rs, unsigned off)
{
    const v2f_t v = __builtin_bit_cast(v2f_t, __builtin_amdgcn_raw_buffer_load_b64(rs, off, 0, 0));
    return make_float2(v.x, v.y);
}

struct GroupVis {
    int nu, nv, lu0, lv0;   // footprint size and first cell relative to the tile
    int R;                  // footprint rows per step of the group's walk
    float rcp_nu;           // 1 / max(nu, 1)
    int2 pt;                // (table offset, h + 1)
    float conj, gx, gy;
    uint32_t o;             // caller index
};

// A record's walk geometry, formed once by the lane that holds it: nu, nv,
// R = L / nu rows per step, the step count (largest over the C records the
// groups take together) and the conjugation flag, 7 bits each (all <= 65).
__device__ __forceinline__ uint32_t pack_geo(int nu, int nv, int R, int steps, bool conj)
{
    return (uint32_t)nu | (uint32_t)nv << 7 | (uint32_t)R << 14 | (uint32_t)steps << 21 | (conj ? 1u << 31 : 0u);
}

// A group's visibility, taken by lane shuffles from the lane that holds its
// record (byte address src = 4 * lane): the wavefront loads kWave records at
// once, one per lane, so their memory latency is paid once per kWave
// visibilities instead of once per visibility.
__device__ __forceinline__ int lane_take(int src, int x) { return __builtin_amdgcn_ds_bpermute(src, x); }
__device__ __forceinline__ float lane_take(int src, float x)
{
    return __int_as_float(__builtin_amdgcn_ds_bpermute(src, __float_as_int(x)));
}

// With one group per wavefront (ONE) the holding lane `slot` is the same for
// every lane: its fields are read into scalar registers, and what is formed
// from them is scalar work.
template <bool ONE> __device__ __forceinline__ int take(int src, int slot, int x)
{
    return ONE ? __builtin_amdgcn_readlane(x, slot) : lane_take(src, x);
}
template <bool ONE> __device__ __forceinline__ float take(int src, int slot, float x)
{
    return __int_as_float(take<ONE>(src, slot, __float_as_int(x)));
}

template <bool ONE>
__device__ __forceinline__ GroupVis group_vis(const VisRec &hr, uint32_t hgeo, float hrcp, int2 hpt, uint32_t ho,
                                              int src, int slot)
{
    GroupVis v;
    const uint32_t start = (uint32_t)take<ONE>(src, slot, (int)hr.start);   // tile-relative
    const uint32_t geo = (uint32_t)take<ONE>(src, slot, (int)hgeo);
    v.o = (uint32_t)take<ONE>(src, slot, (int)ho);
    v.lu0 = (int)(start & 0xffffu);
    v.lv0 = (int)(start >> 16);
    v.nu = (int)(geo & 0x7fu);
    v.nv = (int)((geo >> 7) & 0x7fu);
    v.R = (int)((geo >> 14) & 0x7fu);
    v.rcp_nu = take<ONE>(src, slot, hrcp);
    v.pt = make_int2(take<ONE>(src, slot, hpt.x), take<ONE>(src, slot, hpt.y));
    v.conj = (geo & 0x80000000u) ? -1.0f : 1.0f;
    v.gx = take<ONE>(src, slot, hr.gx);
    v.gy = take<ONE>(src, slot, hr.gy);
    return v;
}

// n / d for 0 <= n <= 1024, 1 <= d <= 64: the float quotient is within 3e-4
// of n / d, and a non-integer n / d sits at least 1 / 64 below the next
// integer, so the added 1 / 128 floors correctly
// x plus lane x ^ m's x by a DPP move (no LDS): CTRL is a quad permutation
// or a row (16-lane) mirror
template <int CTRL> __device__ __forceinline__ float dpp_add(float x)
{
    return x + __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), CTRL, 0xf, 0xf, false));
}

// Butterfly sums over lane groups of 2^lane_shift lanes: within 16 lanes by
// DPP (quad swaps 1 and 2, half-row and row mirrors), beyond by shuffles.
// Every lane of a group ends with the group's sums, in a fixed order.
__device__ __forceinline__ void group_sum(float &x, float &y, int lane_shift)
{
    if (lane_shift >= 1) { x = dpp_add<0xb1>(x); y = dpp_add<0xb1>(y); }
    if (lane_shift >= 2) { x = dpp_add<0x4e>(x); y = dpp_add<0x4e>(y); }
    if (lane_shift >= 3) { x = dpp_add<0x141>(x); y = dpp_add<0x141>(y); }
    if (lane_shift >= 4) { x = dpp_add<0x140>(x); y = dpp_add<0x140>(y); }
    for (int off = 16; off < (1 << lane_shift); off <<= 1) {
        x += __shfl_xor(x, off, kWave);
        y += __shfl_xor(y, off, kWave);
    }
}

// LDS byte addresses: loads and stores through them need no base addition
typedef __attribute__((address_space(3))) char lds_char_t;
__device__ __forceinline__ uint32_t lds_addr(const void *p)
{
    return (uint32_t)(uintptr_t)(const lds_char_t *)p;
}
typedef float v4f_lds_t __attribute__((ext_vector_type(4)));
typedef __attribute__((address_space(3))) v2f_t lds_v2f_t;
typedef __attribute__((address_space(3))) v4f_lds_t lds_v4f_t;
template <typename T> __device__ __forceinline__ T lds_load(int a);
template <> __device__ __forceinline__ float2 lds_load<float2>(int a)
{
    const v2f_t v = *(const lds_v2f_t *)(uintptr_t)(uint32_t)a;
    return make_float2(v.x, v.y);
}
template <> __device__ __forceinline__ float4 lds_load<float4>(int a)
{
    const v4f_lds_t v = *(const lds_v4f_t *)(uintptr_t)(uint32_t)a;
    return make_float4(v.x, v.y, v.z, v.w);
}
__device__ __forceinline__ void lds_store(int a, const float2 &v)
{
    v2f_t t;
    t.x = v.x;
    t.y = v.y;
    *(lds_v2f_t *)(uintptr_t)(uint32_t)a = t;
}
__device__ __forceinline__ void lds_store(int a, const float4 &v)
{
    v4f_lds_t t;
    t.x = v.x;
    t.y = v.y;
    t.z = v.z;
    t.w = v.w;
    *(lds_v4f_t *)(uintptr_t)(uint32_t)a = t;
}

__device__ __forceinline__ int small_div(int n, float rcp_d)
{
    return (int)((float)n * rcp_d + 0.0078125f);
}

template <bool COMP> struct AccT;
template <> struct AccT<true> { using T = float4; };    // sum and TwoSum error per component
template <> struct AccT<false> { using T = float2; };

__device__ __forceinline__ void acc_add(float4 &c, float px, float py)
{
    two_sum_acc(c.x, c.z, px);
    two_sum_acc(c.y, c.w, py);
}
__device__ __forceinline__ void acc_add(float2 &c, float px, float py)
{
    c.x += px;
    c.y += py;
}
__device__ __forceinline__ void acc_zero(float4 &c) { c = make_float4(0.f, 0.f, 0.f, 0.f); }
__device__ __forceinline__ void acc_zero(float2 &c) { c = make_float2(0.f, 0.f); }

template <int MODE, bool COMP, bool ONE>
__global__ void __launch_bounds__(kWave)
chunk_kernel(const VisRec *__restrict__ recs, const uint32_t *__restrict__ perm, const int2 *__restrict__ plane_tab,
             const int4 *__restrict__ chunks, const long long *__restrict__ slab_off, ChunkArgs a, ChunkIO io)
{
    using Acc = typename AccT<COMP>::T;
    extern __shared__ float4 lds4[];
    constexpr bool kGrid = MODE != MODE_DEGRID;
    constexpr bool kDegrid = MODE != MODE_GRID;
    const int4 ch = chunks[a.chunk_base + blockIdx.x];   // tile (over all grids), first record, count, window S
    const int S = ch.w, stride = S + 1, cells = S * stride;
    const int lane = threadIdx.x;
    const int L = ONE ? kWave : 1 << a.lane_shift, C = ONE ? 1 : a.copies;
    const int g = ONE ? 0 : lane >> a.lane_shift, c = ONE ? lane : lane & (L - 1);
    const int grid_no = ch.x / a.ntiles, tl = ch.x - grid_no * a.ntiles;
    const int tile_u0 = (tl % a.ntx) * a.tile, tile_v0 = (tl / a.ntx) * a.tile;
    const int gc = a.G / 2;
    const float os = (float)a.OS;
    // raw (stride 0) buffer resources over the tables; word 3 is gfx9's 32-bit data format
    const __amdgpu_buffer_rsrc_t kg_rsrc =
        __builtin_amdgcn_make_buffer_rsrc((void *)io.kg, (short)0, a.tab_bytes, 0x00020000);
    const __amdgpu_buffer_rsrc_t kd_rsrc =
        __builtin_amdgcn_make_buffer_rsrc((void *)io.kd, (short)0, a.tab_bytes, 0x00020000);

    // LDS: [model window, S x stride float2][C sub-grid copies][one spare cell per lane]
    const int win_cells = kDegrid ? (cells + 2) & ~1 : 0;   // the window and a zero cell
    float2 *win = reinterpret_cast<float2 *>(lds4);
    Acc *acc = reinterpret_cast<Acc *>(win + win_cells);
    const int cstride = cells + 2;
    if (kGrid) {
        for (int i = lane; i < C * cstride + kWave; i += kWave) acc_zero(acc[i]);
    }
    // the plane table in LDS: a visibility's plane lookup then costs no memory latency
    int2 *lds_pt = reinterpret_cast<int2 *>(reinterpret_cast<char *>(lds4) + a.pt_lds_off);
    const bool pt_in_lds = a.lds_planes > 0;
    for (int i = lane; i < a.lds_planes; i += kWave) lds_pt[i] = plane_tab[i];
    if (kDegrid) {
        for (int i = lane; i < S * S; i += kWave) {
            const int row = i / S, col = i - row * S;
            const int gy = tile_v0 + row, gx = tile_u0 + col;
            float2 v = make_float2(0.f, 0.f);
            if (gy < a.G && gx < a.G) v = io.grid[((size_t)grid_no * a.G + gy) * a.G + gx];
            win[row * stride + col] = v;
        }
        if (lane == 0) win[cells] = make_float2(0.f, 0.f);
    }
    __syncthreads();

    // the value each visibility brings: the visibility to grid, or the
    // measured value to subtract from (absent: degrid only)
    const float2 *val_src = MODE == MODE_GRID ? io.vis : io.measured;
    VisRec hr;            // this lane's record of the current kWave (start tile-relative)
    uint32_t hgeo;        // its walk geometry
    float hrcp;           // 1 / max(nu, 1)
    int2 hpt;             // its plane's (table offset, h + 1)
    float2 hval;          // its value
    uint32_t ho;          // its caller index
    const int iters = (ch.z + C - 1) / C;
    for (int it = 0; it < iters; ++it) {
        const int first = it * C;                  // group 0's visibility in the chunk
        const int slot = first & (kWave - 1);      // C divides kWave
        if (slot == 0) {
            const int j = first + lane;
            hr.gx = hr.gy = 0.f;
            hr.start = hr.meta = 0u;
            hpt = make_int2(0, 0);
            hval = make_float2(0.f, 0.f);
            ho = 0u;
            if (j < ch.z) {
                hr = recs[ch.y + j];
                ho = perm ? perm[ch.y + j] : (uint32_t)(ch.y + j);
                if (val_src) hval = val_src[ho];
                const int plane = (hr.meta >> 16) & 0x7fff;
                hpt = pt_in_lds ? lds_pt[plane] : plane_tab[plane];
                hr.start -= (uint32_t)tile_u0 | (uint32_t)tile_v0 << 16;   // the first cell lies in the tile
            }
            // The group's L lanes cover R = L / nu footprint rows at a time
            // (nu <= 2h + 1 <= L); the groups step together, as often as the
            // largest ceil(nv / R) of the C records taken at once.
            const int nu = (int)(hr.meta & 0xffu), nv = (int)((hr.meta >> 8) & 0xffu);
            hrcp = __builtin_amdgcn_rcpf((float)max(nu, 1));
            const int R = max(small_div(L, hrcp), 1);
            int st = small_div(nv + R - 1, __builtin_amdgcn_rcpf((float)R));
            for (int off = 1; off < C; off <<= 1) st = max(st, __shfl_xor(st, off, kWave));
            hgeo = pack_geo(nu, nv, R, st, (hr.meta & 0x80000000u) != 0u);
        }
        const int src = (slot + g) << 2;
        const GroupVis v = group_vis<ONE>(hr, hgeo, hrcp, hpt, ho, src, slot);
        const bool ok = v.nu > 0;                  // records past the chunk are empty
        // the step count, equal on every lane
        const int steps = (int)(((uint32_t)__builtin_amdgcn_readfirstlane(take<ONE>(src, slot, (int)hgeo)) >> 21) & 0x7fu);
        const float2 cur_val = make_float2(take<ONE>(src, slot, hval.x), take<ONE>(src, slot, hval.y));
        const int cell_u0 = tile_u0 + v.lu0 - gc, cell_v0 = tile_v0 + v.lv0 - gc;
        // Lane r (r < nv) of the group holds the kernel-table byte offset of
        // footprint row r (std_gridding's abs((int)ROUND((v - grid_pos) * OS)),
        // degridgrid.c:110, in the phase-major layout) and lane u (u < nu) that
        // of column u.  Offsets of absent rows and columns lie past the table,
        // so their loads read zeros.
        const int ky_own = c < v.nv ? 8 * row_part(kernel_tap(cell_v0 + c, v.gy, os), a.OS, a.os_shift, v.pt.y)
                                    : kOob;
        const int kx_own = c < v.nu
                               ? 8 * (v.pt.x + col_part(kernel_tap(cell_u0 + c, v.gx, os), a.OS, a.os_shift, v.pt.y))
                               : kOob;
        // The group's L lanes cover R = L / nu footprint rows at a time: lane
        // c takes column c % nu of rows c / nu, c / nu + R, c / nu + 2R, ...
        // (lanes past R * nu idle), so its column, and that column's table
        // offset, stay fixed and each step advances the row by R.  The row
        // offset comes by a lane shuffle from the row's lane (byte address ra).
        const int gbase = g << a.lane_shift;
        const int R = v.R;
        const int r0 = small_div(c, v.rcp_nu), col = c - __mul24(r0, max(v.nu, 1));
        const int rend = (gbase + v.nv) << 2;
        const int ra0 = r0 < R ? (gbase + r0) << 2 : rend;
        const int kx = __builtin_amdgcn_ds_bpermute((gbase + col) << 2, kx_own);
        const int dra = R << 2;
        const int ix0 = __mul24(v.lv0 + r0, stride) + v.lu0 + col;   // window cell of the first tap
        const uint32_t lds0 = lds_addr(lds4);

        float2 res = make_float2(0.f, 0.f);
        if (kDegrid) {
            // window byte offsets; dead taps read the zero cell past the window
            const int dwb = __mul24(R, stride) * (int)sizeof(float2);
            const int zero_b = (int)lds0 + cells * (int)sizeof(float2);
            float sxx = 0.f, syy = 0.f, syx = 0.f, sxy = 0.f;
            int ra = ra0, wb = (int)lds0 + ix0 * (int)sizeof(float2);
            auto batch = [&](auto nb) {
                constexpr int N = decltype(nb)::value;
                float2 kk[N];
                int wi[N];
#pragma unroll
                for (int q = 0; q < N; ++q) {
                    const bool live = ra < rend;
                    const int ky = __builtin_amdgcn_ds_bpermute(ra, ky_own);
                    kk[q] = tab_load(kd_rsrc, live ? (unsigned)(kx + ky) : (unsigned)kOob);
                    wi[q] = live ? wb : zero_b;
                    ra += dra;
                    wb += dwb;
                }
                float2 w[N];
#pragma unroll
                for (int q = 0; q < N; ++q) w[q] = lds_load<float2>(wi[q]);
#pragma unroll
                for (int q = 0; q < N; ++q) {
                    // the four real products of grid * k; idle taps add exact zeros
                    sxx = fmaf(w[q].x, kk[q].x, sxx);
                    syy = fmaf(w[q].y, kk[q].y, syy);
                    syx = fmaf(w[q].y, kk[q].x, syx);
                    sxy = fmaf(w[q].x, kk[q].y, sxy);
                }
            };
            int j = 0;
            for (; j + 4 <= steps; j += 4) batch(std::integral_constant<int, 4>{});
            if (j + 2 <= steps) {
                batch(std::integral_constant<int, 2>{});
                j += 2;
            }
            if (j < steps) batch(std::integral_constant<int, 1>{});
            // grid * (k.x, conj * k.y), this lane's share
            float sr = fmaf(-v.conj, syy, sxx), si = fmaf(v.conj, sxy, syx);
            // sum over the group's lanes; every lane ends with the sum
            group_sum(sr, si, a.lane_shift);
            if (MODE == MODE_DEGRID) {
                if (ok && c == 0) {
                    float2 out = make_float2(sr, si);
                    if (io.measured) out = make_float2(cur_val.x - sr, cur_val.y - si);
                    io.vis_out[v.o] = out;
                }
                continue;
            }
            if (ok) {
                res = make_float2(cur_val.x - sr, cur_val.y - si);
                if (io.vis_out && c == 0) io.vis_out[v.o] = res;
            }
        }
        if (kGrid) {
            if (MODE == MODE_GRID && ok) res = cur_val;
            const float vr = res.x, vi = res.y, vrc = v.conj * res.x, vic = v.conj * res.y;
            // byte offsets from the start of LDS; dead taps use the lane's
            // spare cell, which nothing reads
            const int acc0 = (int)lds_addr(acc);
            const int spare_b = acc0 + (C * cstride + lane) * (int)sizeof(Acc);
            const int dab = __mul24(R, stride) * (int)sizeof(Acc);
            int ra = ra0, ab = acc0 + (g * cstride + ix0) * (int)sizeof(Acc);
            auto batch = [&](auto nb) {
                constexpr int N = decltype(nb)::value;
                float2 kk[N];
                int ci[N];
#pragma unroll
                for (int q = 0; q < N; ++q) {
                    const bool live = ra < rend;
                    const int ky = __builtin_amdgcn_ds_bpermute(ra, ky_own);
                    kk[q] = tab_load(kg_rsrc, live ? (unsigned)(kx + ky) : (unsigned)kOob);
                    ci[q] = live ? ab : spare_b;
                    ra += dra;
                    ab += dab;
                }
                // distinct cells per lane (or the lane's spare cell): reads
                // first, then writes
                Acc cell[N];
#pragma unroll
                for (int q = 0; q < N; ++q) cell[q] = lds_load<Acc>(ci[q]);
#pragma unroll
                for (int q = 0; q < N; ++q) {
                    // vis * (k.x, conj * k.y), std_gridding (degridgrid.c:112-118),
                    // rounded before it is added: the running sum takes one
                    // rounding per tap
                    const float px = fmaf(vr, kk[q].x, -(vic * kk[q].y));
                    const float py = fmaf(vi, kk[q].x, vrc * kk[q].y);
                    acc_add(cell[q], px, py);
                }
#pragma unroll
                for (int q = N - 1; q >= 0; --q) lds_store(ci[q], cell[q]);
            };
            int j = 0;
            for (; j + 4 <= steps; j += 4) batch(std::integral_constant<int, 4>{});
            if (j + 2 <= steps) {
                batch(std::integral_constant<int, 2>{});
                j += 2;
            }
            if (j < steps) batch(std::integral_constant<int, 1>{});
        }
    }
    if (kGrid) {
        __syncthreads();
        float2 *slab = io.slabs + slab_off[a.chunk_base + blockIdx.x];
        for (int i = lane; i < S * S; i += kWave) {
            const int row = i / S, col = i - row * S;
            const Acc *p0 = acc + row * stride + col;
            if (COMP) {
                float4 t = *reinterpret_cast<const float4 *>(p0);
                for (int q = 1; q < C; ++q) {
                    const float4 u = *reinterpret_cast<const float4 *>(p0 + q * cstride);
                    two_sum_acc(t.x, t.z, u.x);
                    two_sum_acc(t.y, t.w, u.y);
                    t.z += u.z;
                    t.w += u.w;
                }
                slab[i] = make_float2(t.x + t.z, t.y + t.w);
            } else {
                float2 t = *reinterpret_cast<const float2 *>(p0);
                for (int q = 1; q < C; ++q) {
                    const float2 u = *reinterpret_cast<const float2 *>(p0 + q * cstride);
                    t.x += u.x;
                    t.y += u.y;
                }
                slab[i] = t;
            }
        }
    }
}

// ---- the register gridder of the exact-footprint classes -----------------

// The unfolded copy of one plane; t = (quadrant offset, h, unfolded offset,
// -).  Output sample ((fy OS + fx) n + jy + h) n + jx + h, n = 2h + 1, is the
// quadrant sample (|jy OS + fy|, |jx OS + fx|): a visibility's taps are then
// a (2h) x (2h) block of rows n samples apart.
__device__ void unfold_plane(const float2 *__restrict__ in, int4 t, int OS, float2 *__restrict__ out, int bx,
                             int nbx)
{
    const int h = t.y, n = 2 * h + 1, row = (h + 1) * OS;
    const int total = OS * OS * n * n;
    for (int i = bx * blockDim.x + threadIdx.x; i < total; i += nbx * blockDim.x) {
        const int jx = i % n - h;
        int r = i / n;
        const int jy = r % n - h;
        r /= n;
        const int fx = r % OS, fy = r / OS;
        const int kx = abs(jx * OS + fx), ky = abs(jy * OS + fy);   // < (h + 1) OS
        out[t.z + i] = in[t.x + ky * row + kx];
    }
}

// a sample at byte offset voff + soff of the unfolded table (soff uniform:
// the visibility's base); offsets past the table read zero
__device__ __forceinline__ float2 utab_load(__amdgpu_buffer_rsrc_t rs, unsigned voff, unsigned soff)
{
    const v2f_t v = __builtin_bit_cast(v2f_t, __builtin_amdgcn_raw_buffer_load_b64(rs, voff, soff, 0));
    return make_float2(v.x, v.y);
}

// vis * (k.x, conj k.y) added to a lane's running sum: re += vr kx - cvi ky,
// im += vi kx + cvr ky (cvr, cvi: the conjugation sign applied to vr, vi)
__device__ __forceinline__ void tap_add(float2 &t, float2 k, float vr, float vi, float cvr, float cvi)
{
    t.x = fmaf(k.x, vr, t.x);
    t.x = fmaf(k.y, -cvi, t.x);
    t.y = fmaf(k.x, vi, t.y);
    t.y = fmaf(k.y, cvr, t.y);
}
__device__ __forceinline__ void tap_add(float4 &t, float2 k, float vr, float vi, float cvr, float cvi)
{
    // the product rounded first, then a TwoSum into (sum, error)
    two_sum_acc(t.x, t.z, fmaf(vr, k.x, -(cvi * k.y)));
    two_sum_acc(t.y, t.w, fmaf(vi, k.x, cvr * k.y));
}
__device__ __forceinline__ void acc_merge(float2 &c, const float2 &t)
{
    c.x += t.x;
    c.y += t.y;
}
__device__ __forceinline__ void acc_merge(float4 &c, const float4 &t)
{
    two_sum_acc(c.x, c.z, t.x);
    two_sum_acc(c.y, c.w, t.y);
    c.z += t.z;
    c.w += t.w;
}
__device__ __forceinline__ float2 acc_value(const float2 &c) { return c; }
__device__ __forceinline__ float2 acc_value(const float4 &c) { return make_float2(c.x + c.z, c.y + c.w); }

// vis * (k.x, conj k.y) for conj = -1: re += vr kx + vi ky, im += vi kx - vr ky
__device__ __forceinline__ void tap_add_conj(float2 &t, float2 k, float vr, float vi)
{
    t.x = fmaf(k.x, vr, t.x);
    t.x = fmaf(k.y, vi, t.x);
    t.y = fmaf(k.x, vi, t.y);
    t.y = fmaf(k.y, -vr, t.y);
}
__device__ __forceinline__ void tap_add_conj(float4 &t, float2 k, float vr, float vi)
{
    two_sum_acc(t.x, t.z, fmaf(vr, k.x, vi * k.y));
    two_sum_acc(t.y, t.w, fmaf(vi, k.x, -(vr * k.y)));
}

// len records of one run: per record its base (bytes, the scalar offset of
// the table load) and its visibility.  Records go in groups of 8 whose table
// loads are issued one group ahead of the group being summed.
template <bool CONJ, bool PERM, typename Acc>
__device__ __forceinline__ void fast_run(Acc &run, __amdgpu_buffer_rsrc_t rs, unsigned voff, const uint2 *fr,
                                         const float2 *vv, const uint32_t *pp, const float2 *vis, int len)
{
    constexpr int U = 8;
    auto add = [&](float2 k, float2 v) {
        if (CONJ) tap_add_conj(run, k, v.x, v.y);
        else tap_add(run, k, v.x, v.y, v.x, v.y);
    };
    const int groups = len / U;
    if (groups > 0) {
        float2 k[U];
#pragma unroll
        for (int q = 0; q < U; ++q) k[q] = utab_load(rs, voff, fr[q].x);
        for (int g = 0; g < groups; ++g) {
            const int q0 = g * U;
            float2 kn[U];
            if (g + 1 < groups) {
#pragma unroll
                for (int q = 0; q < U; ++q) kn[q] = utab_load(rs, voff, fr[q0 + U + q].x);
            }
            float2 v[U];
#pragma unroll
            for (int q = 0; q < U; ++q) v[q] = PERM ? vis[pp[q0 + q]] : vv[q0 + q];
#pragma unroll
            for (int q = 0; q < U; ++q) add(k[q], v[q]);
            if (g + 1 < groups) {
#pragma unroll
                for (int q = 0; q < U; ++q) k[q] = kn[q];
            }
        }
    }
    for (int q0 = groups * U; q0 < len; ++q0) add(utab_load(rs, voff, fr[q0].x), PERM ? vis[pp[q0]] : vv[q0]);
}

// One wavefront per item (work item, part w): footprint positions
// p = 64 w + lane < W^2, (a, b) = (p % W, p / W).  kFastWaves independent
// wavefronts share a workgroup (no barriers; each its own LDS cells), so the
// CU's workgroup slots do not cap the resident wavefronts.
constexpr int kFastWaves = 4;
// The items of the exact-footprint classes, class by class: item i of
// segment s (item_begin[s] <= i < item_begin[s + 1]) is part
// (i - item_begin[s]) % parts[s] of chunk chunk_begin[s] + (i - item_begin[s]) / parts[s].
struct FastItems {
    int n_seg, n_items;
    int item_begin[kFastMaxH + 1], chunk_begin[kFastMaxH], parts[kFastMaxH];
};
template <bool COMP, bool PERM>
__global__ void __launch_bounds__(kWave * kFastWaves)
fast_grid_kernel(const uint2 *__restrict__ frecs, const uint32_t *__restrict__ perm, const int4 *__restrict__ chunks,
                 FastItems fi, const long long *__restrict__ slab_off,
                 const float2 *__restrict__ vis, const float2 *__restrict__ utab, int tab_bytes, int tile,
                 float2 *__restrict__ slabs)
{
    using Acc = typename AccT<COMP>::T;
    __shared__ Acc wave_cells[kFastWaves][4 * kWave];   // lane-private: cell q of lane l at [q * 64 + l]
    // the wavefront's index, uniform: its item, records and visibilities are
    // then read through scalar loads
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
    Acc *cells = wave_cells[wave];
    const int it = blockIdx.x * kFastWaves + wave;
    if (it >= fi.n_items) return;       // past the last item (whole workgroups)
    int seg = 0;
    while (seg + 1 < fi.n_seg && it >= fi.item_begin[seg + 1]) ++seg;
    const int2 item = make_int2(fi.chunk_begin[seg] + (it - fi.item_begin[seg]) / fi.parts[seg],
                                (it - fi.item_begin[seg]) % fi.parts[seg]);
    const int4 ch = chunks[item.x];     // tile, first record, count, window S
    const int S = ch.w, W = S - tile + 1, n = W + 1;
    const int lane = threadIdx.x % kWave;
    const int p = item.y * kWave + lane;
    const bool active = p < W * W;
    const int a = active ? p % W : 0, b = active ? p / W : 0;
#pragma unroll
    for (int q = 0; q < 4; ++q) acc_zero(cells[q * kWave + lane]);
    const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc((void *)utab, (short)0, tab_bytes, 0x00020000);
    const uint2 *fr = frecs + ch.y;
    const float2 *vv = PERM ? nullptr : vis + ch.y;
    const uint32_t *pp = PERM ? perm + ch.y : nullptr;
    Acc run;
    acc_zero(run);
    bool started = false;    // a run has been summed into `run`
    unsigned voff = kOob;    // this lane's tap in the current run (bytes from the visibility's base)
    int cell = 0;
    int j = 0;
    while (j < ch.z) {
        // a run: records j .. j + len - 1 share (t_u, t_v) and the conjugation
        const uint2 r0 = fr[j];
        const int len = max(1, min((int)(r0.y >> 11), ch.z - j));
        if (started) acc_merge(cells[cell * kWave + lane], run);
        {
            const int tu = (int)(r0.y & 0x1fu), tv = (int)((r0.y >> 5) & 0x1fu);
            const int du = a >= tu ? a - tu : a - tu + W, dv = b >= tv ? b - tv : b - tv + W;
            voff = active ? (unsigned)(8 * (dv * n + du)) : (unsigned)kOob;
            cell = (a < tu ? 1 : 0) + (b < tv ? 2 : 0);
            started = true;
            acc_zero(run);
        }
        if (r0.y & 0x400u)
            fast_run<true, PERM>(run, rs, voff, fr + j, PERM ? nullptr : vv + j, PERM ? pp + j : nullptr, vis, len);
        else
            fast_run<false, PERM>(run, rs, voff, fr + j, PERM ? nullptr : vv + j, PERM ? pp + j : nullptr, vis, len);
        j += len;
    }
    if (started) acc_merge(cells[cell * kWave + lane], run);
    // cell q of lane (a, b) is window cell (b + (q >> 1) W, a + (q & 1) W)
    if (active) {
        float2 *slab = slabs + slab_off[item.x];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int u = a + (q & 1) * W, v = b + (q >> 1) * W;
            if (u < S && v < S) slab[v * S + u] = acc_value(cells[q * kWave + lane]);
        }
    }
}

// ---- fused degrid + subtract + grid of the exact-footprint classes --------

// The kernel-table copies a pass needs, in one launch: workgroup rows
// (blockIdx.y) are planes of the gridding and degridding tables' phase-major
// copies, then planes of their unfolded copies, then one row whose first
// kConjBlocks workgroups compare the two caller tables sample for sample:
// workgroup b writes mismatch[b] = 1 when some degridding sample is not the
// conjugate of its gridding sample, else 0, and the last of them to finish
// (a counter at mismatch[kConjBlocks + 1], back to 0 after each call) writes
// their OR to mismatch[kConjBlocks].  Every flag is written every call.
constexpr int kPrepBlocks = 128, kConjBlocks = 32;
struct PrepArgs {
    const float2 *kg, *kd;          // the caller's tables (reference layout)
    float2 *rel_g, *rel_d;          // phase-major copies
    float2 *unf_g, *unf_d;          // unfolded copies
    int rows_rel_g, rows_rel_d, rows_unf_g, rows_unf_d, conj;
    int OS;
    long long table_len;            // samples compared
    int *mismatch;                  // kConjBlocks flags, their OR, the arrival counter
};

__global__ void __launch_bounds__(256) prep_tables_kernel(const int2 *__restrict__ plane_tab,
                                                          const int4 *__restrict__ unfold_tab, PrepArgs a)
{
    int y = blockIdx.y;
    if (y < a.rows_rel_g) return relayout_plane(a.kg, plane_tab[y], a.OS, a.rel_g, blockIdx.x, gridDim.x);
    y -= a.rows_rel_g;
    if (y < a.rows_rel_d) return relayout_plane(a.kd, plane_tab[y], a.OS, a.rel_d, blockIdx.x, gridDim.x);
    y -= a.rows_rel_d;
    if (y < a.rows_unf_g) return unfold_plane(a.kg, unfold_tab[y], a.OS, a.unf_g, blockIdx.x, gridDim.x);
    y -= a.rows_unf_g;
    if (y < a.rows_unf_d) return unfold_plane(a.kd, unfold_tab[y], a.OS, a.unf_d, blockIdx.x, gridDim.x);
    if (!a.conj || blockIdx.x >= kConjBlocks) return;
    __shared__ int bad;
    if (threadIdx.x == 0) bad = 0;
    __syncthreads();
    bool ok = true;
    for (long long i = blockIdx.x * 256LL + threadIdx.x; i < a.table_len; i += kConjBlocks * 256LL)
        ok = ok && a.kd[i].x == a.kg[i].x && a.kd[i].y == -a.kg[i].y;
    if (!ok) bad = 1;   // any lane: the same value
    __syncthreads();
    if (threadIdx.x == 0) {
        a.mismatch[blockIdx.x] = bad;
        __threadfence();
        if (atomicAdd(&a.mismatch[kConjBlocks + 1], 1) == kConjBlocks - 1) {   // the last to arrive
            __threadfence();
            int any = 0;
            for (int q = 0; q < kConjBlocks; ++q) any |= __atomic_load_n(&a.mismatch[q], __ATOMIC_RELAXED);
            a.mismatch[kConjBlocks] = any;
            a.mismatch[kConjBlocks + 1] = 0;
        }
    }
}

// ---- fused degrid + subtract + grid: tap ownership -------------------------
// Lanes own footprint taps.  The W^2 taps (W = 2h) of a visibility are taken
// row-major in "pair parts" and at most one "single part": in a pair part
// lane l holds taps 2m and 2m + 1 (m = 64 k + l; W is even, so both sit in
// one footprint row and in adjacent samples of the unfolded table) and reads
// them with one 16-byte load; in the single part lane l holds one tap.  The
// vector L1 serves a 64-lane load instruction in about the same time whether
// each lane reads 8 or 16 bytes, so pairs halve the load time of a tap.  A
// class uses q = W^2 / 128 pair parts and, for the r = W^2 mod 128 taps left,
// one single part when r <= 64, else one more pair part.  Shape <Q, H64>: up
// to Q pair parts and a single part.
template <int Q, bool H64> struct FusedShape {
    static constexpr int taps = 128 * Q + (H64 ? 64 : 0);   // the largest W^2 the shape holds
    // waves per workgroup; each holds its work item's S x S window in LDS.  An
    // exact-footprint class has 2h + 1 >= tile, so S = tile + W - 1 <= 2 W and
    // S^2 <= 4 W^2 <= 4 taps
    static constexpr int waves = taps <= 256 ? 4 : taps <= 512 ? 2 : 1;
    template <bool COMP> static constexpr size_t lds()
    {
        return (size_t)waves * 4 * taps * (COMP ? sizeof(float4) : sizeof(float2));
    }
    // visibilities per step (their table samples stay in registers until
    // the residuals are known; half as many when a degridding table of its
    // own doubles the samples)
    static constexpr int ub = Q <= 1 ? 4 : Q <= 4 ? 2 : 1;
    // resident wavefronts per SIMD the register allocation aims at
    template <bool COMP> static constexpr int occupancy() { return Q <= 1 ? 8 : Q <= 2 ? 6 : Q <= 4 ? (COMP ? 4 : 5) : (COMP ? 2 : 3); }
};

struct FusedArgs {
    int chunk_b, chunk_e, tile, G, ntx, ntiles, tab_bytes;
    int wave_cells;   // LDS cells per wave: the largest S^2 of the launch's classes
};

// ---- cross-lane sums of the fused kernel ----------------------------------
// A step's U visibilities need 2U wavefront sums (real and imaginary
// prediction).  Rather than one 64-lane butterfly per value, the values are
// folded together: v_permlane32_swap and v_permlane16_swap exchange half of
// two registers at a time, so each fold halves the lanes a value spans and
// the number of registers, until every 16-lane row holds one value; a
// 16-lane butterfly of DPP adds then leaves each row's total in all of its
// lanes.
// fold32(a, b): lanes 0-31 get a(l) + a(l + 32), lanes 32-63 b(l - 32) + b(l)
__device__ __forceinline__ float fold32(float a, float b)
{
    const auto r = __builtin_amdgcn_permlane32_swap(__float_as_uint(a), __float_as_uint(b), false, false);
    return __uint_as_float(r[0]) + __uint_as_float(r[1]);
}
// fold16(a, b): rows 0, 1, 2, 3 get a's rows 0 + 1, b's rows 0 + 1, a's rows
// 2 + 3, b's rows 2 + 3
__device__ __forceinline__ float fold16(float a, float b)
{
    const auto r = __builtin_amdgcn_permlane16_swap(__float_as_uint(a), __float_as_uint(b), false, false);
    return __uint_as_float(r[0]) + __uint_as_float(r[1]);
}
// the total of the lane's 16-lane row, in every lane of the row
__device__ __forceinline__ float row_total(float x)
{
    x = dpp_add<0xb1>(x);
    x = dpp_add<0x4e>(x);
    x = dpp_add<0x141>(x);
    return dpp_add<0x140>(x);
}
__device__ __forceinline__ float lane_value(float x, int l)
{
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(x), l));
}

// vis * k (CONJ: vis * conj k) added to a lane's running sum; the same
// roundings as tap_add with the conjugation sign applied to vr, vi
template <bool CONJ> __device__ __forceinline__ void tap_add_s(float2 &t, float2 k, float vr, float vi)
{
    t.x = fmaf(k.x, vr, t.x);
    t.x = fmaf(k.y, CONJ ? vi : -vi, t.x);
    t.y = fmaf(k.x, vi, t.y);
    t.y = fmaf(k.y, CONJ ? -vr : vr, t.y);
}
template <bool CONJ> __device__ __forceinline__ void tap_add_s(float4 &t, float2 k, float vr, float vi)
{
    two_sum_acc(t.x, t.z, fmaf(vr, k.x, (CONJ ? vi : -vi) * k.y));
    two_sum_acc(t.y, t.w, fmaf(vi, k.x, (CONJ ? -vr : vr) * k.y));
}
// model * (degridding sample) added to the prediction partials; the sample
// is the degridding table's own, or the conjugate of the gridding sample
template <bool CONJ> __device__ __forceinline__ void pred_add(float &pr, float &pim, float2 g, float2 q)
{
    pr = fmaf(g.x, q.x, pr);
    pr = fmaf(CONJ ? g.y : -g.y, q.y, pr);
    pim = fmaf(g.y, q.x, pim);
    pim = fmaf(CONJ ? -g.x : g.x, q.y, pim);
}

// two adjacent unfolded-table samples at byte offset voff + soff
__device__ __forceinline__ float4 utab_load2(__amdgpu_buffer_rsrc_t rs, unsigned voff, unsigned soff)
{
    const auto v = __builtin_amdgcn_raw_buffer_load_b128(rs, voff, soff, 0);
    return make_float4(__uint_as_float(v[0]), __uint_as_float(v[1]), __uint_as_float(v[2]), __uint_as_float(v[3]));
}

// One work item of the fused pass: for each of its visibilities (in plan
// order) the prediction sum over the footprint -- each lane's taps read the
// model grid at their window cells, held in registers for the run -- is
// reduced over the wavefront, subtracted from the measured value, and the
// residual gridded with the same table samples into the lanes' running sums.
// OWN_D: the degridding table is not the conjugate of the gridding table,
// and each tap reads a sample of both.
//
// Runs and cells.  A run's visibilities share their first window cell
// (t_u, t_v), so tap (d_u, d_v) of each of them lands on window cell
// (t_u + d_u, t_v + d_v): a lane's taps keep their cells, model values and
// table offsets for the run and sum into registers, which are added into
// the LDS window when the run ends.  Within a run the taps of a lane are
// distinct cells from every other lane's, and runs are added one after the
// other, so every cell receives its runs' sums in run order.
//
// Visibilities go U at a time.  After the folds the lanes of row r hold
// (U = 4) visibility r's real part in one register and its imaginary part in
// another; (U = 2) one register with rows (re 0, im 0, re 1, im 1); (U = 1)
// rows (re, re, im, im).  Each lane loads the measured value of its row, the
// row's first lane stores the residual, and the residuals are then read into
// scalar registers for the gridding products.
template <bool COMP, int Q, bool H64, bool OWN_D>
__device__ __forceinline__ void fused_work_item(const uint2 *__restrict__ frecs, const uint32_t *__restrict__ perm,
                                                const int4 ch, int ci, const long long *__restrict__ slab_off,
                                                const float2 *__restrict__ measured,
                                                const float2 *__restrict__ model, float2 *__restrict__ vis_out,
                                                __amdgpu_buffer_rsrc_t rg, __amdgpu_buffer_rsrc_t rd,
                                                const FusedArgs &fa, float2 *__restrict__ slabs,
                                                typename AccT<COMP>::T *cells)
{
    using Acc = typename AccT<COMP>::T;
    constexpr int NP = Q + (H64 ? 1 : 0);   // parts; part Q is the single part
    const int S = ch.w, tile = fa.tile, W = S - tile + 1, n = W + 1, WW = W * W;
    const int grid_no = ch.x / fa.ntiles, tl = ch.x - grid_no * fa.ntiles;   // ch.x: tile over all grids
    const int tile_u0 = (tl % fa.ntx) * tile, tile_v0 = (tl / fa.ntx) * tile;
    model += (size_t)grid_no * fa.G * fa.G;
    const int lane = threadIdx.x % kWave;
    const int row = lane >> 4;
    const uint2 *fr = frecs + ch.y;
    const float *meas1 = reinterpret_cast<const float *>(measured);
    float *out1 = reinterpret_cast<float *>(vis_out);
    // taps in pair parts; the rest (<= 64) in the single part
    const int r128 = WW & 127;
    const int n_pair = (r128 == 0 || r128 > 64) ? WW / 2 : WW / 128 * 64;
    const int n_single = WW - 2 * n_pair;
    // Row mode (the one-pair-part + single-part shape, h = 6: 16 single
    // taps): in a step of four visibilities the single part serves all four
    // at once, row r of the wavefront holding visibility r's single taps.
    // Row r's prediction partials join the folded sums of the same row, and
    // its gridding products use the row's residual; the four rows' running
    // sums are added together when the run ends.  Steps of fewer
    // visibilities use row 0 only.
    constexpr bool ROWS = !COMP && !OWN_D && Q == 1 && H64 && FusedShape<Q, H64>::ub == 4;
    Acc run[NP][2];
    unsigned voff[NP];
    float2 g[NP][2];
    int tp[NP];   // the lane's first tap of each part, d_u | d_v << 8; -1 when unused
#pragma unroll
    for (int k = 0; k < NP; ++k) {
        int t = -1;
        if (k < Q) {
            const int m = k * kWave + lane;
            if (m < n_pair) t = 2 * m;
        } else if (ROWS ? (lane & 15) < n_single : lane < n_single) {
            t = 2 * n_pair + (ROWS ? lane & 15 : lane);
        }
        tp[k] = t >= 0 ? (t % W) | (t / W) << 8 : -1;
        voff[k] = (unsigned)kOob;
        acc_zero(run[k][0]);
        acc_zero(run[k][1]);
        g[k][0] = g[k][1] = make_float2(0.f, 0.f);
    }
    for (int c = lane; c < S * S; c += kWave) acc_zero(cells[c]);
    int ptu = -1, ptv = 0;   // the previous run's (t_u, t_v); -1: none yet
    // the run's sums of a lane's taps into their window cells at (tu, tv)
    auto merge_run = [&](int tu, int tv) {
        if constexpr (ROWS) {
            // the four rows' single-part sums of each tap, added (every lane takes part)
            Acc &t = run[Q][0];
            t.x = fold16(fold32(t.x, t.x), fold32(t.x, t.x));
            t.y = fold16(fold32(t.y, t.y), fold32(t.y, t.y));
        }
#pragma unroll
        for (int k = 0; k < NP; ++k) {
            if (ROWS && k == Q && row != 0) continue;
            if (tp[k] >= 0) {
                const int c = (tv + (tp[k] >> 8)) * S + tu + (tp[k] & 0xff);
                acc_merge(cells[c], run[k][0]);
                if (k < Q) acc_merge(cells[c + 1], run[k][1]);
            }
        }
    };
    int j = 0;
    while (j < ch.z) {
        const uint2 r0 = fr[j];
        const int len = max(1, min((int)(r0.y >> 11), ch.z - j));
        const int tu = (int)(r0.y & 0x1fu), tv = (int)((r0.y >> 5) & 0x1fu);
        const bool conj = (r0.y & 0x400u) != 0u;
        if (ptu >= 0) merge_run(ptu, ptv);
#pragma unroll
        for (int k = 0; k < NP; ++k) {
            // opaque to the optimiser: what is formed from the tap is formed
            // again per run, not hoisted into registers held all along
            asm volatile("" : "+v"(tp[k]));
            const bool act = tp[k] >= 0;
            const int du = act ? tp[k] & 0xff : 0, dv = act ? tp[k] >> 8 : 0;
            voff[k] = act ? (unsigned)(8 * (dv * n + du)) : (unsigned)kOob;
            // the model values at the taps' cells for this run
            const int gu = tile_u0 + tu + du, gv = tile_v0 + tv + dv;
            const float2 *mrow = model + (size_t)gv * fa.G;
            const bool in_v = act && gv < fa.G;
            g[k][0] = (in_v && gu < fa.G) ? mrow[gu] : make_float2(0.f, 0.f);
            if (k < Q) g[k][1] = (in_v && gu + 1 < fa.G) ? mrow[gu + 1] : make_float2(0.f, 0.f);
            acc_zero(run[k][0]);
            acc_zero(run[k][1]);
        }
        ptu = tu;
        ptv = tv;
        auto run_body = [&](auto conj_tag) {
            constexpr bool CONJ = decltype(conj_tag)::value;
            auto step = [&](auto ucount, int i0) {
                constexpr int U = decltype(ucount)::value;
                const int q0 = j + i0;   // the step's first record, from the work item's first
                // the visibility (rv) and component (rc: 0 real, 1 imaginary) of this lane's row
                const int rv = U == 4 ? row : U == 2 ? row >> 1 : 0;
                const int rc = U == 2 ? (row & 1) : U == 1 ? (row >> 1) : 0;
                const uint32_t o = perm ? perm[ch.y + q0 + rv] : (uint32_t)(ch.y + q0 + rv);
                // the single part serves the step's four rows at once (row mode)
                constexpr bool RS = ROWS && U == 4;
                float4 kk[U][NP], kd[U][NP];
                uint32_t base[U];
#pragma unroll
                for (int u = 0; u < U; ++u) base[u] = fr[q0 + u].x;
                float2 krow = make_float2(0.f, 0.f);   // row mode: the lane's row's single-part sample
                if constexpr (RS) {
                    const uint32_t brow = row == 0 ? base[0] : row == 1 ? base[1] : row == 2 ? base[2] : base[3];
                    const auto v = __builtin_amdgcn_raw_buffer_load_b64(rg, voff[Q] + brow, 0, 0);
                    krow = make_float2(__uint_as_float(v[0]), __uint_as_float(v[1]));
                }
                // outside row mode, a row-mode lane of rows 1-3 has no single tap
                const unsigned vs = ROWS && !RS && row != 0 ? (unsigned)kOob : voff[NP - 1];
#pragma unroll
                for (int u = 0; u < U; ++u) {
#pragma unroll
                    for (int k = 0; k < NP; ++k) {
                        if (k < Q) {
                            kk[u][k] = utab_load2(rg, voff[k], base[u]);
                            if (OWN_D) kd[u][k] = utab_load2(rd, voff[k], base[u]);
                        } else if (!RS) {
                            const float2 a = utab_load(rg, vs, base[u]);
                            kk[u][k] = make_float4(a.x, a.y, 0.f, 0.f);
                            if (OWN_D) {
                                const float2 b = utab_load(rd, vs, base[u]);
                                kd[u][k] = make_float4(b.x, b.y, 0.f, 0.f);
                            }
                        }
                    }
                }
                float2 m2 = make_float2(0.f, 0.f);
                float m1 = 0.f;
                if (U == 4) m2 = measured[o];
                else m1 = meas1[2 * (size_t)o + rc];
                float pr[U], pim[U];
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    pr[u] = 0.f;
                    pim[u] = 0.f;
#pragma unroll
                    for (int k = 0; k < NP; ++k) {
                        if (RS && k == Q) continue;
                        const float4 q = OWN_D ? kd[u][k]
                                               : make_float4(kk[u][k].x, -kk[u][k].y, kk[u][k].z, -kk[u][k].w);
                        pred_add<CONJ>(pr[u], pim[u], g[k][0], make_float2(q.x, q.y));
                        if (k < Q) pred_add<CONJ>(pr[u], pim[u], g[k][1], make_float2(q.z, q.w));
                    }
                }
                float vr[U], vi[U];
                if constexpr (U == 4) {
                    float s0 = fold16(fold32(pr[0], pr[2]), fold32(pr[1], pr[3]));
                    float s1 = fold16(fold32(pim[0], pim[2]), fold32(pim[1], pim[3]));
                    if constexpr (RS) {
                        float xr = 0.f, xi = 0.f;   // row mode: the row's single-part partials
                        pred_add<CONJ>(xr, xi, g[Q][0], make_float2(krow.x, -krow.y));
                        s0 += xr;
                        s1 += xi;
                    }
                    const float b0 = m2.x - row_total(s0);
                    const float b1 = m2.y - row_total(s1);
                    if (vis_out && (lane & 15) == 0) vis_out[o] = make_float2(b0, b1);
#pragma unroll
                    for (int u = 0; u < U; ++u) {
                        vr[u] = lane_value(b0, 16 * u);
                        vi[u] = lane_value(b1, 16 * u);
                    }
                    if constexpr (RS) tap_add_s<CONJ>(run[Q][0], krow, b0, b1);
                } else if constexpr (U == 2) {
                    const float b = m1 - row_total(fold16(fold32(pr[0], pr[1]), fold32(pim[0], pim[1])));
                    if (vis_out && (lane & 15) == 0) out1[2 * (size_t)o + rc] = b;
                    vr[0] = lane_value(b, 0);
                    vi[0] = lane_value(b, 16);
                    vr[1] = lane_value(b, 32);
                    vi[1] = lane_value(b, 48);
                } else {
                    const float a = fold32(pr[0], pim[0]);
                    const float b = m1 - row_total(fold16(a, a));
                    if (vis_out && (lane & 31) == 0) out1[2 * (size_t)o + rc] = b;
                    vr[0] = lane_value(b, 0);
                    vi[0] = lane_value(b, 32);
                }
#pragma unroll
                for (int u = 0; u < U; ++u) {
#pragma unroll
                    for (int k = 0; k < NP; ++k) {
                        if (RS && k == Q) continue;
                        tap_add_s<CONJ>(run[k][0], make_float2(kk[u][k].x, kk[u][k].y), vr[u], vi[u]);
                        if (k < Q) tap_add_s<CONJ>(run[k][1], make_float2(kk[u][k].z, kk[u][k].w), vr[u], vi[u]);
                    }
                }
            };
            constexpr int UB = FusedShape<Q, H64>::ub;
            constexpr int UU = OWN_D && UB > 1 ? UB / 2 : UB;
            int i = 0;
            for (; i + UU <= len; i += UU) step(std::integral_constant<int, UU>{}, i);
            for (; i < len; ++i) step(std::integral_constant<int, 1>{}, i);
        };
        if (conj) run_body(std::true_type{});
        else run_body(std::false_type{});
        j += len;
    }
    if (ptu >= 0) merge_run(ptu, ptv);
    float2 *slab = slabs + slab_off[ci];
    for (int c = lane; c < S * S; c += kWave) slab[c] = acc_value(cells[c]);
}

// One wavefront per work item; each holds its window in LDS.  Whether the
// degridding table is the conjugate of the gridding table is a device flag
// (set by prep_tables_kernel), read once: the branch is uniform.
template <bool COMP, int Q, bool H64>
__global__ void __launch_bounds__((kWave * FusedShape<Q, H64>::waves))
__attribute__((amdgpu_waves_per_eu(FusedShape<Q, H64>::template occupancy<COMP>())))
fast_degrid_grid_kernel(const uint2 *__restrict__ frecs, const uint32_t *__restrict__ perm,
                        const int4 *__restrict__ chunks, const long long *__restrict__ slab_off,
                        const float2 *__restrict__ measured, const float2 *__restrict__ model,
                        float2 *__restrict__ vis_out, const float2 *__restrict__ utab_g,
                        const float2 *__restrict__ utab_d, const int *__restrict__ mismatch, FusedArgs fa,
                        float2 *__restrict__ slabs)
{
    using Acc = typename AccT<COMP>::T;
    extern __shared__ float4 lds4[];
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
    const int ci = fa.chunk_b + blockIdx.x * FusedShape<Q, H64>::waves + wave;
    if (ci >= fa.chunk_e) return;
    Acc *cells = reinterpret_cast<Acc *>(lds4) + (size_t)wave * fa.wave_cells;   // the window, row-major
    const int4 ch = chunks[ci];
    const __amdgpu_buffer_rsrc_t rg =
        __builtin_amdgcn_make_buffer_rsrc((void *)utab_g, (short)0, fa.tab_bytes, 0x00020000);
    const __amdgpu_buffer_rsrc_t rd =
        __builtin_amdgcn_make_buffer_rsrc((void *)utab_d, (short)0, fa.tab_bytes, 0x00020000);
    if (mismatch[kConjBlocks] != 0)
        fused_work_item<COMP, Q, H64, true>(frecs, perm, ch, ci, slab_off, measured, model, vis_out, rg, rd, fa,
                                            slabs, cells);
    else
        fused_work_item<COMP, Q, H64, false>(frecs, perm, ch, ci, slab_off, measured, model, vis_out, rg, rd, fa,
                                             slabs, cells);
}

constexpr int kMergeThreads = 256;
// A key's pieces summed by one workgroup of the first reduce pass: a key with
// more (a dense tile can have a thousand) is split into groups whose partial
// slabs a second pass sums, so no workgroup's serial chain of loads is long.
constexpr int kGroupPieces = 64;

// Sum the slabs of a (class, tile) key's pieces, in piece order, into the
// key's slab.  One workgroup per key with more than one piece; each thread
// owns kC cells per pass and issues the loads of kP pieces x kC cells
// together.
template <int kC, int kP>
__global__ void __launch_bounds__(kMergeThreads)
reduce_pieces_kernel(const float2 *__restrict__ slabs, const long long *__restrict__ slab_off,
                     const int4 *__restrict__ multi, float2 *__restrict__ key_slabs)
{
    const int4 k = multi[blockIdx.x];   // first chunk, piece count, S*S, key slab start (cells, / 1)
    const long long out0 = ((const long long *)(multi + gridDim.x))[blockIdx.x];
    for (int i0 = 0; i0 < k.z; i0 += kMergeThreads * kC) {
        float sr[kC], si[kC], cr[kC], ci[kC];
#pragma unroll
        for (int c = 0; c < kC; ++c) sr[c] = si[c] = cr[c] = ci[c] = 0.f;
        for (int p0 = 0; p0 < k.y; p0 += kP) {
            float2 v[kP][kC];
#pragma unroll
            for (int q = 0; q < kP; ++q) {
                const bool pq = p0 + q < k.y;
                const long long off = pq ? slab_off[k.x + p0 + q] : 0;
#pragma unroll
                for (int c = 0; c < kC; ++c) {
                    const int i = i0 + c * kMergeThreads + threadIdx.x;
                    v[q][c] = (pq && i < k.z) ? slabs[off + i] : make_float2(0.f, 0.f);
                }
            }
            // absent values are zeros: adding them leaves sum and compensation unchanged
#pragma unroll
            for (int q = 0; q < kP; ++q)
#pragma unroll
                for (int c = 0; c < kC; ++c) {
                    two_sum_acc(sr[c], cr[c], v[q][c].x);
                    two_sum_acc(si[c], ci[c], v[q][c].y);
                }
        }
#pragma unroll
        for (int c = 0; c < kC; ++c) {
            const int i = i0 + c * kMergeThreads + threadIdx.x;
            if (i < k.z) key_slabs[out0 + i] = make_float2(sr[c] + cr[c], si[c] + ci[c]);
        }
    }
}

// Merge by output region: one workgroup per R x R block of grid cells
// (R = RG, 16 or 32, or G when smaller) that some tile's window reaches (the
// plan lists them; the grid is zeroed first, so the others stay zero).  Its
// cells are covered by the key slabs (one per (class, tile), pieces already
// summed) of the tiles whose windows [tile origin, tile origin + S) reach
// into the block; every cell of the block is the compensated sum of those
// slabs' values in (tile row, tile column, class) order, and every cell of
// it is written.  Consecutive lanes read consecutive cells of a slab row,
// and the loads of 8 slabs are in flight together.
constexpr int kRegion = 16;         // default region side (cells)
constexpr int kRegionList = 512;    // slab descriptors staged per pass
constexpr int kRegionCover = 1024;  // covering tiles per region (checked by the plan)

struct RegionArgs {
    int G, tile, ntx, nty, ntiles, smax, R;   // blockIdx.z: the grid
};

template <int RG, int kRegionBatch>   // kRegionBatch: slabs whose loads are issued together
__global__ void __launch_bounds__(kMergeThreads)
merge_region_kernel(const float2 *__restrict__ kslabs, const int *__restrict__ key_S,
                    const long long *__restrict__ key_off, const int *__restrict__ tile_key_begin,
                    const int2 *__restrict__ regions, RegionArgs a, float2 *__restrict__ grid)
{
    constexpr int kCells = RG * RG / kMergeThreads;   // cells per thread at R = RG
    static_assert(kCells >= 1 && RG % 4 == 0 && kWave % RG == 0, "region side");
    __shared__ long long s_off[kRegionList];
    __shared__ int s_geo[kRegionList];          // S | (by + 512) << 10 | (bx + 512) << 20
    __shared__ int s_base[kRegionCover + 1];    // first list position of each covering tile
    __shared__ int s_k0[kRegionCover];          // its first key
    __shared__ int s_b[kRegionCover];           // (by + 512) | (bx + 512) << 10
    const int R = a.R;
    const int2 reg = regions[blockIdx.x];   // workgroup -> region (x | y << 16, grid)
    const int oy0 = (reg.x >> 16) * R, ox0 = (reg.x & 0xffff) * R;
    tile_key_begin += (size_t)reg.y * a.ntiles;
    grid += (size_t)reg.y * a.G * a.G;
    // At R = RG each of the 4 wavefronts owns a band of RG / 4 rows (64 / RG
    // rows of RG lanes at a time, its cells 64 / RG rows apart) and skips the
    // additions of slabs that miss its band; smaller regions (G < RG) take
    // rows in passes of kMergeThreads / R.
    const bool banded = R == RG;
    const int wave = threadIdx.x / 64;
    const int rows_per_pass = banded ? kWave / RG : kMergeThreads / R;
    const int col = threadIdx.x % R;
    const int row0 = banded ? (RG / 4) * wave + (threadIdx.x % 64) / RG : threadIdx.x / R;
    const int band_lo = banded ? (RG / 4) * wave : 0, band_hi = banded ? band_lo + RG / 4 : R;
    float sr[kCells], si[kCells], cr[kCells], ci[kCells];
#pragma unroll
    for (int q = 0; q < kCells; ++q) sr[q] = si[q] = cr[q] = ci[q] = 0.f;

    // covering tiles: ty * tile < oy0 + R and ty * tile + smax > oy0
    const int ty_lo = oy0 - a.smax < 0 ? 0 : (oy0 - a.smax) / a.tile + 1;
    const int tx_lo = ox0 - a.smax < 0 ? 0 : (ox0 - a.smax) / a.tile + 1;
    const int ty_hi = min(a.nty - 1, (oy0 + R - 1) / a.tile), tx_hi = min(a.ntx - 1, (ox0 + R - 1) / a.tile);
    const int ntx_c = tx_hi - tx_lo + 1;
    const int nt = (ty_hi - ty_lo + 1) * ntx_c;
    for (int k = threadIdx.x; k < nt; k += kMergeThreads) {
        const int ty = ty_lo + k / ntx_c, tx = tx_lo + k % ntx_c;
        const int t = ty * a.ntx + tx;
        s_k0[k] = tile_key_begin[t];
        s_base[k + 1] = tile_key_begin[t + 1] - tile_key_begin[t];
        s_b[k] = (oy0 - ty * a.tile + 512) | ((ox0 - tx * a.tile + 512) << 10);
    }
    __syncthreads();
    {
        // s_base[1 .. nt] hold the covering tiles' key counts: turn them into
        // running totals, 4 entries per thread then a scan over the threads
        static_assert(kRegionCover <= 4 * kMergeThreads, "4 covering tiles per thread");
        __shared__ int s_wave[kMergeThreads / 64];
        const int t = threadIdx.x, lane = t & 63, w = t >> 6;
        int loc[4], sum = 0;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int k = 4 * t + j;
            sum += k < nt ? s_base[k + 1] : 0;
            loc[j] = sum;
        }
        int x = sum;   // inclusive scan over the wavefront
        for (int off = 1; off < 64; off <<= 1) {
            const int y = __shfl_up(x, off, 64);
            if (lane >= off) x += y;
        }
        if (lane == 63) s_wave[w] = x;
        __syncthreads();
        int before = x - sum;
        for (int v = 0; v < w; ++v) before += s_wave[v];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int k = 4 * t + j;
            if (k < nt) s_base[k + 1] = before + loc[j];
        }
        if (t == 0) s_base[0] = 0;
    }
    __syncthreads();
    const int total = s_base[nt];
    for (int base = 0; base < total; base += kRegionList) {
        const int n = min(kRegionList, total - base);
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
        for (int e0 = 0; e0 < n; e0 += kRegionBatch) {
            float2 v[kRegionBatch][kCells];
            unsigned hits = 0;   // slabs reaching this wavefront's band (uniform)
#pragma unroll
            for (int u = 0; u < kRegionBatch; ++u) {
                const int e = e0 + u;
                const int geo = e < n ? s_geo[e] : 0;
                const int S = geo & 1023, by = ((geo >> 10) & 1023) - 512, bx = ((geo >> 20) & 1023) - 512;
                // slab rows ly = by + r in [0, S): region rows r in [-by, S - by)
                if (e < n && max(-by, band_lo) < min(S - by, band_hi)) hits |= 1u << u;
                const int lx = bx + col;
                const bool colok = e < n && lx >= 0 && lx < S;
                const float2 *slab = kslabs + (e < n ? s_off[e] : 0) + lx;
#pragma unroll
                for (int q = 0; q < kCells; ++q) {
                    const int r = row0 + q * rows_per_pass, ly = by + r;
                    v[u][q] = (colok && r < R && ly >= 0 && ly < S) ? slab[ly * S] : make_float2(0.f, 0.f);
                }
            }
            // absent values are zeros: adding them leaves sum and compensation unchanged
            hits = __builtin_amdgcn_readfirstlane(hits);
#pragma unroll
            for (int u = 0; u < kRegionBatch; ++u) {
                if (!(hits & (1u << u))) continue;
#pragma unroll
                for (int q = 0; q < kCells; ++q) {
                    two_sum_acc(sr[q], cr[q], v[u][q].x);
                    two_sum_acc(si[q], ci[q], v[u][q].y);
                }
            }
        }
        __syncthreads();
    }
#pragma unroll
    for (int q = 0; q < kCells; ++q) {
        const int r = row0 + q * rows_per_pass;
        if (r < R && oy0 + r < a.G && ox0 + col < a.G)
            grid[(size_t)(oy0 + r) * a.G + ox0 + col] = make_float2(sr[q] + cr[q], si[q] + ci[q]);
    }
}

// n cells to zero (grids only 8-byte aligned)
__global__ void zero_cells_kernel(float2 *g, size_t n)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        g[i] = make_float2(0.f, 0.f);
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
    int n_grids = 1, ntiles = 0;   // grids of the plan (consecutive G x G grids), tiles per grid
    int n_chunks = 0, n_tiles_used = 0, n_class_launches = 0;
    long long total_taps = 0;
    int class_S[kAllClasses] = {};
    int class_h[kAllClasses] = {};
    int class_chunk_begin[kAllClasses + 1] = {};
    VisRec *recs = nullptr;        // plan order
    uint2 *frecs = nullptr;        // plan order: exact-footprint records (unfolded base, t_u | t_v << 8 | conj)
    int n_unfold = 0;              // planes with an unfolded table
    int4 *unfold_tab = nullptr;    // per such plane: quadrant offset, h, unfolded offset
    long long utab_len = 0;        // samples of the unfolded gridding table
    float2 *utab = nullptr;
    float2 *utab_d = nullptr;      // the unfolded degridding table (fused degrid + grid)
    int *conj_flag = nullptr;      // per-block and overall "some degridding sample is not the conjugate
                                   // gridding sample" flags, and an arrival counter (prep_tables_kernel)
    int n_items = 0;               // register-gridder items: (chunk, 64-position part)
    FastItems fast_items = {};
    long long n_fast = 0;          // visibilities in exact-footprint classes
    uint32_t *perm = nullptr;      // plan position -> caller index
    int2 *plane_tab = nullptr;     // (sample offset, h+1) per plane
    long long table_len = 0;       // samples in a kernel table (all planes)
    float2 *ktab = nullptr;        // phase-major copies: gridding table, then degridding table
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
    int region = 0;                    // merge region edge (min(region_rg, G))
    int region_rg = kRegion;           // merge kernel's region side, 16 or 32
    int n_regions = 0;                 // merge regions some tile's window reaches (over all grids)
    int2 *region_list = nullptr;       // their (x | y << 16, grid): the merge runs on these only
    int4 *multi = nullptr;         // first reduce pass: first chunk, pieces, S*S, -, then their output starts
    int n_multi2 = 0;
    int4 *multi2 = nullptr;        // second pass (keys of more than kGroupPieces pieces): first partial, count,
                                   // S*S, -, then their output starts
    long long *part_off = nullptr; // the partial slabs the second pass reads
    float2 *key_slabs = nullptr;   // contiguous with slabs: slabs followed by the reduced key slabs
};

namespace {

// Set-up buffers of plan creation, kept per device from one plan to the
// next (hipFree synchronises the device and costs about a quarter of a
// millisecond a call): grown when a plan needs more, released after a plan
// whose set-up holds more than kSetupKeep bytes.  Plan creations on one
// device take turns (the cache's mutex).
constexpr int kSetupSlots = 10, kHostSlots = 3;
constexpr size_t kSetupKeep = 4ull << 30, kHostKeep = 256ull << 20;
enum { SET_SUPPORTS, SET_PCLASS, SET_PUTAB, SET_RECS, SET_FRECS, SET_KEYS, SET_KEYS_SORTED, SET_VALS, SET_TEMP,
       SET_RUNS };
// pinned host staging of the work-item tables (their uploads then run at
// full link speed)
enum { HOST_CHUNKS, HOST_SLAB_OFF, HOST_TILE_KEYS };
struct SetupCache {
    std::mutex m;
    void *buf[kSetupSlots] = {};
    size_t cap[kSetupSlots] = {};
    void *hbuf[kHostSlots] = {};
    size_t hcap[kHostSlots] = {};
    hipError_t hget(int slot, size_t bytes, void **out)
    {
        if (hcap[slot] < bytes) {
            if (hbuf[slot]) (void)hipHostFree(hbuf[slot]);
            hbuf[slot] = nullptr;
            hcap[slot] = 0;
            const size_t want = bytes + bytes / 4;
            if (hipError_t e = hipHostMalloc(&hbuf[slot], want, hipHostMallocDefault)) {
                hbuf[slot] = nullptr;
                return e;
            }
            hcap[slot] = want;
        }
        *out = hbuf[slot];
        return hipSuccess;
    }
    hipError_t get(int slot, size_t bytes, void **out)
    {
        if (cap[slot] < bytes) {
            if (buf[slot]) (void)hipFree(buf[slot]);   // waits for the work that used it
            buf[slot] = nullptr;
            cap[slot] = 0;
            const size_t want = bytes + bytes / 4;     // headroom for the next, slightly larger plan
            if (hipError_t e = hipMalloc(&buf[slot], want)) {
                buf[slot] = nullptr;
                return e;
            }
            cap[slot] = want;
        }
        *out = buf[slot];
        return hipSuccess;
    }
    void trim()
    {
        size_t total = 0, htotal = 0;
        for (size_t c : cap) total += c;
        for (size_t c : hcap) htotal += c;
        if (total > kSetupKeep)
            for (int i = 0; i < kSetupSlots; ++i) {
                if (buf[i]) (void)hipFree(buf[i]);
                buf[i] = nullptr;
                cap[i] = 0;
            }
        if (htotal > kHostKeep)
            for (int i = 0; i < kHostSlots; ++i) {
                if (hbuf[i]) (void)hipHostFree(hbuf[i]);
                hbuf[i] = nullptr;
                hcap[i] = 0;
            }
    }
};
constexpr int kCacheDevices = 64;
SetupCache g_setup_cache[kCacheDevices];

void plan_free(sdp_grid_plan *p)
{
    if (!p) return;
    (void)hipFree(p->recs);
    (void)hipFree(p->frecs);
    (void)hipFree(p->unfold_tab);
    (void)hipFree(p->utab);
    (void)hipFree(p->utab_d);
    (void)hipFree(p->conj_flag);
    (void)hipFree(p->perm);
    (void)hipFree(p->plane_tab);
    (void)hipFree(p->ktab);
    (void)hipFree(p->chunks);
    (void)hipFree(p->slab_off);
    (void)hipFree(p->slabs);
    (void)hipFree(p->key_S);
    (void)hipFree(p->key_off);
    (void)hipFree(p->tile_key_begin);
    (void)hipFree(p->region_list);
    (void)hipFree(p->multi);
    (void)hipFree(p->multi2);
    (void)hipFree(p->part_off);
    delete p;
}

// lanes per visibility for a class: a power of two >= the largest footprint
// extent, at least 16, at most 64.  The extent is 2h + 1 in both snap modes:
// std_gridding's [ceilf(p - h), ceilf(p + h)) spans 2h cells in exact
// arithmetic, but p - h rounded to float can fall on the integer below and
// make it 2h + 1 (the reference loops over those cells too).
int class_lane_shift(int h, int snap_mode)
{
    (void)snap_mode;
    const int extent = 2 * h + 1;
    int shift = 4;
    while (shift < 6 && (1 << shift) < extent) ++shift;
    return shift;
}

// LDS of a chunk workgroup without the plane table (8-byte aligned)
size_t lds_window_bytes(int mode, int S, int copies, bool comp)
{
    const size_t cells = (size_t)S * (S + 1);
    const size_t win = mode == MODE_GRID ? 0 : ((cells + 2) & ~(size_t)1) * sizeof(float2);
    if (mode == MODE_DEGRID) return win;
    const size_t acc = comp ? sizeof(float4) : sizeof(float2);
    return win + ((size_t)copies * (cells + 2) + kWave) * acc;
}

// planes held in LDS: all of them when there are at most kLdsPlanes
int lds_planes(int n_kernels) { return n_kernels <= kLdsPlanes ? n_kernels : 0; }

size_t lds_bytes(int mode, int S, int copies, bool comp, int n_kernels)
{
    return lds_window_bytes(mode, S, copies, comp) + sizeof(int2) * (size_t)lds_planes(n_kernels);
}

// sub-grid copies that fit LDS for a class (a power of two)
int class_copies(int mode, int S, int col_shift, bool comp, int n_kernels)
{
    int c = kWave >> col_shift;
    while (c > 1 && lds_bytes(mode, S, c, comp, n_kernels) > kMaxLds) c >>= 1;
    return c;
}

// A per-class override for tuning: environment variable `name` holds one
// value (1, 2, 4 or 8) for every class or a comma-separated list, one value
// per class; 0 when unset.
int class_env(const char *name, int c)
{
    const char *e = getenv(name);
    int v = 0;
    for (int k = 0; e && *e && k <= c; ++k) {
        v = atoi(e);
        e = strchr(e, ',');
        if (e) ++e;
    }
    return (v == 1 || v == 2 || v == 4 || v == 8) ? v : 0;
}

// lane groups per wavefront for class c: more groups take more visibilities
// at once (less per-visibility work per group), fewer give each visibility
// more lanes (more footprint rows per step) and, when gridding, leave LDS
// for more resident wavefronts.  Measured at 16.7M visibilities, 32
// w-planes, h 4-16: gridding takes 2 groups for h <= 7 and 1 above, degridding
// as many as the class's footprint width allows.  SDP_GRID_COPIES and
// SDP_DEGRID_GROUPS override these for tuning.
int class_groups(int mode, bool comp, int c, int col_shift)
{
    const int widest = kWave >> col_shift;
    if (mode == MODE_DEGRID) {
        const int v = class_env("SDP_DEGRID_GROUPS", c);
        return v ? std::min(v, widest) : widest;
    }
    const int v = class_env("SDP_GRID_COPIES", c);
    if (v) return std::min(v, widest);
    return std::min(comp || c > 0 ? 1 : 2, widest);
}

// a contiguous range of work items launched together: chunks [b, e), the
// largest window S and half support h among them, and the support class
// whose tuning overrides apply
struct LaunchRange {
    int b, e, S, h, env_c;
    bool fast;
};

// The launches of the LDS kernels: one per support class, and the
// exact-footprint classes grouped by the lane-group width their largest
// footprint needs (h <= 7, <= 15, 16).
std::vector<LaunchRange> lds_ranges(const sdp_grid_plan *plan, bool with_fast)
{
    std::vector<LaunchRange> out;
    for (int c = 0; c < kAllClasses; ++c) {
        const int b = plan->class_chunk_begin[c], e = plan->class_chunk_begin[c + 1];
        if (e <= b || (is_fast(c) && !with_fast)) continue;
        const int h = plan->class_h[c], S = plan->class_S[c];
        if (is_fast(c) && !out.empty() && out.back().fast && out.back().e == b &&
            class_lane_shift(out.back().h, plan->p.snap_mode) == class_lane_shift(h, plan->p.snap_mode)) {
            LaunchRange &r = out.back();
            r.e = e;
            r.S = std::max(r.S, S);
            r.h = std::max(r.h, h);
            continue;
        }
        out.push_back(LaunchRange{b, e, S, h, is_fast(c) ? class_of(h) : c, is_fast(c)});
    }
    return out;
}

template <int MODE, bool COMP>
int launch_class(const sdp_grid_plan *plan, const LaunchRange &rg, const ChunkIO &io, const uint32_t *perm,
                 hipStream_t st)
{
    static bool attr_set = false;
    if (!attr_set) {
        (void)hipFuncSetAttribute((const void *)chunk_kernel<MODE, COMP, false>,
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)kMaxLds);
        (void)hipFuncSetAttribute((const void *)chunk_kernel<MODE, COMP, true>,
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)kMaxLds);
        attr_set = true;
    }
    const int b = rg.b, e = rg.e, S = rg.S, c = rg.env_c;
    const int col_shift = class_lane_shift(rg.h, plan->p.snap_mode);
    int copies = class_groups(MODE, COMP, c, col_shift);
    if (MODE != MODE_DEGRID) copies = std::min(copies, class_copies(MODE, S, col_shift, COMP, plan->p.n_kernels));
    int lane_shift = 0;
    while ((kWave >> lane_shift) > copies) ++lane_shift;   // lanes per group = 64 / copies (copies a power of 2)
    copies = kWave >> lane_shift;
    const int OS = plan->p.oversampling;
    int os_shift = -1;
    for (int k = 0; k < 12; ++k)
        if ((1 << k) == OS) os_shift = k;
    ChunkArgs a{plan->p.grid_size, OS, plan->tile, plan->ntx, plan->ntiles, b, lane_shift, copies, os_shift,
                (int)(8 * plan->table_len), lds_planes(plan->p.n_kernels),
                (int)lds_window_bytes(MODE, S, copies, COMP)};
    const size_t lds = lds_bytes(MODE, S, copies, COMP, plan->p.n_kernels);
    if (copies == 1)
        chunk_kernel<MODE, COMP, true><<<e - b, kWave, lds, st>>>(plan->recs, perm, plan->plane_tab, plan->chunks,
                                                                  plan->slab_off, a, io);
    else
        chunk_kernel<MODE, COMP, false><<<e - b, kWave, lds, st>>>(plan->recs, perm, plan->plane_tab, plan->chunks,
                                                                   plan->slab_off, a, io);
    SDP_LAUNCH_CHECK();
    return SDP_OK;
}

// the register gridder over every exact-footprint work item: one launch
int launch_fast(const sdp_grid_plan *plan, const float2 *vis, const uint32_t *perm, hipStream_t st)
{
    const bool comp = plan->p.accumulate == SDP_ACC_COMPENSATED;
    const int tb = (int)(8 * plan->utab_len);
    const int nb = sdp_div_up(plan->n_items, kFastWaves);
    const FastItems &fi = plan->fast_items;
    if (comp && perm)
        fast_grid_kernel<true, true><<<nb, kWave * kFastWaves, 0, st>>>(plan->frecs, perm, plan->chunks, fi,
                                                                       plan->slab_off, vis, plan->utab, tb,
                                                                       plan->tile, plan->slabs);
    else if (comp)
        fast_grid_kernel<true, false><<<nb, kWave * kFastWaves, 0, st>>>(plan->frecs, perm, plan->chunks, fi,
                                                                        plan->slab_off, vis, plan->utab, tb,
                                                                        plan->tile, plan->slabs);
    else if (perm)
        fast_grid_kernel<false, true><<<nb, kWave * kFastWaves, 0, st>>>(plan->frecs, perm, plan->chunks, fi,
                                                                        plan->slab_off, vis, plan->utab, tb,
                                                                        plan->tile, plan->slabs);
    else
        fast_grid_kernel<false, false><<<nb, kWave * kFastWaves, 0, st>>>(plan->frecs, perm, plan->chunks, fi,
                                                                         plan->slab_off, vis, plan->utab, tb,
                                                                         plan->tile, plan->slabs);
    SDP_LAUNCH_CHECK();
    return SDP_OK;
}


template <bool COMP, int Q, bool H64>
int launch_fused_group(const sdp_grid_plan *plan, int b, int e, const ChunkIO &io, const uint32_t *perm,
                       hipStream_t st)
{
    if (e <= b) return SDP_OK;
    using Shape = FusedShape<Q, H64>;
    constexpr int waves = Shape::waves;
    constexpr size_t lds_max = Shape::template lds<COMP>();
    static bool attr_set = false;
    if (!attr_set) {
        (void)hipFuncSetAttribute((const void *)fast_degrid_grid_kernel<COMP, Q, H64>,
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_max);
        attr_set = true;
    }
    int smax = 0;   // the largest window of the classes in [b, e)
    for (int c = kNumClasses; c < kAllClasses; ++c)
        if (plan->class_chunk_begin[c + 1] > plan->class_chunk_begin[c] && plan->class_chunk_begin[c] < e &&
            plan->class_chunk_begin[c + 1] > b)
            smax = std::max(smax, plan->class_S[c]);
    const size_t lds = (size_t)waves * smax * smax * (COMP ? sizeof(float4) : sizeof(float2));
    if (lds > lds_max)
        return sdp::set_error(SDP_E_UNSUPPORTED, "window of %d cells exceeds the fused kernel's LDS", smax);
    FusedArgs fa{b, e, plan->tile, plan->p.grid_size, plan->ntx, plan->ntiles, (int)(8 * plan->utab_len),
                 smax * smax};
    fast_degrid_grid_kernel<COMP, Q, H64><<<(e - b + waves - 1) / waves, kWave * waves, lds, st>>>(
        plan->frecs, perm, plan->chunks, plan->slab_off, io.measured, io.grid, io.vis_out, plan->utab, plan->utab_d,
        plan->conj_flag, fa, plan->slabs);
    SDP_LAUNCH_CHECK();
    return SDP_OK;
}

// the fused kernel's launches: exact-footprint classes grouped by their
// part shape (pair parts, single part) for W^2 = 4 h^2 taps: h <= 4 one
// single part, 5 one pair part, 6 one pair part + single, 7-8 two pair
// parts, 9 three, 10 three + single, 11 four, 12 four + single, 13-16 up to
// eight + single (a group's kernel runs all its parts; unused ones read
// zeros, so only the rare h >= 13 group has any)
constexpr int kFusedGroups = 9;
constexpr int kFusedLo[kFusedGroups] = {1, 5, 6, 7, 9, 10, 11, 12, 13};
constexpr int kFusedHi[kFusedGroups] = {4, 5, 6, 8, 9, 10, 11, 12, 16};
template <bool COMP>
int launch_fused(const sdp_grid_plan *plan, int group, const ChunkIO &io, const uint32_t *perm, hipStream_t st)
{
    const int b = plan->class_chunk_begin[fast_class(kFusedLo[group])];
    const int e = plan->class_chunk_begin[fast_class(kFusedHi[group]) + 1];
    switch (group) {
    case 0: return launch_fused_group<COMP, 0, true>(plan, b, e, io, perm, st);
    case 1: return launch_fused_group<COMP, 1, false>(plan, b, e, io, perm, st);
    case 2: return launch_fused_group<COMP, 1, true>(plan, b, e, io, perm, st);
    case 3: return launch_fused_group<COMP, 2, false>(plan, b, e, io, perm, st);
    case 4: return launch_fused_group<COMP, 3, false>(plan, b, e, io, perm, st);
    case 5: return launch_fused_group<COMP, 3, true>(plan, b, e, io, perm, st);
    case 6: return launch_fused_group<COMP, 4, false>(plan, b, e, io, perm, st);
    case 7: return launch_fused_group<COMP, 4, true>(plan, b, e, io, perm, st);
    default: return launch_fused_group<COMP, 8, true>(plan, b, e, io, perm, st);
    }
}

int fused_group_chunks(const sdp_grid_plan *plan, int group)
{
    return plan->class_chunk_begin[fast_class(kFusedHi[group]) + 1] -
           plan->class_chunk_begin[fast_class(kFusedLo[group])];
}

// The kernel-table copies a pass reads, in one launch (prep_tables_kernel):
// the phase-major copies when the LDS kernels run (slot 0 gridding, 1
// degridding, into plan->ktab), the unfolded copies (and the conjugate
// check, with both tables) when exact-footprint classes are present.
struct Prepared {
    const float2 *rel_g = nullptr, *rel_d = nullptr;
};
Prepared prepare_tables(const sdp_grid_plan *plan, const float *d_kg, const float *d_kd, bool lds_kernels,
                        bool fast_kernels, hipStream_t st)
{
    Prepared out;
    PrepArgs a{};
    a.kg = (const float2 *)d_kg;
    a.kd = (const float2 *)d_kd;
    a.OS = plan->p.oversampling;
    if (lds_kernels && d_kg) {
        a.rel_g = plan->ktab;
        a.rows_rel_g = plan->p.n_kernels;
        out.rel_g = a.rel_g;
    }
    if (lds_kernels && d_kd) {
        a.rel_d = plan->ktab + (size_t)plan->table_len;
        a.rows_rel_d = plan->p.n_kernels;
        out.rel_d = a.rel_d;
    }
    if (fast_kernels && plan->n_items > 0) {
        if (d_kg) {
            a.unf_g = plan->utab;
            a.rows_unf_g = plan->n_unfold;
        }
        if (d_kd) {
            a.unf_d = plan->utab_d;
            a.rows_unf_d = plan->n_unfold;
        }
        a.conj = d_kg && d_kd;
        a.table_len = plan->table_len;
        a.mismatch = plan->conj_flag;
    }
    const int rows = a.rows_rel_g + a.rows_rel_d + a.rows_unf_g + a.rows_unf_d + (a.conj ? 1 : 0);
    if (rows > 0)
        prep_tables_kernel<<<dim3(kPrepBlocks, rows), 256, 0, st>>>(plan->plane_tab, plan->unfold_tab, a);
    return out;
}

// This thread's helper stream for the current device (default priority) and
// its fork and join events; created on first use.
int side_stream(hipStream_t *side, hipEvent_t *fork, hipEvent_t *join)
{
    struct Side {
        int dev = -1;
        hipStream_t s = nullptr;
        hipEvent_t fork = nullptr, join = nullptr;
    };
    thread_local Side t;
    int dev = 0;
    SDP_HIP_TRY(hipGetDevice(&dev));
    if (t.s == nullptr || t.dev != dev) {
        Side n;
        n.dev = dev;
        SDP_HIP_TRY(hipStreamCreateWithFlags(&n.s, hipStreamNonBlocking));
        SDP_HIP_TRY(hipEventCreateWithFlags(&n.fork, hipEventDisableTiming));
        SDP_HIP_TRY(hipEventCreateWithFlags(&n.join, hipEventDisableTiming));
        t = n;   // a previous device's objects stay with that device
    }
    *side = t.s;
    *fork = t.fork;
    *join = t.join;
    return SDP_OK;
}

template <int MODE>
int launch_chunks(const sdp_grid_plan *plan, const ChunkIO &io, const uint32_t *perm, hipStream_t st)
{
    const bool comp = plan->p.accumulate == SDP_ACC_COMPENSATED;
    // gridding takes the exact-footprint classes in one register-gridder
    // launch, degrid + grid in up to four fused launches
    const bool fast = MODE == MODE_GRID && plan->n_items > 0;
    int fused[kFusedGroups] = {};
    int n_fused = 0;
    if (MODE == MODE_DEGRID_GRID && plan->n_items > 0)
        for (int q = 0; q < kFusedGroups; ++q) n_fused += (fused[q] = fused_group_chunks(plan, q)) > 0;
    const std::vector<LaunchRange> ranges = lds_ranges(plan, MODE == MODE_DEGRID || (MODE == MODE_DEGRID_GRID &&
                                                                                     plan->n_items == 0));
    // The launch with the most work runs on the caller's stream (the
    // register gridder when it runs); the others, too small to fill the GPU
    // on their own, run beside it on a helper stream forked from and joined
    // back to it (slabs of different launches are disjoint).
    int big = -1;
    if (n_fused == 0 && !fast) {
        for (size_t i = 0; i < ranges.size(); ++i)
            if (big < 0 || ranges[i].e - ranges[i].b > ranges[big].e - ranges[big].b) big = (int)i;
    }
    const size_t n_launch = ranges.size() + (fast ? 1 : 0) + n_fused;
    hipStream_t side = st;
    hipEvent_t join = nullptr;
    if (n_launch > 1) {
        hipEvent_t fork = nullptr;
        if (side_stream(&side, &fork, &join) == SDP_OK) {
            SDP_HIP_TRY(hipEventRecord(fork, st));
            SDP_HIP_TRY(hipStreamWaitEvent(side, fork, 0));
        } else {
            side = st;
            join = nullptr;
        }
    }
    int rc = SDP_OK;
    if (fast) rc = launch_fast(plan, io.vis, perm, st);
    // the fused launches alternate between the caller's stream and the
    // helper stream, so that one launch's last wavefronts overlap the next
    // launch's first instead of leaving the GPU part-idle at every boundary
    for (int q = 0, nth = 0; q < kFusedGroups && rc == SDP_OK; ++q)
        if (fused[q] > 0) {
            hipStream_t s = (nth++ & 1) ? side : st;
            rc = comp ? launch_fused<true>(plan, q, io, perm, s) : launch_fused<false>(plan, q, io, perm, s);
        }
    for (size_t i = 0; i < ranges.size() && rc == SDP_OK; ++i) {
        hipStream_t s = (int)i == big ? st : side;
        rc = (comp || MODE == MODE_DEGRID) ? launch_class<MODE, true>(plan, ranges[i], io, perm, s)
                                           : launch_class<MODE, false>(plan, ranges[i], io, perm, s);
    }
    // joined on the error path too, so a capturing caller's stream is never
    // left with an unjoined fork
    if (join) {
        SDP_HIP_TRY(hipEventRecord(join, side));
        SDP_HIP_TRY(hipStreamWaitEvent(st, join, 0));
    }
    return rc;
}

}  // namespace

extern "C" {

int sdp_grid_plan_create(sdp_grid_plan **out, const sdp_grid_params *pp, const int *supports,
                         const float *d_uvw, int n_vis, sdp_stream s)
{
    return sdp_grid_plan_create_multi(out, pp, supports, 1, &d_uvw, &n_vis, s);
}

int sdp_grid_plan_create_multi(sdp_grid_plan **out, const sdp_grid_params *pp, const int *supports, int n_grids,
                               const float *const *d_uvw_of, const int *n_vis_of, sdp_stream s)
{
    SDP_REQUIRE(out && pp && supports && d_uvw_of && n_vis_of, "sdp_grid_plan_create: null argument");
    *out = nullptr;
    SDP_REQUIRE(n_grids >= 1, "grid count %d must be at least 1", n_grids);
    std::vector<int> grid_first(n_grids + 1, 0);
    for (int g = 0; g < n_grids; ++g) {
        SDP_REQUIRE(n_vis_of[g] >= 0, "negative visibility count for grid %d", g);
        SDP_REQUIRE(n_vis_of[g] == 0 || d_uvw_of[g], "null coordinates for grid %d", g);
        SDP_REQUIRE((long long)grid_first[g] + n_vis_of[g] < (1LL << 31), "more than 2^31 - 1 visibilities");
        grid_first[g + 1] = grid_first[g] + n_vis_of[g];
    }
    const int n_vis = grid_first[n_grids];
    const sdp_grid_params p = *pp;
    SDP_REQUIRE(p.grid_size >= 2 && p.grid_size % 2 == 0 && p.grid_size <= 32768,
                "grid size %d must be even and in [2, 32768]", p.grid_size);
    SDP_REQUIRE(p.oversampling >= 1 && p.n_kernels >= 1 && p.n_kernels <= 8192,
                "bad oversampling / n_kernels (1 to 8192 planes)");
    SDP_REQUIRE(p.snap_mode == SDP_SNAP_STD || p.snap_mode == SDP_SNAP_GRIDDING_ACTOR, "bad snap mode");
    SDP_REQUIRE(p.tile >= 0 && p.tile <= 32 && p.chunk >= 0, "tile must be in [1, 32] (0 = default)");
    SDP_REQUIRE(p.accumulate == SDP_ACC_BLOCKED || p.accumulate == SDP_ACC_COMPENSATED, "bad accumulate mode");
    hipStream_t st = sdp::stream_of(s);

    std::vector<int2> tab(p.n_kernels), sup(p.n_kernels);
    std::vector<int> pclass(p.n_kernels);
    int class_h[kAllClasses];
    for (int c = 0; c < kAllClasses; ++c) class_h[c] = -1;
    int H = 0;
    long long table_len = 0;
    for (int k = 0; k < p.n_kernels; ++k) {
        const int h = supports[2 * k], off = supports[2 * k + 1];
        SDP_REQUIRE(h >= 0 && off >= 0, "negative half support or offset for plane %d", k);
        if (h > kClassMaxH[kNumClasses - 1])
            return sdp::set_error(SDP_E_UNSUPPORTED, "half support %d of plane %d exceeds %d", h, k,
                                  kClassMaxH[kNumClasses - 1]);
        // a footprint row or column (up to 2h + 1 cells) must fit one
        // wavefront's lanes
        if (2 * h + 1 > kWave)
            return sdp::set_error(SDP_E_UNSUPPORTED, "half support %d of plane %d: footprint wider than %d cells", h,
                                  k, kWave);
        H = std::max(H, h);
        pclass[k] = class_of(h);
        class_h[pclass[k]] = std::max(class_h[pclass[k]], h);
        tab[k] = make_int2(off, h + 1);
        table_len = std::max(table_len, (long long)off + (long long)(h + 1) * p.oversampling * (h + 1) * p.oversampling);
        sup[k] = make_int2(h, off);
    }

    SDP_REQUIRE(table_len < (1LL << 26), "kernel table of %lld samples is too large", table_len);
    {
        // the phase-major copy permutes samples inside each plane's region,
        // so regions must not overlap (planes sharing one region are fine)
        std::vector<int> order(p.n_kernels);
        for (int k = 0; k < p.n_kernels; ++k) order[k] = k;
        std::sort(order.begin(), order.end(), [&](int x, int y) { return tab[x].x < tab[y].x; });
        for (int i = 1; i < p.n_kernels; ++i) {
            const int2 a = tab[order[i - 1]], b = tab[order[i]];
            const long long end = (long long)a.x + (long long)a.y * p.oversampling * a.y * p.oversampling;
            if (b.x < end && !(b.x == a.x && b.y == a.y))
                return sdp::set_error(SDP_E_UNSUPPORTED, "kernel planes %d and %d overlap in the table", order[i - 1],
                                      order[i]);
        }
    }

    sdp_grid_plan *plan = new sdp_grid_plan();
    plan->p = p;
    plan->n = n_vis;
    plan->n_grids = n_grids;
    plan->H = H;
    // Default tile: 8, or the largest power of two within 2h + 1 of the
    // smallest support when that is narrower, so that every plane's
    // footprint spans the tile and can take the exact-footprint classes.
    int tile_default = 8;
    if (p.snap_mode == SDP_SNAP_STD)
        for (int k = 0; k < p.n_kernels; ++k)
            while (supports[2 * k] >= 1 && tile_default > 2 * supports[2 * k] + 1) tile_default >>= 1;
    plan->tile = std::min(p.tile > 0 ? p.tile : tile_default, p.grid_size);
    plan->chunk_taps = p.chunk > 0 ? p.chunk : (1 << 17);   // the default is settled once the taps are counted
    plan->ntx = (p.grid_size + plan->tile - 1) / plan->tile;
    plan->nty = plan->ntx;
    // Exact-footprint classes (register gridder, std_gridding footprints
    // only): planes with 1 <= h <= 16 whose 2h-wide footprint spans the tile
    // (2h + 1 >= tile), each with an unfolded table of OS^2 (2h + 1)^2
    // samples.  SDP_GRID_FAST=0 turns them off (every visibility then takes
    // the LDS kernels).
    std::vector<int> plane_utab(p.n_kernels, -1);
    std::vector<int4> unfold_tab;
    {
        const char *e = getenv("SDP_GRID_FAST");
        const bool on = !(e && atoi(e) == 0) && p.snap_mode == SDP_SNAP_STD;
        long long len = 0;
        for (int k = 0; on && k < p.n_kernels; ++k) {
            const int h = supports[2 * k];
            if (h < 1 || h > kFastMaxH || 2 * h + 1 < plan->tile) continue;
            const long long nn = 2LL * h + 1;
            if ((len + (long long)p.oversampling * p.oversampling * nn * nn) * 8 >= (1LL << 31)) break;
            plane_utab[k] = (int)len;
            unfold_tab.push_back(make_int4(supports[2 * k + 1], h, (int)len, 0));
            len += (long long)p.oversampling * p.oversampling * nn * nn;
            class_h[fast_class(h)] = h;
        }
        plan->utab_len = len;
        plan->n_unfold = (int)unfold_tab.size();
    }
    for (int c = 0; c < kAllClasses; ++c) {
        plan->class_S[c] = class_h[c] >= 0 ? plan->tile + 2 * class_h[c] - (is_fast(c) ? 1 : 0) : 0;
        plan->class_h[c] = std::max(class_h[c], 0);
        plan->smax = std::max(plan->smax, plan->class_S[c]);
        if (class_h[c] >= 0 &&
            lds_bytes(MODE_DEGRID_GRID, plan->class_S[c], 1, p.accumulate == SDP_ACC_COMPENSATED, p.n_kernels) >
                kMaxLds) {
            const int S = plan->class_S[c], tile = plan->tile;
            delete plan;
            return sdp::set_error(SDP_E_UNSUPPORTED, "window of %d cells (tile %d, half support %d) exceeds LDS", S,
                                  tile, class_h[c]);
        }
    }
    // sort keys rank the support classes present (in class order), so that
    // absent classes take no key space
    int class_rank[kAllClasses], n_ranked = 0;
    for (int c = 0; c < kAllClasses; ++c) class_rank[c] = class_h[c] >= 0 ? n_ranked++ : -1;
    if ((long long)plan->ntx * plan->nty * n_grids * std::max(n_ranked, 1) >= (1LL << 31)) {
        const int tile = plan->tile;
        delete plan;
        return sdp::set_error(SDP_E_UNSUPPORTED, "tile %d too small for %d grids of size %d", tile, n_grids,
                              p.grid_size);
    }
    const int ntiles = plan->ntx * plan->nty;
    plan->ntiles = ntiles;
    const int ntiles_all = ntiles * n_grids;
    const int nkeys = std::max(n_ranked, 1) * ntiles_all;
    {
        // SDP_MERGE_REGION (16 or 32) overrides the region side for tuning
        const char *e = getenv("SDP_MERGE_REGION");
        const int rg = e ? atoi(e) : 0;
        plan->region_rg = (rg == 16 || rg == 32) ? rg : kRegion;
        plan->region = std::min(plan->region_rg, p.grid_size);
        const int rspan = (plan->smax + plan->region + 2 * plan->tile - 1) / plan->tile;
        if (rspan * rspan > kRegionCover) {
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

    plan->table_len = table_len;
    PLAN_TRY(hipMalloc(&plan->ktab, sizeof(float2) * 2 * (size_t)std::max(table_len, 1LL)));
    PLAN_TRY(hipMalloc(&plan->plane_tab, sizeof(int2) * p.n_kernels));
    PLAN_TRY(hipMemcpyAsync(plan->plane_tab, tab.data(), sizeof(int2) * p.n_kernels, hipMemcpyHostToDevice, st));
    if (plan->n_unfold > 0) {
        PLAN_TRY(hipMalloc(&plan->unfold_tab, sizeof(int4) * plan->n_unfold));
        PLAN_TRY(hipMemcpyAsync(plan->unfold_tab, unfold_tab.data(), sizeof(int4) * plan->n_unfold,
                                hipMemcpyHostToDevice, st));
        PLAN_TRY(hipMalloc(&plan->utab, sizeof(float2) * (size_t)plan->utab_len));
        PLAN_TRY(hipMalloc(&plan->utab_d, sizeof(float2) * (size_t)plan->utab_len));
        PLAN_TRY(hipMalloc(&plan->conj_flag, sizeof(int) * (kConjBlocks + 2)));
        PLAN_TRY(hipMemsetAsync(plan->conj_flag, 0, sizeof(int) * (kConjBlocks + 2), st));
    }
    int2 *d_supports = nullptr;
    int *d_pclass = nullptr, *d_putab = nullptr;
    VisRec *recs_tmp = nullptr;
    uint2 *frecs_tmp = nullptr;
    uint32_t *keys = nullptr, *keys_sorted = nullptr, *vals = nullptr;
    KeyRun *d_runs = nullptr;
    // the set-up buffers above come from this device's cache, held until the
    // plan is built (or an error returns)
    int dev = 0;
    PLAN_TRY(hipGetDevice(&dev));
    if (dev < 0 || dev >= kCacheDevices)
        return fail(sdp::set_error(SDP_E_UNSUPPORTED, "device %d: at most %d devices", dev, kCacheDevices));
    SetupCache &cache = g_setup_cache[dev];
    std::lock_guard<std::mutex> cache_lock(cache.m);
    struct TrimGuard {
        SetupCache &c;
        ~TrimGuard() { c.trim(); }
    } trim_guard{cache};
    auto setup_buf = [&](int slot, size_t bytes, auto **ptr) {
        void *q = nullptr;
        hipError_t e = cache.get(slot, bytes, &q);
        using T = std::remove_pointer_t<std::remove_reference_t<decltype(*ptr)>>;
        *ptr = static_cast<T *>(q);
        return e;
    };
    PLAN_TRY(setup_buf(SET_SUPPORTS, sizeof(int2) * p.n_kernels, &d_supports));
    PLAN_TRY(hipMemcpyAsync(d_supports, sup.data(), sizeof(int2) * p.n_kernels, hipMemcpyHostToDevice, st));
    PLAN_TRY(setup_buf(SET_PCLASS, sizeof(int) * p.n_kernels, &d_pclass));
    PLAN_TRY(hipMemcpyAsync(d_pclass, pclass.data(), sizeof(int) * p.n_kernels, hipMemcpyHostToDevice, st));
    PLAN_TRY(setup_buf(SET_PUTAB, sizeof(int) * p.n_kernels, &d_putab));
    PLAN_TRY(hipMemcpyAsync(d_putab, plane_utab.data(), sizeof(int) * p.n_kernels, hipMemcpyHostToDevice, st));

    const size_t n = (size_t)std::max(n_vis, 1);
    PLAN_TRY(setup_buf(SET_RECS, sizeof(VisRec) * n, &recs_tmp));
    PLAN_TRY(setup_buf(SET_FRECS, sizeof(uint2) * n, &frecs_tmp));
    PLAN_TRY(hipMalloc(&plan->recs, sizeof(VisRec) * n));
    PLAN_TRY(hipMalloc(&plan->frecs, sizeof(uint2) * n));
    PLAN_TRY(hipMalloc(&plan->perm, sizeof(uint32_t) * n));
    PLAN_TRY(setup_buf(SET_KEYS, sizeof(uint32_t) * n, &keys));
    PLAN_TRY(setup_buf(SET_KEYS_SORTED, sizeof(uint32_t) * n, &keys_sorted));
    PLAN_TRY(setup_buf(SET_VALS, sizeof(uint32_t) * n, &vals));

    std::vector<KeyRun> runs;
    if (n_vis > 0) {
        PlanArgs a{n_vis, p.grid_size, p.oversampling, p.n_kernels, p.snap_mode, plan->tile, plan->ntx, plan->nty,
                   ntiles, nkeys, 0, p.uv_scale, p.w_scale, 0, 0, ntiles_all, {}};
        for (int c = 0; c < kAllClasses; ++c) a.class_rank[c] = class_rank[c];
        // The low field of the sort key: the plane (slow classes; neighbours
        // then share a kernel table) and (t_u, t_v) (exact-footprint classes,
        // whose lanes keep their tap while it stays the same), when the
        // combined key, dropped records' nkeys << bits included, fits 32
        // bits; (t_u, t_v) alone when only that fits.
        int pb = 0, tb = 0;
        while ((1 << pb) < p.n_kernels) ++pb;
        if (plan->n_unfold > 0) {
            while ((1 << tb) < plan->tile * plan->tile) ++tb;
            ++tb;   // and the conjugation flag
        }
        auto fits = [&](int bits) { return ((unsigned long long)nkeys << bits) < (1ull << 32); };
        if (fits(std::max(pb, tb))) a.plane_bits = std::max(pb, tb);
        else if (tb > 0 && fits(tb)) a.plane_bits = tb;
        // one launch per grid, each writing its visibilities' slots
        for (int g = 0; g < n_grids; ++g) {
            if (n_vis_of[g] == 0) continue;
            PlanArgs ag = a;
            ag.n = n_vis_of[g];
            ag.grid = g;
            ag.first = grid_first[g];
            plan_records_kernel<<<sdp_div_up(ag.n, 256), 256, 0, st>>>(d_uvw_of[g], d_supports, d_pclass, d_putab, ag,
                                                                       recs_tmp, frecs_tmp, keys, vals);
            PLAN_TRY(hipGetLastError());
        }
        int end_bit = 1;
        while (end_bit < 32 && (1ull << end_bit) <= ((unsigned long long)nkeys << a.plane_bits)) ++end_bit;
        size_t temp_bytes = 0;
        PLAN_TRY(rocprim::radix_sort_pairs(nullptr, temp_bytes, keys, keys_sorted, vals, plan->perm, (size_t)n_vis,
                                           0, (unsigned)end_bit, st));
        void *temp = nullptr;
        PLAN_TRY(setup_buf(SET_TEMP, temp_bytes, &temp));
        hipError_t se = rocprim::radix_sort_pairs(temp, temp_bytes, keys, keys_sorted, vals, plan->perm,
                                                  (size_t)n_vis, 0, (unsigned)end_bit, st);
        PLAN_TRY(se);
        gather_records_kernel<<<sdp_div_up(n_vis, 256), 256, 0, st>>>(recs_tmp, frecs_tmp, plan->perm, n_vis,
                                                                      plan->recs, plan->frecs);
        PLAN_TRY(hipGetLastError());
        if (plan->n_unfold > 0) {
            // run lengths of the sorted exact-footprint records; the unsorted
            // keys, the values and the unsorted records' buffer are free
            uint32_t *fl = keys, *rid = vals, *rend = reinterpret_cast<uint32_t *>(recs_tmp);
            fast_run_flags_kernel<<<sdp_div_up(n_vis, 256), 256, 0, st>>>(keys_sorted, plan->frecs, n_vis,
                                                                          a.plane_bits, fl);
            PLAN_TRY(hipGetLastError());
            size_t bb = 0;
            PLAN_TRY(rocprim::inclusive_scan(nullptr, bb, fl, rid, (size_t)n_vis, rocprim::plus<uint32_t>(), st));
            PLAN_TRY(setup_buf(SET_TEMP, bb, &temp));
            se = rocprim::inclusive_scan(temp, bb, fl, rid, (size_t)n_vis, rocprim::plus<uint32_t>(), st);
            PLAN_TRY(se);
            fast_run_ends_kernel<<<sdp_div_up(n_vis, 256), 256, 0, st>>>(fl, rid, n_vis, rend);
            PLAN_TRY(hipGetLastError());
            fast_run_rem_kernel<<<sdp_div_up(n_vis, 256), 256, 0, st>>>(rid, rend, n_vis, plan->frecs);
            PLAN_TRY(hipGetLastError());
        }

        // recs_tmp (16 bytes a record) is free again: its two halves hold
        // the tap counts and their inclusive scan; the unsorted keys and
        // values hold the run flags and run numbers
        unsigned long long *tap_count = reinterpret_cast<unsigned long long *>(recs_tmp);
        unsigned long long *tap_scan = tap_count + n_vis;
        uint32_t *flags = keys, *run_no = vals;
        run_flags_kernel<<<sdp_div_up(n_vis, 256), 256, 0, st>>>(keys_sorted, plan->recs, n_vis, nkeys, a.plane_bits,
                                                                 flags, tap_count);
        PLAN_TRY(hipGetLastError());
        size_t b1 = 0, b2 = 0;
        PLAN_TRY(rocprim::inclusive_scan(nullptr, b1, tap_count, tap_scan, (size_t)n_vis,
                                         rocprim::plus<unsigned long long>(), st));
        PLAN_TRY(rocprim::exclusive_scan(nullptr, b2, flags, run_no, 0u, (size_t)n_vis, rocprim::plus<uint32_t>(),
                                         st));
        PLAN_TRY(setup_buf(SET_TEMP, std::max(b1, b2), &temp));
        se = rocprim::inclusive_scan(temp, b1, tap_count, tap_scan, (size_t)n_vis,
                                     rocprim::plus<unsigned long long>(), st);
        if (se == hipSuccess)
            se = rocprim::exclusive_scan(temp, b2, flags, run_no, 0u, (size_t)n_vis, rocprim::plus<uint32_t>(), st);
        uint32_t last[2] = {0, 0};
        if (se == hipSuccess) se = hipMemcpyAsync(&last[0], run_no + n_vis - 1, 4, hipMemcpyDeviceToHost, st);
        if (se == hipSuccess) se = hipMemcpyAsync(&last[1], flags + n_vis - 1, 4, hipMemcpyDeviceToHost, st);
        (void)hipStreamSynchronize(st);   // the run count is read below
        PLAN_TRY(se);
        const size_t n_runs = (size_t)last[0] + last[1];
        if (n_runs > 0) {
            PLAN_TRY(setup_buf(SET_RUNS, sizeof(KeyRun) * n_runs, &d_runs));
            run_start_kernel<<<sdp_div_up(n_vis, 256), 256, 0, st>>>(keys_sorted, flags, run_no, n_vis,
                                                                     a.plane_bits, d_runs);
            PLAN_TRY(hipGetLastError());
            run_end_kernel<<<sdp_div_up(n_vis, 256), 256, 0, st>>>(keys_sorted, flags, run_no, n_vis, nkeys,
                                                                   a.plane_bits, tap_scan, d_runs);
            PLAN_TRY(hipGetLastError());
            runs.resize(n_runs);
            PLAN_TRY(hipMemcpyAsync(runs.data(), d_runs, sizeof(KeyRun) * n_runs, hipMemcpyDeviceToHost, st));
        }
    }
    PLAN_TRY(hipStreamSynchronize(st));

    // work items, class-major then tile-major, each (class, tile) list cut
    // into pieces of about chunk_taps taps
    // chunks and slab_off: pinned staging holding at most
    // sum ceil(taps / chunk_taps) <= runs + total taps / chunk_taps + 1 entries
    int4 *chunks = nullptr;
    long long *slab_off = nullptr;
    size_t n_ch = 0, max_ch = 0;
    std::vector<int> key_S, key_tile;
    std::vector<long long> key_first_chunk, key_pieces;
    {
        unsigned long long all_taps = 0;
        for (const KeyRun &r : runs) all_taps += r.taps;
        if (p.chunk <= 0) {
            // default work-item size: about eight work items per CU, within
            // 4096 and 131072 taps
            static int cus = 0;
            if (cus == 0) {
                int dev = 0, v = 0;
                cus = (hipGetDevice(&dev) == hipSuccess &&
                       hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && v > 0)
                          ? v
                          : 256;
            }
            plan->chunk_taps =
                (int)std::min<unsigned long long>(1ull << 17, std::max<unsigned long long>(4096, all_taps / (8ull * cus)));
        }
        max_ch = runs.size() + (size_t)(all_taps / (unsigned long long)plan->chunk_taps) + 1;
        PLAN_TRY(cache.hget(HOST_CHUNKS, sizeof(int4) * max_ch, (void **)&chunks));
        PLAN_TRY(cache.hget(HOST_SLAB_OFF, sizeof(long long) * max_ch, (void **)&slab_off));
        for (auto *v : {&key_S, &key_tile}) v->reserve(runs.size());
        for (auto *v : {&key_first_chunk, &key_pieces}) v->reserve(runs.size());
    }
    long long cells = 0;
    int kept = 0;
    size_t ri = 0;
    for (int c = 0; c < kAllClasses; ++c) {
        plan->class_chunk_begin[c] = (int)n_ch;
        for (; class_rank[c] >= 0 && ri < runs.size() && (int)(runs[ri].key / (uint32_t)ntiles_all) == class_rank[c];
             ++ri) {
            const KeyRun &r = runs[ri];
            // t: the tile over all grids (grid * ntiles + tile)
            const int t = (int)(r.key % (uint32_t)ntiles_all), cnt = (int)r.count, first = (int)r.first;
            long long pieces = (long long)((r.taps + plan->chunk_taps - 1) / plan->chunk_taps);
            pieces = std::max(1LL, std::min(pieces, (long long)cnt));
            key_tile.push_back(t);
            key_S.push_back(plan->class_S[c]);
            key_first_chunk.push_back((long long)n_ch);
            if (n_ch + (size_t)pieces > max_ch)   // cannot happen (the bound above)
                return fail(sdp::set_error(SDP_E_UNSUPPORTED, "work-item count past its bound"));
            key_pieces.push_back(pieces);
            for (long long q = 0; q < pieces; ++q) {
                const int b = first + (int)(cnt * q / pieces), e = first + (int)(cnt * (q + 1) / pieces);
                chunks[n_ch] = make_int4(t, b, e - b, plan->class_S[c]);
                slab_off[n_ch++] = cells;
                cells += (long long)plan->class_S[c] * plan->class_S[c];
            }
            kept += cnt;
            plan->total_taps += (long long)r.taps;
            if (is_fast(c)) plan->n_fast += cnt;
        }
    }
    plan->class_chunk_begin[kAllClasses] = (int)n_ch;
    for (int c = 0; c < kAllClasses; ++c)
        plan->n_class_launches += plan->class_chunk_begin[c + 1] > plan->class_chunk_begin[c];
    plan->n_kept = kept;
    plan->n_chunks = (int)n_ch;
    // register-gridder items: each exact-footprint work item in parts of 64
    // footprint positions, one segment per class present
    {
        FastItems &fi = plan->fast_items;
        long long total = 0;
        for (int c = kNumClasses; c < kAllClasses; ++c) {
            const int b = plan->class_chunk_begin[c], e = plan->class_chunk_begin[c + 1];
            if (e <= b) continue;
            const int W = 2 * plan->class_h[c], parts = (W * W + kWave - 1) / kWave;
            fi.item_begin[fi.n_seg] = (int)total;
            fi.chunk_begin[fi.n_seg] = b;
            fi.parts[fi.n_seg] = parts;
            ++fi.n_seg;
            total += (long long)(e - b) * parts;
        }
        if (total >= (1LL << 31)) return fail(sdp::set_error(SDP_E_UNSUPPORTED, "%lld register-gridder items", total));
        fi.item_begin[fi.n_seg] = (int)total;
        fi.n_items = (int)total;
        plan->n_items = (int)total;
    }

    // key slabs: single-piece keys use their chunk's slab; multi-piece keys
    // get a reduced slab placed after the chunk slabs.  A key of more than
    // kGroupPieces pieces is reduced in two passes: its pieces in groups of
    // kGroupPieces into partial slabs (after its key slab), then the partials
    // in group order into the key slab.
    const int nk = (int)key_S.size();
    std::vector<long long> key_off(nk);
    std::vector<int4> multi, multi2;
    std::vector<long long> multi_out, multi2_out, part_off;
    long long extra = 0;
    // SDP_GROUP_PIECES (8 to 1024) overrides kGroupPieces for tuning
    const char *gp_env = getenv("SDP_GROUP_PIECES");
    const long long gp = gp_env && atoi(gp_env) >= 8 && atoi(gp_env) <= 1024 ? atoi(gp_env) : kGroupPieces;
    for (int k = 0; k < nk; ++k) {
        const long long S2 = (long long)key_S[k] * key_S[k];
        const long long pieces = key_pieces[k], c0 = key_first_chunk[k];
        if (pieces == 1) {
            key_off[k] = slab_off[c0];
            continue;
        }
        key_off[k] = cells + extra;
        const long long out = cells + extra;
        extra += S2;
        if (pieces <= gp) {
            multi.push_back(make_int4((int)c0, (int)pieces, (int)S2, 0));
            multi_out.push_back(out);
            continue;
        }
        const long long groups = (pieces + gp - 1) / gp;
        multi2.push_back(make_int4((int)part_off.size(), (int)groups, (int)S2, 0));
        multi2_out.push_back(out);
        for (long long g = 0; g < groups; ++g) {
            const long long b = g * gp;
            multi.push_back(make_int4((int)(c0 + b), (int)std::min<long long>(gp, pieces - b), (int)S2, 0));
            multi_out.push_back(cells + extra);
            part_off.push_back(cells + extra);
            extra += S2;
        }
    }
    // each tile's keys in class order (runs are class-major), tiles (grid-major) in order
    int *tile_key_begin = nullptr;
    PLAN_TRY(cache.hget(HOST_TILE_KEYS, sizeof(int) * ((size_t)ntiles_all + 1), (void **)&tile_key_begin));
    std::fill(tile_key_begin, tile_key_begin + ntiles_all + 1, 0);
    std::vector<int> tile_key_list(nk);
    for (int k = 0; k < nk; ++k) ++tile_key_begin[key_tile[k] + 1];
    for (int t = 0; t < ntiles_all; ++t) {
        plan->n_tiles_used += tile_key_begin[t + 1] > 0;
        tile_key_begin[t + 1] += tile_key_begin[t];
    }
    {
        std::vector<int> fill(tile_key_begin, tile_key_begin + ntiles_all);
        for (int k = 0; k < nk; ++k) tile_key_list[fill[key_tile[k]]++] = k;
    }
    // key tables in tile order so that a tile's keys are contiguous
    std::vector<int> kS_t(nk);
    std::vector<long long> kOff_t(nk);
    for (int i = 0; i < nk; ++i) {
        kS_t[i] = key_S[tile_key_list[i]];
        kOff_t[i] = key_off[tile_key_list[i]];
    }
    plan->n_keys_used = nk;
    plan->n_multi = (int)multi.size();
    plan->n_multi2 = (int)multi2.size();

    PLAN_TRY(hipMalloc(&plan->tile_key_begin, sizeof(int) * ((size_t)ntiles_all + 1)));
    PLAN_TRY(hipMemcpyAsync(plan->tile_key_begin, tile_key_begin, sizeof(int) * ((size_t)ntiles_all + 1),
                            hipMemcpyHostToDevice, st));
    {
        // the merge regions some tile's window [origin, origin + smax) reaches
        // (the merge kernel's own covering test); the others are zero
        const int G = p.grid_size, R = plan->region, nr = (G + R - 1) / R, sm = std::max(plan->smax, 1);
        std::vector<unsigned char> hit((size_t)plan->n_grids * nr * nr, 0);
        for (int z = 0; z < plan->n_grids; ++z)
            for (int t = 0; t < plan->ntiles; ++t) {
                const int ta = z * plan->ntiles + t;
                if (tile_key_begin[ta + 1] == tile_key_begin[ta]) continue;
                const int x0 = (t % plan->ntx) * plan->tile, y0 = (t / plan->ntx) * plan->tile;
                const int x1 = std::min(x0 + sm, G) - 1, y1 = std::min(y0 + sm, G) - 1;
                for (int ry = y0 / R; ry <= y1 / R; ++ry)
                    for (int rx = x0 / R; rx <= x1 / R; ++rx) hit[((size_t)z * nr + ry) * nr + rx] = 1;
            }
        std::vector<int2> list;
        for (int z = 0; z < plan->n_grids; ++z)
            for (int ry = 0; ry < nr; ++ry)
                for (int rx = 0; rx < nr; ++rx)
                    if (hit[((size_t)z * nr + ry) * nr + rx]) list.push_back(make_int2(rx | ry << 16, z));
        plan->n_regions = (int)list.size();
        if (!list.empty()) {
            PLAN_TRY(hipMalloc(&plan->region_list, sizeof(int2) * list.size()));
            PLAN_TRY(hipMemcpyAsync(plan->region_list, list.data(), sizeof(int2) * list.size(),
                                    hipMemcpyHostToDevice, st));
        }
    }
    if (plan->n_chunks > 0) {
        PLAN_TRY(hipMalloc(&plan->chunks, sizeof(int4) * plan->n_chunks));
        PLAN_TRY(hipMemcpyAsync(plan->chunks, chunks, sizeof(int4) * plan->n_chunks, hipMemcpyHostToDevice, st));
        PLAN_TRY(hipMalloc(&plan->slab_off, sizeof(long long) * plan->n_chunks));
        PLAN_TRY(hipMemcpyAsync(plan->slab_off, slab_off, sizeof(long long) * plan->n_chunks,
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
        if (plan->n_multi2 > 0) {
            const size_t bytes = sizeof(int4) * plan->n_multi2 + sizeof(long long) * plan->n_multi2;
            PLAN_TRY(hipMalloc(&plan->multi2, bytes));
            PLAN_TRY(hipMemcpyAsync(plan->multi2, multi2.data(), sizeof(int4) * plan->n_multi2,
                                    hipMemcpyHostToDevice, st));
            PLAN_TRY(hipMemcpyAsync((char *)plan->multi2 + sizeof(int4) * plan->n_multi2, multi2_out.data(),
                                    sizeof(long long) * plan->n_multi2, hipMemcpyHostToDevice, st));
            PLAN_TRY(hipMalloc(&plan->part_off, sizeof(long long) * part_off.size()));
            PLAN_TRY(hipMemcpyAsync(plan->part_off, part_off.data(), sizeof(long long) * part_off.size(),
                                    hipMemcpyHostToDevice, st));
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
    hipStream_t st = sdp::stream_of(s);
    ChunkIO io{};
    // the LDS kernels read the phase-major table, the register gridder the unfolded one
    io.kg = prepare_tables(plan, d_kernels, nullptr, plan->class_chunk_begin[kNumClasses] > 0, true, st).rel_g;
    SDP_LAUNCH_CHECK();
    io.vis = (const float2 *)d_vis;
    io.slabs = plan->slabs;
    return launch_chunks<MODE_GRID>(plan, io, order == SDP_ORDER_ORIGINAL ? plan->perm : nullptr, st);
}

int sdp_grid_merge(const sdp_grid_plan *plan, float *d_grid, sdp_stream s)
{
    SDP_REQUIRE(plan && d_grid, "sdp_grid_merge: null argument");
    hipStream_t st = sdp::stream_of(s);
    const int G = plan->p.grid_size;
    if (plan->n_multi > 0) {
        // Cells x pieces per thread: 2 x 16 measured fastest at 16.7M
        // visibilities (4 x 8 and 1 x 32 within 1.5% of the step);
        // SDP_REDUCE_SHAPE = 4 or 1 selects those for tuning.  Every shape sums
        // a cell's pieces in the same order, so the results are identical.
        static const int shape = [] {
            const char *e = getenv("SDP_REDUCE_SHAPE");
            return e ? atoi(e) : 2;
        }();
        if (shape == 1)
            reduce_pieces_kernel<1, 32><<<plan->n_multi, kMergeThreads, 0, st>>>(plan->slabs, plan->slab_off,
                                                                                plan->multi, plan->key_slabs);
        else if (shape == 4)
            reduce_pieces_kernel<4, 8><<<plan->n_multi, kMergeThreads, 0, st>>>(plan->slabs, plan->slab_off,
                                                                               plan->multi, plan->key_slabs);
        else
            reduce_pieces_kernel<2, 16><<<plan->n_multi, kMergeThreads, 0, st>>>(plan->slabs, plan->slab_off,
                                                                                plan->multi, plan->key_slabs);
        SDP_LAUNCH_CHECK();
        if (plan->n_multi2 > 0) {   // the partial slabs of keys split into groups
            reduce_pieces_kernel<2, 16><<<plan->n_multi2, kMergeThreads, 0, st>>>(plan->slabs, plan->part_off,
                                                                                 plan->multi2, plan->key_slabs);
            SDP_LAUNCH_CHECK();
        }
    }
    const int R = plan->region;
    RegionArgs ra{G, plan->tile, plan->ntx, plan->nty, plan->ntiles, std::max(plan->smax, 1), R};
    // every cell zero, then the regions some window reaches summed over it
    // (a clear beside the piece reduction, on the helper stream, measured no
    // faster: both are bound by HBM)
    const size_t cells = (size_t)G * G * plan->n_grids;
    if (((uintptr_t)d_grid & 15) == 0) {
        if (int e = sdp_memset_zero(d_grid, sizeof(float2) * cells, s)) return e;
    } else {
        zero_cells_kernel<<<(unsigned)std::min<size_t>(sdp_div_up(cells, 256), 8192), 256, 0, st>>>(
            (float2 *)d_grid, cells);
        SDP_LAUNCH_CHECK();
    }
    if (plan->n_regions == 0) return SDP_OK;
    // 8 slabs' loads in flight per batch: 16 and 32 measured no faster
    if (plan->region_rg == 32)
        merge_region_kernel<32, 8><<<plan->n_regions, kMergeThreads, 0, st>>>(
            plan->key_slabs, plan->key_S, plan->key_off, plan->tile_key_begin, plan->region_list, ra,
            (float2 *)d_grid);
    else
        merge_region_kernel<16, 8><<<plan->n_regions, kMergeThreads, 0, st>>>(
            plan->key_slabs, plan->key_S, plan->key_off, plan->tile_key_begin, plan->region_list, ra,
            (float2 *)d_grid);
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
        io.kd = prepare_tables(plan, nullptr, d_kernels, true, false, st).rel_d;
        SDP_LAUNCH_CHECK();
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
        // the LDS kernels' phase-major tables, the fused kernel's unfolded
        // tables and whether one of those serves both
        const Prepared pt = prepare_tables(plan, d_grid_kernels, d_degrid_kernels,
                                           plan->class_chunk_begin[kNumClasses] > 0, true, st);
        io.kg = pt.rel_g;
        io.kd = pt.rel_d;
        SDP_LAUNCH_CHECK();
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
                float ms = -1.f;
                // one untimed pass, then the major cycle's fused degrid +
                // subtract + grid and the merge timed together, the best of
                // three passes (one pass alone picked a 3% slower shape)
                rc = sdp_degrid_grid_accumulate(plan, d_degrid_kernels, d_grid_kernels, (const float *)grid,
                                                (const float *)vis, SDP_ORDER_PLAN, (float *)vis, s);
                if (rc == SDP_OK) rc = sdp_grid_merge(plan, (float *)grid, s);
                for (int rep = 0; rep < 3 && rc == SDP_OK; ++rep) {
                    float t = 0.f;
                    (void)hipEventRecord(e0, st);
                    rc = sdp_degrid_grid_accumulate(plan, d_degrid_kernels, d_grid_kernels, (const float *)grid,
                                                    (const float *)vis, SDP_ORDER_PLAN, (float *)vis, s);
                    if (rc == SDP_OK) rc = sdp_grid_merge(plan, (float *)grid, s);
                    (void)hipEventRecord(e1, st);
                    (void)hipEventSynchronize(e1);
                    (void)hipEventElapsedTime(&t, e0, e1);
                    if (ms < 0.f || t < ms) ms = t;
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
