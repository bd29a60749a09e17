// sdp_cnn.hip -- the learned minor-cycle proposal: an encoder-decoder CNN
// that maps (residual patch, PSF patch) to a model-image update, trained
// and run by the kernels in this file.  The reference has no learned
// component; the network is defined by its FP64 restatement used in the
// tests (encoder levels of two 3x3 convolutions + ReLU with 2x2 average
// pooling, decoder levels of upsampling, a 3x3 convolution, an additive skip
// from the encoder, ReLU and a second convolution, then a 3x3 head).
//
// Every 3x3 convolution, forward and backward, is one implicit GEMM:
//   forward       M = pixels,      N = C_out, K = 9 C_in
//   input grad    M = pixels,      N = C_in,  K = 9 C_out  (taps reflected)
//   weight grad   M = 9 C_in,      N = C_out, K = pixels   (split over K, the
//                 partial products summed in a fixed order afterwards)
// Tensors are NHWC in HBM, so a GEMM row's K elements are contiguous within
// each tap.  Two arithmetic modes: FP32 FMA on the vector ALUs (exact enough
// for gradient checks), and BF16 operands on the gfx950 matrix cores
// (v_mfma_f32_16x16x32_bf16) with FP32 accumulation.  Weights are kept in
// FP32 (master copy for Adam) and rounded to BF16 as tiles are staged; the
// BF16 forward convolutions run on a halo-tiled kernel instead
// (conv_fwd_tile_kernel), from a BF16 copy of the weights formed once per
// parameter change.
//
// Element-wise epilogues are fused: bias, the decoder's skip addition and
// ReLU after a forward GEMM; the ReLU mask of the producing layer after an
// input-gradient GEMM.
#include <algorithm>
#include <cmath>
#include <vector>

#include "sdp_internal.h"

namespace {

constexpr int kT = 256;        // threads per GEMM workgroup
constexpr int kBM = 64, kBN = 64, kBK = 32;
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

enum { G_FWD = 0, G_BWD_DATA = 1, G_BWD_W = 2 };

struct GemmArgs {
    int B, H, W, Cin, Cout;    // convolution geometry (input Cin, output Cout channels)
    int M, N, K;               // GEMM sizes
    int k_chunk;               // BWD_W: K range of one split (multiple of kBK)
    const float *a_src;        // FWD/BWD_W: layer input; BWD_DATA: output gradient
    const float *b_src;        // FWD/BWD_DATA: weights [3][3][Cin][Cout]; BWD_W: output gradient
    float *out;                // FWD: activation; BWD_DATA: input gradient; BWD_W: split partials
    const float *bias;         // FWD: bias (or null)
    const float *skip;         // FWD: added before ReLU (or null)
    const float *mask;         // BWD_DATA: gradient zeroed where mask <= 0 (or null)
    int relu;                  // FWD: apply ReLU
};

// Index tables for A's im2col reads, so that the staging loops do no
// divisions.  A[m][k] = src[b][y + s ky][x + s kx][c] with a pixel index p
// (A's row for the forward and input-gradient products, its column for the
// weight gradient) and a tap index q = t C + c (the other one); s = -1 for
// the input gradient.  Pixel entries hold (b H, y, x), tap entries
// (s ky, s kx, c); absent entries put y far outside the image.
struct PixIdx {
    int bh, y, x;
};
struct TapIdx {
    int dy, dx, c;
};

__device__ __forceinline__ PixIdx pix_index(const GemmArgs &g, int p, int np)
{
    if (p >= np) return PixIdx{0, -(1 << 20), 0};
    const int x = p % g.W, rest = p / g.W, y = rest % g.H, b = rest / g.H;
    return PixIdx{b * g.H, y, x};
}

__device__ __forceinline__ TapIdx tap_index(const GemmArgs &g, int mode, int q, int nq)
{
    if (q >= nq) return TapIdx{-(1 << 20), 0, 0};   // with an absent pixel too, still far outside
    const int C = mode == G_BWD_DATA ? g.Cout : g.Cin;
    const int t = q / C, c = q - t * C;
    const int ky = t / 3 - 1, kx = t % 3 - 1;
    return mode == G_BWD_DATA ? TapIdx{-ky, -kx, c} : TapIdx{ky, kx, c};
}

__device__ __forceinline__ float load_a_idx(const GemmArgs &g, int mode, const PixIdx &p, const TapIdx &t)
{
    const int yy = p.y + t.dy, xx = p.x + t.dx;
    if (yy < 0 || yy >= g.H || xx < 0 || xx >= g.W) return 0.f;
    const int C = mode == G_BWD_DATA ? g.Cout : g.Cin;
    return g.a_src[((size_t)(p.bh + yy) * g.W + xx) * C + t.c];
}

// B[k][n] of the GEMM
__device__ __forceinline__ float load_b(const GemmArgs &g, int mode, int k, int n)
{
    if (k >= g.K || n >= g.N) return 0.f;
    if (mode == G_FWD) return g.b_src[(size_t)k * g.Cout + n];                 // w[t][ci][co], k = t Cin + ci
    if (mode == G_BWD_W) return g.b_src[(size_t)k * g.Cout + n];               // gout[pix][co]
    const int t = k / g.Cout, co = k - t * g.Cout;                             // w[t][n = ci][co]
    return g.b_src[((size_t)t * g.Cin + n) * g.Cout + co];
}

__device__ __forceinline__ void epilogue(const GemmArgs &g, int mode, int m, int n, float v, int split)
{
    if (m >= g.M || n >= g.N) return;
    if (mode == G_FWD) {
        if (g.bias) v += g.bias[n];
        if (g.skip) v += g.skip[(size_t)m * g.N + n];
        if (g.relu) v = fmaxf(v, 0.f);
        g.out[(size_t)m * g.N + n] = v;
    } else if (mode == G_BWD_DATA) {
        if (g.mask && !(g.mask[(size_t)m * g.N + n] > 0.f)) v = 0.f;
        g.out[(size_t)m * g.N + n] = v;
    } else {
        g.out[((size_t)split * g.M + m) * g.N + n] = v;
    }
}

// K range of this workgroup
__device__ __forceinline__ void k_range(const GemmArgs &g, int mode, int &k0, int &k1)
{
    if (mode == G_BWD_W) {
        k0 = blockIdx.z * g.k_chunk;
        k1 = min(g.K, k0 + g.k_chunk);
    } else {
        k0 = 0;
        k1 = g.K;
    }
}

// FP32 vector-ALU GEMM: a 64 x 64 tile per workgroup, 4 x 4 outputs per thread.
template <int MODE>
__global__ void __launch_bounds__(kT) gemm_f32_kernel(GemmArgs g)
{
    __shared__ float As[kBK][kBM + 4], Bs[kBK][kBN + 4];
    const int m0 = blockIdx.x * kBM, n0 = blockIdx.y * kBN;
    const int tx = threadIdx.x % 16, ty = threadIdx.x / 16;
    int k0, k1;
    k_range(g, MODE, k0, k1);
    float acc[4][4] = {};
    constexpr bool kRowPix = MODE != G_BWD_W;
    __shared__ PixIdx s_pix[kRowPix ? kBM : kBK];
    __shared__ TapIdx s_tap[kRowPix ? kBK : kBM];
    if (kRowPix) {
        if (threadIdx.x < kBM) s_pix[threadIdx.x] = pix_index(g, m0 + threadIdx.x, g.M);
    } else {
        if (threadIdx.x < kBM) s_tap[threadIdx.x] = tap_index(g, MODE, m0 + threadIdx.x, g.M);
    }
    for (int kb = k0; kb < k1; kb += kBK) {
        if (threadIdx.x < kBK) {
            if (kRowPix) s_tap[threadIdx.x] = tap_index(g, MODE, kb + threadIdx.x, k1);
            else s_pix[threadIdx.x] = pix_index(g, kb + threadIdx.x, k1);
        }
        __syncthreads();
        for (int i = threadIdx.x; i < kBM * kBK; i += kT) {
            const int mm = i / kBK, kk = i % kBK;
            As[kk][mm] = kRowPix ? load_a_idx(g, MODE, s_pix[mm], s_tap[kk]) : load_a_idx(g, MODE, s_pix[kk], s_tap[mm]);
        }
        for (int i = threadIdx.x; i < kBN * kBK; i += kT) {
            const int kk = i / kBN, nn = i % kBN;
            Bs[kk][nn] = (kb + kk < k1) ? load_b(g, MODE, kb + kk, n0 + nn) : 0.f;
        }
        __syncthreads();
#pragma unroll 4
        for (int kk = 0; kk < kBK; ++kk) {
            float a[4], b[4];
#pragma unroll
            for (int r = 0; r < 4; ++r) a[r] = As[kk][ty * 4 + r];
#pragma unroll
            for (int c = 0; c < 4; ++c) b[c] = Bs[kk][tx * 4 + c];
#pragma unroll
            for (int r = 0; r < 4; ++r)
#pragma unroll
                for (int c = 0; c < 4; ++c) acc[r][c] = fmaf(a[r], b[c], acc[r][c]);
        }
        __syncthreads();
    }
#pragma unroll
    for (int r = 0; r < 4; ++r)
#pragma unroll
        for (int c = 0; c < 4; ++c) epilogue(g, MODE, m0 + ty * 4 + r, n0 + tx * 4 + c, acc[r][c], blockIdx.z);
}

// BF16 matrix-core GEMM: a 64 x 64 tile per workgroup, four wavefronts of
// 32 x 32, each a 2 x 2 block of 16 x 16 x 32 MFMAs per K step.  Operand
// fragments: lane l holds row (A) or column (B) l % 16, K elements
// 8 (l / 16) .. + 7; the accumulator holds column l % 16, rows 4 (l / 16) .. + 3.
template <int MODE>
__global__ void __launch_bounds__(kT) gemm_bf16_kernel(GemmArgs g)
{
    __shared__ __attribute__((aligned(16))) __bf16 As[kBM][kBK + 8];
    __shared__ __attribute__((aligned(16))) __bf16 Bs[kBN][kBK + 8];
    const int m0 = blockIdx.x * kBM, n0 = blockIdx.y * kBN;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int wr = (wave >> 1) * 32, wc = (wave & 1) * 32;
    int k0, k1;
    k_range(g, MODE, k0, k1);
    f32x4 acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
    // A's row indices are fixed for the tile; its column indices per K step
    constexpr bool kRowPix = MODE != G_BWD_W;
    __shared__ PixIdx s_pix[kRowPix ? kBM : kBK];
    __shared__ TapIdx s_tap[kRowPix ? kBK : kBM];
    if (kRowPix) {
        if (threadIdx.x < kBM) s_pix[threadIdx.x] = pix_index(g, m0 + threadIdx.x, g.M);
    } else {
        if (threadIdx.x < kBM) s_tap[threadIdx.x] = tap_index(g, MODE, m0 + threadIdx.x, g.M);
    }
    for (int kb = k0; kb < k1; kb += kBK) {
        if (threadIdx.x < kBK) {
            if (kRowPix) s_tap[threadIdx.x] = tap_index(g, MODE, kb + threadIdx.x, k1);
            else s_pix[threadIdx.x] = pix_index(g, kb + threadIdx.x, k1);
        }
        __syncthreads();
        for (int i = threadIdx.x; i < kBM * kBK; i += kT) {
            const int mm = i / kBK, kk = i % kBK;
            As[mm][kk] = (__bf16)(kRowPix ? load_a_idx(g, MODE, s_pix[mm], s_tap[kk])
                                          : load_a_idx(g, MODE, s_pix[kk], s_tap[mm]));
        }
        for (int i = threadIdx.x; i < kBN * kBK; i += kT) {
            const int nn = i % kBN, kk = i / kBN;
            Bs[nn][kk] = (__bf16)((kb + kk < k1) ? load_b(g, MODE, kb + kk, n0 + nn) : 0.f);
        }
        __syncthreads();
        const int kq = (lane >> 4) * 8;
#pragma unroll
        for (int i = 0; i < 2; ++i) {
            const bf16x8 a = *reinterpret_cast<const bf16x8 *>(&As[wr + i * 16 + (lane & 15)][kq]);
#pragma unroll
            for (int j = 0; j < 2; ++j) {
                const bf16x8 b = *reinterpret_cast<const bf16x8 *>(&Bs[wc + j * 16 + (lane & 15)][kq]);
                acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc[i][j], 0, 0, 0);
            }
        }
        __syncthreads();
    }
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int r = 0; r < 4; ++r)
                epilogue(g, MODE, m0 + wr + i * 16 + (lane >> 4) * 4 + r, n0 + wc + j * 16 + (lane & 15),
                         acc[i][j][r], blockIdx.z);
}

// Forward convolution with at most BN = 16 or 32 output channels (the
// upper levels and the head): a 64 x 64 tile would leave half or more of
// each product empty, so a workgroup takes 256 pixels x BN channels, each
// wavefront 64 pixels.  The input has a power-of-two channel count >= 8: a
// K step's 8-channel groups are 32 contiguous bytes of one tap, read as two
// 16-byte loads and stored to LDS as one 16-byte BF16 vector.
constexpr int kNarrowM = 256;
static_assert(kNarrowM == kT, "one pixel index per thread");

template <int BN>
__global__ void __launch_bounds__(kT) gemm_fwd_narrow_kernel(GemmArgs g, int cin_shift)
{
    constexpr int NB = BN / 16;   // 16-column blocks
    __shared__ __attribute__((aligned(16))) __bf16 As[kNarrowM][kBK + 8];
    __shared__ __attribute__((aligned(16))) __bf16 Bs[BN][kBK + 8];
    __shared__ PixIdx s_pix[kNarrowM];
    const int m0 = blockIdx.x * kNarrowM;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    s_pix[threadIdx.x] = pix_index(g, m0 + threadIdx.x, g.M);
    f32x4 acc[4][NB];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < NB; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
    for (int kb = 0; kb < g.K; kb += kBK) {
        __syncthreads();   // the pixel table is ready; the previous step's products have read LDS
#pragma unroll
        for (int j = 0; j < kBK / 8; ++j) {
            const int q = j * kT + threadIdx.x;           // (pixel, 8-channel group) of this K step
            const int m = q >> 2, kg = (q & 3) * 8, k = kb + kg;
            bf16x8 v;
#pragma unroll
            for (int e = 0; e < 8; ++e) v[e] = (__bf16)0.f;
            if (k < g.K) {
                const int t = k >> cin_shift, c0 = k & (g.Cin - 1);   // tap (ky, kx) = (t / 3, t % 3)
                const int dy = t < 3 ? -1 : (t < 6 ? 0 : 1), dx = t - 3 * (dy + 1) - 1;
                const PixIdx p = s_pix[m];
                const int yy = p.y + dy, xx = p.x + dx;
                if (yy >= 0 && yy < g.H && xx >= 0 && xx < g.W) {
                    const float4 *src =
                        reinterpret_cast<const float4 *>(g.a_src + ((size_t)(p.bh + yy) * g.W + xx) * g.Cin + c0);
                    const float4 u0 = src[0], u1 = src[1];
                    v[0] = (__bf16)u0.x; v[1] = (__bf16)u0.y; v[2] = (__bf16)u0.z; v[3] = (__bf16)u0.w;
                    v[4] = (__bf16)u1.x; v[5] = (__bf16)u1.y; v[6] = (__bf16)u1.z; v[7] = (__bf16)u1.w;
                }
            }
            *reinterpret_cast<bf16x8 *>(&As[m][kg]) = v;
        }
        for (int i = threadIdx.x; i < BN * kBK; i += kT) {
            const int nn = i % BN, kk = i / BN;
            Bs[nn][kk] = (__bf16)((kb + kk < g.K && nn < g.N) ? g.b_src[(size_t)(kb + kk) * g.Cout + nn] : 0.f);
        }
        __syncthreads();
        const int kq = (lane >> 4) * 8;
        bf16x8 b[NB];
#pragma unroll
        for (int j = 0; j < NB; ++j) b[j] = *reinterpret_cast<const bf16x8 *>(&Bs[j * 16 + (lane & 15)][kq]);
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const bf16x8 a = *reinterpret_cast<const bf16x8 *>(&As[wave * 64 + i * 16 + (lane & 15)][kq]);
#pragma unroll
            for (int j = 0; j < NB; ++j)
                acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b[j], acc[i][j], 0, 0, 0);
        }
    }
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < NB; ++j)
#pragma unroll
            for (int r = 0; r < 4; ++r)
                epilogue(g, G_FWD, m0 + wave * 64 + i * 16 + (lane >> 4) * 4 + r, j * 16 + (lane & 15), acc[i][j][r],
                         0);
}

// Forward 3x3 convolution on the matrix cores with the input tile staged
// once: a workgroup computes a TH-row x 16-column tile of output pixels for
// BN output channels.  Per 32-channel chunk of the input it stages the
// tile's (TH + 2) x 18 halo (BF16, channels contiguous) and the chunk's
// weights for all nine taps in LDS, then every tap's product reads its A
// fragments from the halo at the tap's offset, so an input value is read
// from HBM once per chunk instead of once per tap.  Wave w takes image rows
// 2w and 2w + 1 of the tile: two 16-pixel row blocks x BN / 16 column
// blocks of v_mfma_f32_16x16x32_bf16 per tap.  Weights come pre-rounded to
// BF16 in [Cout_pad][9][Cin_pad] order (refresh_bf16_weights; rows from Cout
// to Cout_pad and channels from Cin to Cin_pad are zero, padding rows'
// outputs are not stored); the epilogue adds the bias and the skip input and
// applies the ReLU, as the general kernels do.  Needs Cin % 32 == 0 or
// Cin < 32 (the halo's channels from Cin on are zero), H % TH == 0,
// W % 16 == 0 and BN | Cout_pad.
//
// IN: the input as stored.  IN_DIRECT: H x W.  IN_UPSAMPLE: H/2 x W/2, read
// through the nearest 2x upsampling (upsample2_vec_kernel's map).
// IN_POOL: 2H x 2W, read through the 2x2 average (avgpool2_vec_kernel's
// arithmetic, so the values are the same as the pooled tensor's).
enum { IN_DIRECT = 0, IN_UPSAMPLE = 1, IN_POOL = 2 };
constexpr int kHaloW = 18;
constexpr int kLdA = 40;   // LDS row of one pixel's (or weight column's) 32 BF16 K values, padded

// two floats rounded to BF16 (nearest even), the first in the low half
__device__ __forceinline__ uint32_t pack_bf16x2(float a, float b)
{
    const __bf16 lo = (__bf16)a, hi = (__bf16)b;
    return (uint32_t)__builtin_bit_cast(uint16_t, lo) | (uint32_t)__builtin_bit_cast(uint16_t, hi) << 16;
}

struct ConvArgs {
    int B, H, W, Cin, Cout;
    const float *x;          // NHWC input
    const __bf16 *wt;        // the layer's weights, [Cout_pad][9][Cin_pad] (Cin_pad: Cin rounded up to 32)
    const float *bias, *skip;
    float *out;
    int relu;
    int cinw;                // Cin_pad
};

// four channels ch .. ch + 3 of a pixel's Cin (zero beyond Cin)
__device__ __forceinline__ float4 load_ch4(const float *px, int ch, int cin)
{
    if ((cin & 3) == 0) return ch < cin ? *reinterpret_cast<const float4 *>(px + ch) : make_float4(0.f, 0.f, 0.f, 0.f);
    return make_float4(ch < cin ? px[ch] : 0.f, ch + 1 < cin ? px[ch + 1] : 0.f, ch + 2 < cin ? px[ch + 2] : 0.f,
                       ch + 3 < cin ? px[ch + 3] : 0.f);
}

template <int BN, int TH, int IN>
__global__ void __launch_bounds__(32 * TH) conv_fwd_tile_kernel(ConvArgs c)
{
    constexpr int NJ = BN / 16, NT = 32 * TH, HALO = (TH + 2) * kHaloW;
    // the halo and weight tiles, and after the last chunk each wave's
    // accumulators in pixel-major order (epilogue)
    constexpr size_t kAsBytes = sizeof(__bf16) * HALO * kLdA, kBsBytes = sizeof(__bf16) * 9 * BN * kLdA;
    constexpr int BNP = BN + 4;   // a pixel's row of accumulators, padded (4 pixels 16 banks apart)
    static_assert((size_t)(NT / 64) * 32 * BNP * sizeof(float) <= kAsBytes + kBsBytes, "epilogue tile exceeds LDS");
    __shared__ __attribute__((aligned(16))) unsigned char smem[kAsBytes + kBsBytes];
    auto &As = *reinterpret_cast<__bf16(*)[HALO][kLdA]>(smem);
    auto &Bs = *reinterpret_cast<__bf16(*)[9][BN][kLdA]>(smem + kAsBytes);
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int tiles_x = c.W >> 4, tiles_y = c.H / TH;
    int t = blockIdx.x;
    const int tx = t % tiles_x;
    t /= tiles_x;
    const int ty = t % tiles_y, b = t / tiles_y;
    const int x0 = tx * 16, y0 = ty * TH, n0 = blockIdx.y * BN;
    f32x4 acc[2][NJ];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < NJ; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
    // the stored input's width and its image's first element
    const int sw = IN == IN_UPSAMPLE ? c.W / 2 : IN == IN_POOL ? 2 * c.W : c.W;
    const size_t img = (size_t)c.H * c.W * c.Cin;
    const float *xb = c.x + (IN == IN_UPSAMPLE ? img / 4 : IN == IN_POOL ? 4 * img : img) * b;
    // Halo item i: pixel i / 8 of the (TH + 2) x 18 tile, channels 4 (i % 8)
    // .. + 3 of the chunk; weight item i: output channel (i / 4) % BN, tap
    // i / (4 BN), K values 8 (i % 4) .. + 7.  For 32 or more output
    // channels a thread issues all its weight loads of a chunk, and all its
    // halo loads of direct and upsampled inputs, before writing any (one
    // memory latency per chunk instead of one per item); pooled inputs (four
    // loads per item) go two items at a time and the 16-channel head item by
    // item, which keeps their registers, and so the resident workgroups, as
    // they are.
    constexpr int HN = HALO * 8, HI = (HN + NT - 1) / NT, WN = 9 * BN * 4, WI = (WN + NT - 1) / NT;
    constexpr bool BATCH = IN != IN_POOL && BN >= 32;
    constexpr int HB = BATCH ? HI : (IN == IN_POOL && BN >= 32) ? 2 : 1, WB = BN >= 32 ? WI : 1;
    auto halo_item = [&](int i, int c0) __attribute__((always_inline)) {
        const int h = i >> 3, ch = c0 + 4 * (i & 7);
        const int yy = y0 - 1 + h / kHaloW, xx = x0 - 1 + h % kHaloW;
        float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
        if (i < HN && yy >= 0 && yy < c.H && xx >= 0 && xx < c.W && ch < c.Cin) {
            if (IN == IN_DIRECT) {
                v = load_ch4(xb + ((size_t)yy * sw + xx) * c.Cin, ch, c.Cin);
            } else if (IN == IN_UPSAMPLE) {
                v = load_ch4(xb + ((size_t)(yy >> 1) * sw + (xx >> 1)) * c.Cin, ch, c.Cin);
            } else {
                const float *p = xb + ((size_t)(2 * yy) * sw + 2 * xx) * c.Cin;
                const float4 a = load_ch4(p, ch, c.Cin);
                const float4 e = load_ch4(p + c.Cin, ch, c.Cin);
                const float4 q = load_ch4(p + (size_t)sw * c.Cin, ch, c.Cin);
                const float4 w = load_ch4(p + (size_t)sw * c.Cin + c.Cin, ch, c.Cin);
                v = make_float4(0.25f * ((a.x + e.x) + (q.x + w.x)), 0.25f * ((a.y + e.y) + (q.y + w.y)),
                                0.25f * ((a.z + e.z) + (q.z + w.z)), 0.25f * ((a.w + e.w) + (q.w + w.w)));
            }
        }
        return v;
    };
    for (int c0 = 0; c0 < c.cinw; c0 += kBK) {
#pragma unroll 1
        for (int u0 = 0; u0 < HI; u0 += HB) {
            float4 hv[HB];
#pragma unroll
            for (int u = 0; u < HB; ++u) hv[u] = halo_item(tid + (u0 + u) * NT, c0);
#pragma unroll
            for (int u = 0; u < HB; ++u) {
                const int i = tid + (u0 + u) * NT;
                if (i < HN)
                    *reinterpret_cast<uint2 *>(&As[i >> 3][4 * (i & 7)]) =
                        make_uint2(pack_bf16x2(hv[u].x, hv[u].y), pack_bf16x2(hv[u].z, hv[u].w));
            }
        }
#pragma unroll 1
        for (int u0 = 0; u0 < WI; u0 += WB) {
            u32x4 wv[WB];
#pragma unroll
            for (int u = 0; u < WB; ++u) {
                const int i = tid + (u0 + u) * NT;
                if (i < WN) {
                    const int p4 = i & 3, nk = i >> 2, n = nk % BN, k = nk / BN;
                    wv[u] = *reinterpret_cast<const u32x4 *>(c.wt + ((size_t)(n0 + n) * 9 + k) * c.cinw + c0 + 8 * p4);
                }
            }
#pragma unroll
            for (int u = 0; u < WB; ++u) {
                const int i = tid + (u0 + u) * NT;
                if (i < WN) {
                    const int p4 = i & 3, nk = i >> 2, n = nk % BN, k = nk / BN;
                    *reinterpret_cast<u32x4 *>(&Bs[k][n][8 * p4]) = wv[u];
                }
            }
        }
        __syncthreads();
        const int kq = (lane >> 4) * 8;
#pragma unroll
        for (int k = 0; k < 9; ++k) {
            const int ky = k / 3, kx = k % 3;   // halo offsets of the tap (the halo starts one pixel out)
            bf16x8 bw[NJ];
#pragma unroll
            for (int j = 0; j < NJ; ++j) bw[j] = *reinterpret_cast<const bf16x8 *>(&Bs[k][16 * j + (lane & 15)][kq]);
#pragma unroll
            for (int i = 0; i < 2; ++i) {
                const int h = (2 * wave + i + ky) * kHaloW + (lane & 15) + kx;
                const bf16x8 a = *reinterpret_cast<const bf16x8 *>(&As[h][kq]);
#pragma unroll
                for (int j = 0; j < NJ; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, bw[j], acc[i][j], 0, 0, 0);
            }
        }
        __syncthreads();
    }
    // accumulator (i, j): output channel n0 + 16 j + lane % 16, pixels 4 (lane / 16) .. + 3 of row 2 wave + i.
    // Where a skip input is added and the layer's BN channels are all real
    // and 16-byte aligned, each wave turns its two rows' accumulators into
    // pixel-major order through LDS, then reads the skip input and writes
    // four channels per lane (16-byte accesses instead of 4-byte ones; for
    // the stores alone the round trip through LDS measured slower).
    if (BN >= 32 && c.skip && c.Cout % 4 == 0 && n0 + BN <= c.Cout &&
        ((reinterpret_cast<uintptr_t>(c.out) | reinterpret_cast<uintptr_t>(c.skip)) & 15) == 0) {
        float *Cs = reinterpret_cast<float *>(smem) + wave * 32 * BNP;
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int j = 0; j < NJ; ++j)
#pragma unroll
                for (int r = 0; r < 4; ++r) Cs[(16 * i + 4 * (lane >> 4) + r) * BNP + 16 * j + (lane & 15)] = acc[i][j][r];
        __syncthreads();
        constexpr int QG = BN / 4;
#pragma unroll
        for (int it = lane; it < 32 * QG; it += 64) {
            const int px = it / QG, q = it % QG, n = n0 + 4 * q;
            const float4 a = *reinterpret_cast<const float4 *>(&Cs[px * BNP + 4 * q]);
            const size_t m = ((size_t)(b * c.H + y0 + 2 * wave + (px >> 4)) * c.W + x0 + (px & 15)) * c.Cout + n;
            float v[4] = {a.x, a.y, a.z, a.w};
#pragma unroll
            for (int e = 0; e < 4; ++e) v[e] += c.bias ? c.bias[n + e] : 0.f;
            const float4 sk = *reinterpret_cast<const float4 *>(c.skip + m);
            v[0] += sk.x; v[1] += sk.y; v[2] += sk.z; v[3] += sk.w;
            if (c.relu)
#pragma unroll
                for (int e = 0; e < 4; ++e) v[e] = fmaxf(v[e], 0.f);
            *reinterpret_cast<float4 *>(c.out + m) = make_float4(v[0], v[1], v[2], v[3]);
        }
        return;
    }
#pragma unroll
    for (int i = 0; i < 2; ++i) {
        const int y = y0 + 2 * wave + i;
#pragma unroll
        for (int j = 0; j < NJ; ++j) {
            const int n = n0 + 16 * j + (lane & 15);
            if (n >= c.Cout) continue;   // a padding row of the weights
            const float bn = c.bias ? c.bias[n] : 0.f;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const size_t m = ((size_t)(b * c.H + y) * c.W + x0 + 4 * (lane >> 4) + r);
                float v = acc[i][j][r] + bn;
                if (c.skip) v += c.skip[m * c.Cout + n];
                if (c.relu) v = fmaxf(v, 0.f);
                c.out[m * c.Cout + n] = v;
            }
        }
    }
}

// The layers' weights rounded to BF16 in [Cout_pad][9][Cin_pad] order, from
// the FP32 parameters ([3][3][Cin][Cout] per layer), for conv_fwd_tile_kernel
// (the padding rows and channels stay as created: zero).
constexpr int kMaxLayers = 32;   // 4 levels - 1 layers at most 6 levels
struct Bf16Layers {
    int n;
    int cin[kMaxLayers], cout[kMaxLayers], cinw[kMaxLayers];
    long long off[kMaxLayers + 1];   // first weight of each layer in the parameter vector; off[n]: the end
    long long dst[kMaxLayers];       // its first element in the BF16 array
};

__global__ void bf16_weights_kernel(const float *params, Bf16Layers L, __bf16 *wt)
{
    const long long total = L.off[L.n];
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < total; i += (long long)gridDim.x * blockDim.x) {
        int l = 0;
        while (i >= L.off[l + 1]) ++l;
        if (i >= L.off[l] + 9LL * L.cin[l] * L.cout[l]) continue;   // a bias
        const long long r = i - L.off[l];   // (k * Cin + ci) * Cout + co
        const int co = (int)(r % L.cout[l]);
        const long long kc = r / L.cout[l];
        const int ci = (int)(kc % L.cin[l]), k = (int)(kc / L.cin[l]);
        wt[L.dst[l] + ((long long)co * 9 + k) * L.cinw[l] + ci] = (__bf16)params[i];
    }
}

// weight gradient: sum the K-split partials in split order
__global__ void sum_splits_kernel(const float *part, int splits, size_t n, float *out)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        float s = 0.f;
        for (int q = 0; q < splits; ++q) s += part[(size_t)q * n + i];
        out[i] = s;
    }
}

// bias gradient: column sums of the M x N output gradient, in two passes
// with a fixed order.  Pass 1: workgroup b sums rows [b rows, (b + 1) rows)
// into part[b][c]; when N divides the workgroup, thread t keeps column
// t % N and steps kT / N rows at a time, so a wavefront reads whole
// consecutive rows.  Pass 2: the partials in workgroup order.
constexpr int kBiasBlocks = 256;
__global__ void bias_part_kernel(const float *gout, int M, int N, int rows, float *part)
{
    __shared__ float red[kT];
    const int m0 = blockIdx.x * rows, m1 = min(M, m0 + rows);
    if (N <= kT && kT % N == 0) {
        const int c = threadIdx.x % N, step = kT / N;
        float s = 0.f;
        for (int m = m0 + threadIdx.x / N; m < m1; m += step) s += gout[(size_t)m * N + c];
        red[threadIdx.x] = s;
        __syncthreads();
        if (threadIdx.x < N) {
            float t = 0.f;
            for (int j = 0; j < step; ++j) t += red[j * N + threadIdx.x];
            part[(size_t)blockIdx.x * N + threadIdx.x] = t;
        }
    } else {
        for (int c = threadIdx.x; c < N; c += kT) {
            float s = 0.f;
            for (int m = m0; m < m1; ++m) s += gout[(size_t)m * N + c];
            part[(size_t)blockIdx.x * N + c] = s;
        }
    }
}

__global__ void bias_sum_kernel(const float *part, int blocks, int N, float *gb)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= N) return;
    float s = 0.f;
    for (int b = 0; b < blocks; ++b) s += part[(size_t)b * N + c];
    gb[c] = s;
}

// 2x2 average pool (NHWC), input H x W -> H/2 x W/2
__global__ void avgpool2_kernel(const float *in, int B, int H, int W, int C, float *out)
{
    const size_t n = (size_t)B * (H / 2) * (W / 2) * C;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const int c = i % C;
        size_t r = i / C;
        const int x = r % (W / 2);
        r /= (W / 2);
        const int y = r % (H / 2);
        const int b = r / (H / 2);
        const float *p = in + ((size_t)(b * H + 2 * y) * W + 2 * x) * C + c;
        out[i] = 0.25f * ((p[0] + p[C]) + (p[(size_t)W * C] + p[(size_t)W * C + C]));
    }
}

// ge = (skip_grad + avgpool2 backward of g_low) masked by act > 0, at H x W
__global__ void pool_back_add_mask_kernel(const float *g_low, const float *skip_grad, const float *act, int B, int H,
                                          int W, int C, float *out)
{
    const size_t n = (size_t)B * H * W * C;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const int c = i % C;
        size_t r = i / C;
        const int x = r % W;
        r /= W;
        const int y = r % H;
        const int b = r / H;
        float v = 0.25f * g_low[((size_t)(b * (H / 2) + y / 2) * (W / 2) + x / 2) * C + c];
        if (skip_grad) v += skip_grad[i];
        out[i] = act[i] > 0.f ? v : 0.f;
    }
}

// nearest 2x upsample (NHWC), input H x W -> 2H x 2W
// The same two maps for C % 4 == 0 and fewer than 2^31 16-byte groups: one
// group of 4 channels per thread, 32-bit index arithmetic.
__global__ void avgpool2_vec_kernel(const float4 *in, unsigned B, unsigned H, unsigned W, unsigned C4, float4 *out)
{
    const unsigned W2 = W / 2, H2 = H / 2, n = B * H2 * W2 * C4;
    for (unsigned i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const unsigned c = i % C4, px = i / C4, x = px % W2, r = px / W2, y = r % H2, b = r / H2;
        const float4 *p = in + ((b * H + 2 * y) * W + 2 * x) * C4 + c;
        const float4 a = p[0], e = p[C4], u = p[W * C4], v = p[W * C4 + C4];
        out[i] = make_float4(0.25f * ((a.x + e.x) + (u.x + v.x)), 0.25f * ((a.y + e.y) + (u.y + v.y)),
                             0.25f * ((a.z + e.z) + (u.z + v.z)), 0.25f * ((a.w + e.w) + (u.w + v.w)));
    }
}
__global__ void upsample2_vec_kernel(const float4 *in, unsigned B, unsigned H, unsigned W, unsigned C4, float4 *out)
{
    const unsigned W2 = 2 * W, H2 = 2 * H, n = B * H2 * W2 * C4;
    for (unsigned i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const unsigned c = i % C4, px = i / C4, x = px % W2, r = px / W2, y = r % H2, b = r / H2;
        out[i] = in[((b * H + y / 2) * W + x / 2) * C4 + c];
    }
}
__global__ void upsample2_kernel(const float *in, int B, int H, int W, int C, float *out)
{
    const size_t n = (size_t)B * (2 * H) * (2 * W) * C;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const int c = i % C;
        size_t r = i / C;
        const int x = r % (2 * W);
        r /= (2 * W);
        const int y = r % (2 * H);
        const int b = r / (2 * H);
        out[i] = in[((size_t)(b * H + y / 2) * W + x / 2) * C + c];
    }
}

// upsample2 backward: each low-resolution cell sums its 2x2 block, then the
// ReLU mask of the low-resolution activation (if any)
__global__ void upsample2_back_kernel(const float *g, int B, int H, int W, int C, const float *mask, float *out)
{
    const size_t n = (size_t)B * H * W * C;   // low resolution
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const int c = i % C;
        size_t r = i / C;
        const int x = r % W;
        r /= W;
        const int y = r % H;
        const int b = r / H;
        const size_t W2 = 2 * (size_t)W;
        const float *p = g + ((size_t)(b * 2 * H + 2 * y) * W2 + 2 * x) * C + c;
        float v = (p[0] + p[C]) + (p[W2 * C] + p[W2 * C + C]);
        if (mask && !(mask[i] > 0.f)) v = 0.f;
        out[i] = v;
    }
}

// mean squared error: grad = 2 (y - t) / n, per-block partial sums of (y - t)^2
__global__ void mse_kernel(const float *y, const float *t, size_t n, float *grad, double *part)
{
    __shared__ double red[kT];
    double s = 0.0;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const float d = y[i] - t[i];
        grad[i] = (float)(2.0 * (double)d / (double)n);
        s += (double)d * (double)d;
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for (int off = kT / 2; off > 0; off >>= 1) {
        if (threadIdx.x < off) red[threadIdx.x] += red[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// Adam (Kingma and Ba 2015, Algorithm 1) with bias corrections c1 = 1 - b1^t, c2 = 1 - b2^t
__global__ void adam_kernel(float *p, const float *g, float *m, float *v, size_t n, float lr, float b1, float b2,
                            float eps, float c1, float c2)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const float gi = g[i];
        const float mi = b1 * m[i] + (1.f - b1) * gi;
        const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
        m[i] = mi;
        v[i] = vi;
        p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    }
}

// network input from G x G images: per patch p with origin (x0, y0),
// channel 0 = residual[y0 + i][x0 + j] / peak[p] (0 outside the image),
// channel 1 = psf[G/2 - T/2 + i][G/2 - T/2 + j]
__global__ void gather_patches_kernel(int G, int T, int n, const int2 *origins, const float *res, const float *peak,
                                      const float *psf, float *batch)
{
    const size_t total = (size_t)n * T * T;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (size_t)gridDim.x * blockDim.x) {
        const int x = i % T, y = (i / T) % T, p = (int)(i / ((size_t)T * T));
        const int2 o = origins[p];
        const int gy = o.y + y, gx = o.x + x;
        const float r = (gy >= 0 && gy < G && gx >= 0 && gx < G) ? res[(size_t)gy * G + gx] / peak[p] : 0.f;
        const int py = G / 2 - T / 2 + y, px = G / 2 - T / 2 + x;
        const float q = (py >= 0 && py < G && px >= 0 && px < G) ? psf[(size_t)py * G + px] : 0.f;
        batch[2 * i] = r;
        batch[2 * i + 1] = q;
    }
}

// one channel of patches: out[p][i][j] = img[y0 + i][x0 + j] / peak[p]
// (0 outside the image)
__global__ void crop_patches_kernel(int G, int T, int n, const int2 *origins, const float *img, const float *peak,
                                    float *out)
{
    const size_t total = (size_t)n * T * T;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (size_t)gridDim.x * blockDim.x) {
        const int x = i % T, y = (i / T) % T, p = (int)(i / ((size_t)T * T));
        const int2 o = origins[p];
        const int gy = o.y + y, gx = o.x + x;
        out[i] = (gy >= 0 && gy < G && gx >= 0 && gx < G) ? img[(size_t)gy * G + gx] / peak[p] : 0.f;
    }
}

// A sky image for training pairs: per pixel, in double, the point sources
// (y, x, flux) at it in order, then each elliptical Gaussian (cy, cx,
// fwhm_a, fwhm_b, theta, peak) whose stamp -- the square of half side
// r = ceil(3 max(fwhm_a, fwhm_b)) around (floor cy, floor cx) -- holds it:
// peak exp(-(u^2 / (fwhm_a^2 k) + v^2 / (fwhm_b^2 k)) / 2), k = 1 / (8 ln 2),
// (u, v) the offset rotated by theta.  The parameters sit in LDS.
constexpr int kSkyMax = 1024;   // points + Gaussians x 6 doubles per launch
__global__ void sky_render_kernel(int G, int n_points, const double *points, int n_gauss, const double *gauss,
                                  float *sky)
{
    __shared__ double sp[3 * kSkyMax], sg[6 * kSkyMax / 2];
    for (int i = threadIdx.x; i < 3 * n_points; i += blockDim.x) sp[i] = points[i];
    for (int i = threadIdx.x; i < 6 * n_gauss; i += blockDim.x) sg[i] = gauss[i];
    __syncthreads();
    const double k = 1.0 / (8.0 * log(2.0));
    const size_t npix = (size_t)G * G;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < npix; i += (size_t)gridDim.x * blockDim.x) {
        const int y = (int)(i / G), x = (int)(i % G);
        double v = 0.0;
        for (int q = 0; q < n_points; ++q)
            if ((int)sp[3 * q] == y && (int)sp[3 * q + 1] == x) v += sp[3 * q + 2];
        for (int q = 0; q < n_gauss; ++q) {
            const double *g = sg + 6 * q;
            const double cy = g[0], cx = g[1], fa = g[2], fb = g[3], th = g[4];
            const int r = (int)ceil(3.0 * fmax(fa, fb));
            const int iy = (int)cy, ix = (int)cx;
            if (y < iy - r || y > iy + r || x < ix - r || x > ix + r) continue;
            const double dy = (double)y - cy, dx = (double)x - cx;
            const double u = dx * cos(th) + dy * sin(th);
            const double w = -dx * sin(th) + dy * cos(th);
            v += g[5] * exp(-0.5 * (u * u / (fa * fa * k) + w * w / (fb * fb * k)));
        }
        sky[i] = (float)v;
    }
}

// model[y0 + i][x0 + j] += scale * peak[p] * out[p][i][j] over each patch's
// centre [margin, T - margin)^2 (centres of distinct patches do not overlap),
// for outputs above floor only
__global__ void scatter_add_kernel(int G, int T, int margin, int n, const int2 *origins, const float *out,
                                   const float *peak, float scale, float floor, const float *batch_in, float in_floor,
                                   float *model)
{
    const int C = T - 2 * margin;
    const size_t total = (size_t)n * C * C;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (size_t)gridDim.x * blockDim.x) {
        const int x = margin + (int)(i % C), y = margin + (int)((i / C) % C), p = (int)(i / ((size_t)C * C));
        const int2 o = origins[p];
        const int gy = o.y + y, gx = o.x + x;
        const size_t k = ((size_t)p * T + y) * T + x;
        const float v = out[k];
        const bool near = !batch_in || batch_in[2 * k] > in_floor;   // input residual / peak above in_floor
        if (gy >= 0 && gy < G && gx >= 0 && gx < G && v > floor && near)
            model[(size_t)gy * G + gx] += scale * peak[p] * v;
    }
}

// largest |residual| over each candidate patch (the whole T x T tile): a
// 1024-thread workgroup per patch, thread (r, c) taking columns c, c + 256,
// ... of rows r, r + 4, ... (row-contiguous reads, no divisions)
constexpr int kPeakThreads = 1024;
__global__ void __launch_bounds__(kPeakThreads)
patch_peak_kernel(int G, int T, int margin, int stride, int npx, const float *res, float *peak)
{
    __shared__ float red[kPeakThreads / 64];
    const int p = blockIdx.x;
    const int x0 = (p % npx) * stride - margin, y0 = (p / npx) * stride - margin;
    const int c = threadIdx.x & 255, r = threadIdx.x >> 8;
    float m = 0.f;
    for (int y = r; y < T; y += kPeakThreads / 256) {
        const int gy = y0 + y;
        if (gy < 0 || gy >= G) continue;
        const float *row = res + (size_t)gy * G;
        for (int x = c; x < T; x += 256) {
            const int gx = x0 + x;
            if (gx >= 0 && gx < G) m = fmaxf(m, fabsf(row[gx]));
        }
    }
    for (int off = 32; off > 0; off >>= 1) m = fmaxf(m, __shfl_xor(m, off));
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        float v = red[0];
        for (int w = 1; w < kPeakThreads / 64; ++w) v = fmaxf(v, red[w]);
        peak[p] = v;
    }
}

unsigned int ew_grid(size_t n)
{
    size_t b = (n + kT - 1) / kT;
    return (unsigned int)std::max<size_t>(1, std::min<size_t>(b, 16384));
}

}  // namespace

// ---- network object ---------------------------------------------------------

struct sdp_cnn {
    sdp_cnn_params p;
    struct Layer {
        int cin, cout;
        size_t w_off, b_off;   // into the parameter vector
        int cout_pad;          // output channels of its BF16 weights (16 when cout < 16)
        int cin_pad;           // input channels of its BF16 weights (cin rounded up to 32)
        size_t w16_off;        // into wt16
    };
    std::vector<Layer> layers;   // enc0a, enc0b, ..., dec(L-2)u, dec(L-2)d, ..., dec0u, dec0d, head
    size_t n_params = 0;
    float *params = nullptr, *grads = nullptr, *adam_m = nullptr, *adam_v = nullptr;
    __bf16 *wt16 = nullptr;       // BF16 weights in [Cout_pad][9][Cin_pad] order (conv_fwd_tile_kernel)
    size_t n_wt16 = 0;
    int wt16_stale = 1;           // params changed since wt16 was formed
    int adam_t = 0;
    // activations per level (B_max x H_l x W_l x C)
    std::vector<float *> x, a, e, up, s, d, genc, g1, g2;
    float *work = nullptr;        // weight-gradient split partials
    size_t work_floats = 0;
    double *loss_part = nullptr;
    int batch = 0;                // batch of the last forward
    int pooled_mem[8] = {};       // the last forward formed x[l] in memory
    int upsampled_mem[8] = {};    // ... and up[l]
    const float *last_in = nullptr;

    int ch(int l) const { return p.base << l; }
    int hl(int l) const { return p.height >> l; }
    int wl(int l) const { return p.width >> l; }
};

namespace {

int enc_a(int l) { return 2 * l; }
int enc_b(int l) { return 2 * l + 1; }
int dec_u(const sdp_cnn *n, int l) { return 2 * n->p.levels + 2 * (n->p.levels - 2 - l); }
int dec_d(const sdp_cnn *n, int l) { return dec_u(n, l) + 1; }
int head(const sdp_cnn *n) { return (int)n->layers.size() - 1; }

// K splits of a weight-gradient GEMM: enough workgroups to fill the GPU
int wgrad_splits(const GemmArgs &g, int *chunk)
{
    const int tiles = ((g.M + kBM - 1) / kBM) * ((g.N + kBN - 1) / kBN);
    int splits = std::max(1, std::min(1024 / std::max(tiles, 1), (g.K + 4 * kBK - 1) / (4 * kBK)));
    *chunk = ((g.K + splits - 1) / splits + kBK - 1) / kBK * kBK;
    return (g.K + *chunk - 1) / *chunk;
}

int run_gemm(const sdp_cnn *n, int mode, GemmArgs g, hipStream_t st)
{
    const bool bf16 = n->p.precision == SDP_CNN_BF16;
    dim3 grid((g.M + kBM - 1) / kBM, (g.N + kBN - 1) / kBN, 1);
    if (mode == G_BWD_W) {
        const int splits = wgrad_splits(g, &g.k_chunk);
        SDP_REQUIRE((size_t)splits * g.M * g.N <= n->work_floats, "weight-gradient workspace too small");
        grid.z = splits;
    }
    switch (mode) {
    case G_FWD:
        if (bf16 && g.N <= 32 && g.Cin >= 8 && (g.Cin & (g.Cin - 1)) == 0) {
            const int blocks = (g.M + kNarrowM - 1) / kNarrowM, sh = __builtin_ctz(g.Cin);
            if (g.N <= 16) gemm_fwd_narrow_kernel<16><<<blocks, kT, 0, st>>>(g, sh);
            else gemm_fwd_narrow_kernel<32><<<blocks, kT, 0, st>>>(g, sh);
        } else if (bf16) {
            gemm_bf16_kernel<G_FWD><<<grid, kT, 0, st>>>(g);
        } else {
            gemm_f32_kernel<G_FWD><<<grid, kT, 0, st>>>(g);
        }
        break;
    case G_BWD_DATA:
        if (bf16) gemm_bf16_kernel<G_BWD_DATA><<<grid, kT, 0, st>>>(g);
        else gemm_f32_kernel<G_BWD_DATA><<<grid, kT, 0, st>>>(g);
        break;
    default:
        if (bf16) gemm_bf16_kernel<G_BWD_W><<<grid, kT, 0, st>>>(g);
        else gemm_f32_kernel<G_BWD_W><<<grid, kT, 0, st>>>(g);
        break;
    }
    SDP_LAUNCH_CHECK();
    return SDP_OK;
}

GemmArgs geom(int B, int H, int W, int cin, int cout)
{
    GemmArgs g{};
    g.B = B; g.H = H; g.W = W; g.Cin = cin; g.Cout = cout;
    return g;
}

// whether layer li at H x W takes conv_fwd_tile_kernel (which can also read
// its input through the 2x upsampling or the 2x2 pooling)
bool tile_path(const sdp_cnn *n, int li, int H, int W)
{
    const auto &L = n->layers[li];
    return n->p.precision == SDP_CNN_BF16 && (L.cin % kBK == 0 || L.cin < kBK) &&
           (L.cout_pad == 16 || L.cout_pad == 32 || L.cout_pad % 64 == 0) && H % 8 == 0 && W % 16 == 0;
}

template <int IN>
void launch_tile(const ConvArgs &c, int BN, int TH, dim3 grid, hipStream_t st)
{
    if (TH == 16) {
        if (BN == 16) conv_fwd_tile_kernel<16, 16, IN><<<grid, 512, 0, st>>>(c);
        else if (BN == 32) conv_fwd_tile_kernel<32, 16, IN><<<grid, 512, 0, st>>>(c);
        else conv_fwd_tile_kernel<64, 16, IN><<<grid, 512, 0, st>>>(c);
    } else {
        if (BN == 16) conv_fwd_tile_kernel<16, 8, IN><<<grid, 256, 0, st>>>(c);
        else if (BN == 32) conv_fwd_tile_kernel<32, 8, IN><<<grid, 256, 0, st>>>(c);
        else conv_fwd_tile_kernel<64, 8, IN><<<grid, 256, 0, st>>>(c);
    }
}

// out = act(conv(in) + bias (+ skip)) at H x W; `in_mode` (tile path only):
// in is H x W (IN_DIRECT), H/2 x W/2 read upsampled, or 2H x 2W read pooled
int conv_forward(const sdp_cnn *n, int li, int B, int H, int W, const float *in, float *out, const float *skip,
                 int relu, hipStream_t st, int in_mode = IN_DIRECT)
{
    const auto &L = n->layers[li];
    if (tile_path(n, li, H, W)) {
        ConvArgs c{B, H, W, L.cin, L.cout, in, n->wt16 + L.w16_off, n->params + L.b_off, skip, out, relu, L.cin_pad};
        const int BN = std::min(L.cout_pad, 64), TH = H % 16 == 0 ? 16 : 8;
        const dim3 grid((unsigned)(B * (H / TH) * (W / 16)), (unsigned)(L.cout_pad / BN));
        if (in_mode == IN_UPSAMPLE) launch_tile<IN_UPSAMPLE>(c, BN, TH, grid, st);
        else if (in_mode == IN_POOL) launch_tile<IN_POOL>(c, BN, TH, grid, st);
        else launch_tile<IN_DIRECT>(c, BN, TH, grid, st);
        SDP_LAUNCH_CHECK();
        return SDP_OK;
    }
    SDP_REQUIRE(in_mode == IN_DIRECT, "conv_forward: a resampled input needs the tiled kernel");
    GemmArgs g = geom(B, H, W, L.cin, L.cout);
    g.M = B * H * W; g.N = L.cout; g.K = 9 * L.cin;
    g.a_src = in; g.b_src = n->params + L.w_off; g.out = out;
    g.bias = n->params + L.b_off; g.skip = skip; g.relu = relu;
    return run_gemm(n, G_FWD, g, st);
}

// x[l + 1] = avgpool2(e[l]) and up[l] = upsample2(d[l + 1]) (e at the top)
int pool_level(const sdp_cnn *n, int l, int B, hipStream_t st)
{
    const int H = n->hl(l), W = n->wl(l);
    const size_t np = (size_t)B * (H / 2) * (W / 2) * n->ch(l);
    if (n->ch(l) % 4 == 0 && (size_t)B * H * W * n->ch(l) / 4 < (1u << 31))
        avgpool2_vec_kernel<<<ew_grid(np / 4), kT, 0, st>>>((const float4 *)n->e[l], B, H, W, n->ch(l) / 4,
                                                           (float4 *)n->x[l + 1]);
    else
        avgpool2_kernel<<<ew_grid(np), kT, 0, st>>>(n->e[l], B, H, W, n->ch(l), n->x[l + 1]);
    SDP_LAUNCH_CHECK();
    return SDP_OK;
}

int upsample_level(const sdp_cnn *n, int l, int B, const float *low, hipStream_t st)
{
    const int H = n->hl(l), W = n->wl(l);
    const size_t nu = (size_t)B * H * W * n->ch(l + 1);
    if (n->ch(l + 1) % 4 == 0 && nu / 4 < (1u << 31))
        upsample2_vec_kernel<<<ew_grid(nu / 4), kT, 0, st>>>((const float4 *)low, B, H / 2, W / 2, n->ch(l + 1) / 4,
                                                            (float4 *)n->up[l]);
    else
        upsample2_kernel<<<ew_grid(nu), kT, 0, st>>>(low, B, H / 2, W / 2, n->ch(l + 1), n->up[l]);
    SDP_LAUNCH_CHECK();
    return SDP_OK;
}

// gin = conv backward data of gout (masked by mask > 0 when mask is given);
// weight and bias gradients of the layer from (in, gout)
int conv_backward(const sdp_cnn *n, int li, int B, int H, int W, const float *in, const float *gout, float *gin,
                  const float *mask, hipStream_t st)
{
    const auto &L = n->layers[li];
    if (gin) {
        GemmArgs g = geom(B, H, W, L.cin, L.cout);
        g.M = B * H * W; g.N = L.cin; g.K = 9 * L.cout;
        g.a_src = gout; g.b_src = n->params + L.w_off; g.out = gin; g.mask = mask;
        if (int e = run_gemm(n, G_BWD_DATA, g, st)) return e;
    }
    GemmArgs g = geom(B, H, W, L.cin, L.cout);
    g.M = 9 * L.cin; g.N = L.cout; g.K = B * H * W;
    g.a_src = in; g.b_src = gout; g.out = n->work;
    if (int e = run_gemm(n, G_BWD_W, g, st)) return e;
    int chunk = 0;
    const int sp = wgrad_splits(g, &chunk);
    const size_t nw = (size_t)g.M * g.N;
    sum_splits_kernel<<<ew_grid(nw), kT, 0, st>>>(n->work, sp, nw, n->grads + L.w_off);
    SDP_LAUNCH_CHECK();
    // the weight partials are consumed: the bias partials reuse the workspace
    // (at most kBiasBlocks * cout floats; it holds 1024 times the largest
    // layer's 9 * cin * cout weights)
    const int M = B * H * W, nb = std::min(kBiasBlocks, (M + 1023) / 1024), rows = (M + nb - 1) / nb;
    bias_part_kernel<<<nb, kT, 0, st>>>(gout, M, L.cout, rows, n->work);
    SDP_LAUNCH_CHECK();
    bias_sum_kernel<<<(L.cout + kT - 1) / kT, kT, 0, st>>>(n->work, nb, L.cout, n->grads + L.b_off);
    SDP_LAUNCH_CHECK();
    return SDP_OK;
}

// wt16 from the FP32 parameters when they changed since it was formed
int refresh_bf16_weights(sdp_cnn *n, hipStream_t st)
{
    if (!n->wt16_stale) return SDP_OK;
    Bf16Layers L{};
    SDP_REQUIRE(n->layers.size() <= (size_t)kMaxLayers, "too many layers for the BF16 weight table");
    L.n = (int)n->layers.size();
    for (int i = 0; i < L.n; ++i) {
        L.cin[i] = n->layers[i].cin;
        L.cout[i] = n->layers[i].cout;
        L.cinw[i] = n->layers[i].cin_pad;
        L.off[i] = (long long)n->layers[i].w_off;
        L.dst[i] = (long long)n->layers[i].w16_off;
    }
    L.off[L.n] = (long long)n->n_params;
    bf16_weights_kernel<<<ew_grid(n->n_params), kT, 0, st>>>(n->params, L, n->wt16);
    SDP_LAUNCH_CHECK();
    n->wt16_stale = 0;
    return SDP_OK;
}

}  // namespace

extern "C" {

int sdp_cnn_create(sdp_cnn **out, const sdp_cnn_params *pp)
{
    SDP_REQUIRE(out && pp, "sdp_cnn_create: null argument");
    *out = nullptr;
    const sdp_cnn_params p = *pp;
    SDP_REQUIRE(p.levels >= 1 && p.levels <= 6 && p.base >= 1 && p.base <= 512 && p.in_ch >= 1 && p.out_ch >= 1,
                "bad network shape");
    SDP_REQUIRE(p.max_batch >= 1 && p.height >= 1 && p.width >= 1, "bad activation size");
    SDP_REQUIRE(p.height % (1 << (p.levels - 1)) == 0 && p.width % (1 << (p.levels - 1)) == 0,
                "height and width must be divisible by 2^(levels-1)");
    SDP_REQUIRE(p.precision == SDP_CNN_FP32 || p.precision == SDP_CNN_BF16, "bad precision");
    SDP_REQUIRE((long long)p.max_batch * p.height * p.width < (1LL << 31) / std::max(1, p.base << (p.levels - 1)),
                "activations too large");
    if (int e = sdp_init()) return e;
    sdp_cnn *n = new sdp_cnn();
    n->p = p;
    auto add = [&](int cin, int cout) {
        const int cp = cout < 16 ? 16 : cout, cip = (cin + 31) / 32 * 32;
        sdp_cnn::Layer L{cin, cout, n->n_params, n->n_params + (size_t)9 * cin * cout, cp, cip, n->n_wt16};
        n->n_params += (size_t)9 * cin * cout + cout;
        n->n_wt16 += (size_t)9 * cip * cp;
        n->layers.push_back(L);
    };
    int prev = p.in_ch;
    for (int l = 0; l < p.levels; ++l) {
        add(prev, n->ch(l));
        add(n->ch(l), n->ch(l));
        prev = n->ch(l);
    }
    for (int l = p.levels - 2; l >= 0; --l) {
        add(n->ch(l + 1), n->ch(l));
        add(n->ch(l), n->ch(l));
    }
    add(n->ch(0), p.out_ch);

    bool ok = true;
    auto alloc = [&](float **ptr, size_t count) {
        if (!ok) return;
        if (hipMalloc((void **)ptr, sizeof(float) * std::max<size_t>(count, 1)) != hipSuccess) ok = false;
    };
    alloc(&n->params, n->n_params);
    if (ok && hipMalloc((void **)&n->wt16, sizeof(__bf16) * std::max<size_t>(n->n_wt16, 1)) != hipSuccess) ok = false;
    if (ok && hipMemset(n->wt16, 0, sizeof(__bf16) * std::max<size_t>(n->n_wt16, 1)) != hipSuccess) ok = false;
    alloc(&n->grads, n->n_params);
    alloc(&n->adam_m, n->n_params);
    alloc(&n->adam_v, n->n_params);
    const int L = p.levels;
    n->x.assign(L, nullptr); n->a.assign(L, nullptr); n->e.assign(L, nullptr); n->up.assign(L, nullptr);
    n->s.assign(L, nullptr); n->d.assign(L, nullptr); n->genc.assign(L, nullptr);
    n->g1.assign(L, nullptr); n->g2.assign(L, nullptr);
    size_t wmax = 0;
    for (int l = 0; l < L; ++l) {
        const size_t pix = (size_t)p.max_batch * n->hl(l) * n->wl(l);
        const int c = n->ch(l), cu = l + 1 < L ? n->ch(l + 1) : c;
        const int cmax = std::max({c, cu, p.in_ch, p.out_ch});
        if (l > 0) alloc(&n->x[l], pix * c);   // pooled input of level l (channels of level l - 1 <= c)
        alloc(&n->a[l], pix * c);
        alloc(&n->e[l], pix * c);
        if (l < L - 1) {
            alloc(&n->up[l], pix * cu);
            alloc(&n->s[l], pix * c);
            alloc(&n->d[l], pix * c);
            alloc(&n->genc[l], pix * c);
        }
        alloc(&n->g1[l], pix * cmax);
        alloc(&n->g2[l], pix * cmax);
        for (const auto &ly : n->layers) wmax = std::max(wmax, (size_t)9 * ly.cin * ly.cout);
    }
    n->work_floats = wmax * 1024;
    alloc(&n->work, n->work_floats);
    if (ok && hipMalloc((void **)&n->loss_part, sizeof(double) * 1024) != hipSuccess) ok = false;
    if (ok) {
        ok = hipMemset(n->grads, 0, sizeof(float) * n->n_params) == hipSuccess &&
             hipMemset(n->adam_m, 0, sizeof(float) * n->n_params) == hipSuccess &&
             hipMemset(n->adam_v, 0, sizeof(float) * n->n_params) == hipSuccess &&
             hipMemset(n->params, 0, sizeof(float) * n->n_params) == hipSuccess;
    }
    if (!ok) {
        sdp_cnn_destroy(n);
        return sdp::set_error(SDP_E_NOMEM, "sdp_cnn_create: device allocation failed");
    }
    *out = n;
    return SDP_OK;
}

int sdp_cnn_destroy(sdp_cnn *n)
{
    if (!n) return SDP_OK;
    float *bufs[] = {n->params, n->grads, n->adam_m, n->adam_v, n->work};
    for (float *b : bufs) (void)hipFree(b);
    (void)hipFree(n->wt16);
    for (auto *v : {&n->x, &n->a, &n->e, &n->up, &n->s, &n->d, &n->genc, &n->g1, &n->g2})
        for (float *b : *v) (void)hipFree(b);
    (void)hipFree(n->loss_part);
    delete n;
    return SDP_OK;
}

size_t sdp_cnn_param_count(const sdp_cnn *n) { return n ? n->n_params : 0; }

int sdp_cnn_describe(const sdp_cnn *n, sdp_cnn_params *out)
{
    SDP_REQUIRE(n && out, "sdp_cnn_describe: null argument");
    *out = n->p;
    return SDP_OK;
}

int sdp_cnn_set_params(sdp_cnn *n, const float *h_params)
{
    SDP_REQUIRE(n && h_params, "sdp_cnn_set_params: null argument");
    SDP_HIP_TRY(hipMemcpy(n->params, h_params, sizeof(float) * n->n_params, hipMemcpyHostToDevice));
    n->wt16_stale = 1;
    n->adam_t = 0;
    SDP_HIP_TRY(hipMemset(n->adam_m, 0, sizeof(float) * n->n_params));
    SDP_HIP_TRY(hipMemset(n->adam_v, 0, sizeof(float) * n->n_params));
    return SDP_OK;
}

int sdp_cnn_get_params(const sdp_cnn *n, float *h_params)
{
    SDP_REQUIRE(n && h_params, "sdp_cnn_get_params: null argument");
    SDP_HIP_TRY(hipMemcpy(h_params, n->params, sizeof(float) * n->n_params, hipMemcpyDeviceToHost));
    return SDP_OK;
}

int sdp_cnn_get_grads(const sdp_cnn *n, float *h_grads)
{
    SDP_REQUIRE(n && h_grads, "sdp_cnn_get_grads: null argument");
    SDP_HIP_TRY(hipMemcpy(h_grads, n->grads, sizeof(float) * n->n_params, hipMemcpyDeviceToHost));
    return SDP_OK;
}

int sdp_cnn_forward(sdp_cnn *n, const float *d_in, float *d_out, int B, sdp_stream s)
{
    SDP_REQUIRE(n && d_in && d_out, "sdp_cnn_forward: null argument");
    SDP_REQUIRE(B >= 1 && B <= n->p.max_batch, "batch %d outside [1, %d]", B, n->p.max_batch);
    hipStream_t st = sdp::stream_of(s);
    if (n->p.precision == SDP_CNN_BF16)
        if (int e = refresh_bf16_weights(n, st)) return e;
    const int L = n->p.levels;
    // A level's pooled input and a decoder's upsampled input are read
    // through the resampling by the tiled kernel when it takes the layer;
    // they are formed in memory otherwise, and for the backward pass when it
    // needs them (pooled_mem / upsampled_mem say which are formed).
    const float *xin = d_in;
    for (int l = 0; l < L; ++l) {
        const int H = n->hl(l), W = n->wl(l);
        int mode = IN_DIRECT;
        if (l > 0) {
            n->pooled_mem[l] = !tile_path(n, enc_a(l), H, W);
            if (n->pooled_mem[l]) {
                if (int e = pool_level(n, l - 1, B, st)) return e;
                xin = n->x[l];
            } else {
                xin = n->e[l - 1];
                mode = IN_POOL;
            }
        }
        if (int e = conv_forward(n, enc_a(l), B, H, W, xin, n->a[l], nullptr, 1, st, mode)) return e;
        if (int e = conv_forward(n, enc_b(l), B, H, W, n->a[l], n->e[l], nullptr, 1, st)) return e;
    }
    const float *dcur = n->e[L - 1];
    for (int l = L - 2; l >= 0; --l) {
        const int H = n->hl(l), W = n->wl(l);
        n->upsampled_mem[l] = !tile_path(n, dec_u(n, l), H, W);
        if (n->upsampled_mem[l]) {
            if (int e = upsample_level(n, l, B, dcur, st)) return e;
            if (int e = conv_forward(n, dec_u(n, l), B, H, W, n->up[l], n->s[l], n->e[l], 1, st)) return e;
        } else {
            if (int e = conv_forward(n, dec_u(n, l), B, H, W, dcur, n->s[l], n->e[l], 1, st, IN_UPSAMPLE)) return e;
        }
        if (int e = conv_forward(n, dec_d(n, l), B, H, W, n->s[l], n->d[l], nullptr, 1, st)) return e;
        dcur = n->d[l];
    }
    if (int e = conv_forward(n, head(n), B, n->hl(0), n->wl(0), dcur, d_out, nullptr, 0, st)) return e;
    n->batch = B;
    n->last_in = d_in;
    return SDP_OK;
}

int sdp_cnn_backward(sdp_cnn *n, const float *d_grad_out, sdp_stream s)
{
    SDP_REQUIRE(n && d_grad_out, "sdp_cnn_backward: null argument");
    SDP_REQUIRE(n->batch >= 1, "sdp_cnn_backward: no forward pass to differentiate");
    hipStream_t st = sdp::stream_of(s);
    const int L = n->p.levels, B = n->batch;
    // head: gradient into d_0 (= e_0 when there is no decoder), masked by its ReLU
    const float *d0 = L > 1 ? n->d[0] : n->e[0];
    float *gcur = n->g1[0];
    if (int e = conv_backward(n, head(n), B, n->hl(0), n->wl(0), d0, d_grad_out, gcur, d0, st)) return e;
    // decoder, level 0 upwards; gcur holds dL/d(d_l) already masked
    for (int l = 0; l < L - 1; ++l) {
        const int H = n->hl(l), W = n->wl(l);
        // d_l = relu(conv(s_l)): gradient into s_l, masked by s_l > 0 (s_l = relu(u_l + e_l))
        if (int e = conv_backward(n, dec_d(n, l), B, H, W, n->s[l], gcur, n->genc[l], n->s[l], st)) return e;
        // u_l = conv(up_l): gradient into up_l (up_l formed now if the forward read it upsampled)
        if (!n->upsampled_mem[l])
            if (int e = upsample_level(n, l, B, l + 1 <= L - 2 ? n->d[l + 1] : n->e[L - 1], st)) return e;
        if (int e = conv_backward(n, dec_u(n, l), B, H, W, n->up[l], n->genc[l], n->g2[l], nullptr, st)) return e;
        // up_l = upsample2(d_{l+1}); d_{l+1} is the next decoder output or, at the top, e_top
        const float *mask = (l + 1 <= L - 2) ? n->d[l + 1] : n->e[L - 1];
        float *gnext = n->g1[l + 1];
        upsample2_back_kernel<<<ew_grid((size_t)B * n->hl(l + 1) * n->wl(l + 1) * n->ch(l + 1)), kT, 0, st>>>(
            n->g2[l], B, n->hl(l + 1), n->wl(l + 1), n->ch(l + 1), mask, gnext);
        SDP_LAUNCH_CHECK();
        gcur = gnext;
    }
    // encoder, top down.  Per level: ge = dL/d(e_l) masked by e_l > 0 in g1[l]
    // (at the top it is already there), ga = dL/d(a_l) in g2[l], and the
    // input gradient dL/d(x_l) back in g1[l] for the level below.
    for (int l = L - 1; l >= 0; --l) {
        const int H = n->hl(l), W = n->wl(l);
        if (l < L - 1) {
            // e_l feeds the skip of decoder level l and the pool into level l + 1
            pool_back_add_mask_kernel<<<ew_grid((size_t)B * H * W * n->ch(l)), kT, 0, st>>>(
                n->g1[l + 1], n->genc[l], n->e[l], B, H, W, n->ch(l), n->g1[l]);
            SDP_LAUNCH_CHECK();
        }
        // e_l = relu(conv(a_l)): gradient into a_l, masked by a_l > 0
        if (int e = conv_backward(n, enc_b(l), B, H, W, n->a[l], n->g1[l], n->g2[l], n->a[l], st)) return e;
        // a_l = relu(conv(x_l)): gradient into x_l (not needed at level 0); x_l
        // formed now if the forward read it pooled
        if (l > 0 && !n->pooled_mem[l])
            if (int e = pool_level(n, l - 1, B, st)) return e;
        const float *xl = l == 0 ? n->last_in : n->x[l];
        if (int e = conv_backward(n, enc_a(l), B, H, W, xl, n->g2[l], l == 0 ? nullptr : n->g1[l], nullptr, st))
            return e;
    }
    return SDP_OK;
}

int sdp_cnn_mse(size_t count, const float *d_out, const float *d_target, float *d_grad, double *loss, sdp_stream s)
{
    SDP_REQUIRE(count > 0 && d_out && d_target && d_grad, "sdp_cnn_mse: bad argument");
    hipStream_t st = sdp::stream_of(s);
    // per-thread, per-device partial sums, allocated on the first call (the
    // call is synchronous, so a thread never has two in flight)
    static thread_local double *parts[64] = {};
    int dev = 0;
    SDP_HIP_TRY(hipGetDevice(&dev));
    SDP_REQUIRE(dev >= 0 && dev < 64, "sdp_cnn_mse: device index out of range");
    if (!parts[dev]) SDP_HIP_TRY(hipMalloc(&parts[dev], sizeof(double) * 1024));
    double *part = parts[dev];
    const unsigned int nb = std::min<unsigned int>(1024, ew_grid(count));
    mse_kernel<<<nb, kT, 0, st>>>(d_out, d_target, count, d_grad, part);
    std::vector<double> h(nb);
    hipError_t e = hipMemcpyAsync(h.data(), part, sizeof(double) * nb, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) return sdp::set_error(SDP_E_HIP, "sdp_cnn_mse: %s", hipGetErrorString(e));
    double sum = 0.0;
    for (double v : h) sum += v;
    if (loss) *loss = sum / (double)count;
    return SDP_OK;
}

int sdp_cnn_gather_patches(int G, int T, int n, const int *d_origins, const float *d_peaks, const float *d_residual,
                           const float *d_psf, float *d_batch, sdp_stream s)
{
    SDP_REQUIRE(G >= 1 && T >= 1 && n >= 0 && d_origins && d_peaks && d_residual && d_psf && d_batch,
                "sdp_cnn_gather_patches: bad argument");
    if (n == 0) return SDP_OK;
    gather_patches_kernel<<<ew_grid((size_t)n * T * T), kT, 0, sdp::stream_of(s)>>>(
        G, T, n, (const int2 *)d_origins, d_residual, d_peaks, d_psf, d_batch);
    SDP_LAUNCH_CHECK();
    return SDP_OK;
}

int sdp_cnn_crop_patches(int G, int T, int n, const int *d_origins, const float *d_peaks, const float *d_image,
                         float *d_out, sdp_stream s)
{
    SDP_REQUIRE(G >= 1 && T >= 1 && n >= 0 && d_origins && d_peaks && d_image && d_out,
                "sdp_cnn_crop_patches: bad argument");
    if (n == 0) return SDP_OK;
    crop_patches_kernel<<<ew_grid((size_t)n * T * T), kT, 0, sdp::stream_of(s)>>>(
        G, T, n, (const int2 *)d_origins, d_image, d_peaks, d_out);
    SDP_LAUNCH_CHECK();
    return SDP_OK;
}

int sdp_sky_render(int G, int n_points, const double *d_points, int n_gauss, const double *d_gauss, float *d_sky,
                   sdp_stream s)
{
    SDP_REQUIRE(G >= 1 && n_points >= 0 && n_gauss >= 0 && d_sky && (n_points == 0 || d_points) &&
                    (n_gauss == 0 || d_gauss), "sdp_sky_render: bad argument");
    SDP_REQUIRE(n_points <= kSkyMax && n_gauss <= kSkyMax / 2, "sdp_sky_render: at most %d points and %d Gaussians",
                kSkyMax, kSkyMax / 2);
    sky_render_kernel<<<ew_grid((size_t)G * G), kT, 0, sdp::stream_of(s)>>>(G, n_points, d_points, n_gauss, d_gauss,
                                                                          d_sky);
    SDP_LAUNCH_CHECK();
    return SDP_OK;
}

int sdp_cnn_scatter_add(int G, int T, int margin, int n, const int *d_origins, const float *d_peaks,
                        const float *d_out, float scale, float floor, const float *d_batch_in, float in_floor,
                        float *d_model, sdp_stream s)
{
    SDP_REQUIRE(G >= 1 && T >= 1 && margin >= 0 && 2 * margin < T && n >= 0 && d_origins && d_peaks && d_out &&
                    d_model, "sdp_cnn_scatter_add: bad argument");
    if (n == 0) return SDP_OK;
    const int C = T - 2 * margin;
    scatter_add_kernel<<<ew_grid((size_t)n * C * C), kT, 0, sdp::stream_of(s)>>>(
        G, T, margin, n, (const int2 *)d_origins, d_out, d_peaks, scale, floor, d_batch_in, in_floor, d_model);
    SDP_LAUNCH_CHECK();
    return SDP_OK;
}

int sdp_cnn_patch_peaks(int G, int T, int margin, const float *d_residual, float *d_peaks, int *n_patches,
                        sdp_stream s)
{
    SDP_REQUIRE(G >= 1 && T >= 1 && margin >= 0 && 2 * margin < T && d_residual && n_patches,
                "sdp_cnn_patch_peaks: bad argument");
    const int stride = T - 2 * margin;
    const int npx = (G + stride - 1) / stride;
    *n_patches = npx * npx;
    if (!d_peaks) return SDP_OK;
    patch_peak_kernel<<<npx * npx, kPeakThreads, 0, sdp::stream_of(s)>>>(G, T, margin, stride, npx, d_residual, d_peaks);
    SDP_LAUNCH_CHECK();
    return SDP_OK;
}

int sdp_cnn_adam(sdp_cnn *n, float lr, float beta1, float beta2, float eps, sdp_stream s)
{
    SDP_REQUIRE(n, "sdp_cnn_adam: null network");
    n->adam_t += 1;
    const float c1 = (float)(1.0 - std::pow((double)beta1, n->adam_t));
    const float c2 = (float)(1.0 - std::pow((double)beta2, n->adam_t));
    adam_kernel<<<ew_grid(n->n_params), kT, 0, sdp::stream_of(s)>>>(n->params, n->grads, n->adam_m, n->adam_v,
                                                                     n->n_params, lr, beta1, beta2, eps, c1, c2);
    SDP_LAUNCH_CHECK();
    n->wt16_stale = 1;
    return SDP_OK;
}

}  // extern "C"
