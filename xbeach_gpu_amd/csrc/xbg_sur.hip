// Learned wave surrogate: BF16 MFMA kernels for gfx950 (MI355X) and the
// encoder-decoder that uses them, forward, backward and Adam, all on the
// device.  No counterpart in the reference: the surrogate stands in for the
// wave action solve (wavestep, Wavestep.cu:755-1167) between full-solver
// checkpoints and predicts the wave-group-averaged fields the flow and
// sediment steps read from it (H, wave and roller dissipation, Fx, Fy).
//
// Tensors are NHWC.  Activations and their gradients are BF16, weights are
// kept as FP32 masters (Adam) with BF16 copies for the matrix cores, and
// every product accumulates in FP32 (v_mfma_f32_16x16x32_bf16).
//
// Convolutions are implicit GEMMs (M = pixels, N = output channels, K =
// taps x input channels):
//   k_conv_fwd    forward (also the data gradient, through the flipped and
//                 transposed weights: dX = conv(dZ, W'))
//   k_conv_wgrad  weight gradient (M = output channels, N = input
//                 channels, K = pixels), split over pixel ranges into FP32
//                 partials that k_reduce_parts sums in a fixed order
// A 16 x 16 x 32 MFMA takes from each lane 8 consecutive K values of one A
// row (lane % 16) and of one B column (lane % 16), the lanes' K slices
// being lane / 16; it leaves in each lane four rows 4*(lane/16)+r of column
// lane % 16 of the 16 x 16 result.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/xbg.h"

namespace {

typedef __bf16 bf16;
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ f32x4 mfma16(bf16x8 a, bf16x8 b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ bf16 to_bf16(float x) { return (bf16)x; }

constexpr int KB = 32;        // input channels per K block
constexpr int LROW = KB + 8;  // LDS row (pixel) stride in bf16: 80 B, keeps 16-B reads aligned and spread

// ---------------------------------------------------------------------------
// Forward convolution, 'same' padding (TAPS 9: 3x3, TAPS 1: 1x1).
// Workgroup: 4 waves, TW pixels along x times 64 MB / TW rows (64 MB pixels)
// and BN output channels; wave w computes MB blocks of 16 pixels (pixels
// 64 m + 16 w + 0..15 of the tile, m < MB) x BN channels, so each staged
// weight block feeds MB times the matrix work.  Per K block the input tile
// (+1-pixel halo for 3x3) and the block's weights are staged in LDS.  The
// sum over K blocks and taps runs in the same order for any MB.  Epilogue:
// + bias, optional ReLU, BF16 or FP32 store.
// Requires Cin % 32 == 0, Cout % BN == 0, W % TW == 0, H % (64 MB / TW) == 0.
// ---------------------------------------------------------------------------
template <int TAPS, int TW, int BN, int MB, bool RELU, bool OUTF32>
__global__ __launch_bounds__(256) void k_conv_fwd(const bf16* __restrict__ X, const bf16* __restrict__ Wt,
                                                  const float* __restrict__ bias, void* __restrict__ Yv, int H,
                                                  int W, int Cin, int Cout) {
    constexpr int TR = 64 * MB / TW;            // rows per tile
    constexpr int HALO = TAPS == 9 ? 1 : 0;
    constexpr int SR = TR + 2 * HALO, SC = TW + 2 * HALO;
    __shared__ __attribute__((aligned(16))) bf16 sX[SR * SC * LROW];
    __shared__ __attribute__((aligned(16))) bf16 sW[TAPS * BN * LROW];
    const int tid = threadIdx.x, w = tid >> 6, lane = tid & 63;
    const int ntx = W / TW;
    const int xt = blockIdx.x % ntx, cb = blockIdx.x / ntx;
    const int x0 = xt * TW, y0 = blockIdx.y * TR, n = blockIdx.z;
    const int co0 = cb * BN;
    f32x4 acc[MB][BN / 16];
#pragma unroll
    for (int m = 0; m < MB; m++)
#pragma unroll
        for (int j = 0; j < BN / 16; j++) acc[m][j] = f32x4{0.f, 0.f, 0.f, 0.f};
    const int s = lane >> 4;  // K slice of 8 channels
    const size_t img = (size_t)n * H * W;
    for (int c0 = 0; c0 < Cin; c0 += KB) {
        __syncthreads();
        // input tile: SR x SC pixels x 32 channels, 16-B chunks of 8 channels
        for (int k = tid; k < SR * SC * 4; k += 256) {
            int ch = k & 3, pix = k >> 2;
            int r = pix / SC, c = pix % SC;
            int yy = y0 + r - HALO, xx = x0 + c - HALO;
            bf16x8 v = {};
            if (yy >= 0 && yy < H && xx >= 0 && xx < W)
                v = *(const bf16x8*)(X + (img + (size_t)yy * W + xx) * Cin + c0 + ch * 8);
            *(bf16x8*)(sX + pix * LROW + ch * 8) = v;
        }
        // weights of the block: TAPS x BN x 32
        for (int k = tid; k < TAPS * BN * 4; k += 256) {
            int ch = k & 3, row = k >> 2;  // row = tap * BN + j
            int tap = row / BN, j = row % BN;
            *(bf16x8*)(sW + row * LROW + ch * 8) =
                *(const bf16x8*)(Wt + ((size_t)tap * Cout + co0 + j) * Cin + c0 + ch * 8);
        }
        __syncthreads();
#pragma unroll
        for (int tap = 0; tap < TAPS; tap++) {
            const int dy = TAPS == 9 ? tap / 3 : 0, dx = TAPS == 9 ? tap % 3 : 0;
            bf16x8 a[MB];
#pragma unroll
            for (int m = 0; m < MB; m++) {
                const int q = 64 * m + 16 * w + (lane & 15);  // this lane's pixel (A row) within the tile
                a[m] = *(const bf16x8*)(sX + ((q / TW + dy) * SC + q % TW + dx) * LROW + s * 8);
            }
#pragma unroll
            for (int j = 0; j < BN / 16; j++) {
                bf16x8 b = *(const bf16x8*)(sW + (tap * BN + 16 * j + (lane & 15)) * LROW + s * 8);
#pragma unroll
                for (int m = 0; m < MB; m++) acc[m][j] = mfma16(a[m], b, acc[m][j]);
            }
        }
    }
    // epilogue: lane holds rows (pixels) 64 m + 16 w + 4 s + r, column
    // (channel) lane & 15
#pragma unroll
    for (int j = 0; j < BN / 16; j++) {
        const int co = co0 + 16 * j + (lane & 15);
        const float bv = bias ? bias[co] : 0.0f;
#pragma unroll
        for (int m = 0; m < MB; m++)
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int qq = 64 * m + 16 * w + 4 * s + r;
                const int yy = y0 + qq / TW, xx = x0 + qq % TW;
                float v = acc[m][j][r] + bv;
                if (RELU) v = v > 0.0f ? v : 0.0f;
                const size_t o = (img + (size_t)yy * W + xx) * Cout + co;
                if (OUTF32) ((float*)Yv)[o] = v;
                else ((bf16*)Yv)[o] = to_bf16(v);
            }
    }
}

// ---------------------------------------------------------------------------
// Weight gradient of a convolution: dW[tap][co][ci] = sum over pixels of
// dZ[p][co] * X[p + tap offset][ci].  A workgroup takes a 64 x 64 (or
// smaller) block of (co, ci), one row of taps (dy) and a range of image
// rows; per 64-pixel segment dZ and the three x-shifted copies of X are
// staged transposed (channel-major) in LDS so that a lane's 8 consecutive
// pixels are one 16-B read; each thread stages four consecutive pixels of
// eight channels and writes each channel's four values as one 8-B store.
// Waves take the output-channel tiles in turn; with fewer than four tiles
// (32 output channels) two waves share a tile and split its input-channel
// tiles.  Partials go to part[split][tap][co][ci] (FP32).
// Requires W % 64 == 0 or W == 32 (TW), Cout % CBO == 0, Cin % CBI == 0.
// ---------------------------------------------------------------------------
constexpr int WPX = 64;           // pixels per staged segment
constexpr int WLD = WPX + 8;      // LDS stride of a transposed channel row
typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));
template <int TAPS, int CBO, int CBI>
__global__ __launch_bounds__(256) void k_conv_wgrad(const bf16* __restrict__ X, const bf16* __restrict__ dZ,
                                                    float* __restrict__ part, int N, int H, int W, int Cin,
                                                    int Cout, int rows_per_split) {
    constexpr int NDX = TAPS == 9 ? 3 : 1;
    __shared__ __attribute__((aligned(16))) bf16 sD[CBO * WLD];
    __shared__ __attribute__((aligned(16))) bf16 sXt[NDX * CBI * WLD];
    const int tid = threadIdx.x, w = tid >> 6, lane = tid & 63, s = lane >> 4;
    const int nco = Cout / CBO, nci = Cin / CBI;
    int b = blockIdx.x;
    const int cob = b % nco;
    b /= nco;
    const int cib = b % nci;
    const int dy = blockIdx.y;  // 0..2 (TAPS 9) or 0
    const int split = blockIdx.z;
    const int co0 = cob * CBO, ci0 = cib * CBI;
    constexpr int MT = CBO / 16;  // output-channel tiles
    constexpr int NT = CBI / 16;  // input-channel tiles
    constexpr int WPM = MT >= 4 ? 1 : 4 / MT;        // waves per output-channel tile
    constexpr int MT_PER_W = MT >= 4 ? MT / 4 : 1;
    constexpr int NT_W = NT / WPM;                   // input-channel tiles per wave
    static_assert(MT >= 4 ? MT % 4 == 0 : 4 % MT == 0, "output-channel tiles over four waves");
    static_assert(NT % WPM == 0, "input-channel tiles over the waves of an output tile");
    const int t0 = MT >= 4 ? 0 : (w / MT) * NT_W;
    f32x4 acc[MT_PER_W][NDX][NT_W];
#pragma unroll
    for (int a = 0; a < MT_PER_W; a++)
#pragma unroll
        for (int d = 0; d < NDX; d++)
#pragma unroll
            for (int t = 0; t < NT_W; t++) acc[a][d][t] = f32x4{0.f, 0.f, 0.f, 0.f};
    const int total_rows = N * H;
    const int r0 = split * rows_per_split, r1 = min(total_rows, r0 + rows_per_split);
    const int segs = (W + WPX - 1) / WPX;
    for (int row = r0; row < r1; row++) {
        const int n = row / H, y = row % H;
        const int ys = y + (TAPS == 9 ? dy - 1 : 0);
        for (int sg = 0; sg < segs; sg++) {
            const int x0 = sg * WPX;
            const int np = min(WPX, W - x0);
            __syncthreads();
            // dZ segment, transposed: sD[co][p], four pixels per thread
            for (int k = tid; k < (WPX / 4) * (CBO / 8); k += 256) {
                const int p = (k / (CBO / 8)) * 4, ch = (k % (CBO / 8)) * 8;
                bf16x8 v[4];
#pragma unroll
                for (int u = 0; u < 4; u++) {
                    v[u] = bf16x8{};
                    if (p + u < np) v[u] = *(const bf16x8*)(dZ + (((size_t)n * H + y) * W + x0 + p + u) * Cout + co0 + ch);
                }
#pragma unroll
                for (int e = 0; e < 8; e++)
                    *(bf16x4*)(sD + (ch + e) * WLD + p) = bf16x4{v[0][e], v[1][e], v[2][e], v[3][e]};
            }
            // X of the shifted row, one copy per dx, transposed: sXt[dx][ci][p]
            for (int k = tid; k < NDX * (WPX / 4) * (CBI / 8); k += 256) {
                const int d = k / ((WPX / 4) * (CBI / 8));
                const int r = k % ((WPX / 4) * (CBI / 8));
                const int p = (r / (CBI / 8)) * 4, ch = (r % (CBI / 8)) * 8;
                bf16x8 v[4];
#pragma unroll
                for (int u = 0; u < 4; u++) {
                    const int xs = x0 + p + u + (TAPS == 9 ? d - 1 : 0);
                    v[u] = bf16x8{};
                    if (p + u < np && ys >= 0 && ys < H && xs >= 0 && xs < W)
                        v[u] = *(const bf16x8*)(X + (((size_t)n * H + ys) * W + xs) * Cin + ci0 + ch);
                }
#pragma unroll
                for (int e = 0; e < 8; e++)
                    *(bf16x4*)(sXt + (d * CBI + ch + e) * WLD + p) = bf16x4{v[0][e], v[1][e], v[2][e], v[3][e]};
            }
            __syncthreads();
#pragma unroll
            for (int kk = 0; kk < WPX / 32; kk++) {
#pragma unroll
                for (int a = 0; a < MT_PER_W; a++) {
                    const int mt = MT >= 4 ? w + 4 * a : w % MT;
                    bf16x8 av = *(const bf16x8*)(sD + (mt * 16 + (lane & 15)) * WLD + kk * 32 + s * 8);
#pragma unroll
                    for (int d = 0; d < NDX; d++)
#pragma unroll
                        for (int t = 0; t < NT_W; t++) {
                            bf16x8 bv = *(const bf16x8*)(sXt + (d * CBI + (t0 + t) * 16 + (lane & 15)) * WLD +
                                                         kk * 32 + s * 8);
                            acc[a][d][t] = mfma16(av, bv, acc[a][d][t]);
                        }
                }
            }
        }
    }
    // lane holds rows (co) 4s + r of tile mt, column (ci) lane & 15 of tile t0 + t
#pragma unroll
    for (int a = 0; a < MT_PER_W; a++)
#pragma unroll
        for (int d = 0; d < NDX; d++)
#pragma unroll
            for (int t = 0; t < NT_W; t++)
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    const int mt = MT >= 4 ? w + 4 * a : w % MT;
                    const int co = co0 + mt * 16 + 4 * s + r;
                    const int ci = ci0 + (t0 + t) * 16 + (lane & 15);
                    const int tap = TAPS == 9 ? dy * 3 + d : 0;
                    part[(((size_t)split * TAPS + tap) * Cout + co) * Cin + ci] = acc[a][d][t][r];
                }
}

// out[i] = sum over splits of part[s*n + i], in split order
__global__ __launch_bounds__(256) void k_reduce_parts(const float* __restrict__ part, int nsplit, size_t n,
                                                      float* __restrict__ out) {
    size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    float a = 0.0f;
    // unrolled: many loads in flight, the sum still in split order
#pragma unroll 16
    for (int sp = 0; sp < nsplit; sp++) a += part[(size_t)sp * n + i];
    out[i] = a;
}

// Flipped, transposed weights for the data gradient: Wd[tap][ci][co] =
// W[TAPS-1-tap][co][ci] (the 180-degree rotation of a 3x3 kernel).
__global__ __launch_bounds__(256) void k_flip_weights(const bf16* __restrict__ Wt, bf16* __restrict__ Wd, int taps,
                                                      int Cout, int Cin) {
    size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    size_t n = (size_t)taps * Cout * Cin;
    if (i >= n) return;
    int ci = i % Cin;
    size_t r = i / Cin;
    int co = r % Cout;
    int tap = r / Cout;
    Wd[((size_t)(taps - 1 - tap) * Cin + ci) * Cout + co] = Wt[i];
}

// Per-channel sums of a [P][C] BF16 tensor (bias gradient), partials per
// workgroup over a range of pixels: part[blockIdx.x][c].  With C dividing
// 256 the workgroup's threads cover 256 / C pixel lanes of all C channels;
// each lane sums its pixels in order and the lanes are added in lane order.
__global__ __launch_bounds__(256) void k_colsum(const bf16* __restrict__ G, size_t P, int C, int ppb,
                                                float* __restrict__ part) {
    __shared__ float red[256];
    const size_t p0 = (size_t)blockIdx.x * ppb;
    const size_t p1 = p0 + ppb < P ? p0 + ppb : P;
    if (C <= 256 && 256 % C == 0) {
        const int lanes = 256 / C, c = threadIdx.x % C, l = threadIdx.x / C;
        float a = 0.0f;
#pragma unroll 8
        for (size_t p = p0 + l; p < p1; p += lanes) a += (float)G[p * C + c];
        red[threadIdx.x] = a;
        __syncthreads();
        if (l == 0) {
            float t = 0.0f;
            for (int k = 0; k < lanes; k++) t += red[k * C + c];
            part[(size_t)blockIdx.x * C + c] = t;
        }
        return;
    }
    for (int c = threadIdx.x; c < C; c += 256) {
        float a = 0.0f;
#pragma unroll 16
        for (size_t p = p0; p < p1; p++) a += (float)G[p * C + c];
        part[(size_t)blockIdx.x * C + c] = a;
    }
}

// ReLU backward in place: g = g where y > 0 else 0.
__global__ __launch_bounds__(256) void k_relu_bwd(bf16* __restrict__ g, const bf16* __restrict__ y, size_t n) {
    size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n && !((float)y[i] > 0.0f)) g[i] = to_bf16(0.0f);
}

// 2x2 average pooling (NHWC) and its gradient.
__global__ __launch_bounds__(256) void k_pool_fwd(const bf16* __restrict__ X, bf16* __restrict__ Y, int N, int H,
                                                  int W, int C) {
    size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    const int h = H / 2, w = W / 2;
    size_t n_out = (size_t)N * h * w * C;
    if (i >= n_out) return;
    int c = i % C;
    size_t r = i / C;
    int x = r % w;
    r /= w;
    int y = r % h;
    int n = r / h;
    auto at = [&](int yy, int xx) { return (float)X[(((size_t)n * H + yy) * W + xx) * C + c]; };
    float v = 0.25f * (at(2 * y, 2 * x) + at(2 * y, 2 * x + 1) + at(2 * y + 1, 2 * x) + at(2 * y + 1, 2 * x + 1));
    Y[i] = to_bf16(v);
}
// dX (+)= dY / 4 broadcast over each 2x2 cell; accumulate adds to dX
__global__ __launch_bounds__(256) void k_pool_bwd(const bf16* __restrict__ dY, bf16* __restrict__ dX, int N, int H,
                                                  int W, int C, int accumulate) {
    size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    size_t n_in = (size_t)N * H * W * C;
    if (i >= n_in) return;
    int c = i % C;
    size_t r = i / C;
    int x = r % W;
    r /= W;
    int y = r % H;
    int n = r / H;
    float g = 0.25f * (float)dY[(((size_t)n * (H / 2) + y / 2) * (W / 2) + x / 2) * C + c];
    if (accumulate) g += (float)dX[i];
    dX[i] = to_bf16(g);
}

// Nearest 2x upsampling of X [N][h][w][C] into channels [c0, c0+C) of Y
// [N][2h][2w][Ctot] (the decoder's concatenation writes both halves of Y
// in place), and its gradient (2x2 sums).
__global__ __launch_bounds__(256) void k_up_fwd(const bf16* __restrict__ X, bf16* __restrict__ Y, int N, int h,
                                                int w, int C, int Ctot, int c0) {
    size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    size_t n_out = (size_t)N * 2 * h * 2 * w * C;
    if (i >= n_out) return;
    int c = i % C;
    size_t r = i / C;
    int x = r % (2 * w);
    r /= 2 * w;
    int y = r % (2 * h);
    int n = r / (2 * h);
    Y[(((size_t)n * 2 * h + y) * 2 * w + x) * Ctot + c0 + c] = X[(((size_t)n * h + y / 2) * w + x / 2) * C + c];
}
__global__ __launch_bounds__(256) void k_up_bwd(const bf16* __restrict__ dY, bf16* __restrict__ dX, int N, int h,
                                                int w, int C, int Ctot, int c0) {
    size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    size_t n_in = (size_t)N * h * w * C;
    if (i >= n_in) return;
    int c = i % C;
    size_t r = i / C;
    int x = r % w;
    r /= w;
    int y = r % h;
    int n = r / h;
    const int W2 = 2 * w;
    auto at = [&](int yy, int xx) { return (float)dY[(((size_t)n * 2 * h + yy) * W2 + xx) * Ctot + c0 + c]; };
    dX[i] = to_bf16(at(2 * y, 2 * x) + at(2 * y, 2 * x + 1) + at(2 * y + 1, 2 * x) + at(2 * y + 1, 2 * x + 1));
}

// Copy channels [c0, c0+C) between a [P][Ca] and a [P][Cb] tensor:
// dst[p][dc0 + c] = src[p][sc0 + c] (skip concatenation and its gradient
// split), accumulate adds to dst.
__global__ __launch_bounds__(256) void k_chan_copy(const bf16* __restrict__ src, int Cs, int sc0,
                                                   bf16* __restrict__ dst, int Cd, int dc0, int C, size_t P,
                                                   int accumulate) {
    size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= P * C) return;
    int c = i % C;
    size_t p = i / C;
    float v = (float)src[p * Cs + sc0 + c];
    if (accumulate) v += (float)dst[p * Cd + dc0 + c];
    dst[p * Cd + dc0 + c] = to_bf16(v);
}

// The output head: a 1x1 convolution C -> K (K small) in FP32 on the
// vector units, and its gradients.  Y[p][k] = b[k] + sum_c X[p][c] W[k][c].
template <int K>
__global__ __launch_bounds__(256) void k_head_fwd(const bf16* __restrict__ X, const float* __restrict__ Wh,
                                                  const float* __restrict__ bh, float* __restrict__ Y, size_t P,
                                                  int C) {
    size_t p = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (p >= P) return;
    float a[K];
#pragma unroll
    for (int k = 0; k < K; k++) a[k] = bh[k];
    for (int c = 0; c < C; c++) {
        float x = (float)X[p * C + c];
#pragma unroll
        for (int k = 0; k < K; k++) a[k] += x * Wh[k * C + c];
    }
#pragma unroll
    for (int k = 0; k < K; k++) Y[p * K + k] = a[k];
}
// dX[p][c] = sum_k dY[p][k] W[k][c] (BF16 out)
template <int K>
__global__ __launch_bounds__(256) void k_head_dgrad(const float* __restrict__ dY, const float* __restrict__ Wh,
                                                    bf16* __restrict__ dX, size_t P, int C) {
    size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= P * C) return;
    int c = i % C;
    size_t p = i / C;
    float a = 0.0f;
#pragma unroll
    for (int k = 0; k < K; k++) a += dY[p * K + k] * Wh[k * C + c];
    dX[i] = to_bf16(a);
}
// partials of dW[k][c] = sum_p dY[p][k] X[p][c] and db[k] = sum_p dY[p][k]
// over ppb pixels per workgroup: part[blk][K*C + K].  With C dividing 256
// the threads cover 256 / C pixel lanes of all C channels; each lane sums
// its pixels in order and the lanes are added in lane order.
template <int K>
__global__ __launch_bounds__(256) void k_head_wgrad(const float* __restrict__ dY, const bf16* __restrict__ X,
                                                    size_t P, int C, int ppb, float* __restrict__ part) {
    const size_t p0 = (size_t)blockIdx.x * ppb;
    const size_t p1 = p0 + ppb < P ? p0 + ppb : P;
    const int n = K * C + K;
    if (C <= 256 && 256 % C == 0) {
        __shared__ float red[K][256];
        __shared__ float redb[K][256];
        const int lanes = 256 / C, c = threadIdx.x % C, l = threadIdx.x / C;
        float a[K], b[K];
#pragma unroll
        for (int k = 0; k < K; k++) a[k] = b[k] = 0.0f;
#pragma unroll 4
        for (size_t p = p0 + l; p < p1; p += lanes) {
            const float x = (float)X[p * C + c];
#pragma unroll
            for (int k = 0; k < K; k++) {
                const float d = dY[p * K + k];
                a[k] += d * x;
                b[k] += d;
            }
        }
#pragma unroll
        for (int k = 0; k < K; k++) {
            red[k][threadIdx.x] = a[k];
            redb[k][threadIdx.x] = b[k];
        }
        __syncthreads();
        if (l == 0) {
#pragma unroll
            for (int k = 0; k < K; k++) {
                float t = 0.0f;
                for (int q = 0; q < lanes; q++) t += red[k][q * C + c];
                part[(size_t)blockIdx.x * n + k * C + c] = t;
            }
        }
        if ((int)threadIdx.x < K) {
            // the bias sums of the lanes' channel-0 threads
            const int k = threadIdx.x;
            float t = 0.0f;
            for (int q = 0; q < lanes; q++) t += redb[k][q * C];
            part[(size_t)blockIdx.x * n + K * C + k] = t;
        }
        return;
    }
    for (int j = threadIdx.x; j < n; j += 256) {
        float a = 0.0f;
        if (j < K * C) {
            int k = j / C, c = j % C;
#pragma unroll 16
            for (size_t p = p0; p < p1; p++) a += dY[p * K + k] * (float)X[p * C + c];
        } else {
            int k = j - K * C;
#pragma unroll 16
            for (size_t p = p0; p < p1; p++) a += dY[p * K + k];
        }
        part[(size_t)blockIdx.x * n + j] = a;
    }
}

// Mean squared error over the normalised targets: per-workgroup partial
// sums of (y - t)^2 and the gradient g = 2 (y - t) / n.
__global__ __launch_bounds__(256) void k_mse(const float* __restrict__ Y, const bf16* __restrict__ T, size_t n,
                                             float* __restrict__ G, float* __restrict__ part) {
    __shared__ float red[256];
    size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    float e = 0.0f;
    if (i < n) {
        float d = Y[i] - (float)T[i];
        G[i] = 2.0f * d / (float)n;
        e = d * d;
    }
    red[threadIdx.x] = e;
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
        if ((int)threadIdx.x < st) red[threadIdx.x] += red[threadIdx.x + st];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// Adam (Kingma and Ba 2015) on FP32 masters; the BF16 copy (may be NULL)
// follows the update.  bc1 = 1 - beta1^t, bc2 = 1 - beta2^t.
__global__ __launch_bounds__(256) void k_adam(float* __restrict__ w, bf16* __restrict__ w16,
                                              const float* __restrict__ g, float* __restrict__ m,
                                              float* __restrict__ v, size_t n, float lr, float b1, float b2,
                                              float eps, float bc1, float bc2) {
    size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    float gi = g[i];
    float mi = b1 * m[i] + (1.0f - b1) * gi;
    float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    float mh = mi / bc1, vh = vi / bc2;
    float wi = w[i] - lr * mh / (sqrtf(vh) + eps);
    w[i] = wi;
    if (w16) w16[i] = to_bf16(wi);
}

__global__ __launch_bounds__(256) void k_f32_to_bf16(const float* __restrict__ a, bf16* __restrict__ b, size_t n) {
    size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) b[i] = to_bf16(a[i]);
}

}  // namespace

// ===========================================================================
// The network: a 4-level convolutional encoder-decoder (U-Net shape)
//   level 0 (H x W):       32 ch   conv3 x2          skip0
//   level 1 (H/2 x W/2):   64 ch   pool, conv3 x2    skip1
//   level 2 (H/4 x W/4):  128 ch   pool, conv3 x2    skip2
//   level 3 (H/8 x W/8):  256 ch   pool, conv3 x2
//   decoder: conv1 (halve channels) at the coarse level, nearest 2x
//   upsampling, concatenation with the skip, conv3 x2; head conv1 32 -> 5
// ReLU after every 3x3 convolution; the 1x1 convolutions are linear.
// Inputs: 8 channels (zb, zs, uu, vv, Hm0, Tp, dir, 0), zero-padded to 32
// on the device; outputs: 5 channels (H, Dw, Dr, Fx, Fy), normalised.
// 1.8 M parameters.  H must be even and W/8 either 32 or a multiple of 64.
// ===========================================================================
namespace {

struct ConvL {
    int cin, cout, taps, relu;
    size_t w, b;  // offsets of weights [taps][cout][cin] and bias [cout] in the parameter vector
};

}  // namespace

struct xbg_sur {
    int nmax = 0, H = 0, W = 0;
    hipStream_t st = nullptr;
    std::vector<ConvL> L;  // 17 convolutions, then the head (index 17)
    size_t nparam = 0;
    float *P = nullptr, *G = nullptr, *M = nullptr, *V = nullptr;  // FP32 params, grads, Adam moments
    bf16* P16 = nullptr;                                           // BF16 copy of the parameter vector
    bf16* Wd = nullptr;                                            // flipped weights for a data gradient
    float* part = nullptr;                                         // split-K / reduction partials
    size_t npart = 0;
    float* lossp = nullptr;
    int adam_t = 0;
    // activations (A) and their gradients (D), by name
    std::vector<void*> allocs;
    bf16 *x32, *e0a, *e0, *p1, *e1a, *e1, *p2, *e2a, *e2, *p3, *b3a, *b3, *u2, *c2, *d2a, *d2, *u1, *c1, *d1a, *d1,
        *u0, *c0, *d0a, *d0;
    bf16 *g_e0a, *g_e0, *g_p1, *g_e1a, *g_e1, *g_p2, *g_e2a, *g_e2, *g_p3, *g_b3a, *g_b3, *g_u2, *g_c2, *g_d2a,
        *g_d2, *g_u1, *g_c1, *g_d1a, *g_d1, *g_u0, *g_c0, *g_d0a, *g_d0;
    float *y = nullptr, *gy = nullptr;
    std::string err;
};

namespace {

thread_local std::string g_sur_error;

int sfail(xbg_sur* s, int code, const std::string& m) {
    if (s) s->err = m;
    else g_sur_error = m;
    return code;
}
#define SHIP(s, call)                                                                            \
    do {                                                                                         \
        hipError_t e_ = (call);                                                                  \
        if (e_ != hipSuccess) return sfail(s, XBG_E_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

inline unsigned nblk(size_t n) { return (unsigned)((n + 255) / 256); }

// the forward convolution for any supported shape
int conv_fwd(hipStream_t st, const bf16* X, const bf16* Wt, const float* bias, void* Y, int n, int H, int W, int Cin,
             int Cout, int taps, int relu, int outf32) {
    if (Cin % 32 || Cout % 32 || !(taps == 1 || taps == 9) || outf32) return XBG_E_ARG;
    int TW = W % 64 == 0 ? 64 : (W == 32 ? 32 : 0);
    if (!TW || H % (64 / TW)) return XBG_E_ARG;
    int BN = Cout % 64 == 0 ? 64 : 32;
    // four 16-pixel blocks per wave where the rows divide, else one
    const int MB = H % (4 * 64 / TW) == 0 ? 4 : 1;
    dim3 g((W / TW) * (Cout / BN), H / (MB * 64 / TW), n);
#define CF(T_, TW_, BN_, R_)                                                                                       \
    do {                                                                                                       \
        if (MB == 4) k_conv_fwd<T_, TW_, BN_, 4, R_, false><<<g, dim3(256), 0, st>>>(X, Wt, bias, Y, H, W, Cin, Cout); \
        else k_conv_fwd<T_, TW_, BN_, 1, R_, false><<<g, dim3(256), 0, st>>>(X, Wt, bias, Y, H, W, Cin, Cout);    \
    } while (0)
#define CF_R(T_, TW_, BN_) do { if (relu) CF(T_, TW_, BN_, true); else CF(T_, TW_, BN_, false); } while (0)
#define CF_BN(T_, TW_) do { if (BN == 64) CF_R(T_, TW_, 64); else CF_R(T_, TW_, 32); } while (0)
#define CF_TW(T_) do { if (TW == 64) CF_BN(T_, 64); else CF_BN(T_, 32); } while (0)
    if (taps == 9) CF_TW(9);
    else CF_TW(1);
#undef CF
#undef CF_R
#undef CF_BN
#undef CF_TW
    return hipGetLastError() == hipSuccess ? 0 : XBG_E_HIP;
}

// dW (FP32, [taps][Cout][Cin]) of one convolution from its input X and the
// gradient dZ of its pre-activation output; partials in part (npart floats)
int conv_wgrad(hipStream_t st, const bf16* X, const bf16* dZ, float* dW, float* part, size_t npart, int n, int H,
               int W, int Cin, int Cout, int taps) {
    if (Cin % 32 || Cout % 32 || !(taps == 1 || taps == 9)) return XBG_E_ARG;
    const int CBO = Cout % 64 == 0 ? 64 : 32, CBI = Cin % 64 == 0 ? 64 : 32;
    const int gx = (Cout / CBO) * (Cin / CBI), gy = taps == 9 ? 3 : 1;
    const size_t per = (size_t)taps * Cout * Cin;
    int rows = n * H;
    int want = 2048 / (gx * gy);
    if (want < 1) want = 1;
    if ((size_t)want * per > npart) want = (int)(npart / per);
    if (want < 1) return XBG_E_ARG;
    if (want > rows) want = rows;
    int rps = (rows + want - 1) / want;
    int nsplit = (rows + rps - 1) / rps;
    dim3 g(gx, gy, nsplit);
#define CW(T_, O_, I_) k_conv_wgrad<T_, O_, I_><<<g, dim3(256), 0, st>>>(X, dZ, part, n, H, W, Cin, Cout, rps)
#define CW_I(T_, O_) do { if (CBI == 64) CW(T_, O_, 64); else CW(T_, O_, 32); } while (0)
#define CW_O(T_) do { if (CBO == 64) CW_I(T_, 64); else CW_I(T_, 32); } while (0)
    if (taps == 9) CW_O(9);
    else CW_O(1);
#undef CW
#undef CW_I
#undef CW_O
    k_reduce_parts<<<nblk(per), dim3(256), 0, st>>>(part, nsplit, per, dW);
    return hipGetLastError() == hipSuccess ? 0 : XBG_E_HIP;
}

// per-channel sums of a [P][C] gradient into out (bias gradient)
int colsum(hipStream_t st, const bf16* Gm, size_t P, int C, float* out, float* part, size_t npart) {
    int nb = 1024;
    if ((size_t)nb * C > npart) nb = (int)(npart / C);
    int ppb = (int)((P + nb - 1) / nb);
    nb = (int)((P + ppb - 1) / ppb);
    k_colsum<<<nb, 256, 0, st>>>(Gm, P, C, ppb, part);
    k_reduce_parts<<<nblk(C), 256, 0, st>>>(part, nb, (size_t)C, out);
    return hipGetLastError() == hipSuccess ? 0 : XBG_E_HIP;
}

// data gradient dX = conv(dZ, flipped W) for a convolution Cin -> Cout
int conv_dgrad(hipStream_t st, const bf16* dZ, const bf16* Wt, bf16* Wd, bf16* dX, int n, int H, int W, int Cin,
               int Cout, int taps) {
    size_t nw = (size_t)taps * Cout * Cin;
    k_flip_weights<<<nblk(nw), 256, 0, st>>>(Wt, Wd, taps, Cout, Cin);
    return conv_fwd(st, dZ, Wd, nullptr, dX, n, H, W, Cout, Cin, taps, 0, 0);
}

// deterministic initialisation: He-normal weights from a counter-based hash
// (Box-Muller on two 24-bit uniforms), zero biases
unsigned long long mix64(unsigned long long x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}
double gauss(unsigned seed, size_t i) {
    unsigned long long h = mix64(((unsigned long long)seed << 40) ^ (unsigned long long)i);
    double u1 = ((h >> 40) + 0.5) / 16777216.0, u2 = (((h >> 16) & 0xFFFFFF) + 0.5) / 16777216.0;
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
}

}  // namespace

extern "C" {

const char* xbg_sur_last_error(const xbg_sur* s) { return s ? s->err.c_str() : g_sur_error.c_str(); }

int xbg_sur_destroy(xbg_sur* s) {
    if (!s) return XBG_OK;
    if (s->st) hipStreamSynchronize(s->st);
    for (void* p : s->allocs) hipFree(p);
    if (s->st) hipStreamDestroy(s->st);
    delete s;
    return XBG_OK;
}

int xbg_sur_create(int nmax, int H, int W, unsigned seed, xbg_sur** out) {
    if (!out || nmax < 1 || H < 16 || H % 16 || !((W / 8 == 32 && W % 8 == 0) || (W % 512 == 0)))
        return sfail(nullptr, XBG_E_ARG, "xbg_sur_create: need H a multiple of 16 and W = 256 or a multiple of 512");
    *out = nullptr;
    xbg_sur* s = new xbg_sur();
    s->nmax = nmax;
    s->H = H;
    s->W = W;
    auto bail = [&](int rc) {
        xbg_sur_destroy(s);
        return rc;
    };
    const int spec[17][4] = {{32, 32, 9, 1},   {32, 32, 9, 1},   {32, 64, 9, 1},   {64, 64, 9, 1},
                             {64, 128, 9, 1},  {128, 128, 9, 1}, {128, 256, 9, 1}, {256, 256, 9, 1},
                             {256, 128, 1, 0}, {256, 128, 9, 1}, {128, 128, 9, 1}, {128, 64, 1, 0},
                             {128, 64, 9, 1},  {64, 64, 9, 1},   {64, 32, 1, 0},   {64, 32, 9, 1},
                             {32, 32, 9, 1}};
    size_t off = 0;
    for (auto& q : spec) {
        ConvL c{q[0], q[1], q[2], q[3], off, 0};
        off += (size_t)c.taps * c.cout * c.cin;
        c.b = off;
        off += c.cout;
        s->L.push_back(c);
    }
    ConvL head{32, 5, 1, 0, off, 0};
    off += 5 * 32;
    head.b = off;
    off += 5;
    s->L.push_back(head);
    s->nparam = off;
    if (hipStreamCreateWithFlags(&s->st, hipStreamNonBlocking) != hipSuccess)
        return bail(sfail(nullptr, XBG_E_HIP, "xbg_sur_create: stream"));
    auto A = [&](void** p, size_t bytes) -> bool {
        if (hipMalloc(p, bytes ? bytes : 16) != hipSuccess) return false;
        s->allocs.push_back(*p);  // freed by xbg_sur_destroy on any failure below
        return hipMemset(*p, 0, bytes) == hipSuccess;
    };
    const size_t P0 = (size_t)nmax * H * W, P1 = P0 / 4, P2 = P0 / 16, P3 = P0 / 64;
    bool ok = A((void**)&s->P, off * 4) && A((void**)&s->G, off * 4) && A((void**)&s->M, off * 4) &&
              A((void**)&s->V, off * 4) && A((void**)&s->P16, off * 2) && A((void**)&s->Wd, 9 * 256 * 256 * 2);
    s->npart = (size_t)32 << 20;
    ok = ok && A((void**)&s->part, s->npart * 4) && A((void**)&s->lossp, nblk(P0 * 5) * 4);
    struct B {
        bf16** p;
        size_t n;
    } acts[] = {{&s->x32, P0 * 32}, {&s->e0a, P0 * 32}, {&s->e0, P0 * 32},  {&s->p1, P1 * 32},
                {&s->e1a, P1 * 64}, {&s->e1, P1 * 64},  {&s->p2, P2 * 64},  {&s->e2a, P2 * 128},
                {&s->e2, P2 * 128}, {&s->p3, P3 * 128}, {&s->b3a, P3 * 256}, {&s->b3, P3 * 256},
                {&s->u2, P3 * 128}, {&s->c2, P2 * 256}, {&s->d2a, P2 * 128}, {&s->d2, P2 * 128},
                {&s->u1, P2 * 64},  {&s->c1, P1 * 128}, {&s->d1a, P1 * 64}, {&s->d1, P1 * 64},
                {&s->u0, P1 * 32},  {&s->c0, P0 * 64},  {&s->d0a, P0 * 32}, {&s->d0, P0 * 32},
                {&s->g_e0a, P0 * 32}, {&s->g_e0, P0 * 32}, {&s->g_p1, P1 * 32}, {&s->g_e1a, P1 * 64},
                {&s->g_e1, P1 * 64}, {&s->g_p2, P2 * 64}, {&s->g_e2a, P2 * 128}, {&s->g_e2, P2 * 128},
                {&s->g_p3, P3 * 128}, {&s->g_b3a, P3 * 256}, {&s->g_b3, P3 * 256}, {&s->g_u2, P3 * 128},
                {&s->g_c2, P2 * 256}, {&s->g_d2a, P2 * 128}, {&s->g_d2, P2 * 128}, {&s->g_u1, P2 * 64},
                {&s->g_c1, P1 * 128}, {&s->g_d1a, P1 * 64}, {&s->g_d1, P1 * 64}, {&s->g_u0, P1 * 32},
                {&s->g_c0, P0 * 64}, {&s->g_d0a, P0 * 32}, {&s->g_d0, P0 * 32}};
    for (auto& b : acts) ok = ok && A((void**)b.p, b.n * 2);
    ok = ok && A((void**)&s->y, P0 * 5 * 4) && A((void**)&s->gy, P0 * 5 * 4);
    if (!ok) return bail(sfail(nullptr, XBG_E_HIP, "xbg_sur_create: out of device memory"));
    // He-normal initialisation
    std::vector<float> h(off, 0.0f);
    for (auto& c : s->L) {
        double sd = std::sqrt(2.0 / (c.taps * c.cin));
        if (!c.relu) sd = std::sqrt(1.0 / (c.taps * c.cin));
        for (size_t i = 0; i < (size_t)c.taps * c.cout * c.cin; i++) h[c.w + i] = (float)(sd * gauss(seed, c.w + i));
    }
    if (hipMemcpy(s->P, h.data(), off * 4, hipMemcpyHostToDevice) != hipSuccess)
        return bail(sfail(nullptr, XBG_E_HIP, "xbg_sur_create: upload"));
    k_f32_to_bf16<<<nblk(off), 256, 0, s->st>>>(s->P, s->P16, off);
    if (hipStreamSynchronize(s->st) != hipSuccess) return bail(sfail(nullptr, XBG_E_HIP, "xbg_sur_create: init"));
    *out = s;
    return XBG_OK;
}

size_t xbg_sur_nparams(const xbg_sur* s) { return s ? s->nparam : 0; }

int xbg_sur_get_params(xbg_sur* s, float* host) {
    if (!s || !host) return XBG_E_ARG;
    SHIP(s, hipStreamSynchronize(s->st));
    SHIP(s, hipMemcpy(host, s->P, s->nparam * 4, hipMemcpyDeviceToHost));
    return XBG_OK;
}

int xbg_sur_set_params(xbg_sur* s, const float* host) {
    if (!s || !host) return XBG_E_ARG;
    SHIP(s, hipStreamSynchronize(s->st));
    SHIP(s, hipMemcpy(s->P, host, s->nparam * 4, hipMemcpyHostToDevice));
    k_f32_to_bf16<<<nblk(s->nparam), 256, 0, s->st>>>(s->P, s->P16, s->nparam);
    SHIP(s, hipMemsetAsync(s->M, 0, s->nparam * 4, s->st));
    SHIP(s, hipMemsetAsync(s->V, 0, s->nparam * 4, s->st));
    s->adam_t = 0;
    SHIP(s, hipStreamSynchronize(s->st));
    return XBG_OK;
}

int xbg_sur_get_grads(xbg_sur* s, float* host) {
    if (!s || !host) return XBG_E_ARG;
    SHIP(s, hipStreamSynchronize(s->st));
    SHIP(s, hipMemcpy(host, s->G, s->nparam * 4, hipMemcpyDeviceToHost));
    return XBG_OK;
}

}  // extern "C"

namespace {

int sur_forward(xbg_sur* s, const bf16* x8, int n) {
    hipStream_t st = s->st;
    const int H = s->H, W = s->W;
    const size_t P0 = (size_t)n * H * W, P1 = P0 / 4, P2 = P0 / 16, P3 = P0 / 64;
    auto Wt = [&](int l) { return s->P16 + s->L[l].w; };
    auto Bs = [&](int l) { return s->P + s->L[l].b; };
    int rc = 0;
    // the 8 input channels into the first 8 of 32 (channels 8..31 stay the
    // zeros written at creation)
    k_chan_copy<<<nblk(P0 * 8), 256, 0, st>>>(x8, 8, 0, s->x32, 32, 0, 8, P0, 0);
#define CONV(l, X, Y, h, w)                                                                                    \
    if ((rc = conv_fwd(st, X, Wt(l), Bs(l), Y, n, h, w, s->L[l].cin, s->L[l].cout, s->L[l].taps, s->L[l].relu, \
                       0)))                                                                                    \
        return sfail(s, rc, "conv " #l);
    CONV(0, s->x32, s->e0a, H, W);
    CONV(1, s->e0a, s->e0, H, W);
    k_pool_fwd<<<nblk(P1 * 32), 256, 0, st>>>(s->e0, s->p1, n, H, W, 32);
    CONV(2, s->p1, s->e1a, H / 2, W / 2);
    CONV(3, s->e1a, s->e1, H / 2, W / 2);
    k_pool_fwd<<<nblk(P2 * 64), 256, 0, st>>>(s->e1, s->p2, n, H / 2, W / 2, 64);
    CONV(4, s->p2, s->e2a, H / 4, W / 4);
    CONV(5, s->e2a, s->e2, H / 4, W / 4);
    k_pool_fwd<<<nblk(P3 * 128), 256, 0, st>>>(s->e2, s->p3, n, H / 4, W / 4, 128);
    CONV(6, s->p3, s->b3a, H / 8, W / 8);
    CONV(7, s->b3a, s->b3, H / 8, W / 8);
    CONV(8, s->b3, s->u2, H / 8, W / 8);
    k_up_fwd<<<nblk(P2 * 128), 256, 0, st>>>(s->u2, s->c2, n, H / 8, W / 8, 128, 256, 0);
    k_chan_copy<<<nblk(P2 * 128), 256, 0, st>>>(s->e2, 128, 0, s->c2, 256, 128, 128, P2, 0);
    CONV(9, s->c2, s->d2a, H / 4, W / 4);
    CONV(10, s->d2a, s->d2, H / 4, W / 4);
    CONV(11, s->d2, s->u1, H / 4, W / 4);
    k_up_fwd<<<nblk(P1 * 64), 256, 0, st>>>(s->u1, s->c1, n, H / 4, W / 4, 64, 128, 0);
    k_chan_copy<<<nblk(P1 * 64), 256, 0, st>>>(s->e1, 64, 0, s->c1, 128, 64, 64, P1, 0);
    CONV(12, s->c1, s->d1a, H / 2, W / 2);
    CONV(13, s->d1a, s->d1, H / 2, W / 2);
    CONV(14, s->d1, s->u0, H / 2, W / 2);
    k_up_fwd<<<nblk(P0 * 32), 256, 0, st>>>(s->u0, s->c0, n, H / 2, W / 2, 32, 64, 0);
    k_chan_copy<<<nblk(P0 * 32), 256, 0, st>>>(s->e0, 32, 0, s->c0, 64, 32, 32, P0, 0);
    CONV(15, s->c0, s->d0a, H, W);
    CONV(16, s->d0a, s->d0, H, W);
#undef CONV
    const ConvL& hd = s->L[17];
    k_head_fwd<5><<<nblk(P0), 256, 0, st>>>(s->d0, s->P + hd.w, s->P + hd.b, s->y, P0, 32);
    SHIP(s, hipGetLastError());
    return 0;
}

// one convolution's backward: g (gradient of its output, overwritten by
// the ReLU mask), dW, db, and dX (NULL: not needed)
int conv_bwd(xbg_sur* s, int l, const bf16* X, const bf16* Y, bf16* g, bf16* dX, int n, int H, int W) {
    hipStream_t st = s->st;
    const ConvL& c = s->L[l];
    const size_t P = (size_t)n * H * W;
    int rc;
    if (c.relu) k_relu_bwd<<<nblk(P * c.cout), 256, 0, st>>>(g, Y, P * c.cout);
    if ((rc = conv_wgrad(st, X, g, s->G + c.w, s->part, s->npart, n, H, W, c.cin, c.cout, c.taps)))
        return sfail(s, rc, "wgrad " + std::to_string(l));
    if ((rc = colsum(st, g, P, c.cout, s->G + c.b, s->part, s->npart))) return sfail(s, rc, "colsum");
    if (dX && (rc = conv_dgrad(st, g, s->P16 + c.w, s->Wd, dX, n, H, W, c.cin, c.cout, c.taps)))
        return sfail(s, rc, "dgrad " + std::to_string(l));
    return 0;
}

int sur_loss_grad(xbg_sur* s, const bf16* t5, int n, double* loss) {
    hipStream_t st = s->st;
    const size_t P0 = (size_t)n * s->H * s->W, ne = P0 * 5;
    const unsigned nb = nblk(ne);
    k_mse<<<nb, 256, 0, st>>>(s->y, t5, ne, s->gy, s->lossp);
    std::vector<float> h(nb);
    SHIP(s, hipMemcpyAsync(h.data(), s->lossp, nb * 4, hipMemcpyDeviceToHost, st));
    SHIP(s, hipStreamSynchronize(st));
    double a = 0.0;
    for (float v : h) a += v;
    if (loss) *loss = a / (double)ne;
    return 0;
}

int sur_backward(xbg_sur* s, int n) {
    hipStream_t st = s->st;
    const int H = s->H, W = s->W;
    const size_t P0 = (size_t)n * H * W, P1 = P0 / 4, P2 = P0 / 16, P3 = P0 / 64;
    int rc;
    const ConvL& hd = s->L[17];
    {
        int nb = 1024;
        int ppb = (int)((P0 + nb - 1) / nb);
        nb = (int)((P0 + ppb - 1) / ppb);
        k_head_wgrad<5><<<nb, 256, 0, st>>>(s->gy, s->d0, P0, 32, ppb, s->part);
        k_reduce_parts<<<nblk(5 * 32 + 5), 256, 0, st>>>(s->part, nb, (size_t)(5 * 32 + 5), s->G + hd.w);
        k_head_dgrad<5><<<nblk(P0 * 32), 256, 0, st>>>(s->gy, s->P + hd.w, s->g_d0, P0, 32);
    }
    if ((rc = conv_bwd(s, 16, s->d0a, s->d0, s->g_d0, s->g_d0a, n, H, W))) return rc;
    if ((rc = conv_bwd(s, 15, s->c0, s->d0a, s->g_d0a, s->g_c0, n, H, W))) return rc;
    k_up_bwd<<<nblk(P1 * 32), 256, 0, st>>>(s->g_c0, s->g_u0, n, H / 2, W / 2, 32, 64, 0);
    if ((rc = conv_bwd(s, 14, s->d1, s->u0, s->g_u0, s->g_d1, n, H / 2, W / 2))) return rc;
    if ((rc = conv_bwd(s, 13, s->d1a, s->d1, s->g_d1, s->g_d1a, n, H / 2, W / 2))) return rc;
    if ((rc = conv_bwd(s, 12, s->c1, s->d1a, s->g_d1a, s->g_c1, n, H / 2, W / 2))) return rc;
    k_up_bwd<<<nblk(P2 * 64), 256, 0, st>>>(s->g_c1, s->g_u1, n, H / 4, W / 4, 64, 128, 0);
    if ((rc = conv_bwd(s, 11, s->d2, s->u1, s->g_u1, s->g_d2, n, H / 4, W / 4))) return rc;
    if ((rc = conv_bwd(s, 10, s->d2a, s->d2, s->g_d2, s->g_d2a, n, H / 4, W / 4))) return rc;
    if ((rc = conv_bwd(s, 9, s->c2, s->d2a, s->g_d2a, s->g_c2, n, H / 4, W / 4))) return rc;
    k_up_bwd<<<nblk(P3 * 128), 256, 0, st>>>(s->g_c2, s->g_u2, n, H / 8, W / 8, 128, 256, 0);
    if ((rc = conv_bwd(s, 8, s->b3, s->u2, s->g_u2, s->g_b3, n, H / 8, W / 8))) return rc;
    if ((rc = conv_bwd(s, 7, s->b3a, s->b3, s->g_b3, s->g_b3a, n, H / 8, W / 8))) return rc;
    if ((rc = conv_bwd(s, 6, s->p3, s->b3a, s->g_b3a, s->g_p3, n, H / 8, W / 8))) return rc;
    // e2 feeds the pool and the skip: its gradient is the sum of both
    k_chan_copy<<<nblk(P2 * 128), 256, 0, st>>>(s->g_c2, 256, 128, s->g_e2, 128, 0, 128, P2, 0);
    k_pool_bwd<<<nblk(P2 * 128), 256, 0, st>>>(s->g_p3, s->g_e2, n, H / 4, W / 4, 128, 1);
    if ((rc = conv_bwd(s, 5, s->e2a, s->e2, s->g_e2, s->g_e2a, n, H / 4, W / 4))) return rc;
    if ((rc = conv_bwd(s, 4, s->p2, s->e2a, s->g_e2a, s->g_p2, n, H / 4, W / 4))) return rc;
    k_chan_copy<<<nblk(P1 * 64), 256, 0, st>>>(s->g_c1, 128, 64, s->g_e1, 64, 0, 64, P1, 0);
    k_pool_bwd<<<nblk(P1 * 64), 256, 0, st>>>(s->g_p2, s->g_e1, n, H / 2, W / 2, 64, 1);
    if ((rc = conv_bwd(s, 3, s->e1a, s->e1, s->g_e1, s->g_e1a, n, H / 2, W / 2))) return rc;
    if ((rc = conv_bwd(s, 2, s->p1, s->e1a, s->g_e1a, s->g_p1, n, H / 2, W / 2))) return rc;
    k_chan_copy<<<nblk(P0 * 32), 256, 0, st>>>(s->g_c0, 64, 32, s->g_e0, 32, 0, 32, P0, 0);
    k_pool_bwd<<<nblk(P0 * 32), 256, 0, st>>>(s->g_p1, s->g_e0, n, H, W, 32, 1);
    if ((rc = conv_bwd(s, 1, s->e0a, s->e0, s->g_e0, s->g_e0a, n, H, W))) return rc;
    if ((rc = conv_bwd(s, 0, s->x32, s->e0a, s->g_e0a, nullptr, n, H, W))) return rc;
    SHIP(s, hipGetLastError());
    return 0;
}

}  // namespace

extern "C" {

int xbg_sur_forward(xbg_sur* s, const void* x8, int n, float* y) {
    if (!s || !x8 || n < 1 || n > s->nmax) return sfail(s, XBG_E_ARG, "xbg_sur_forward: bad argument");
    int rc = sur_forward(s, (const bf16*)x8, n);
    if (rc) return rc;
    if (y) SHIP(s, hipMemcpyAsync(y, s->y, (size_t)n * s->H * s->W * 5 * 4, hipMemcpyDeviceToDevice, s->st));
    SHIP(s, hipStreamSynchronize(s->st));
    return XBG_OK;
}

int xbg_sur_loss(xbg_sur* s, const void* x8, const void* t5, int n, double* loss) {
    if (!s || !x8 || !t5 || n < 1 || n > s->nmax) return sfail(s, XBG_E_ARG, "xbg_sur_loss: bad argument");
    int rc = sur_forward(s, (const bf16*)x8, n);
    if (!rc) rc = sur_loss_grad(s, (const bf16*)t5, n, loss);
    return rc;
}

int xbg_sur_backward(xbg_sur* s, const void* x8, const void* t5, int n, double* loss) {
    int rc = xbg_sur_loss(s, x8, t5, n, loss);
    if (!rc) rc = sur_backward(s, n);
    if (!rc) SHIP(s, hipStreamSynchronize(s->st));
    return rc;
}

int xbg_sur_train_step(xbg_sur* s, const void* x8, const void* t5, int n, double lr, double* loss) {
    int rc = xbg_sur_loss(s, x8, t5, n, loss);
    if (!rc) rc = sur_backward(s, n);
    if (rc) return rc;
    s->adam_t++;
    const float b1 = 0.9f, b2 = 0.999f;
    const float bc1 = 1.0f - std::pow(b1, (float)s->adam_t), bc2 = 1.0f - std::pow(b2, (float)s->adam_t);
    k_adam<<<nblk(s->nparam), 256, 0, s->st>>>(s->P, s->P16, s->G, s->M, s->V, s->nparam, (float)lr, b1, b2, 1e-8f,
                                               bc1, bc2);
    SHIP(s, hipGetLastError());
    return XBG_OK;
}

int xbg_sur_sync(xbg_sur* s) {
    if (!s) return XBG_E_ARG;
    SHIP(s, hipStreamSynchronize(s->st));
    return XBG_OK;
}

// Layer-level entry points on caller device buffers (tests, benchmarks).
int xbg_sur_conv_fwd(const void* X, const void* Wt, const float* bias, void* Y, int n, int H, int W, int Cin,
                     int Cout, int taps, int relu) {
    int rc = conv_fwd(nullptr, (const bf16*)X, (const bf16*)Wt, bias, Y, n, H, W, Cin, Cout, taps, relu, 0);
    if (rc) return sfail(nullptr, rc, "xbg_sur_conv_fwd: unsupported shape");
    if (hipDeviceSynchronize() != hipSuccess) return sfail(nullptr, XBG_E_HIP, "xbg_sur_conv_fwd: run");
    return XBG_OK;
}

int xbg_sur_conv_dgrad(const void* dZ, const void* Wt, void* dX, int n, int H, int W, int Cin, int Cout, int taps) {
    bf16* Wd = nullptr;
    if (hipMalloc((void**)&Wd, (size_t)taps * Cin * Cout * 2) != hipSuccess) return XBG_E_HIP;
    int rc = conv_dgrad(nullptr, (const bf16*)dZ, (const bf16*)Wt, Wd, (bf16*)dX, n, H, W, Cin, Cout, taps);
    hipError_t e = hipDeviceSynchronize();
    hipFree(Wd);
    if (rc) return sfail(nullptr, rc, "xbg_sur_conv_dgrad: unsupported shape");
    return e == hipSuccess ? XBG_OK : XBG_E_HIP;
}

int xbg_sur_conv_wgrad(const void* X, const void* dZ, float* dW, float* db, int n, int H, int W, int Cin, int Cout,
                       int taps) {
    size_t np = (size_t)16 << 20;
    float* part = nullptr;
    if (hipMalloc((void**)&part, np * 4) != hipSuccess) return XBG_E_HIP;
    int rc = conv_wgrad(nullptr, (const bf16*)X, (const bf16*)dZ, dW, part, np, n, H, W, Cin, Cout, taps);
    if (!rc && db) rc = colsum(nullptr, (const bf16*)dZ, (size_t)n * H * W, Cout, db, part, np);
    hipError_t e = hipDeviceSynchronize();
    hipFree(part);
    if (rc) return sfail(nullptr, rc, "xbg_sur_conv_wgrad: unsupported shape");
    return e == hipSuccess ? XBG_OK : XBG_E_HIP;
}

}  // extern "C"

namespace {

struct InNorm {
    float mean[7], std[7];  // zb, zs, uu, vv, Hm0, Tp, dir
};

// Surrogate inputs of one domain: x8[i] = the normalised zb, zs, uu, vv of
// cell i and the offshore Hm0, Tp, mean direction, then a zero.
__global__ __launch_bounds__(256) void k_pack_inputs(const float* __restrict__ zb, const float* __restrict__ zs,
                                                     const float* __restrict__ uu, const float* __restrict__ vv,
                                                     size_t n, float Hm0, float Tp, float dir, InNorm nm,
                                                     bf16* __restrict__ x8) {
    size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const float v[7] = {zb[i], zs[i], uu[i], vv[i], Hm0, Tp, dir};
    bf16x8 o;
#pragma unroll
    for (int k = 0; k < 7; k++) o[k] = to_bf16((v[k] - nm.mean[k]) / nm.std[k]);
    o[7] = to_bf16(0.0f);
    *(bf16x8*)(x8 + i * 8) = o;
}

}  // namespace

extern "C" int xbg_sur_pack_inputs(const float* zb, const float* zs, const float* uu, const float* vv, size_t n,
                                   double Hm0, double Tp, double dir, const float* norm14, void* x8, void* stream) {
    if (!zb || !zs || !uu || !vv || !norm14 || !x8) return sfail(nullptr, XBG_E_ARG, "xbg_sur_pack_inputs: NULL");
    InNorm nm;
    for (int k = 0; k < 7; k++) {
        nm.mean[k] = norm14[k];
        nm.std[k] = norm14[7 + k];
    }
    // on the stream that produced the fields (a domain's stream is
    // non-blocking: the null stream would not wait for its window means)
    hipStream_t st = (hipStream_t)stream;
    k_pack_inputs<<<nblk(n), 256, 0, st>>>(zb, zs, uu, vv, n, (float)Hm0, (float)Tp, (float)dir, nm, (bf16*)x8);
    if (hipGetLastError() != hipSuccess || hipStreamSynchronize(st) != hipSuccess)
        return sfail(nullptr, XBG_E_HIP, "xbg_sur_pack_inputs: run");
    return XBG_OK;
}
