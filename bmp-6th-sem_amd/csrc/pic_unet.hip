// pic_unet.hip — learned initial guess for the Poisson solve.
//
// A 4-level U-Net over the N x N node image (3x3 convolutions, 16/32/64/64
// channels, ReLU, 2x2 average pooling, nearest upsampling with skip
// concatenation, a 1x1 head) maps (rho_n, phi_{n-1}) to the change
// phi_n - phi_{n-1}.  Weights and activations are BF16, every convolution
// accumulates in FP32 on the MFMA units (v_mfma_f32_16x16x32_bf16 as an
// implicit GEMM: M = 16 output channels, N = 16 pixels, K = two taps x 16
// input channels; pooling, upsampling and the head are fused into the
// convolutions).  Training is online: forward, backward (input gradients by
// the same kernel with flipped, transposed weights; weight gradients by MFMA
// with K = 32 pixels, all nine taps per wave) and Adam with FP32 master
// weights.  Every
// gradient and loss reduction writes per-wave or per-workgroup partials that
// k_sum_slots adds in a fixed order (no floating-point atomics), so training
// is reproducible bit for bit.
//
// Activations are NHWC with one zero pixel of padding on every side, so the
// 3x3 taps need no bounds checks; gradients are FP32 in the same layout.
#include "pic_internal.h"
#include "pic_unet.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

namespace pic {

typedef uint16_t bf16;
typedef short s4v __attribute__((ext_vector_type(4)));
typedef float f4v __attribute__((ext_vector_type(4)));

__device__ __forceinline__ float bf2f(bf16 v) { return __uint_as_float((uint32_t)v << 16); }
// round to nearest even (finite inputs)
__device__ __forceinline__ bf16 f2bf(float f) {
    uint32_t u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (bf16)(u >> 16);
}
static inline bf16 f2bf_host(float f) {
    uint32_t u;
    std::memcpy(&u, &f, 4);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (bf16)(u >> 16);
}

static constexpr int kT = 256;
static inline unsigned nb(int64_t n) { return (unsigned)((n + kT - 1) / kT); }

// ------------------------------------------------------------- MFMA probe
// D = A B for 16 x 16 operands given row-major in `a`, `b` (as float, exact in
// BF16); D row-major in `d`.  Lane l holds A[l%16][4(l/16)+e], B[4(l/16)+e][l%16]
// and receives D[4(l/16)+e][l%16].
__global__ void k_mfma_probe(const float* a, const float* b, float* d) {
    const int l = threadIdx.x, i = l & 15, kq = l >> 4;
    s4v av, bv;
    for (int e = 0; e < 4; ++e) {
        av[e] = (short)f2bf(a[i * 16 + 4 * kq + e]);
        bv[e] = (short)f2bf(b[(4 * kq + e) * 16 + i]);
    }
    f4v acc = {0.f, 0.f, 0.f, 0.f};
    acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(av, bv, acc, 0, 0, 0);
    for (int e = 0; e < 4; ++e) d[(4 * kq + e) * 16 + i] = acc[e];
}

// ------------------------------------------------ forward convolution (x32)
// The forward pass's 3x3 convolutions on v_mfma_f32_16x16x32_bf16 with the
// roles of the operands chosen for the epilogue: A = weights (M = 16 output
// channels), B = input (N = 16 pixels: 2 rows x 8 columns), K = 32 = two taps
// x 16 input channels (taps (0,1), (2,3), (4,5), (6,7), (8, none): five K
// steps per 16 channels).  Lane (i, q) (i = lane % 16, q = lane / 16) holds
// the 8 K values 8q .. 8q + 7 of its A row and B column, i.e. tap 2j + q/2 and
// channels 8(q%2) .. 8(q%2) + 7 of K step j, and receives D[4q + e][i]: four
// consecutive output channels of one pixel, stored as one 8-byte word.
//
// A workgroup (4 waves) computes a tile of 4 rows x 64 pixels for NT*16
// output channels; wave w takes rows 2(w/2), 2(w/2) + 1 and columns
// 32(w%2) .. + 31 as four pixel tiles.  Per fill of up to 32 input channels
// the padded input tile (6 rows x 66 pixels) goes to LDS as 16-byte channel
// groups in [row][group][pixel] order with a row stride of 8 (mod 16) slots,
// so the 16 lanes of a quarter-wave (8 pixels of one row, 8 of the next) read
// 16 distinct slots of every bank set; the weights go to LDS in the lane order
// of the A operand.  Epilogue: bias, ReLU, BF16 rounding, then any of: the
// output (8 bytes per lane), the 2x2 average pool of the rounded values
// (lanes i, i^1, i^8, i^9 hold a pooling window; (a + b) + (c + d) in FP32,
// a, b the upper pair), and the 1x1 head (cout = 16: the pixel's 16 channels gathered in
// one lane and summed in channel order onto the head bias).  The input may be the
// concatenation of a 2x nearest upsampling of U (first cu channels) and X, read
// in place (no upsampled tensor).
typedef int i4v __attribute__((ext_vector_type(4)));
typedef __bf16 bf8v __attribute__((ext_vector_type(8)));

static constexpr int kFTH = 4;    // tile rows
static constexpr int kFPX = 66;   // tile columns with the halo
static constexpr int kFCK = 32;   // input channels per LDS fill

struct ConvFwdArgs {
    const bf16* X;                 // input (padded NHWC); with U: the channels after the first cu
    int xC, xo;
    const bf16* U;                 // null, or the (H/2 x W/2) tensor whose upsampling gives channels 0 .. cu-1
    int uC, uo, cu;
    int cin;                       // input channels (multiple of 16)
    const bf16* Wt;                // [9][coutT][cin]
    const float* bias;
    int coutT;
    bf16* Y;                       // output (padded NHWC), or null
    int yC, yo;
    bf16* P;                       // 2x2 average-pooled output (H/2 x W/2, padded NHWC), or null
    int pC, po;
    const float* wh;               // 1x1 head (16 weights, bias bh) into out[H][W] (FP32), or out null
    const float* bh;
    float* out;
    float* Yf;                     // FP32 output without bias or ReLU (input gradients), padded NHWC
    int H, W;
};

// LDS slots (16 bytes) per tile row for g channel groups: g * 66 rounded up to 8 (mod 16)
__host__ __device__ constexpr int fwd_row_stride(int g) { return g * kFPX + ((8 - (g * kFPX) % 16) + 16) % 16; }

__device__ __forceinline__ int fwd_tile(int H, int W, int& y0, int& x0) {
    const int nxt = (W + 63) >> 6;
    const int ntile = (H / kFTH) * nxt;
    int tile = (int)blockIdx.x;
    // tiles that share halo rows go to one XCD (workgroups are dealt to the 8 XCDs in turn)
    if ((ntile & 7) == 0) tile = (tile & 7) * (ntile >> 3) + (tile >> 3);
    y0 = (tile / nxt) * kFTH;
    x0 = (tile % nxt) * 64;
    return tile;
}

// hardware round-to-nearest-even conversion (v_cvt_pk_bf16_f32)
__device__ __forceinline__ bf16 f2bf_hw(float f) { return __builtin_bit_cast(bf16, (__bf16)f); }
// lane l receives v of lane l^1 (quad_perm [1,0,3,2]) / of lane l^8 (rotate a row of 16 by 8)
__device__ __forceinline__ float dpp_xor1(float v) {
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0xB1, 0xF, 0xF, false));
}
__device__ __forceinline__ float dpp_xor8(float v) {
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x128, 0xF, 0xF, false));
}

// Epilogue of one pixel tile m and channel tile t: v = the FP32 sums of lane
// (i, q) for channels co .. co+3 of pixel (y, x).  EPI 0: output only; 1:
// output and pooled output; 2: head (and the output when a.Y is set); 3: the
// sums as FP32 into a.Yf (stride a.yC, first channel a.yo), no bias or ReLU.
template <int EPI>
__device__ __forceinline__ void fwd_epilogue(const ConvFwdArgs& a, f4v acc, int co, int y, int x, int lane) {
    if constexpr (EPI == 3) {
        if (x < a.W)
            *(float4*)(a.Yf + ((size_t)(y + 1) * (a.W + 2) + x + 1) * a.yC + a.yo + co) =
                make_float4(acc[0], acc[1], acc[2], acc[3]);
        return;
    }
    const float4 bb = *(const float4*)(a.bias + co);
    float v[4] = {acc[0] + bb.x, acc[1] + bb.y, acc[2] + bb.z, acc[3] + bb.w};
    uint2 packed;
    bf16* h = (bf16*)&packed;
#pragma unroll
    for (int e = 0; e < 4; ++e) {
        h[e] = f2bf_hw(fmaxf(v[e], 0.f));
        v[e] = bf2f(h[e]);
    }
    const bool in = x < a.W;
    const int Wp = a.W + 2;
    if ((EPI != 2 || a.Y) && in) *(uint2*)(a.Y + ((size_t)(y + 1) * Wp + x + 1) * a.yC + a.yo + co) = packed;
    if (EPI == 1) {
        uint2 pk;
        bf16* ph = (bf16*)&pk;
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const float r = v[e] + dpp_xor1(v[e]);
            const float s = r + dpp_xor8(r);
            ph[e] = f2bf_hw(0.25f * s);
        }
        if (in && (lane & 9) == 0)
            *(uint2*)(a.P + ((size_t)(y / 2 + 1) * (a.W / 2 + 2) + x / 2 + 1) * a.pC + a.po + co) = pk;
    }
    if (EPI == 2) {
        // head: each lane's four channels, then the sum over the four lanes of the pixel
        // (lanes i, i+16, i+32, i+48) as ((q0 + q1) + (q2 + q3)) plus the bias
        float s = a.wh[co] * v[0];
#pragma unroll
        for (int e = 1; e < 4; ++e) s += a.wh[co + e] * v[e];
        s += __shfl_xor(s, 16, 64);
        s += __shfl_xor(s, 32, 64);
        if (in && lane < 16) a.out[(size_t)y * a.W + x] = s + a.bh[0];
    }
}

// LDS slot offset of tap t relative to the centre tap (row stride rs)
__host__ __device__ constexpr int tap_off(int t, int rs) { return t > 8 ? 0 : (t / 3 - 1) * rs + (t % 3 - 1); }

// CK: input channels per LDS fill (16 or 32); cin must be a multiple of CK.
template <int CK, int NT, int EPI>
__global__ __launch_bounds__(256, CK == 16 ? 8 : 4) void k_conv_fwd(ConvFwdArgs a) {
    constexpr int NG = CK / 8;                 // 16-byte channel groups per fill
    constexpr int RS = fwd_row_stride(NG);
    constexpr int NW = (CK / 16) * 5 * NT * 64;
    __shared__ i4v xs[(kFTH + 2) * RS];
    __shared__ i4v ws[NW];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int i = lane & 15, q = lane >> 4;
    int y0, x0;
    fwd_tile(a.H, a.W, y0, x0);
    const int cob = (int)blockIdx.y * NT * 16;
    const int Wp = a.W + 2;
    const int lrow = 2 * (wv >> 1) + (i >> 3), lcol = 32 * (wv & 1) + (i & 7);
    const int tq = q >> 1;               // which tap of the pair
    const int lbase = (lrow + 1) * RS + (q & 1) * kFPX + lcol + 1;
    f4v acc[4][NT];
#pragma unroll
    for (int m = 0; m < 4; ++m)
#pragma unroll
        for (int t = 0; t < NT; ++t) acc[m][t] = (f4v){0.f, 0.f, 0.f, 0.f};
    for (int c0 = 0; c0 < a.cin; c0 += CK) {
        __syncthreads();
        // padded rows y0 .. y0+5, padded columns x0 .. x0+65 of channels c0 .. c0+CK-1: the tile's
        // own columns x0+1 .. x0+64 (thread: column x0+1+p, one (row, group) per wave and pass), then
        // the two halo columns
        {
            const int p = threadIdx.x & 63;
            const bool colin = x0 + p < a.W;          // an image column
            const bool colpad = x0 + p <= a.W;        // an image column or the right halo column
            const bf16* xb = a.X + (size_t)(x0 + 1 + p) * a.xC + a.xo + c0 - a.cu;
            const bf16* ub = a.U ? a.U + (size_t)(((x0 + p) >> 1) + 1) * a.uC + a.uo + c0 : nullptr;
#pragma unroll
            for (int k = 0; k < (kFTH + 2) * NG / 4; ++k) {
                const int rg = 4 * k + (int)(threadIdx.x >> 6);
                const int r = rg / NG, g = rg % NG;
                const int Yp = y0 + r;
                i4v v = {0, 0, 0, 0};
                if (c0 + 8 * g < a.cu) {
                    if (colin && Yp >= 1 && Yp <= a.H)
                        v = *(const i4v*)(ub + (size_t)(((Yp - 1) >> 1) + 1) * ((a.W >> 1) + 2) * a.uC + 8 * g);
                } else if (colpad) {
                    v = *(const i4v*)(xb + (size_t)Yp * Wp * a.xC + 8 * g);
                }
                xs[r * RS + g * kFPX + 1 + p] = v;
            }
            if (threadIdx.x < (kFTH + 2) * NG * 2) {
                const int h = threadIdx.x & 1, rg = threadIdx.x >> 1;
                const int r = rg / NG, g = rg % NG;
                const int Yp = y0 + r, Xp = x0 + 65 * h;
                const int ch = c0 + 8 * g;
                i4v v = {0, 0, 0, 0};
                if (Xp <= a.W + 1) {
                    if (ch < a.cu) {
                        if (Yp >= 1 && Yp <= a.H && Xp >= 1 && Xp <= a.W)
                            v = *(const i4v*)(a.U + ((size_t)(((Yp - 1) >> 1) + 1) * ((a.W >> 1) + 2) +
                                                     ((Xp - 1) >> 1) + 1) * a.uC + a.uo + ch);
                    } else {
                        v = *(const i4v*)(a.X + ((size_t)Yp * Wp + Xp) * a.xC + a.xo + ch - a.cu);
                    }
                }
                xs[r * RS + g * kFPX + 65 * h] = v;
            }
        }
#pragma unroll
        for (int e0 = 0; e0 < NW; e0 += 256) {
            const int e = e0 + (int)threadIdx.x;
            if (NW % 256 == 0 || e < NW) {
                const int l = e & 63;
                const int t = (e >> 6) % NT, j = (e >> 6) / NT % 5, s = (e >> 6) / NT / 5;
                const int tap = 2 * j + ((l >> 4) >> 1);
                i4v v = {0, 0, 0, 0};
                if (tap < 9)
                    v = *(const i4v*)(a.Wt + ((size_t)tap * a.coutT + cob + 16 * t + (l & 15)) * a.cin + c0 + 16 * s +
                                      8 * ((l >> 4) & 1));
                ws[e] = v;
            }
        }
        __syncthreads();
#pragma unroll
        for (int s = 0; s < CK / 16; ++s) {
#pragma unroll
            for (int j = 0; j < 5; ++j) {
                const int base = lbase + 2 * s * kFPX + (tq ? tap_off(2 * j + 1, RS) : tap_off(2 * j, RS));
                bf8v wa[NT];
#pragma unroll
                for (int t = 0; t < NT; ++t) wa[t] = __builtin_bit_cast(bf8v, ws[((s * 5 + j) * NT + t) * 64 + lane]);
#pragma unroll
                for (int m = 0; m < 4; ++m) {
                    i4v xv = xs[base + 8 * m];
                    if (j == 4 && tq) xv = (i4v){0, 0, 0, 0};   // the pair's missing tenth tap
                    const bf8v xb = __builtin_bit_cast(bf8v, xv);
#pragma unroll
                    for (int t = 0; t < NT; ++t)
                        acc[m][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wa[t], xb, acc[m][t], 0, 0, 0);
                }
            }
        }
    }
    const int y = y0 + lrow;
#pragma unroll
    for (int t = 0; t < NT; ++t)
#pragma unroll
        for (int m = 0; m < 4; ++m) fwd_epilogue<EPI>(a, acc[m][t], cob + 16 * t + 4 * q, y, x0 + lcol + 8 * m, lane);
}

// The first layer: 2 input channels (X = the compact input, 2 channels per
// pixel, padded), 16 output channels; K = 9 taps x 2 channels in one K step
// of 32 (lane (i, q) holds taps 4q .. 4q+3, both channels; K 18 .. 31 zero).
// Weights from Wt [9][16][cin] (channels 0 and 1 used).
__global__ __launch_bounds__(256) void k_conv0_fwd(ConvFwdArgs a) {
    constexpr int RS0 = 72;   // 32-bit words per tile row (== 8 mod 64: two rows of 8 pixels hit distinct banks)
    __shared__ uint32_t xs[(kFTH + 2) * RS0];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int i = lane & 15, q = lane >> 4;
    // A: W[co = i][k = 8q + e] = Wt[tap = 4q + e/2][i][e % 2], two channels per 32-bit load
    uint32_t w4[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int tap = 4 * q + k;
        w4[k] = tap < 9 ? *(const uint32_t*)(a.Wt + ((size_t)tap * a.coutT + i) * a.cin) : 0u;
    }
    int y0, x0;
    fwd_tile(a.H, a.W, y0, x0);
    const int Wp = a.W + 2;
    for (int e = threadIdx.x; e < (kFTH + 2) * kFPX; e += 256) {
        const int r = e / kFPX, p = e - r * kFPX;
        const int Xp = x0 + p;
        xs[r * RS0 + p] = Xp <= a.W + 1 ? *(const uint32_t*)(a.X + ((size_t)(y0 + r) * Wp + Xp) * 2) : 0u;
    }
    const bf8v wa = __builtin_bit_cast(bf8v, *(const i4v*)w4);
    __syncthreads();
    const int lrow = 2 * (wv >> 1) + (i >> 3), lcol = 32 * (wv & 1) + (i & 7);
    const int y = y0 + lrow;
    const int lbase = (lrow + 1) * RS0 + lcol + 1;
#pragma unroll
    for (int m = 0; m < 4; ++m) {
        uint32_t x4[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            // tap 4q + k: q = 0: taps 0..3, q = 1: 4..7, q = 2: 8 and zeros, q = 3: zeros
            const int off = q == 0 ? tap_off(k, RS0) : q == 1 ? tap_off(4 + k, RS0) : tap_off(8, RS0);
            const uint32_t v = xs[lbase + off + 8 * m];
            x4[k] = (q < 2 || (q == 2 && k == 0)) ? v : 0u;
        }
        const bf8v xb = __builtin_bit_cast(bf8v, *(const i4v*)x4);
        f4v acc = {0.f, 0.f, 0.f, 0.f};
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wa, xb, acc, 0, 0, 0);
        fwd_epilogue<0>(a, acc, 4 * q, y, x0 + lcol + 8 * m, lane);
    }
}

// 16x16x32 BF16 MFMA lane check: A (16 x 32) and B (32 x 16) row-major; lane
// l holds A[l%16][8(l/16)+e] and B[8(l/16)+e][l%16] and receives
// D[4(l/16)+e][l%16].
__global__ void k_mfma32_probe(const float* a, const float* b, float* d) {
    const int l = threadIdx.x, i = l & 15, q = l >> 4;
    uint16_t av[8], bv[8];
    for (int e = 0; e < 8; ++e) {
        av[e] = f2bf(a[i * 32 + 8 * q + e]);
        bv[e] = f2bf(b[(8 * q + e) * 16 + i]);
    }
    f4v acc = {0.f, 0.f, 0.f, 0.f};
    acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf8v, *(const i4v*)av),
                                                  __builtin_bit_cast(bf8v, *(const i4v*)bv), acc, 0, 0, 0);
    for (int e = 0; e < 4; ++e) d[(4 * q + e) * 16 + i] = acc[e];
}

// Partial of dW[tap][co][ci] = sum_{y,x} dZT[co][y][x] XT[ci][y+dyy][x+dxx]
// (tap = 3 dyy + dxx) over one chunk of rows, on v_mfma_f32_16x16x32_bf16
// with K = 32 pixels of a row: A = dZT (M = 16 output channels), B = the
// input row shifted by the tap (N = 16 input channels).  dZT: [cout][H][W]
// (unpadded, W a multiple of 8), XT: [cin][H+2][xs] (padded columns 0 ..
// W+1, row stride xs a multiple of 8 with at least 8 columns past W+1).  One
// workgroup: one (co, ci) tile pair (blockIdx.y), rows [r0, r0 + rows) of
// chunk blockIdx.x; wave w takes a quarter of each row's columns (in 32-pixel
// steps) and all nine taps (nine accumulators): per 32 pixels it loads its dZ
// chunk once and, for each of the three input rows, two aligned 16-byte
// chunks from which the three column shifts are formed by byte alignment
// (lane (i, q) holds pixels 8q .. 8q+7 of channel i).  The four waves' tiles
// are added in a fixed order through LDS and written to slot `chunk` of part
// (slot stride 9 cout cin); k_sum_slots adds the slots in chunk order.
__device__ __forceinline__ i4v shift_elems(i4v lo, i4v hi, int s) {
    // elements s .. s+7 of the 16 elements lo:hi (s = 0, 1, 2)
    if (s == 0) return lo;
    if (s == 2) return (i4v){lo[1], lo[2], lo[3], hi[0]};
    return (i4v){(int)__builtin_amdgcn_alignbyte((unsigned)lo[1], (unsigned)lo[0], 2),
                 (int)__builtin_amdgcn_alignbyte((unsigned)lo[2], (unsigned)lo[1], 2),
                 (int)__builtin_amdgcn_alignbyte((unsigned)lo[3], (unsigned)lo[2], 2),
                 (int)__builtin_amdgcn_alignbyte((unsigned)hi[0], (unsigned)lo[3], 2)};
}

__global__ __launch_bounds__(kT) void k_wgrad(const bf16* __restrict__ dZT, const bf16* __restrict__ XT, int cout,
                                              int cin, int H, int W, int xs, int rows, float* __restrict__ part) {
    __shared__ f4v red[3][9][64];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int chunk = blockIdx.x, pair = blockIdx.y;
    const int cit = cin >> 4;
    const int cot = pair / cit, cti = pair % cit;
    const int i = lane & 15, q = lane >> 4;
    f4v acc[9];
#pragma unroll
    for (int t = 0; t < 9; ++t) acc[t] = (f4v){0.f, 0.f, 0.f, 0.f};
    const bf16* zc = dZT + (size_t)(cot * 16 + i) * H * W + 8 * q;
    const bf16* xc = XT + (size_t)(cti * 16 + i) * (H + 2) * xs + 8 * q;
    const int r0 = chunk * rows, r1 = min(H, r0 + rows);
    const int seg = (((W + 3) / 4) + 31) / 32 * 32;   // columns per wave, in 32-pixel steps
    const int xa = wv * seg, xb = min(W, xa + seg);
    for (int y = r0; y < r1; ++y) {
        for (int px0 = xa; px0 < xb; px0 += 32) {
            const bool in = px0 + 8 * q < xb;   // W is a multiple of 8
            const i4v zv = in ? *(const i4v*)(zc + (size_t)y * W + px0) : (i4v){0, 0, 0, 0};
            const bf8v a = __builtin_bit_cast(bf8v, zv);
#pragma unroll
            for (int dyy = 0; dyy < 3; ++dyy) {
                const bf16* xr = xc + (size_t)(y + dyy) * xs + px0;
                i4v lo = {0, 0, 0, 0}, hi = {0, 0, 0, 0};
                if (in) {
                    lo = *(const i4v*)xr;
                    hi = *(const i4v*)(xr + 8);
                }
#pragma unroll
                for (int dxx = 0; dxx < 3; ++dxx)
                    acc[3 * dyy + dxx] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
                        a, __builtin_bit_cast(bf8v, shift_elems(lo, hi, dxx)), acc[3 * dyy + dxx], 0, 0, 0);
            }
        }
    }
    if (wv > 0)
#pragma unroll
        for (int t = 0; t < 9; ++t) red[wv - 1][t][lane] = acc[t];
    __syncthreads();
    if (wv > 0) return;
    float* dst = part + (size_t)chunk * 9 * cout * cin;
#pragma unroll
    for (int t = 0; t < 9; ++t) {
        const f4v v = ((acc[t] + red[0][t][lane]) + red[1][t][lane]) + red[2][t][lane];
#pragma unroll
        for (int e = 0; e < 4; ++e) dst[((size_t)t * cout + cot * 16 + 4 * q + e) * cin + cti * 16 + i] = v[e];
    }
}

// dZ = bf16(dY * (Y > 0)); dY FP32 (stride gC, first channel go), Y BF16
// (stride yC, first channel yo), dZ padded NHWC with c channels.  The whole
// padded grid is written (zero halo): the buffer is shared by layers of
// different shapes and the input-gradient convolution reads the halo.
// With it, its CHW transpose dZT [c][H][W] and the bias-gradient partials,
// in one pass.  Workgroup: 64 padded columns (blockIdx.x) of a
// chunk of kRB padded rows (blockIdx.y), 16 channels (blockIdx.z); per row
// each thread forms 4 channels of one pixel (16-byte dY and 8-byte Y loads),
// the tile is transposed through LDS, and thread t writes channel t / 16,
// pixels 4 (t % 16) .. +3 of dZT and adds them to its bias sum (rows in
// order); the 16 sums of a channel are added in a fixed order into slot
// (blockIdx.y * gridDim.x + blockIdx.x) of part (stride c).
static constexpr int kRB = 16;
__global__ __launch_bounds__(kT) void k_relu_bwd_t(const float* __restrict__ dY, int gC, int go,
                                                   const bf16* __restrict__ Y, int yC, int yo, bf16* __restrict__ dZ,
                                                   bf16* __restrict__ dZT, float* __restrict__ part, int c, int H,
                                                   int W) {
    __shared__ bf16 tile[16][68];
    __shared__ float bs[16][17];
    const int t = threadIdx.x;
    const int c0 = blockIdx.z * 16;
    const int X0 = blockIdx.x * 64;            // first padded column
    const int lp = t >> 2, lc = (t & 3) * 4;   // load: pixel, first of 4 channels
    const int wc = t >> 4, wp = (t & 15) * 4;  // transposed write: channel, first of 4 pixels
    float bsum = 0.f;
    const int y0 = blockIdx.y * kRB, y1 = min(H + 2, y0 + kRB);
    const int xx = X0 + lp;
    const bool colin = xx >= 1 && xx <= W;
    // the next row's dY and Y are loaded while the current row is transposed
    auto load = [&](int yy, float4& g, uint2& yb) {
        g = make_float4(0.f, 0.f, 0.f, 0.f);
        yb = make_uint2(0u, 0u);
        if (colin && yy >= 1 && yy <= H) {
            const size_t q = (size_t)yy * (W + 2) + xx;
            g = *(const float4*)(dY + q * gC + go + c0 + lc);
            yb = *(const uint2*)(Y + q * yC + yo + c0 + lc);
        }
    };
    float4 gn;
    uint2 ybn;
    load(y0, gn, ybn);
    for (int yy = y0; yy < y1; ++yy) {
        const float4 g = gn;
        const uint2 yb = ybn;
        if (yy + 1 < y1) load(yy + 1, gn, ybn);
        uint2 zv = make_uint2(0u, 0u);
        if (xx < W + 2) {
            const size_t q = (size_t)yy * (W + 2) + xx;
            const bf16* yv = (const bf16*)&yb;
            bf16* z = (bf16*)&zv;
            z[0] = f2bf(bf2f(yv[0]) > 0.f ? g.x : 0.f);
            z[1] = f2bf(bf2f(yv[1]) > 0.f ? g.y : 0.f);
            z[2] = f2bf(bf2f(yv[2]) > 0.f ? g.z : 0.f);
            z[3] = f2bf(bf2f(yv[3]) > 0.f ? g.w : 0.f);
            *(uint2*)(dZ + q * c + c0 + lc) = zv;
        }
        const bf16* z = (const bf16*)&zv;
#pragma unroll
        for (int k = 0; k < 4; ++k) tile[lc + k][lp] = z[k];
        __syncthreads();
        if (yy >= 1 && yy <= H) {
            // interior columns of this block: padded X0 + wp + k, image column X0 + wp + k - 1
            bf16 o[4];
            bool any = false;
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const int xx2 = X0 + wp + k;
                o[k] = tile[wc][wp + k];
                if (xx2 >= 1 && xx2 <= W) {
                    bsum += bf2f(o[k]);
                    any = true;
                }
            }
            if (any) {
                bf16* dst = dZT + ((size_t)(c0 + wc) * H + (yy - 1)) * W;
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    const int xx2 = X0 + wp + k;
                    if (xx2 >= 1 && xx2 <= W) dst[xx2 - 1] = o[k];
                }
            }
        }
        __syncthreads();
    }
    bs[wc][t & 15] = bsum;
    __syncthreads();
    if (t < 16) {
        float v = 0.f;
        for (int k = 0; k < 16; ++k) v += bs[t][k];
        part[((size_t)blockIdx.y * gridDim.x + blockIdx.x) * c + c0 + t] = v;
    }
}

// out[c][y][x] = in[(y * (W + 2) + x) * C + o + c] over the padded grid (pad = 1)
// or the interior only (pad = 0, out [c][H][W]).  Workgroup: 64 output
// columns of one output row, 16 channels (blockIdx.z); the tile goes through
// LDS so both the NHWC reads (16 bytes per pixel) and the CHW writes (64
// consecutive pixels of a channel) are contiguous.  c, o, C multiples of 8.
__global__ __launch_bounds__(kT) void k_to_chw(const bf16* __restrict__ in, int C, int o, int c, int H, int W, int pad,
                                               int os, bf16* __restrict__ out) {
    __shared__ bf16 tile[16][66];
    const int Ho = H + 2 * pad, Wo = W + 2 * pad;
    const int x0 = blockIdx.x * 64, y = blockIdx.y, c0 = blockIdx.z * 16;
    const int t = threadIdx.x;
    if (t < 128) {
        const int px = t >> 1, half = t & 1;
        const int x = x0 + px;
        int4 v = make_int4(0, 0, 0, 0);
        if (x < Wo && c0 + 8 * half < c)
            v = *(const int4*)(in + ((size_t)(y + 1 - pad) * (W + 2) + x + 1 - pad) * C + o + c0 + 8 * half);
        const bf16* b = (const bf16*)&v;
#pragma unroll
        for (int k = 0; k < 8; ++k) tile[8 * half + k][px] = b[k];
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int idx = t + kT * k;
        const int ch = idx >> 6, px = idx & 63;
        const int x = x0 + px;
        if (x < Wo && c0 + ch < c) out[((size_t)(c0 + ch) * Ho + y) * os + x] = tile[ch][px];
    }
}

// gin (H x W) += 0.25 gout (H/2 x W/2) at each of the four children.
__global__ void k_pool2_bwd(const float* __restrict__ gout, int oC, int oo, float* __restrict__ gin, int iC, int io,
                            int c, int H, int W) {
    const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= (int64_t)H * W * c) return;
    const int ch = (int)(n % c);
    const int64_t p = n / c;
    const int y = (int)(p / W), x = (int)(p - (int64_t)y * W);
    const int w = W / 2;
    const float g = gout[((size_t)(y / 2 + 1) * (w + 2) + x / 2 + 1) * oC + oo + ch];
    gin[((size_t)(y + 1) * (W + 2) + x + 1) * iC + io + ch] += 0.25f * g;
}

// nearest upsampling: out (2h x 2w) <- in (h x w)
// Eight channels (16 bytes) per thread; c, io and oo are multiples of 8.
__global__ void k_up2(const bf16* __restrict__ in, int iC, int io, bf16* __restrict__ out, int oC, int oo, int c,
                      int h, int w) {
    const int H = 2 * h, W = 2 * w, c8 = c / 8;
    const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= (int64_t)H * W * c8) return;
    const int ch = (int)(n % c8) * 8;
    const int64_t p = n / c8;
    const int y = (int)(p / W), x = (int)(p - (int64_t)y * W);
    *(int4*)(out + ((size_t)(y + 1) * (W + 2) + x + 1) * oC + oo + ch) =
        *(const int4*)(in + ((size_t)(y / 2 + 1) * (w + 2) + x / 2 + 1) * iC + io + ch);
}

// gin (h x w) = sum of gout over the four children (2h x 2w)
__global__ void k_up2_bwd(const float* __restrict__ gout, int oC, int oo, float* __restrict__ gin, int iC, int io,
                          int c, int h, int w) {
    const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= (int64_t)h * w * c) return;
    const int ch = (int)(n % c);
    const int64_t p = n / c;
    const int y = (int)(p / w), x = (int)(p - (int64_t)y * w);
    const int W = 2 * w;
    auto at = [&](int yy, int xx) { return gout[((size_t)(yy + 1) * (W + 2) + xx + 1) * oC + oo + ch]; };
    gin[((size_t)(y + 1) * (w + 2) + x + 1) * iC + io + ch] =
        (at(2 * y, 2 * x) + at(2 * y, 2 * x + 1)) + (at(2 * y + 1, 2 * x) + at(2 * y + 1, 2 * x + 1));
}

// Output gradient of the masked mean squared error against `tgt`; the block's
// share of the loss goes to lpart[blockIdx.x].
__global__ __launch_bounds__(kT) void k_loss_mse(const float* __restrict__ out, const float* __restrict__ tgt,
                                                 const uint8_t* __restrict__ mask, const int* __restrict__ count,
                                                 float* __restrict__ gout, float* __restrict__ lpart, int n) {
    __shared__ float sh[kT / 64];
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    const float inv = *count > 0 ? 1.f / (float)*count : 0.f;
    float l = 0.f;
    if (p < n) {
        float g = 0.f;
        if (mask[p]) {
            const float d = out[p] - tgt[p];
            l = d * d * inv;
            g = 2.f * d * inv;
        }
        gout[p] = g;
    }
    for (int off = 32; off > 0; off >>= 1) l += __shfl_down(l, off, 64);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = l;
    __syncthreads();
    if (threadIdx.x == 0) lpart[blockIdx.x] = (sh[0] + sh[1]) + (sh[2] + sh[3]);
}

// Residual loss: the prediction d (= out on unknown nodes, 0 elsewhere) should
// cancel the scaled starting residual, res = r0 / max|r0| - (A / diag) d on
// unknown nodes; loss = mean res^2 (the block's share to lpart[blockIdx.x]).
// (A / diag has unit diagonal and off-diagonals -cxd, -cyd.)
__global__ __launch_bounds__(kT) void k_loss_res(const float* __restrict__ out, const double* __restrict__ r0,
                                                 const unsigned long long* __restrict__ scl,
                                                 const uint8_t* __restrict__ mask, const int* __restrict__ count,
                                                 float cxd, float cyd, int N, float* __restrict__ res,
                                                 float* __restrict__ lpart) {
    __shared__ float sh[kT / 64];
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    const float inv = *count > 0 ? 1.f / (float)*count : 0.f;
    float l = 0.f;
    if (p < N * N) {
        float v = 0.f;
        if (mask[p]) {
            const int y = p / N, x = p - y * N;
            auto d = [&](int yy, int xx) -> float {
                if (yy < 0 || yy >= N || xx < 0 || xx >= N) return 0.f;
                const int q = yy * N + xx;
                return mask[q] ? out[q] : 0.f;
            };
            const double s0 = __longlong_as_double((long long)scl[0]);
            const float rs = (float)(r0[p] / (s0 > 0.0 ? s0 : 1.0));
            const float ad = d(y, x) - cxd * (d(y, x - 1) + d(y, x + 1)) - cyd * (d(y - 1, x) + d(y + 1, x));
            v = rs - ad;
            l = v * v * inv;
        }
        res[p] = v;
    }
    for (int off = 32; off > 0; off >>= 1) l += __shfl_down(l, off, 64);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = l;
    __syncthreads();
    if (threadIdx.x == 0) lpart[blockIdx.x] = (sh[0] + sh[1]) + (sh[2] + sh[3]);
}

// d loss / d out = -2 (A / diag) res / count on unknown nodes (A symmetric)
__global__ void k_loss_res_grad(const float* __restrict__ res, const uint8_t* __restrict__ mask,
                                const int* __restrict__ count, float cxd, float cyd, int N, float* __restrict__ gout) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= N * N) return;
    const float inv = *count > 0 ? 1.f / (float)*count : 0.f;
    float g = 0.f;
    if (mask[p]) {
        const int y = p / N, x = p - y * N;
        auto r = [&](int yy, int xx) -> float {
            if (yy < 0 || yy >= N || xx < 0 || xx >= N) return 0.f;
            return res[yy * N + xx];
        };
        g = -2.f * inv * (r(y, x) - cxd * (r(y, x - 1) + r(y, x + 1)) - cyd * (r(y - 1, x) + r(y + 1, x)));
    }
    gout[p] = g;
}

// Head backward from the output gradient gout: gd0[p][c] = gout wh[c]; the
// block's shares of dwh[c] = sum gout d0[c] and dbh = sum gout go to slot
// blockIdx.x of part (17 floats: dwh then dbh).  c = 16 channels.
__global__ __launch_bounds__(kT) void k_head_bwd(const float* __restrict__ gout, const bf16* __restrict__ d0,
                                                 const float* __restrict__ wh, float* __restrict__ gd0,
                                                 float* __restrict__ part, int H, int W) {
    __shared__ float sh[17][kT / 64];
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    float red[17];
#pragma unroll
    for (int k = 0; k < 17; ++k) red[k] = 0.f;
    if (p < H * W) {
        const int y = p / W, x = p - y * W;
        const size_t q = (size_t)(y + 1) * (W + 2) + x + 1;
        const float g = gout[p];
        red[16] = g;
#pragma unroll
        for (int k = 0; k < 16; ++k) {
            gd0[q * 16 + k] = g * wh[k];
            red[k] = g * bf2f(d0[q * 16 + k]);
        }
    }
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < 17; ++k) {
        float v = red[k];
        for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
        if (lane == 0) sh[k][wid] = v;
    }
    __syncthreads();
    if (threadIdx.x < 17) {
        float s = 0.f;
        for (int w = 0; w < kT / 64; ++w) s += sh[threadIdx.x][w];
        part[(size_t)blockIdx.x * 17 + threadIdx.x] = s;
    }
}

// out[j] = the sum over slots of part[slot * len + j] in a fixed order: wave w
// of a workgroup (64 outputs) adds slots w, w + 4, w + 8, ... in four
// interleaved partial sums (a, b, c, d over slots w + 16k, w + 4 + 16k, ...),
// then out = ((s0 + s1) + (s2 + s3)) over the waves.  The reductions above
// leave their partials in slots, so a gradient or loss does not depend on the
// order the workgroups ran in.
__global__ __launch_bounds__(kT) void k_sum_slots(const float* __restrict__ part, int nslot, int64_t len,
                                                  float* __restrict__ out) {
    __shared__ float sh[4][64];
    const int lx = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int64_t j = (int64_t)blockIdx.x * 64 + lx;
    const int64_t jc = j < len ? j : len - 1;
    float a = 0.f, b = 0.f, c = 0.f, d = 0.f;
    int q = w;
    for (; q + 12 < nslot; q += 16) {
        a += part[(size_t)q * len + jc];
        b += part[(size_t)(q + 4) * len + jc];
        c += part[(size_t)(q + 8) * len + jc];
        d += part[(size_t)(q + 12) * len + jc];
    }
    for (; q < nslot; q += 4) a += part[(size_t)q * len + jc];
    sh[w][lx] = (a + b) + (c + d);
    __syncthreads();
    if (w == 0 && j < len) out[j] = (sh[0][lx] + sh[1][lx]) + (sh[2][lx] + sh[3][lx]);
}
static inline unsigned sum_blocks(int64_t len) { return (unsigned)((len + 63) / 64); }

// The same for a few outputs over many slots (losses, head gradients): one
// workgroup per output j, thread t adds slots t, t + kT, ..., and the
// workgroup's sums are added by a fixed tree.
__global__ __launch_bounds__(kT) void k_sum_slots_few(const float* __restrict__ part, int nslot, int len,
                                                      float* __restrict__ out) {
    __shared__ float sh[kT];
    const int j = blockIdx.x;
    float t = 0.f;
    for (int q = threadIdx.x; q < nslot; q += kT) t += part[(size_t)q * len + j];
    sh[threadIdx.x] = t;
    __syncthreads();
    for (int w = kT / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[j] = sh[0];
}

// count[0] = number of non-zero mask entries (count zeroed by the caller's
// kernel); n a multiple of 16: each thread counts 16 bytes, one integer
// atomic per workgroup
__device__ __forceinline__ int nonzero_bytes(uint32_t x) {
    x |= x >> 4;
    x |= x >> 2;
    x |= x >> 1;
    return __popc(x & 0x01010101u);
}
__global__ __launch_bounds__(kT) void k_mask_count(const uint8_t* __restrict__ mask, int n, int* __restrict__ count) {
    __shared__ int sh[kT / 64];
    const int p = (blockIdx.x * blockDim.x + threadIdx.x) * 16;
    int v = 0;
    if (p < n) {
        const uint4 m = *(const uint4*)(mask + p);
        v = nonzero_bytes(m.x) + nonzero_bytes(m.y) + nonzero_bytes(m.z) + nonzero_bytes(m.w);
    }
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        const int t = (sh[0] + sh[1]) + (sh[2] + sh[3]);
        if (t) atomicAdd(count, t);
    }
}

__global__ void k_zero_i32(int* p) { *p = 0; }

// Adam on FP32 masters; refreshes the BF16 copy (and for conv weights the
// flipped, transposed copy used by the input-gradient convolution).
__global__ void k_adam(float* __restrict__ w, const float* __restrict__ g, float* __restrict__ m,
                       float* __restrict__ v, int n, float lr_t, float b1, float b2, float eps,
                       bf16* __restrict__ wb, bf16* __restrict__ wflip, int cout, int cin) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const float gk = g[k];
    const float mk = b1 * m[k] + (1.f - b1) * gk;
    const float vk = b2 * v[k] + (1.f - b2) * gk * gk;
    m[k] = mk;
    v[k] = vk;
    const float wk = w[k] - lr_t * mk / (sqrtf(vk) + eps);
    w[k] = wk;
    if (wb) {
        const bf16 b = f2bf(wk);
        wb[k] = b;
        if (wflip) {
            // k = (tap * cout + co) * cin + ci  ->  wflip[(8 - tap) * cin + ci][co]
            const int ci = k % cin, r = k / cin, co = r % cout, tap = r / cout;
            wflip[((size_t)(8 - tap) * cin + ci) * cout + co] = b;
        }
    }
}

// Inputs from the PIC grids, scaled per solve so the network sees O(1) values
// whatever the amplitude: channel 0 is the starting residual
// r0 = rhs - A phi_{n-1} (computed in FP64, a function of rho_n and
// phi_{n-1}) over max|r0|, channel 1 is phi_{n-1} over max|phi_{n-1}|.  The
// target and the output are the change of phi in units of max|r0| / diag,
// diag = 2(1/dx^2 + 1/dy^2), so the map the network learns does not depend on
// the solve's amplitude (the change is linear in r0).
// Each workgroup's max|r0| and max|phi| go to mpart[2 b], mpart[2 b + 1];
// k_unet_scale reduces them (one atomic per workgroup on two addresses would
// serialise: 96 us at 1024^2).
__global__ __launch_bounds__(kT) void k_unet_resid(const double* __restrict__ rhs, const double* __restrict__ phi,
                                                   const uint8_t* __restrict__ dir, int gx, int N, double cx,
                                                   double cy, double diag, double* __restrict__ r0,
                                                   double* __restrict__ mpart) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    double v = 0.0, a = 0.0;
    if (p < N * N) {
        const int y = p / N, x = p - y * N;
        const int n = y * gx + x;
        if (!dir[n])
            v = rhs[n] - (diag * phi[n] - cx * (phi[n - 1] + phi[n + 1]) - cy * (phi[n - gx] + phi[n + gx]));
        r0[p] = v;
        a = fabs(phi[n]);
    }
    double m = fabs(v);
    for (int off = 32; off > 0; off >>= 1) {
        m = fmax(m, __shfl_down(m, off, 64));
        a = fmax(a, __shfl_down(a, off, 64));
    }
    __shared__ double sm[kT / 64], sa[kT / 64];
    if ((threadIdx.x & 63) == 0) {
        sm[threadIdx.x >> 6] = m;
        sa[threadIdx.x >> 6] = a;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < kT / 64; ++w) {
            m = fmax(m, sm[w]);
            a = fmax(a, sa[w]);
        }
        mpart[2 * blockIdx.x] = m;
        mpart[2 * blockIdx.x + 1] = a;
    }
}

// scl[0] = max|r0|, scl[1] = max|phi| (as bits) over the nblk partials; one
// workgroup of kT threads.
__global__ __launch_bounds__(kT) void k_unet_scale(const double* __restrict__ mpart, int nblk,
                                                   unsigned long long* __restrict__ scl) {
    double m = 0.0, a = 0.0;
    for (int b = threadIdx.x; b < nblk; b += kT) {
        m = fmax(m, mpart[2 * b]);
        a = fmax(a, mpart[2 * b + 1]);
    }
    for (int off = 32; off > 0; off >>= 1) {
        m = fmax(m, __shfl_down(m, off, 64));
        a = fmax(a, __shfl_down(a, off, 64));
    }
    __shared__ double sm[kT / 64], sa[kT / 64];
    if ((threadIdx.x & 63) == 0) {
        sm[threadIdx.x >> 6] = m;
        sa[threadIdx.x >> 6] = a;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < kT / 64; ++w) {
            m = fmax(m, sm[w]);
            a = fmax(a, sa[w]);
        }
        scl[0] = (unsigned long long)__double_as_longlong(m);
        scl[1] = (unsigned long long)__double_as_longlong(a);
    }
}

__device__ __forceinline__ double scl_get(const unsigned long long* scl, int k) {
    const double v = __longlong_as_double((long long)scl[k]);
    return v > 0.0 ? v : 1.0;
}

// x0c: the compact 2-channel input (padded); x0 (training only, else null):
// the same two values in channels 0, 1 of the 16-channel input tensor read by
// the first layer's weight gradient.
__global__ void k_unet_input(const double* __restrict__ r0, const double* __restrict__ phi, int gx, int N,
                             const unsigned long long* __restrict__ scl, bf16* __restrict__ x0c,
                             bf16* __restrict__ x0) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= N * N) return;
    const int y = p / N, x = p - y * N;
    const double sr = 1.0 / scl_get(scl, 0), sp = 1.0 / scl_get(scl, 1);
    const size_t px = (size_t)(y + 1) * (N + 2) + x + 1;
    const bf16 a = f2bf((float)(r0[p] * sr)), b = f2bf((float)(phi[y * gx + x] * sp));
    *(uint32_t*)(x0c + px * 2) = (uint32_t)a | ((uint32_t)b << 16);
    if (x0) {
        x0[px * 16] = a;
        x0[px * 16 + 1] = b;
    }
}

// target (phi_new - phi_old) diag / max|r0| on the image; mask = not Dirichlet
__global__ void k_unet_target(const double* __restrict__ phi_new, const double* __restrict__ phi_old,
                              const uint8_t* __restrict__ dir, int gx, int N, double diag,
                              const unsigned long long* __restrict__ scl, float* __restrict__ t,
                              uint8_t* __restrict__ mask) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= N * N) return;
    const int y = p / N, x = p - y * N;
    const int n = y * gx + x;
    t[p] = (float)((phi_new[n] - phi_old[n]) * (diag / scl_get(scl, 0)));
    mask[p] = dir[n] ? 0 : 1;
}

// ================================================================= host side
namespace {

struct Tensor {
    bf16* p = nullptr;   // padded NHWC
    int H = 0, W = 0, C = 0;
    float* g = nullptr;  // gradient (same layout, FP32), when needed
    bool grad = true;    // false: forward-only tensor, no gradient buffer
};

struct Conv {
    int in, io, cin;     // input tensor index, first channel, channels (multiple of 16)
    int cin_real;        // channels carrying data (fan-in of the initialisation)
    int out, oo, cout;   // output tensor index, first channel, channels
    size_t woff, boff;   // offsets of W [9][cout][cin] and b [cout] in the parameter vector
};

struct Resample {        // pool (down) or upsample (up)
    bool up;
    int in, io, out, oo, c;
};

} // namespace
} // namespace pic

struct pic_unet {
    int N = 0;
    std::vector<pic::Tensor> t;           // activations
    std::vector<pic::Conv> conv;
    // parameter vector: all conv W and b, then head wh[16], bh
    size_t nparam = 0, head_w = 0, head_b = 0;
    float* w = nullptr;                   // FP32 masters
    float* g = nullptr;                   // gradients
    float* m = nullptr;
    float* v = nullptr;
    pic::bf16* wb = nullptr;              // BF16 copies (conv weights; biases/head unused)
    pic::bf16* wflip = nullptr;           // [9][cin][cout] flipped copies, same offsets as W
    float* out = nullptr;                 // head output N x N
    float* tgt = nullptr;                 // target (MSE) or residual buffer (residual loss)
    float* gout = nullptr;                // d loss / d output
    uint8_t* mask = nullptr;
    float* loss = nullptr;
    int* count = nullptr;
    float* loss_host = nullptr;
    pic::bf16* dz = nullptr;              // relu'ed output gradient (padded NHWC)
    pic::bf16* dzt = nullptr;             // its CHW transpose
    pic::bf16* xt = nullptr;              // CHW padded transpose of a conv input
    int step = 0;
    hipStream_t stream = nullptr;
    bool own_stream = false;
    double* r0 = nullptr;                 // starting residual on the image (FP64)
    unsigned long long* scl = nullptr;    // per-solve scales (max|r0|, max|phi|)
    double* mpart = nullptr;              // per-workgroup maxima for the scales
    float* wpart = nullptr;               // weight-gradient partials, one slot per row chunk
    float* bpart = nullptr;               // bias / head / loss partials, one slot per workgroup
    size_t wpart_n = 0, bpart_n = 0;
};

namespace pic {

static int unet_fail(pic_unet* u, const char* what) {
    set_error("pic_unet: %s", what);
    (void)u;
    return PIC_ERR_ALLOC;
}

// layer table, built for an N x N image (N a multiple of 64)
static int unet_build(pic_unet* u) {
    const int N = u->N;
    auto add = [&](int H, int C, bool grad = true) {
        Tensor x;
        x.H = H;
        x.W = H;
        x.C = C;
        x.grad = grad;
        u->t.push_back(x);
        return (int)u->t.size() - 1;
    };
    const int X0 = add(N, 16), A0 = add(N, 16), CAT0 = add(N, 48);
    const int P0 = add(N / 2, 16), A1 = add(N / 2, 32), CAT1 = add(N / 2, 96);
    const int P1 = add(N / 4, 32), A2 = add(N / 4, 64), CAT2 = add(N / 4, 128);
    const int P2 = add(N / 8, 64), A3 = add(N / 8, 64), B3 = add(N / 8, 64);
    const int A4 = add(N / 4, 64), D2 = add(N / 4, 64);
    const int A5 = add(N / 2, 32), D1 = add(N / 2, 32);
    const int A6 = add(N, 16), D0 = add(N, 16);
    // inference only: compact skip outputs and the 2-channel input of the first layer
    add(N, 16, false);
    add(N / 2, 32, false);
    add(N / 4, 64, false);
    add(N, 2, false);
    size_t off = 0;
    auto conv = [&](int in, int io, int cin, int cin_real, int out, int oo, int cout) {
        Conv c{in, io, cin, cin_real, out, oo, cout, 0, 0};
        c.woff = off;
        off += (size_t)9 * cout * cin;
        c.boff = off;
        off += cout;
        u->conv.push_back(c);
    };
    conv(X0, 0, 16, 2, A0, 0, 16);          // 0
    conv(A0, 0, 16, 16, CAT0, 32, 16);      // 1  s0
    conv(P0, 0, 16, 16, A1, 0, 32);         // 2
    conv(A1, 0, 32, 32, CAT1, 64, 32);      // 3  s1
    conv(P1, 0, 32, 32, A2, 0, 64);         // 4
    conv(A2, 0, 64, 64, CAT2, 64, 64);      // 5  s2
    conv(P2, 0, 64, 64, A3, 0, 64);         // 6
    conv(A3, 0, 64, 64, B3, 0, 64);         // 7
    conv(CAT2, 0, 128, 128, A4, 0, 64);     // 8
    conv(A4, 0, 64, 64, D2, 0, 64);         // 9
    conv(CAT1, 0, 96, 96, A5, 0, 32);       // 10
    conv(A5, 0, 32, 32, D1, 0, 32);         // 11
    conv(CAT0, 0, 48, 48, A6, 0, 16);       // 12
    conv(A6, 0, 16, 16, D0, 0, 16);         // 13
    u->head_w = off;
    off += 16;
    u->head_b = off;
    off += 1;
    u->nparam = off;
    (void)B3;
    (void)D2;
    (void)D1;
    return PIC_OK;
}

// tensor indices used by the fixed schedule below
enum { X0 = 0, A0, CAT0, P0, A1, CAT1, P1, A2, CAT2, P2, A3, B3, A4, D2, A5, D1, A6, D0, S0, S1, S2, X0C };

// Rows per workgroup of the weight gradient: one, or more when the layer has
// many tile pairs (about 2048 workgroups either way).
static int wgrad_rows(int H, int pairs) { return std::max(1, pairs * H / 2048); }
static int wgrad_chunks(int H, int pairs) { return (H + wgrad_rows(H, pairs) - 1) / wgrad_rows(H, pairs); }
// row stride of the transposed, padded input: W + 2 columns rounded up to 8, plus 8
static int xt_stride(int W) { return (W + 2 + 7) / 8 * 8 + 8; }

// dW = the weight gradient: per-chunk partials into `part`, then their sum in
// chunk order.
static int wgrad_any(hipStream_t s, const bf16* dZT, const bf16* XT, int cout, int cin, int H, int W, float* part,
                     float* dW) {
    const int pairs = (cout / 16) * (cin / 16);
    const int rows = wgrad_rows(H, pairs);
    const int nchunk = wgrad_chunks(H, pairs);
    hipLaunchKernelGGL(k_wgrad, dim3((unsigned)nchunk, (unsigned)pairs), dim3(kT), 0, s, dZT, XT, cout, cin, H, W,
                       xt_stride(W), rows, part);
    const int64_t len = (int64_t)9 * cout * cin;
    hipLaunchKernelGGL(k_sum_slots, dim3(sum_blocks(len)), dim3(kT), 0, s, part, nchunk, len, dW);
    PIC_HIP(hipGetLastError());
    return PIC_OK;
}

static size_t tsize(const Tensor& x) { return (size_t)(x.H + 2) * (x.W + 2) * x.C; }
// bias-gradient slots of k_relu_bwd_t: one per (64 padded columns, kRB padded rows)
static int bias_parts(int H, int W) { return ((W + 2 + 63) / 64) * ((H + 2 + kRB - 1) / kRB); }

// nearest 2x upsampling of `in` (channels io ..) into `out` (channels oo ..)
static int upsample(pic_unet* u, int in, int io, int out, int oo, int c) {
    const Tensor& a = u->t[in];
    const Tensor& b = u->t[out];
    const int64_t n = (int64_t)b.H * b.W * (c / 8);
    hipLaunchKernelGGL(k_up2, dim3(nb(n)), dim3(kT), 0, u->stream, a.p, a.C, io, b.p, b.C, oo, c, a.H, a.W);
    PIC_HIP(hipGetLastError());
    return PIC_OK;
}

template <int NT, int EPI>
static void conv_fwd_nt(hipStream_t s, const ConvFwdArgs& a, unsigned tiles) {
    const dim3 grid(tiles, (unsigned)(a.coutT / (16 * NT)));
    // launches of at most 4 workgroups per CU: 32 channels per LDS fill (half the waits for
    // global memory); larger ones: 16 (less LDS and fewer registers: 8 waves per SIMD)
    if (grid.x * grid.y <= 1024 && a.cin % 32 == 0)
        hipLaunchKernelGGL((k_conv_fwd<32, NT, EPI>), grid, dim3(256), 0, s, a);
    else
        hipLaunchKernelGGL((k_conv_fwd<16, NT, EPI>), grid, dim3(256), 0, s, a);
}

// Forward convolution k.  Training keeps every activation the backward pass
// reads (skip outputs inside the concatenated tensors, the upsampled halves,
// the last layer's output); inference writes the skip outputs to compact
// tensors, reads the upsampled halves in place and keeps only the head output.
static int conv_fwd(pic_unet* u, int k, bool train) {
    const Conv& c = u->conv[k];
    const Tensor& in = u->t[c.in];
    const Tensor& out = u->t[c.out];
    ConvFwdArgs a{};
    a.X = in.p;
    a.xC = in.C;
    a.xo = c.io;
    a.cin = c.cin;
    a.Wt = u->wb + c.woff;
    a.bias = u->w + c.boff;
    a.coutT = c.cout;
    a.Y = out.p;
    a.yC = out.C;
    a.yo = c.oo;
    a.H = in.H;
    a.W = in.W;
    auto skip_out = [&](int S, int P) {
        if (!train) {
            a.Y = u->t[S].p;
            a.yC = u->t[S].C;
            a.yo = 0;
        }
        a.P = u->t[P].p;
        a.pC = u->t[P].C;
        a.po = 0;
    };
    auto up_in = [&](int U, int S, int cu) {
        if (train) return;
        a.U = u->t[U].p;
        a.uC = u->t[U].C;
        a.uo = 0;
        a.cu = cu;
        a.X = u->t[S].p;
        a.xC = u->t[S].C;
        a.xo = 0;
    };
    switch (k) {
    case 1: skip_out(S0, P0); break;
    case 3: skip_out(S1, P1); break;
    case 5: skip_out(S2, P2); break;
    case 8: up_in(B3, S2, 64); break;
    case 10: up_in(D2, S1, 64); break;
    case 12: up_in(D1, S0, 32); break;
    case 13:
        a.wh = u->w + u->head_w;
        a.bh = u->w + u->head_b;
        a.out = u->out;
        if (!train) a.Y = nullptr;
        break;
    default: break;
    }
    const unsigned tiles = (unsigned)((a.H / kFTH) * ((a.W + 63) / 64));
    if (k == 0) {
        a.X = u->t[X0C].p;
        a.xC = 2;
        a.xo = 0;
        hipLaunchKernelGGL(k_conv0_fwd, dim3(tiles), dim3(256), 0, u->stream, a);
    } else {
        const int nt = tiles >= 1024 ? c.cout / 16 : 1;
        if (a.out) {
            if (c.cout != 16) {
                set_error("pic_unet: the head needs 16 channels (got %d)", c.cout);
                return PIC_ERR_UNSUPPORTED;
            }
            conv_fwd_nt<1, 2>(u->stream, a, tiles);
        } else if (a.P) {
            switch (nt) {
            case 1: conv_fwd_nt<1, 1>(u->stream, a, tiles); break;
            case 2: conv_fwd_nt<2, 1>(u->stream, a, tiles); break;
            case 4: conv_fwd_nt<4, 1>(u->stream, a, tiles); break;
            default: set_error("pic_unet: unsupported channel count %d", c.cout); return PIC_ERR_UNSUPPORTED;
            }
        } else {
            switch (nt) {
            case 1: conv_fwd_nt<1, 0>(u->stream, a, tiles); break;
            case 2: conv_fwd_nt<2, 0>(u->stream, a, tiles); break;
            case 4: conv_fwd_nt<4, 0>(u->stream, a, tiles); break;
            default: set_error("pic_unet: unsupported channel count %d", c.cout); return PIC_ERR_UNSUPPORTED;
            }
        }
    }
    PIC_HIP(hipGetLastError());
    return PIC_OK;
}

// The forward pass from the compact input (X0C) to u->out.
static int unet_forward(pic_unet* u, bool train) {
    for (int k = 0; k < 14; ++k) {
        if (train && k == 8) PIC_TRY(upsample(u, B3, 0, CAT2, 0, 64));
        if (train && k == 10) PIC_TRY(upsample(u, D2, 0, CAT1, 0, 64));
        if (train && k == 12) PIC_TRY(upsample(u, D1, 0, CAT0, 0, 32));
        PIC_TRY(conv_fwd(u, k, train));
    }
    return PIC_OK;
}

// Backward of conv k given the FP32 gradient of its output (in the output
// tensor's gradient buffer); writes the input gradient (unless k == 0) and
// adds the weight and bias gradients.
static int conv_bwd(pic_unet* u, int k) {
    const Conv& c = u->conv[k];
    const Tensor& a = u->t[c.in];
    const Tensor& b = u->t[c.out];
    const int H = a.H, W = a.W;
    hipStream_t s = u->stream;
    // dZ (padded NHWC, for the input gradient), its CHW transpose and the bias partials in one pass
    const dim3 rg((unsigned)((W + 2 + 63) / 64), (unsigned)((H + 2 + kRB - 1) / kRB), (unsigned)(c.cout / 16));
    hipLaunchKernelGGL(k_relu_bwd_t, rg, dim3(kT), 0, s, b.g, b.C, c.oo, b.p, b.C, c.oo, u->dz, u->dzt, u->bpart,
                       c.cout, H, W);
    hipLaunchKernelGGL(k_to_chw, dim3((unsigned)((W + 2 + 63) / 64), (unsigned)(H + 2), (unsigned)((c.cin + 15) / 16)),
                       dim3(kT), 0, s, a.p, a.C, c.io, c.cin, H, W, 1, xt_stride(W), u->xt);
    PIC_HIP(hipGetLastError());
    PIC_TRY(wgrad_any(s, u->dzt, u->xt, c.cout, c.cin, H, W, u->wpart, u->g + c.woff));
    hipLaunchKernelGGL(k_sum_slots_few, dim3((unsigned)c.cout), dim3(kT), 0, s, u->bpart, (int)(rg.x * rg.y), c.cout,
                       u->g + c.boff);
    PIC_HIP(hipGetLastError());
    if (k > 0) {
        // input gradient: the 3x3 convolution of dZ with the flipped, transposed weights
        ConvFwdArgs d{};
        d.X = u->dz;
        d.xC = c.cout;
        d.cin = c.cout;
        d.Wt = u->wflip + c.woff;
        d.coutT = c.cin;
        d.Yf = a.g;
        d.yC = a.C;
        d.yo = c.io;
        d.H = H;
        d.W = W;
        const unsigned tiles = (unsigned)((H / kFTH) * ((W + 63) / 64));
        const int ct = c.cin / 16;
        const int nt = tiles < 1024 ? 1 : ct % 4 == 0 ? 4 : ct % 2 == 0 ? 2 : 1;
        switch (nt) {
        case 1: conv_fwd_nt<1, 3>(s, d, tiles); break;
        case 2: conv_fwd_nt<2, 3>(s, d, tiles); break;
        case 4: conv_fwd_nt<4, 3>(s, d, tiles); break;
        default: set_error("pic_unet: unsupported channel count %d", c.cin); return PIC_ERR_UNSUPPORTED;
        }
        PIC_HIP(hipGetLastError());
    }
    return PIC_OK;
}

// loss 0: masked mean squared error against u->tgt; 1: the residual loss
// (needs u->r0, u->scl from unet_prepare and the grid constants cxd, cyd).
static int unet_backward(pic_unet* u, int loss_kind = 0, float cxd = 0.f, float cyd = 0.f) {
    hipStream_t s = u->stream;
    const int N = u->N;
    const unsigned gN = nb((int64_t)N * N);
    hipLaunchKernelGGL(k_zero_i32, dim3(1), dim3(1), 0, s, u->count);
    hipLaunchKernelGGL(k_mask_count, dim3(nb((int64_t)N * N / 16)), dim3(kT), 0, s, u->mask, N * N, u->count);
    if (loss_kind == 0) {
        hipLaunchKernelGGL(k_loss_mse, dim3(gN), dim3(kT), 0, s, u->out, u->tgt, u->mask, u->count, u->gout, u->bpart,
                           N * N);
    } else {
        hipLaunchKernelGGL(k_loss_res, dim3(gN), dim3(kT), 0, s, u->out, u->r0, u->scl, u->mask, u->count, cxd, cyd,
                           N, u->tgt, u->bpart);
        hipLaunchKernelGGL(k_loss_res_grad, dim3(gN), dim3(kT), 0, s, u->tgt, u->mask, u->count, cxd, cyd, N, u->gout);
    }
    hipLaunchKernelGGL(k_sum_slots_few, dim3(1), dim3(kT), 0, s, u->bpart, (int)gN, 1, u->loss);
    hipLaunchKernelGGL(k_head_bwd, dim3(gN), dim3(kT), 0, s, u->gout, u->t[D0].p, u->w + u->head_w, u->t[D0].g,
                       u->bpart, N, N);
    // head_b follows head_w's 16 weights in the parameter vector, as in the slots
    hipLaunchKernelGGL(k_sum_slots_few, dim3(17), dim3(kT), 0, s, u->bpart, (int)gN, 17, u->g + u->head_w);
    PIC_HIP(hipGetLastError());
    auto up_bwd = [&](int from, int fo, int to, int c) -> int {
        const Tensor& a = u->t[from];
        const Tensor& b = u->t[to];
        const int64_t n = (int64_t)b.H * b.W * c;
        hipLaunchKernelGGL(k_up2_bwd, dim3(nb(n)), dim3(kT), 0, s, a.g, a.C, fo, b.g, b.C, 0, c, b.H, b.W);
        PIC_HIP(hipGetLastError());
        return PIC_OK;
    };
    auto pool_bwd = [&](int from, int to, int to_o, int c) -> int {
        const Tensor& a = u->t[from];
        const Tensor& b = u->t[to];
        const int64_t n = (int64_t)b.H * b.W * c;
        hipLaunchKernelGGL(k_pool2_bwd, dim3(nb(n)), dim3(kT), 0, s, a.g, a.C, 0, b.g, b.C, to_o, c, b.H, b.W);
        PIC_HIP(hipGetLastError());
        return PIC_OK;
    };
    PIC_TRY(conv_bwd(u, 13));
    PIC_TRY(conv_bwd(u, 12));
    PIC_TRY(up_bwd(CAT0, 0, D1, 32));
    PIC_TRY(conv_bwd(u, 11));
    PIC_TRY(conv_bwd(u, 10));
    PIC_TRY(up_bwd(CAT1, 0, D2, 64));
    PIC_TRY(conv_bwd(u, 9));
    PIC_TRY(conv_bwd(u, 8));
    PIC_TRY(up_bwd(CAT2, 0, B3, 64));
    PIC_TRY(conv_bwd(u, 7));
    PIC_TRY(conv_bwd(u, 6));
    PIC_TRY(pool_bwd(P2, CAT2, 64, 64));
    PIC_TRY(conv_bwd(u, 5));
    PIC_TRY(conv_bwd(u, 4));
    PIC_TRY(pool_bwd(P1, CAT1, 64, 32));
    PIC_TRY(conv_bwd(u, 3));
    PIC_TRY(conv_bwd(u, 2));
    PIC_TRY(pool_bwd(P0, CAT0, 32, 16));
    PIC_TRY(conv_bwd(u, 1));
    PIC_TRY(conv_bwd(u, 0));
    return PIC_OK;
}

// BF16 and flipped copies of the conv weights from the FP32 masters.
__global__ void k_refresh_bf16(const float* __restrict__ w, int n, bf16* __restrict__ wb, bf16* __restrict__ wflip,
                               int cout, int cin) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const bf16 b = f2bf(w[k]);
    wb[k] = b;
    const int ci = k % cin, r = k / cin, co = r % cout, tap = r / cout;
    wflip[((size_t)(8 - tap) * cin + ci) * cout + co] = b;
}

static int refresh_copies(pic_unet* u) {
    for (const Conv& c : u->conv) {
        const int n = 9 * c.cout * c.cin;
        hipLaunchKernelGGL(k_refresh_bf16, dim3(nb(n)), dim3(kT), 0, u->stream, u->w + c.woff, n, u->wb + c.woff,
                           u->wflip + c.woff, c.cout, c.cin);
    }
    PIC_HIP(hipGetLastError());
    return PIC_OK;
}

static int adam_step(pic_unet* u, float lr) {
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    u->step += 1;
    const float lr_t = lr * std::sqrt(1.f - std::pow(b2, (float)u->step)) / (1.f - std::pow(b1, (float)u->step));
    for (const Conv& c : u->conv) {
        const int n = 9 * c.cout * c.cin;
        hipLaunchKernelGGL(k_adam, dim3(nb(n)), dim3(kT), 0, u->stream, u->w + c.woff, u->g + c.woff, u->m + c.woff,
                           u->v + c.woff, n, lr_t, b1, b2, eps, u->wb + c.woff, u->wflip + c.woff, c.cout, c.cin);
        hipLaunchKernelGGL(k_adam, dim3(nb(c.cout)), dim3(kT), 0, u->stream, u->w + c.boff, u->g + c.boff,
                           u->m + c.boff, u->v + c.boff, c.cout, lr_t, b1, b2, eps, (bf16*)nullptr, (bf16*)nullptr, 1,
                           1);
    }
    hipLaunchKernelGGL(k_adam, dim3(1), dim3(kT), 0, u->stream, u->w + u->head_w, u->g + u->head_w, u->m + u->head_w,
                       u->v + u->head_w, 17, lr_t, b1, b2, eps, (bf16*)nullptr, (bf16*)nullptr, 1, 1);
    PIC_HIP(hipGetLastError());
    return PIC_OK;
}

// deterministic initialisation (He normal for convs, small head)
static uint64_t sm64(uint64_t& s) {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
static double unif(uint64_t& s) { return ((sm64(s) >> 11) + 0.5) * (1.0 / 9007199254740992.0); }
static double gauss(uint64_t& s) {
    double a = unif(s), b = unif(s);
    return std::sqrt(-2.0 * std::log(a)) * std::cos(6.283185307179586 * b);
}

// Upload host inputs [N][N][2] into the compact input and, for training, the
// 16-channel input; run the forward pass; copy the output [N][N] to `out`.
static int unet_forward_host(pic_unet* u, const float* x, float* out, bool train) {
    const int N = u->N;
    std::vector<bf16> hc((size_t)(N + 2) * (N + 2) * 2, 0);
    for (int y = 0; y < N; ++y)
        for (int xx = 0; xx < N; ++xx)
            for (int c = 0; c < 2; ++c)
                hc[((size_t)(y + 1) * (N + 2) + xx + 1) * 2 + c] = f2bf_host(x[((size_t)y * N + xx) * 2 + c]);
    PIC_HIP(hipMemcpyAsync(u->t[X0C].p, hc.data(), hc.size() * 2, hipMemcpyHostToDevice, u->stream));
    std::vector<bf16> h;
    if (train) {
        h.assign((size_t)(N + 2) * (N + 2) * 16, 0);
        for (size_t p = 0; p < (size_t)(N + 2) * (N + 2); ++p) {
            h[p * 16] = hc[p * 2];
            h[p * 16 + 1] = hc[p * 2 + 1];
        }
        PIC_HIP(hipMemcpyAsync(u->t[X0].p, h.data(), h.size() * 2, hipMemcpyHostToDevice, u->stream));
    }
    PIC_TRY(unet_forward(u, train));
    PIC_HIP(hipMemcpyAsync(out, u->out, (size_t)N * N * 4, hipMemcpyDeviceToHost, u->stream));
    PIC_HIP(hipStreamSynchronize(u->stream));
    return PIC_OK;
}

} // namespace pic

using namespace pic;

extern "C" {

int pic_unet_mfma_probe(const float* a, const float* b, float* d) {
    if (!a || !b || !d) {
        set_error("pic_unet_mfma_probe: null argument");
        return PIC_ERR_ARG;
    }
    float *da, *db, *dd;
    PIC_HIP(hipMalloc(&da, 256 * sizeof(float)));
    PIC_HIP(hipMalloc(&db, 256 * sizeof(float)));
    PIC_HIP(hipMalloc(&dd, 256 * sizeof(float)));
    PIC_HIP(hipMemcpy(da, a, 256 * sizeof(float), hipMemcpyHostToDevice));
    PIC_HIP(hipMemcpy(db, b, 256 * sizeof(float), hipMemcpyHostToDevice));
    hipLaunchKernelGGL(k_mfma_probe, dim3(1), dim3(64), 0, 0, da, db, dd);
    PIC_HIP(hipGetLastError());
    PIC_HIP(hipMemcpy(d, dd, 256 * sizeof(float), hipMemcpyDeviceToHost));
    (void)hipFree(da);
    (void)hipFree(db);
    (void)hipFree(dd);
    return PIC_OK;
}

int pic_unet_mfma32_probe(const float* a, const float* b, float* d) {
    if (!a || !b || !d) {
        set_error("pic_unet_mfma32_probe: null argument");
        return PIC_ERR_ARG;
    }
    float *da, *db, *dd;
    PIC_HIP(hipMalloc(&da, 512 * sizeof(float)));
    PIC_HIP(hipMalloc(&db, 512 * sizeof(float)));
    PIC_HIP(hipMalloc(&dd, 256 * sizeof(float)));
    PIC_HIP(hipMemcpy(da, a, 512 * sizeof(float), hipMemcpyHostToDevice));
    PIC_HIP(hipMemcpy(db, b, 512 * sizeof(float), hipMemcpyHostToDevice));
    hipLaunchKernelGGL(k_mfma32_probe, dim3(1), dim3(64), 0, 0, da, db, dd);
    PIC_HIP(hipGetLastError());
    PIC_HIP(hipMemcpy(d, dd, 256 * sizeof(float), hipMemcpyDeviceToHost));
    (void)hipFree(da);
    (void)hipFree(db);
    (void)hipFree(dd);
    return PIC_OK;
}

int pic_unet_create(int n, uint64_t seed, pic_unet** out) {
    if (!out || n < 64 || n % 64 != 0) {
        set_error("pic_unet_create: image size must be a positive multiple of 64 (got %d)", n);
        return PIC_ERR_ARG;
    }
    int dev = 0;
    if (hipGetDeviceCount(&dev) != hipSuccess || dev < 1) {
        set_error("pic_unet_create: no HIP device");
        return PIC_ERR_HIP;
    }
    pic_unet* u = new pic_unet();
    u->N = n;
    unet_build(u);
    auto cleanup = [&](int rc) {
        pic_unet_destroy(u);
        return rc;
    };
    if (hipStreamCreateWithFlags(&u->stream, hipStreamNonBlocking) != hipSuccess) return cleanup(unet_fail(u, "stream"));
    u->own_stream = true;
    size_t maxz = 0, maxx = 0;
    for (Tensor& x : u->t) {
        const size_t sz = tsize(x);
        if (hipMalloc(&x.p, sz * sizeof(bf16)) != hipSuccess) return cleanup(unet_fail(u, "activation allocation"));
        if (hipMemset(x.p, 0, sz * sizeof(bf16)) != hipSuccess) return cleanup(unet_fail(u, "memset"));
        if (!x.grad) continue;
        if (hipMalloc(&x.g, sz * sizeof(float)) != hipSuccess) return cleanup(unet_fail(u, "gradient allocation"));
        if (hipMemset(x.g, 0, sz * sizeof(float)) != hipSuccess) return cleanup(unet_fail(u, "memset"));
    }
    for (const Conv& c : u->conv) {
        const Tensor& a = u->t[c.in];
        maxz = std::max(maxz, (size_t)(a.H + 2) * (a.W + 2) * c.cout);
        maxx = std::max(maxx, (size_t)(a.H + 2) * xt_stride(a.W) * c.cin);
    }
    const size_t P = u->nparam;
    for (const Conv& c : u->conv) {
        const Tensor& a = u->t[c.in];
        u->wpart_n = std::max(u->wpart_n, (size_t)wgrad_chunks(a.H, (c.cout / 16) * (c.cin / 16)) * 9 * c.cout * c.cin);
        u->bpart_n = std::max(u->bpart_n, (size_t)bias_parts(a.H, a.W) * c.cout);
    }
    u->bpart_n = std::max(u->bpart_n, (size_t)nb((int64_t)n * n) * 17);
    bool ok = hipMalloc(&u->wpart, u->wpart_n * 4) == hipSuccess && hipMalloc(&u->bpart, u->bpart_n * 4) == hipSuccess &&
              hipMalloc(&u->w, P * 4) == hipSuccess && hipMalloc(&u->g, P * 4) == hipSuccess &&
              hipMalloc(&u->m, P * 4) == hipSuccess && hipMalloc(&u->v, P * 4) == hipSuccess &&
              hipMalloc(&u->wb, P * 2) == hipSuccess && hipMalloc(&u->wflip, P * 2) == hipSuccess &&
              hipMalloc(&u->out, (size_t)n * n * 4) == hipSuccess && hipMalloc(&u->tgt, (size_t)n * n * 4) == hipSuccess &&
              hipMalloc(&u->mask, (size_t)n * n) == hipSuccess && hipMalloc(&u->loss, 4) == hipSuccess &&
              hipMalloc(&u->gout, (size_t)n * n * 4) == hipSuccess &&
              hipMalloc(&u->count, 4) == hipSuccess && hipMalloc(&u->r0, (size_t)n * n * 8) == hipSuccess &&
              hipMalloc(&u->scl, 16) == hipSuccess &&
              hipMalloc(&u->mpart, (size_t)nb((int64_t)n * n) * 2 * sizeof(double)) == hipSuccess &&
              hipMalloc(&u->dz, maxz * 2) == hipSuccess && hipMalloc(&u->dzt, maxz * 2) == hipSuccess &&
              hipMalloc(&u->xt, maxx * 2) == hipSuccess &&
              hipHostMalloc((void**)&u->loss_host, 4, hipHostMallocDefault) == hipSuccess;
    if (!ok) return cleanup(unet_fail(u, "parameter allocation"));
    (void)hipMemset(u->dz, 0, maxz * 2);
    (void)hipMemset(u->m, 0, P * 4);
    (void)hipMemset(u->v, 0, P * 4);
    (void)hipMemset(u->g, 0, P * 4);
    std::vector<float> w(P, 0.f);
    uint64_t s = seed * 0x2545F4914F6CDD1Dull + 1;
    for (const Conv& c : u->conv) {
        const double std_ = std::sqrt(2.0 / (9.0 * c.cin_real));
        for (int tap = 0; tap < 9; ++tap)
            for (int co = 0; co < c.cout; ++co)
                for (int ci = 0; ci < c.cin; ++ci)
                    w[c.woff + ((size_t)tap * c.cout + co) * c.cin + ci] = ci < c.cin_real ? (float)(std_ * gauss(s)) : 0.f;
    }
    for (int k = 0; k < 16; ++k) w[u->head_w + k] = (float)(1e-3 * gauss(s));
    if (hipMemcpy(u->w, w.data(), P * 4, hipMemcpyHostToDevice) != hipSuccess) return cleanup(unet_fail(u, "copy"));
    if (refresh_copies(u) != PIC_OK || hipStreamSynchronize(u->stream) != hipSuccess)
        return cleanup(unet_fail(u, "weight copies"));
    *out = u;
    return PIC_OK;
}

void pic_unet_destroy(pic_unet* u) {
    if (!u) return;
    if (u->stream) (void)hipStreamSynchronize(u->stream);
    for (Tensor& x : u->t) {
        if (x.p) (void)hipFree(x.p);
        if (x.g) (void)hipFree(x.g);
    }
    void* ptrs[] = {u->w, u->g, u->m, u->v, u->wb, u->wflip, u->out, u->tgt, u->gout, u->mask, u->loss, u->count,
                    u->dz, u->dzt, u->xt, u->r0, u->scl, u->mpart, u->wpart, u->bpart};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    if (u->loss_host) (void)hipHostFree(u->loss_host);
    if (u->own_stream && u->stream) (void)hipStreamDestroy(u->stream);
    delete u;
}

int64_t pic_unet_num_params(pic_unet* u) { return u ? (int64_t)u->nparam : 0; }

int pic_unet_get_params(pic_unet* u, float* w) {
    if (!u || !w) {
        set_error("pic_unet_get_params: null argument");
        return PIC_ERR_ARG;
    }
    PIC_HIP(hipStreamSynchronize(u->stream));
    PIC_HIP(hipMemcpy(w, u->w, u->nparam * 4, hipMemcpyDeviceToHost));
    return PIC_OK;
}

int pic_unet_set_params(pic_unet* u, const float* w) {
    if (!u || !w) {
        set_error("pic_unet_set_params: null argument");
        return PIC_ERR_ARG;
    }
    PIC_HIP(hipStreamSynchronize(u->stream));
    PIC_HIP(hipMemcpy(u->w, w, u->nparam * 4, hipMemcpyHostToDevice));
    PIC_TRY(refresh_copies(u));
    PIC_HIP(hipStreamSynchronize(u->stream));
    return PIC_OK;
}

int pic_unet_get_grads(pic_unet* u, float* g) {
    if (!u || !g) {
        set_error("pic_unet_get_grads: null argument");
        return PIC_ERR_ARG;
    }
    PIC_HIP(hipStreamSynchronize(u->stream));
    PIC_HIP(hipMemcpy(g, u->g, u->nparam * 4, hipMemcpyDeviceToHost));
    return PIC_OK;
}

// x: host [N][N][2] inputs (already scaled); out: host [N][N]
int pic_unet_forward(pic_unet* u, const float* x, float* out) {
    if (!u || !x || !out) {
        set_error("pic_unet_forward: null argument");
        return PIC_ERR_ARG;
    }
    return unet_forward_host(u, x, out, false);
}

// One training step on host data: forward, backward, and (lr > 0) Adam.
// tgt: [N][N]; mask: [N][N] (1 = counted).  *loss receives the loss before the update.
int pic_unet_train_step(pic_unet* u, const float* x, const float* tgt, const uint8_t* mask, float lr, float* loss) {
    if (!u || !x || !tgt || !mask) {
        set_error("pic_unet_train_step: null argument");
        return PIC_ERR_ARG;
    }
    const int N = u->N;
    std::vector<float> o((size_t)N * N);
    PIC_TRY(unet_forward_host(u, x, o.data(), true));
    PIC_HIP(hipMemcpyAsync(u->tgt, tgt, (size_t)N * N * 4, hipMemcpyHostToDevice, u->stream));
    PIC_HIP(hipMemcpyAsync(u->mask, mask, (size_t)N * N, hipMemcpyHostToDevice, u->stream));
    PIC_TRY(unet_backward(u));
    PIC_HIP(hipMemcpyAsync(u->loss_host, u->loss, 4, hipMemcpyDeviceToHost, u->stream));
    if (lr > 0.f) PIC_TRY(adam_step(u, lr));
    PIC_HIP(hipStreamSynchronize(u->stream));
    if (loss) *loss = *u->loss_host;
    return PIC_OK;
}

// Training step with the residual loss on host data: r: [n][n] scaled
// residual (max |r| = 1 or less), mask: [n][n] unknown nodes, cxd, cyd: the
// operator's off-diagonals over its diagonal.
int pic_unet_train_step_residual(pic_unet* u, const float* x, const double* r, const uint8_t* mask, float cxd,
                                 float cyd, float lr, float* loss) {
    if (!u || !x || !r || !mask) {
        set_error("pic_unet_train_step_residual: null argument");
        return PIC_ERR_ARG;
    }
    const int N = u->N;
    std::vector<float> o((size_t)N * N);
    PIC_TRY(unet_forward_host(u, x, o.data(), true));
    const double one = 1.0;
    unsigned long long sc[2];
    std::memcpy(&sc[0], &one, 8);
    sc[1] = sc[0];
    PIC_HIP(hipMemcpyAsync(u->r0, r, (size_t)N * N * 8, hipMemcpyHostToDevice, u->stream));
    PIC_HIP(hipMemcpyAsync(u->scl, sc, 16, hipMemcpyHostToDevice, u->stream));
    PIC_HIP(hipMemcpyAsync(u->mask, mask, (size_t)N * N, hipMemcpyHostToDevice, u->stream));
    PIC_TRY(unet_backward(u, 1, cxd, cyd));
    PIC_HIP(hipMemcpyAsync(u->loss_host, u->loss, 4, hipMemcpyDeviceToHost, u->stream));
    if (lr > 0.f) PIC_TRY(adam_step(u, lr));
    PIC_HIP(hipStreamSynchronize(u->stream));
    if (loss) *loss = *u->loss_host;
    return PIC_OK;
}

// Average device time of the forward pass and of a full training step
// (forward, backward, Adam at lr = 0 so the weights stay) over `reps` runs on
// the inputs and target currently on the device.
int pic_unet_benchmark(pic_unet* u, int reps, float* fwd_ms, float* train_ms) {
    if (!u || reps < 1) {
        set_error("pic_unet_benchmark: invalid argument");
        return PIC_ERR_ARG;
    }
    hipEvent_t e0, e1, e2;
    PIC_HIP(hipEventCreate(&e0));
    PIC_HIP(hipEventCreate(&e1));
    PIC_HIP(hipEventCreate(&e2));
    PIC_TRY(unet_forward(u, false));               // warm-up
    PIC_HIP(hipEventRecord(e0, u->stream));
    for (int k = 0; k < reps; ++k) PIC_TRY(unet_forward(u, false));
    PIC_HIP(hipEventRecord(e1, u->stream));
    for (int k = 0; k < reps; ++k) {
        PIC_TRY(unet_forward(u, true));
        PIC_TRY(unet_backward(u));
    }
    PIC_HIP(hipEventRecord(e2, u->stream));
    PIC_HIP(hipEventSynchronize(e2));
    float a = 0, b = 0;
    PIC_HIP(hipEventElapsedTime(&a, e0, e1));
    PIC_HIP(hipEventElapsedTime(&b, e1, e2));
    if (fwd_ms) *fwd_ms = a / reps;
    if (train_ms) *train_ms = b / reps;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipEventDestroy(e2);
    return PIC_OK;
}

// the context's network (null when the learned guess is off)
pic_unet* pic_learned_unet(pic_ctx* c) { return c ? c->unet : nullptr; }

} // extern "C"

// ---------------------------------------------------- use inside the solver
namespace pic {

// r0 and the per-solve scales for (rhs = c->rho after the Dirichlet overwrite, phi_old)
static int unet_prepare(pic_ctx* c, pic_unet* u, const double* phi_old, bool train) {
    const int N = u->N, gx = c->g.gx;
    hipStream_t s = c->stream;
    const double cx = 1.0 / (c->g.dx * c->g.dx), cy = 1.0 / (c->g.dy * c->g.dy);
    const unsigned nblk = nb((int64_t)N * N);
    hipLaunchKernelGGL(k_unet_resid, dim3(nblk), dim3(kT), 0, s, c->rho, phi_old, c->dir, gx, N, cx, cy,
                       2 * (cx + cy), u->r0, u->mpart);
    hipLaunchKernelGGL(k_unet_scale, dim3(1), dim3(kT), 0, s, u->mpart, (int)nblk, u->scl);
    hipLaunchKernelGGL(k_unet_input, dim3(nb((int64_t)N * N)), dim3(kT), 0, s, u->r0, phi_old, gx, N, u->scl,
                       u->t[X0C].p, train ? u->t[X0].p : (bf16*)nullptr);
    PIC_HIP(hipGetLastError());
    return PIC_OK;
}

int unet_predict(pic_ctx* c, pic_unet* u, UnetPrediction* o) {
    PIC_TRY(unet_prepare(c, u, c->phi, false));
    hipStream_t keep = u->stream;
    u->stream = c->stream;
    const int rc = unet_forward(u, false);
    u->stream = keep;
    PIC_TRY(rc);
    o->out = u->out;
    o->scl = u->scl;
    o->N = u->N;
    return PIC_OK;
}

int unet_train_pair(pic_ctx* c, pic_unet* u, const double* phi_old, float lr, float* loss) {
    const int N = u->N, gx = c->g.gx;
    hipStream_t s = c->stream;
    PIC_TRY(unet_prepare(c, u, phi_old, true));
    const double cx = 1.0 / (c->g.dx * c->g.dx), cy = 1.0 / (c->g.dy * c->g.dy);
    const double diag = 2 * (cx + cy);
    // the unknown-node mask (the MSE target it also writes is not used)
    hipLaunchKernelGGL(k_unet_target, dim3(nb((int64_t)N * N)), dim3(kT), 0, s, c->phi, phi_old, c->dir, gx, N, diag,
                       u->scl, u->tgt, u->mask);
    PIC_HIP(hipGetLastError());
    hipStream_t keep = u->stream;
    u->stream = s;
    int rc = unet_forward(u, true);
    if (rc == PIC_OK) rc = unet_backward(u, 1, (float)(cx / diag), (float)(cy / diag));
    if (rc == PIC_OK) rc = adam_step(u, lr);
    u->stream = keep;
    PIC_TRY(rc);
    if (loss) {
        PIC_HIP(hipMemcpyAsync(u->loss_host, u->loss, 4, hipMemcpyDeviceToHost, s));
        PIC_HIP(hipStreamSynchronize(s));
        *loss = *u->loss_host;
    }
    return PIC_OK;
}

} // namespace pic
