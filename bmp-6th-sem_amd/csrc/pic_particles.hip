// pic_particles.hip — gfx950 kernels for the particle stages of the PIC step:
// encode/bin (counting sort), CIC deposit (int64 fixed point, per-cell corner
// sums), gather + Boris push + wall absorption + re-bin.
//
// Floating-point contraction is disabled for this file (-ffp-contract=off in
// the build) so that every FP64 expression rounds exactly as written; the
// oracle is compiled the same way.
//
// Storage is cell-sorted structure-of-arrays; `start[k]` is the first slot of
// key k.  Two layouts:
//   F64    — x, y, vx, vy, vz as double; the push is the reference's FP64
//            expression sequence (mover.cpp:103-176) bit for bit.
//   CELL32 — int32 key + float in-cell offsets (units of dx, dy) + float
//            velocities; the push runs in FP64 in cell units: bilinear weights
//            from the offsets, the reference's Boris step, and the new position
//            t = (cell + offset) + v * dt / dx.
#include "pic_internal.h"

#include <cmath>

namespace pic {

// ---------------------------------------------------------------- layouts
template <int L> struct Lay;
template <> struct Lay<PIC_LAYOUT_F64>    { using P = double; using V = double; };
template <> struct Lay<PIC_LAYOUT_CELL32> { using P = float;  using V = float;  };

static constexpr int kBlock = 256;

static inline unsigned blocks_for(int64_t n, int per = kBlock) {
    return (unsigned)((n + per - 1) / per);
}

// Workgroup -> tile order.  Workgroups are dealt round-robin to the 8 XCDs
// (each with its own L2); giving each XCD a contiguous band of tiles keeps the
// field rows and cell starts a tile touches in that XCD's L2.
static constexpr int kXcd = 8;
__device__ __forceinline__ int xcd_tile(int b, int ntiles) {
    const int per = (ntiles + kXcd - 1) / kXcd;
    const int x = b % kXcd, k = b / kXcd;
    const int full = ntiles - (per - 1) * kXcd;     // XCDs that get `per` tiles
    // XCD x owns tiles [first(x), first(x) + count(x)); workgroup b is its k-th
    const int first = x < full ? x * per : full * per + (x - full) * (per - 1);
    return first + k;
}

// Round to nearest (ties to even) of 0 <= v < 2^51 by the 1.5 * 2^52
// addition: the sum's last mantissa bit is the units digit.  Larger values
// take the library conversion.  Same result as llrint / __double2ll_rn.
__device__ __forceinline__ long long round_fixed(double v) {
    if (v < 0x1p51) {
        double t = v + 0x1.8p52;
        return __double_as_longlong(t) - 0x4338000000000000LL;
    }
    return __double2ll_rn(v);
}

// Largest float below 1: stored offsets stay in [0, 1).
__device__ __forceinline__ float clamp_offset(float o) {
    return o < 1.0f ? o : 0x1.fffffep-1f;
}

// key -> (key % GRID_X, key / GRID_X) exactly: float quotient estimate
// (within one of the true quotient for keys < 2^24) plus an integer fix-up.
__device__ __forceinline__ void split_key(int k, const GridDev& g, int& kx, int& ky) {
    int q = (int)((float)k * g.rcp_gx);
    int r = k - q * g.gx;
    if (r < 0) { q -= 1; r += g.gx; }
    else if (r >= g.gx) { q += 1; r -= g.gx; }
    kx = r;
    ky = q;
}

// ================================================================ encode
// Upload path: absolute FP64 particles -> storage layout + reference key
// (int)(x * inv_dx), (int)(y * inv_dy) (particles.cpp:229-235).
template <int L>
__global__ __launch_bounds__(kBlock) void k_encode(
    int n, const double* __restrict__ x, const double* __restrict__ y,
    const double* __restrict__ vx, const double* __restrict__ vy, const double* __restrict__ vz,
    ParticleBuf out, int32_t* __restrict__ key, GridDev g)
{
    using P = typename Lay<L>::P;
    using V = typename Lay<L>::V;
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double px = x[i], py = y[i];
    double tx = px * g.inv_dx, ty = py * g.inv_dy;
    int gx = (int)tx, gy = (int)ty;
    int k = gy * g.gx + gx;
    if (L == PIC_LAYOUT_F64) {
        ((P*)out.pa)[i] = (P)px;
        ((P*)out.pb)[i] = (P)py;
    } else {
        ((P*)out.pa)[i] = clamp_offset((float)(tx - (double)gx));
        ((P*)out.pb)[i] = clamp_offset((float)(ty - (double)gy));
    }
    ((V*)out.vx)[i] = (V)vx[i];
    ((V*)out.vy)[i] = (V)vy[i];
    ((V*)out.vz)[i] = (V)vz[i];
    out.cell[i] = k;
    key[i] = k;
}

__global__ void k_copy_i32(int n, const int32_t* __restrict__ a, int32_t* __restrict__ b) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = a[i];
}

__global__ void k_set_i32(int32_t* p, int v) { *p = v; }

// ================================================================== bin
// Counting sort of key[] (-1 = absorbed): count (slot inside the cell),
// exclusive scan of the counts (which also resets them), scatter.
__global__ __launch_bounds__(kBlock) void k_count(
    const int32_t* __restrict__ nsrc, const int32_t* __restrict__ key,
    int32_t* __restrict__ counts, int32_t* __restrict__ rank)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= *nsrc) return;
    int k = key[i];
    rank[i] = k >= 0 ? atomicAdd(&counts[k], 1) : -1;
}

__device__ __forceinline__ int wave_incl_scan(int v) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        int t = __shfl_up(v, d, 64);
        if (lane >= d) v += t;
    }
    return v;
}

static constexpr int kScanPer = 8;
static constexpr int kScanChunk = kBlock * kScanPer;   // 2048 counts per block

__global__ __launch_bounds__(kBlock) void k_scan_local(
    int32_t* __restrict__ counts, int G, int32_t* __restrict__ start, int32_t* __restrict__ bsum)
{
    __shared__ int wsum[kBlock / 64];
    const int t = threadIdx.x, lane = t & 63, wid = t >> 6;
    const int base = blockIdx.x * kScanChunk + t * kScanPer;
    int v[kScanPer];
    int s = 0;
#pragma unroll
    for (int j = 0; j < kScanPer; ++j) {
        int idx = base + j;
        v[j] = idx < G ? counts[idx] : 0;
        s += v[j];
    }
#pragma unroll
    for (int j = 0; j < kScanPer; ++j) {
        int idx = base + j;
        if (idx < G) counts[idx] = 0;
    }
    int incl = wave_incl_scan(s);
    if (lane == 63) wsum[wid] = incl;
    __syncthreads();
    int wpre = 0;
    for (int w = 0; w < wid; ++w) wpre += wsum[w];
    int ex = wpre + incl - s;
#pragma unroll
    for (int j = 0; j < kScanPer; ++j) {
        int idx = base + j;
        if (idx < G) start[idx] = ex;
        ex += v[j];
    }
    if (t == kBlock - 1) bsum[blockIdx.x] = wpre + incl;
}

// One workgroup of 1024 threads scans up to 8192 block sums in place.
__global__ __launch_bounds__(1024) void k_scan_top(
    int32_t* __restrict__ bsum, int nb, int32_t* __restrict__ start, int G, int32_t* __restrict__ ndst)
{
    __shared__ int wsum[16];
    const int t = threadIdx.x, lane = t & 63, wid = t >> 6;
    int v[8];
    int s = 0;
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        int idx = t * 8 + j;
        v[j] = idx < nb ? bsum[idx] : 0;
        s += v[j];
    }
    int incl = wave_incl_scan(s);
    if (lane == 63) wsum[wid] = incl;
    __syncthreads();
    int wpre = 0;
    for (int w = 0; w < wid; ++w) wpre += wsum[w];
    int ex = wpre + incl - s;
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        int idx = t * 8 + j;
        if (idx < nb) bsum[idx] = ex;
        ex += v[j];
    }
    if (t == 1023) {
        int total = wpre + incl;
        start[G] = total;
        *ndst = total;
    }
}

__global__ __launch_bounds__(kBlock) void k_scan_add(
    int32_t* __restrict__ start, int G, const int32_t* __restrict__ boff)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < G) start[i] += boff[i / kScanChunk];
}

template <int L>
__global__ __launch_bounds__(kBlock) void k_scatter(
    const int32_t* __restrict__ nsrc, ParticleBuf src, ParticleBuf dst,
    const int32_t* __restrict__ key, const int32_t* __restrict__ rank,
    const int32_t* __restrict__ start)
{
    using P = typename Lay<L>::P;
    using V = typename Lay<L>::V;
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= *nsrc) return;
    int k = key[i];
    if (k < 0) return;
    int d = start[k] + rank[i];
    ((P*)dst.pa)[d] = ((const P*)src.pa)[i];
    ((P*)dst.pb)[d] = ((const P*)src.pb)[i];
    ((V*)dst.vx)[d] = ((const V*)src.vx)[i];
    ((V*)dst.vy)[d] = ((const V*)src.vy)[i];
    ((V*)dst.vz)[d] = ((const V*)src.vz)[i];
    dst.cell[d] = k;
}

// ============================================================== deposit
// CIC weights normalised by dx*dy, rounded to int64 at 2^charge_shift
// (energy weights at 2^energy_shift eV) and summed as integers: the result
// does not depend on summation order.  Weights follow interpolate
// (interpolation.cpp:72-98): for key cell (gx, gy), lx = x - gx*dx,
// ly = y - gy*dy; corners (gx,gy) (dx-lx)(dy-ly), (gx+1,gy) lx(dy-ly),
// (gx,gy+1) (dx-lx)ly, (gx+1,gy+1) lx*ly; CELL32 uses the stored offsets
// (1-ox)(1-oy), ox(1-oy), (1-ox)oy, ox*oy.  Cells on the last node row or
// column deposit nothing (InterReductionKernel returns, :122).
struct DepositArgs {
    double inv_area;   // 1 / (dx*dy)
    double scale_q;    // 2^charge_shift
    double scale_e;    // 2^energy_shift
    double scale_eq;   // 2^(energy_shift - charge_shift)
    double half_x;     // 2^(charge_shift / 2)
    double half_y;     // 2^(charge_shift - charge_shift / 2)
    double denom;      // 1 / |q/m|
    int    energy;
};

template <int L>
__device__ __forceinline__ void cic_weights(const ParticleBuf& b, int i, int gxk, int gyk,
                                            const GridDev& g, const DepositArgs& a, double w[4]) {
    if (L == PIC_LAYOUT_F64) {
        double x = ((const double*)b.pa)[i], y = ((const double*)b.pb)[i];
        double lx = x - gxk * g.dx;
        double ly = y - gyk * g.dy;
        w[0] = ((g.dx - lx) * (g.dy - ly)) * a.inv_area;
        w[1] = (lx * (g.dy - ly)) * a.inv_area;
        w[2] = ((g.dx - lx) * ly) * a.inv_area;
        w[3] = (lx * ly) * a.inv_area;
    } else {
        double fx = (double)((const float*)b.pa)[i], fy = (double)((const float*)b.pb)[i];
        w[0] = (1.0 - fx) * (1.0 - fy);
        w[1] = fx * (1.0 - fy);
        w[2] = (1.0 - fx) * fy;
        w[3] = fx * fy;
    }
}

template <int L>
__device__ __forceinline__ double particle_energy(const ParticleBuf& b, int i, const DepositArgs& a) {
    using V = typename Lay<L>::V;
    double pvx = (double)((const V*)b.vx)[i];
    double pvy = (double)((const V*)b.vy)[i];
    double pvz = (double)((const V*)b.vz)[i];
    double v2 = pvx * pvx + pvy * pvy + pvz * pvz;
    return 0.5 * v2 * a.denom;
}

// One workgroup per tile of kDepC consecutive cells of one key row (the
// reference's per-cell reduction and corner meshes, :56-148).  The tile's
// particles are contiguous in cell-sorted storage; lane t of the workgroup
// takes groups of kVec consecutive particles (one 16-byte load per field, so
// a wave reads whole cache lines), adds their fixed-point corner weights in
// registers while the cell stays the same, and adds the running sums to
// per-cell int64 accumulators in LDS when the cell changes (a cell of ~16
// particles is flushed by one to three lanes).  Thread t then writes cell
// c0 + t's four corner sums to the node-indexed corner meshes: corner c of
// cell k goes to node k, k+1, k+GX, k+GX+1.  Every cell is written by exactly
// one thread, so no global atomics; cells on the last node row or column
// write zeros (the reference skips them, :122).
//
// Weights are formed already scaled by 2^S: for CELL32 (1 - ox) 2^h *
// (1 - oy) 2^(S-h) rounds exactly like ((1 - ox)(1 - oy)) 2^S (scaling by a
// power of two commutes with rounding), and the energy weight
// w 2^S * (en 2^(E-S)) like (w en) 2^E, so the sums equal the oracle's bit
// for bit.
static constexpr int kDepC = 256;
static constexpr int kVec = 4;

template <int L> struct Quad {
    using P = typename Lay<L>::P;
    using V = typename Lay<L>::V;
    int cell[kVec];
    P a[kVec], b[kVec];
    V vx[kVec], vy[kVec], vz[kVec];
};

// kVec consecutive entries starting at a multiple of kVec (16/32-byte aligned).
template <typename T>
__device__ __forceinline__ void load_quad(const T* __restrict__ p, int i, T out[kVec]);
template <>
__device__ __forceinline__ void load_quad<float>(const float* __restrict__ p, int i, float out[kVec]) {
    const float4 v = *(const float4*)(p + i);
    out[0] = v.x; out[1] = v.y; out[2] = v.z; out[3] = v.w;
}
template <>
__device__ __forceinline__ void load_quad<int>(const int* __restrict__ p, int i, int out[kVec]) {
    const int4 v = *(const int4*)(p + i);
    out[0] = v.x; out[1] = v.y; out[2] = v.z; out[3] = v.w;
}
template <>
__device__ __forceinline__ void load_quad<double>(const double* __restrict__ p, int i, double out[kVec]) {
    const double2 u = *(const double2*)(p + i), v = *(const double2*)(p + i + 2);
    out[0] = u.x; out[1] = u.y; out[2] = v.x; out[3] = v.y;
}

template <int L>
__global__ __launch_bounds__(kBlock) void k_deposit_tile(
    ParticleBuf b, GridDev g, DepositArgs a, int tiles_per_row, int ntiles, const int32_t* __restrict__ start,
    long long* __restrict__ cq, long long* __restrict__ ce)
{
    using P = typename Lay<L>::P;
    using V = typename Lay<L>::V;
    __shared__ unsigned long long acc[8][kDepC];
    const int tid = threadIdx.x;
    const int tile = xcd_tile(blockIdx.x, ntiles);
    const int r = tile / tiles_per_row;
    const int c0 = (tile % tiles_per_row) * kDepC;
    const int cw = min(kDepC, g.gx - c0);
    const int k0 = r * g.gx + c0;
#pragma unroll
    for (int c = 0; c < 8; ++c) acc[c][tid] = 0ull;
    __syncthreads();
    const int beg = start[k0], end = start[k0 + cw];
    const int first = beg & ~(kVec - 1);
    for (int base = first + kVec * tid; base < end; base += kVec * kBlock) {
        Quad<L> q;
        load_quad<int>(b.cell, base, q.cell);
        load_quad<P>((const P*)b.pa, base, q.a);
        load_quad<P>((const P*)b.pb, base, q.b);
        if (a.energy) {
            load_quad<V>((const V*)b.vx, base, q.vx);
            load_quad<V>((const V*)b.vy, base, q.vy);
            load_quad<V>((const V*)b.vz, base, q.vz);
        }
        int cur = -1;
        long long s0 = 0, s1 = 0, s2 = 0, s3 = 0, e0 = 0, e1 = 0, e2 = 0, e3 = 0;
        auto flush = [&]() {
            const int lc = cur - k0;
            atomicAdd(&acc[0][lc], (unsigned long long)s0);
            atomicAdd(&acc[1][lc], (unsigned long long)s1);
            atomicAdd(&acc[2][lc], (unsigned long long)s2);
            atomicAdd(&acc[3][lc], (unsigned long long)s3);
            if (a.energy) {
                atomicAdd(&acc[4][lc], (unsigned long long)e0);
                atomicAdd(&acc[5][lc], (unsigned long long)e1);
                atomicAdd(&acc[6][lc], (unsigned long long)e2);
                atomicAdd(&acc[7][lc], (unsigned long long)e3);
            }
        };
#pragma unroll
        for (int i = 0; i < kVec; ++i) {
            const int pi = base + i;
            if (pi < beg || pi >= end) continue;
            const int cl = q.cell[i];
            if (cl != cur) {
                if (cur >= 0) flush();
                cur = cl;
                s0 = s1 = s2 = s3 = e0 = e1 = e2 = e3 = 0;
            }
            double w0, w1, w2, w3;
            if (L == PIC_LAYOUT_F64) {
                int kx, ky;
                split_key(cl, g, kx, ky);
                double lx = (double)q.a[i] - kx * g.dx;
                double ly = (double)q.b[i] - ky * g.dy;
                w0 = (((g.dx - lx) * (g.dy - ly)) * a.inv_area) * a.scale_q;
                w1 = ((lx * (g.dy - ly)) * a.inv_area) * a.scale_q;
                w2 = (((g.dx - lx) * ly) * a.inv_area) * a.scale_q;
                w3 = ((lx * ly) * a.inv_area) * a.scale_q;
            } else {
                const double fx = (double)q.a[i], fy = (double)q.b[i];
                const double ux = (1.0 - fx) * a.half_x, vx_ = fx * a.half_x;
                const double uy = (1.0 - fy) * a.half_y, vy_ = fy * a.half_y;
                w0 = ux * uy;
                w1 = vx_ * uy;
                w2 = ux * vy_;
                w3 = vx_ * vy_;
            }
            s0 += round_fixed(w0);
            s1 += round_fixed(w1);
            s2 += round_fixed(w2);
            s3 += round_fixed(w3);
            if (a.energy) {
                const double pvx = (double)q.vx[i], pvy = (double)q.vy[i], pvz = (double)q.vz[i];
                const double v2 = pvx * pvx + pvy * pvy + pvz * pvz;
                const double en = 0.5 * v2 * a.denom;
                const double es = en * a.scale_eq;    // exact: power of two
                e0 += round_fixed(w0 * es);
                e1 += round_fixed(w1 * es);
                e2 += round_fixed(w2 * es);
                e3 += round_fixed(w3 * es);
            }
        }
        if (cur >= 0) flush();
    }
    __syncthreads();
    if (tid >= cw) return;
    const bool live = (r + 1 < g.gy) && (c0 + tid + 1 < g.gx);
    const int gp = g.gp;
    const int k = k0 + tid;
    cq[k] = live ? (long long)acc[0][tid] : 0;
    cq[gp + k + 1] = live ? (long long)acc[1][tid] : 0;
    cq[2 * gp + k + g.gx] = live ? (long long)acc[2][tid] : 0;
    cq[3 * gp + k + g.gx + 1] = live ? (long long)acc[3][tid] : 0;
    if (a.energy) {
        ce[k] = live ? (long long)acc[4][tid] : 0;
        ce[gp + k + 1] = live ? (long long)acc[5][tid] : 0;
        ce[2 * gp + k + g.gx] = live ? (long long)acc[6][tid] : 0;
        ce[3 * gp + k + g.gx + 1] = live ? (long long)acc[7][tid] : 0;
    }
}

// Storage not cell-sorted (after an in-place push): per-particle 64-bit
// integer atomics into corner mesh 0 (the other three were zeroed).
template <int L>
__global__ __launch_bounds__(kBlock) void k_deposit_atomic(
    const int32_t* __restrict__ nsrc, ParticleBuf b, GridDev g, DepositArgs a,
    unsigned long long* __restrict__ cq, unsigned long long* __restrict__ ce)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= *nsrc) return;
    int k = b.cell[i];
    if (k < 0) return;
    int kx, ky;
    split_key(k, g, kx, ky);
    if (kx + 1 >= g.gx || ky + 1 >= g.gy) return;
    double w[4];
    cic_weights<L>(b, i, kx, ky, g, a, w);
    int node[4] = {k, k + 1, k + g.gx, k + g.gx + 1};
#pragma unroll
    for (int c = 0; c < 4; ++c)
        atomicAdd(cq + node[c], (unsigned long long)__double2ll_rn(w[c] * a.scale_q));
    if (a.energy) {
        double en = particle_energy<L>(b, i, a);
#pragma unroll
        for (int c = 0; c < 4; ++c)
            atomicAdd(ce + node[c], (unsigned long long)__double2ll_rn((w[c] * en) * a.scale_e));
    }
}

__global__ __launch_bounds__(kBlock) void k_zero_u64(unsigned long long* p, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0ull;
}

__global__ __launch_bounds__(kBlock) void k_zero_f64(double* p, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.0;
}

// Node sum of the corner meshes (ReducePartialMeshesKernel :150-163),
// final_mesh = W * dx*dy * 2^-S * Q, rho += final_mesh * (-q/eps)
// (RhoKernel :165-171).
__global__ __launch_bounds__(kBlock) void k_deposit_finalize(
    int G, int gp, const long long* __restrict__ cq, const long long* __restrict__ ce,
    double cqf, double cef, double chrg, int energy,
    double* __restrict__ final_mesh, double* __restrict__ energy_mesh, double* __restrict__ rho)
{
    int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= G) return;
    long long q = cq[n] + cq[gp + n] + cq[2 * gp + n] + cq[3 * gp + n];
    double f = (double)q * cqf;
    final_mesh[n] = f;
    rho[n] += f * chrg;
    if (energy) {
        long long e = ce[n] + ce[gp + n] + ce[2 * gp + n] + ce[3 * gp + n];
        energy_mesh[n] = (double)e * cef;
    }
}

// ================================================================= push
// F64 gather (deterministic_mover, mover.cpp:103-133): gx = (int)(x/dx) — a
// division, unlike the bin key — lx = x - gx*dx, bilinear weights over
// (dx*dy).  The field arrays are padded with zero nodes so the reference's
// flat indexing (gy*GRID_X + gx + 1 etc.) stays in bounds at the last node.
// Boris (mover.cpp:151-166) for Bz; when bx or by is non-zero the same
// rotation is done with the full vector t = (q/m) B dt / 2.
// Boundary (mover.cpp:325,331-352): absorbed when x >= W, x < 0, y >= H or
// y < 0 (CELL32: the same walls in cell units); particles leaving through
// x < 0 are counted separately.

// Particle state after the push, in storage form.
template <int L> struct Pushed {
    using P = typename Lay<L>::P;
    using V = typename Lay<L>::V;
    int key;            // new bin key, -1 when absorbed
    bool left;          // absorbed through x < 0
    P a, b;             // F64: x, y   CELL32: offsets
    V vx, vy, vz;
};

template <int GENB>
__device__ __forceinline__ void boris(double Epx, double Epy, double Bpz, const SpeciesDev& s,
                                      double& pvx, double& pvy, double& pvz) {
    double hx = s.half_qm * Epx * s.dt;     // 0.5 * pctm * Epx * dt
    double hy = s.half_qm * Epy * s.dt;
    double Vxminus = pvx + hx;
    double Vyminus = pvy + hy;
    if (!GENB) {
        double t = s.half_qm * Bpz * s.dt;
        double Vxprime = Vxminus + Vyminus * t;
        double Vyprime = Vyminus - Vxminus * t;
        double sf = 2 * t / (1 + t * t);
        pvx = Vxminus + Vyprime * sf + hx;
        pvy = Vyminus - Vxprime * sf + hy;
    } else {
        double tx = s.half_qm * s.bx * s.dt;
        double ty = s.half_qm * s.by * s.dt;
        double tz = s.half_qm * Bpz * s.dt;
        double Vzminus = pvz;
        double vpx = Vxminus + (Vyminus * tz - Vzminus * ty);
        double vpy = Vyminus + (Vzminus * tx - Vxminus * tz);
        double vpz = Vzminus + (Vxminus * ty - Vyminus * tx);
        double rden = 1 / (1 + (tx * tx + ty * ty + tz * tz));
        double sx = (2 * tx) * rden, sy = (2 * ty) * rden, sz = (2 * tz) * rden;
        pvx = Vxminus + (vpy * sz - vpz * sy) + hx;
        pvy = Vyminus + (vpz * sx - vpx * sz) + hy;
        pvz = Vzminus + (vpx * sy - vpy * sx);
    }
}

// Particle state as loaded from storage.
template <int L> struct PState {
    using P = typename Lay<L>::P;
    using V = typename Lay<L>::V;
    int cell;           // CELL32 only
    P a, b;             // F64: x, y   CELL32: offsets
    V vx, vy, vz;
};

template <int L>
__device__ __forceinline__ PState<L> load_state(const ParticleBuf& b, int i) {
    using P = typename Lay<L>::P;
    using V = typename Lay<L>::V;
    PState<L> p;
    p.cell = L == PIC_LAYOUT_CELL32 ? b.cell[i] : 0;
    p.a = ((const P*)b.pa)[i];
    p.b = ((const P*)b.pb)[i];
    p.vx = ((const V*)b.vx)[i];
    p.vy = ((const V*)b.vy)[i];
    p.vz = ((const V*)b.vz)[i];
    return p;
}

// The four field nodes around a particle (loads only, so that callers can
// order them against other loads).
struct Corners {
    double4 f00, f10, f01, f11;
    int cgx, cgy;       // gather cell
};

template <int L>
__device__ __forceinline__ Corners gather(const PState<L>& p, const GridDev& g, const double4* __restrict__ fld) {
    Corners c;
    int i00;
    if (L == PIC_LAYOUT_F64) {
        c.cgx = (int)((double)p.a / g.dx);
        c.cgy = (int)((double)p.b / g.dy);
        i00 = c.cgy * g.gx + c.cgx;
    } else {
        split_key(p.cell, g, c.cgx, c.cgy);
        i00 = p.cell;
    }
    c.f00 = fld[i00];
    c.f10 = fld[i00 + 1];
    c.f01 = fld[i00 + g.gx];
    c.f11 = fld[i00 + g.gx + 1];
    return c;
}

template <int L, int GENB>
__device__ __forceinline__ Pushed<L> push_state(const PState<L>& p, const Corners& c, const GridDev& g,
                                               const SpeciesDev& s) {
    using P = typename Lay<L>::P;
    using V = typename Lay<L>::V;
    double pvx = (double)p.vx;
    double pvy = (double)p.vy;
    double pvz = (double)p.vz;
    const double4 f00 = c.f00, f10 = c.f10, f01 = c.f01, f11 = c.f11;
    Pushed<L> o;
    if (L == PIC_LAYOUT_F64) {
        double px = (double)p.a;
        double py = (double)p.b;
        double lx = px - c.cgx * g.dx;
        double ly = py - c.cgy * g.dy;
        double wx0 = g.dx - lx, wy0 = g.dy - ly;
        double area = g.dx * g.dy;
        double Epx = (f00.x * wx0 * wy0 + f10.x * lx * wy0 + f01.x * wx0 * ly + f11.x * lx * ly) / area;
        double Epy = (f00.y * wx0 * wy0 + f10.y * lx * wy0 + f01.y * wx0 * ly + f11.y * lx * ly) / area;
        double Bpz = (f00.z * wx0 * wy0 + f10.z * lx * wy0 + f01.z * wx0 * ly + f11.z * lx * ly) / area;
        boris<GENB>(Epx, Epy, Bpz, s, pvx, pvy, pvz);
        px = px + s.dt * pvx;
        py = py + s.dt * pvy;
        bool out = (px >= g.width || px < 0.0 || py >= g.height || py < 0);
        o.left = px < 0.0;
        if (out) {
            o.key = -1;
            return o;
        }
        int ngx = (int)(px * g.inv_dx), ngy = (int)(py * g.inv_dy);
        o.key = ngy * g.gx + ngx;
        o.a = (P)px;
        o.b = (P)py;
    } else {
        double fx = (double)p.a;
        double fy = (double)p.b;
        double w00 = (1.0 - fx) * (1.0 - fy), w10 = fx * (1.0 - fy);
        double w01 = (1.0 - fx) * fy, w11 = fx * fy;
        double Epx = f00.x * w00 + f10.x * w10 + f01.x * w01 + f11.x * w11;
        double Epy = f00.y * w00 + f10.y * w10 + f01.y * w01 + f11.y * w11;
        double Bpz = f00.z * w00 + f10.z * w10 + f01.z * w01 + f11.z * w11;
        boris<GENB>(Epx, Epy, Bpz, s, pvx, pvy, pvz);
        double tx = ((double)c.cgx + fx) + pvx * s.dtdx;
        double ty = ((double)c.cgy + fy) + pvy * s.dtdy;
        bool out = (tx >= g.nx || tx < 0.0 || ty >= g.ny || ty < 0.0);
        o.left = tx < 0.0;
        if (out) {
            o.key = -1;
            return o;
        }
        int ngx = (int)tx, ngy = (int)ty;
        o.key = ngy * g.gx + ngx;
        o.a = clamp_offset((float)(tx - (double)ngx));
        o.b = clamp_offset((float)(ty - (double)ngy));
    }
    o.vx = (V)pvx;
    o.vy = (V)pvy;
    o.vz = (V)pvz;
    return o;
}

template <int L, int GENB>
__device__ __forceinline__ Pushed<L> push_one(const ParticleBuf& b, int i, const GridDev& g, const SpeciesDev& s,
                                             const double4* __restrict__ fld) {
    PState<L> p = load_state<L>(b, i);
    Corners c = gather<L>(p, g, fld);
    return push_state<L, GENB>(p, c, g, s);
}

template <int L>
__device__ __forceinline__ void store_state(const ParticleBuf& d, int j, const Pushed<L>& o) {
    using P = typename Lay<L>::P;
    using V = typename Lay<L>::V;
    ((P*)d.pa)[j] = o.a;
    ((P*)d.pb)[j] = o.b;
    ((V*)d.vx)[j] = o.vx;
    ((V*)d.vy)[j] = o.vy;
    ((V*)d.vz)[j] = o.vz;
}

// In-place push without re-binning (stage API pic_push(ctx, s, 0)); the
// stored cell becomes the new key (-1 when absorbed).
template <int L, int GENB>
__global__ __launch_bounds__(kBlock) void k_push(
    const int32_t* __restrict__ nsrc, ParticleBuf b, GridDev g, SpeciesDev s,
    const double4* __restrict__ fld, int32_t* __restrict__ key,
    unsigned long long* __restrict__ absorbed)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= *nsrc) return;
    Pushed<L> o = push_one<L, GENB>(b, i, g, s, fld);
    key[i] = o.key;
    b.cell[i] = o.key;
    if (o.key < 0) {
        if (o.left) atomicAdd(absorbed, 1ull);
        atomicAdd(absorbed + 1, 1ull);
        return;
    }
    store_state<L>(b, i, o);
}

// ====================================================== tiled push + bin
// One workgroup per tile of kTileC consecutive cells of one key row; the
// tile's particles are the contiguous range [start[k0], start[k0 + w)).  The
// workgroup pushes them (no state is written) and counts the new keys in an
// LDS window over key rows r-1..r+1, columns c0-1..c0+kTileC (a particle
// moves less than a cell per step): one LDS atomic per particle gives its
// place in the window bin; the window then reserves each bin's slots with one
// global atomic, and every particle gets its slot inside its new cell.  Particles that land outside
// the window take a slot with a global atomic.  The per-particle codes wait
// in LDS (the first kCodeCap of the tile; the rest in `rank`) until the
// window's global offsets are known, so `rank` is written once.  While one
// particle is pushed the state of the thread's next particle is already
// being loaded.  k_scatter_push recomputes the identical push and writes each
// particle to start[key] + slot.
static constexpr int kTileC = 256;
static constexpr int kHW = kTileC + 2, kHR = 3;
static constexpr int kRankBits = 21;            // window bin (< 1024) in bits 21..30
static constexpr int kFallback = (int)0x80000000;
static constexpr int kCodeCap = 4096;

template <int L, int GENB>
__global__ __launch_bounds__(kBlock) void k_push_tile(
    ParticleBuf b, GridDev g, SpeciesDev s, int tiles_per_row, int ntiles, const double4* __restrict__ fld,
    const int32_t* __restrict__ start, int32_t* __restrict__ counts, int32_t* __restrict__ rank,
    unsigned long long* __restrict__ absorbed)
{
    __shared__ int hist[kHR * kHW];
    __shared__ int codes[kCodeCap];
    __shared__ unsigned long long nabs[2];
    const int tid = threadIdx.x;
    const int tile = xcd_tile(blockIdx.x, ntiles);
    const int r = tile / tiles_per_row;
    const int c0 = (tile % tiles_per_row) * kTileC;
    const int cw = min(kTileC, g.gx - c0);
    const int k0 = r * g.gx + c0;
    for (int j = tid; j < kHR * kHW; j += kBlock) hist[j] = 0;
    if (tid < 2) nabs[tid] = 0ull;
    __syncthreads();
    const int beg = start[k0], end = start[k0 + cw];
    PState<L> cur;
    if (beg + tid < end) cur = load_state<L>(b, beg + tid);
    for (int base = beg; base < end; base += kBlock) {
        const int i = base + tid;
        const bool have = i < end;
        Corners cn;
        if (have) cn = gather<L>(cur, g, fld);
        PState<L> nxt;
        if (i + kBlock < end) nxt = load_state<L>(b, i + kBlock);
        int bin = -1, code = -1;
        if (have) {
            Pushed<L> o = push_state<L, GENB>(cur, cn, g, s);
            if (o.key < 0) {
                if (o.left) atomicAdd(&nabs[0], 1ull);
                atomicAdd(&nabs[1], 1ull);
            } else {
                int gxn, gyn;
                split_key(o.key, g, gxn, gyn);
                const int dr = gyn - (r - 1), dc = gxn - (c0 - 1);
                if (dr >= 0 && dr < kHR && dc >= 0 && dc < kHW) bin = dr * kHW + dc;
                else code = kFallback | atomicAdd(&counts[o.key], 1);
            }
        }
        if (bin >= 0) code = (bin << kRankBits) | atomicAdd(&hist[bin], 1);
        if (have) {
            if (i - beg < kCodeCap) codes[i - beg] = code;
            else rank[i] = code;
        }
        cur = nxt;
    }
    __syncthreads();
    for (int j = tid; j < kHR * kHW; j += kBlock) {
        int cnt = hist[j];
        if (cnt) {
            int k = (r - 1 + j / kHW) * g.gx + (c0 - 1 + j % kHW);
            hist[j] = atomicAdd(&counts[k], cnt);
        }
    }
    if (tid == 0) {
        if (nabs[0]) atomicAdd(absorbed, nabs[0]);
        if (nabs[1]) atomicAdd(absorbed + 1, nabs[1]);
    }
    __syncthreads();
    for (int i = beg + tid; i < end; i += kBlock) {
        const int li = i - beg;
        const int v = li < kCodeCap ? codes[li] : rank[i];
        if (v == -1) continue;                      // absorbed: the scatter skips it
        if (v & kFallback) rank[i] = v & ~kFallback;
        else rank[i] = hist[v >> kRankBits] + (v & ((1 << kRankBits) - 1));
    }
}

// Recompute the push and write each surviving particle to its sorted slot.
template <int L, int GENB>
__global__ __launch_bounds__(kBlock) void k_scatter_push(
    const int32_t* __restrict__ nsrc, ParticleBuf src, ParticleBuf dst, GridDev g, SpeciesDev s,
    const double4* __restrict__ fld, const int32_t* __restrict__ rank, const int32_t* __restrict__ start)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= *nsrc) return;
    const PState<L> p = load_state<L>(src, i);
    const int rk = rank[i];
    const Corners c = gather<L>(p, g, fld);
    const Pushed<L> o = push_state<L, GENB>(p, c, g, s);
    if (o.key < 0) return;
    const int d = start[o.key] + rk;
    store_state<L>(dst, d, o);
    dst.cell[d] = o.key;
}

// ============================================================ launchers
int launch_zero_f64(pic_ctx* c, double* p, int64_t n) {
    if (n <= 0) return PIC_OK;
    hipLaunchKernelGGL(k_zero_f64, dim3(blocks_for(n)), dim3(kBlock), 0, c->stream, p, n);
    PIC_HIP(hipGetLastError());
    return PIC_OK;
}

int launch_encode_f64(pic_ctx* c, int s, int64_t n, const double* x, const double* y,
                      const double* vx, const double* vy, const double* vz) {
    Species& sp = c->sp[s];
    if (n > 0) {
        if (c->p.layout == PIC_LAYOUT_F64)
            hipLaunchKernelGGL(k_encode<PIC_LAYOUT_F64>, dim3(blocks_for(n)), dim3(kBlock), 0, c->stream,
                               (int)n, x, y, vx, vy, vz, sp.buf[sp.cur], sp.key, c->g);
        else
            hipLaunchKernelGGL(k_encode<PIC_LAYOUT_CELL32>, dim3(blocks_for(n)), dim3(kBlock), 0, c->stream,
                               (int)n, x, y, vx, vy, vz, sp.buf[sp.cur], sp.key, c->g);
        PIC_HIP(hipGetLastError());
    }
    hipLaunchKernelGGL(k_set_i32, dim3(1), dim3(1), 0, c->stream, sp.ndev + sp.cur, (int)n);
    PIC_HIP(hipGetLastError());
    sp.n_upper = n;
    sp.deposit_pending = 0;
    return PIC_OK;
}

int launch_key_from_cell(pic_ctx* c, int s, int64_t n) {
    Species& sp = c->sp[s];
    if (n > 0) {
        hipLaunchKernelGGL(k_copy_i32, dim3(blocks_for(n)), dim3(kBlock), 0, c->stream,
                           (int)n, sp.buf[sp.cur].cell, sp.key);
        PIC_HIP(hipGetLastError());
    }
    hipLaunchKernelGGL(k_set_i32, dim3(1), dim3(1), 0, c->stream, sp.ndev + sp.cur, (int)n);
    PIC_HIP(hipGetLastError());
    sp.n_upper = n;
    return PIC_OK;
}

static int scan_counts(pic_ctx* c, Species& sp, int dst) {
    hipLaunchKernelGGL(k_scan_local, dim3(c->scan_nb), dim3(kBlock), 0, c->stream,
                       sp.counts, c->G, sp.start, c->scan_bsum);
    hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(1024), 0, c->stream,
                       c->scan_bsum, c->scan_nb, sp.start, c->G, sp.ndev + dst);
    hipLaunchKernelGGL(k_scan_add, dim3(blocks_for(c->G)), dim3(kBlock), 0, c->stream,
                       sp.start, c->G, c->scan_bsum);
    PIC_HIP(hipGetLastError());
    return PIC_OK;
}

// Counting sort of the current storage by sp.key (upload / pic_bin / after an
// in-place push).
int launch_bin(pic_ctx* c, int s) {
    Species& sp = c->sp[s];
    hipEvent_t ev;
    PIC_TRY(timer_begin(c, PIC_STAGE_BIN, &ev));
    const int src = sp.cur, dst = sp.cur ^ 1;
    const int64_t n = sp.n_upper;
    if (n > 0) {
        hipLaunchKernelGGL(k_count, dim3(blocks_for(n)), dim3(kBlock), 0, c->stream,
                           sp.ndev + src, sp.key, sp.counts, sp.rank);
        PIC_HIP(hipGetLastError());
    }
    PIC_TRY(scan_counts(c, sp, dst));
    if (n > 0) {
        if (c->p.layout == PIC_LAYOUT_F64)
            hipLaunchKernelGGL(k_scatter<PIC_LAYOUT_F64>, dim3(blocks_for(n)), dim3(kBlock), 0, c->stream,
                               sp.ndev + src, sp.buf[src], sp.buf[dst], sp.key, sp.rank, sp.start);
        else
            hipLaunchKernelGGL(k_scatter<PIC_LAYOUT_CELL32>, dim3(blocks_for(n)), dim3(kBlock), 0, c->stream,
                               sp.ndev + src, sp.buf[src], sp.buf[dst], sp.key, sp.rank, sp.start);
        PIC_HIP(hipGetLastError());
    }
    sp.cur = dst;
    sp.sorted = 1;
    PIC_TRY(timer_end(c, PIC_STAGE_BIN, ev));
    return PIC_OK;
}

static DepositArgs deposit_args(const pic_ctx* c, const Species& sp) {
    DepositArgs a;
    const double dx = c->g.dx, dy = c->g.dy;
    a.inv_area = 1.0 / (dx * dy);
    a.scale_q = std::ldexp(1.0, c->p.charge_shift);
    a.scale_e = std::ldexp(1.0, c->p.energy_shift);
    a.scale_eq = std::ldexp(1.0, c->p.energy_shift - c->p.charge_shift);
    a.half_x = std::ldexp(1.0, c->p.charge_shift / 2);
    a.half_y = std::ldexp(1.0, c->p.charge_shift - c->p.charge_shift / 2);
    a.denom = 1 / std::fabs(sp.qm);
    a.energy = c->p.deposit_energy;
    return a;
}

static SpeciesDev species_dev(const pic_ctx* c, const Species& sp) {
    SpeciesDev d;
    d.qm = sp.qm;
    d.half_qm = 0.5 * sp.qm;
    d.denom = 1 / std::fabs(sp.qm);
    d.dt = c->p.dt;
    d.dtdx = c->p.dt * c->g.inv_dx;
    d.dtdy = c->p.dt * c->g.inv_dy;
    d.bx = c->p.bx;
    d.by = c->p.by;
    d.general_b = (c->p.bx != 0.0 || c->p.by != 0.0);
    return d;
}

static int tiles_per_row(const pic_ctx* c) { return (c->g.gx + kTileC - 1) / kTileC; }

static int finalize_deposit(pic_ctx* c, Species& sp) {
    const double dx = c->g.dx, dy = c->g.dy;
    const double cqf = sp.W * (dx * dy) * std::ldexp(1.0, -c->p.charge_shift);
    const double cef = sp.W * (dx * dy) * std::ldexp(1.0, -c->p.energy_shift);
    const double chrg = -sp.charge / c->p.epsilon;
    hipLaunchKernelGGL(k_deposit_finalize, dim3(blocks_for(c->G)), dim3(kBlock), 0, c->stream,
                       c->G, c->g.gp, (const long long*)sp.q_charge, (const long long*)sp.q_energy,
                       cqf, cef, chrg, c->p.deposit_energy, sp.final_mesh, sp.energy_mesh, c->rho);
    PIC_HIP(hipGetLastError());
    return PIC_OK;
}

// Charge (and energy) deposit of species s for its current particles:
// per-cell corner sums on cell-sorted storage, per-particle atomics
// otherwise, then node sums into final_mesh / energy_mesh and rho.
int launch_deposit(pic_ctx* c, int s) {
    Species& sp = c->sp[s];
    hipEvent_t ev;
    PIC_TRY(timer_begin(c, PIC_STAGE_DEPOSIT, &ev));
    DepositArgs a = deposit_args(c, sp);
    if (!sp.deposit_pending) {
        if (sp.sorted) {
            const int tpr = (c->g.gx + kDepC - 1) / kDepC;
            const unsigned ntiles = (unsigned)(tpr * c->g.gy);
            if (c->p.layout == PIC_LAYOUT_F64)
                hipLaunchKernelGGL(k_deposit_tile<PIC_LAYOUT_F64>, dim3(ntiles), dim3(kBlock), 0,
                                   c->stream, sp.buf[sp.cur], c->g, a, tpr, (int)ntiles, sp.start,
                                   (long long*)sp.q_charge, (long long*)sp.q_energy);
            else
                hipLaunchKernelGGL(k_deposit_tile<PIC_LAYOUT_CELL32>, dim3(ntiles), dim3(kBlock), 0,
                                   c->stream, sp.buf[sp.cur], c->g, a, tpr, (int)ntiles, sp.start,
                                   (long long*)sp.q_charge, (long long*)sp.q_energy);
        } else {
            const int64_t nacc = 4 * (int64_t)c->g.gp;
            hipLaunchKernelGGL(k_zero_u64, dim3(blocks_for(nacc)), dim3(kBlock), 0, c->stream,
                               (unsigned long long*)sp.q_charge, nacc);
            hipLaunchKernelGGL(k_zero_u64, dim3(blocks_for(nacc)), dim3(kBlock), 0, c->stream,
                               (unsigned long long*)sp.q_energy, nacc);
            if (sp.n_upper > 0) {
                if (c->p.layout == PIC_LAYOUT_F64)
                    hipLaunchKernelGGL(k_deposit_atomic<PIC_LAYOUT_F64>, dim3(blocks_for(sp.n_upper)), dim3(kBlock),
                                       0, c->stream, sp.ndev + sp.cur, sp.buf[sp.cur], c->g, a,
                                       (unsigned long long*)sp.q_charge, (unsigned long long*)sp.q_energy);
                else
                    hipLaunchKernelGGL(k_deposit_atomic<PIC_LAYOUT_CELL32>, dim3(blocks_for(sp.n_upper)),
                                       dim3(kBlock), 0, c->stream, sp.ndev + sp.cur, sp.buf[sp.cur], c->g, a,
                                       (unsigned long long*)sp.q_charge, (unsigned long long*)sp.q_energy);
            }
        }
        PIC_HIP(hipGetLastError());
        sp.deposit_pending = 1;
    }
    PIC_TRY(timer_end(c, PIC_STAGE_DEPOSIT, ev));
    PIC_TRY(timer_begin(c, PIC_STAGE_FINALIZE, &ev));
    PIC_TRY(finalize_deposit(c, sp));
    PIC_TRY(timer_end(c, PIC_STAGE_FINALIZE, ev));
    return PIC_OK;
}

// In-place push, no re-bin (keys left in sp.key).
int launch_push(pic_ctx* c, int s) {
    Species& sp = c->sp[s];
    hipEvent_t ev;
    PIC_TRY(timer_begin(c, PIC_STAGE_PUSH, &ev));
    SpeciesDev d = species_dev(c, sp);
    const int64_t n = sp.n_upper;
    if (n > 0) {
        dim3 grid(blocks_for(n)), blk(kBlock);
        const bool f64 = c->p.layout == PIC_LAYOUT_F64;
        if (f64 && !d.general_b)
            hipLaunchKernelGGL((k_push<PIC_LAYOUT_F64, 0>), grid, blk, 0, c->stream,
                               sp.ndev + sp.cur, sp.buf[sp.cur], c->g, d, c->fld, sp.key, sp.absorbed);
        else if (f64)
            hipLaunchKernelGGL((k_push<PIC_LAYOUT_F64, 1>), grid, blk, 0, c->stream,
                               sp.ndev + sp.cur, sp.buf[sp.cur], c->g, d, c->fld, sp.key, sp.absorbed);
        else if (!d.general_b)
            hipLaunchKernelGGL((k_push<PIC_LAYOUT_CELL32, 0>), grid, blk, 0, c->stream,
                               sp.ndev + sp.cur, sp.buf[sp.cur], c->g, d, c->fld, sp.key, sp.absorbed);
        else
            hipLaunchKernelGGL((k_push<PIC_LAYOUT_CELL32, 1>), grid, blk, 0, c->stream,
                               sp.ndev + sp.cur, sp.buf[sp.cur], c->g, d, c->fld, sp.key, sp.absorbed);
        PIC_HIP(hipGetLastError());
    }
    sp.deposit_pending = 0;
    sp.sorted = 0;
    PIC_TRY(timer_end(c, PIC_STAGE_PUSH, ev));
    return PIC_OK;
}

template <int L, int GENB>
static void tile_kernel(pic_ctx* c, Species& sp, const SpeciesDev& d, int tpr, unsigned ntiles) {
    hipLaunchKernelGGL((k_push_tile<L, GENB>), dim3(ntiles), dim3(kBlock), 0, c->stream,
                       sp.buf[sp.cur], c->g, d, tpr, (int)ntiles, c->fld, sp.start, sp.counts, sp.rank, sp.absorbed);
}

template <int L, int GENB>
static void scatter_kernel(pic_ctx* c, Species& sp, const SpeciesDev& d, int src, int dst) {
    hipLaunchKernelGGL((k_scatter_push<L, GENB>), dim3(blocks_for(sp.n_upper)), dim3(kBlock), 0, c->stream,
                       sp.ndev + src, sp.buf[src], sp.buf[dst], c->g, d, c->fld, sp.rank, sp.start);
}

// Gather + Boris + walls + re-bin of species s: tile push/count, prefix
// scan, recomputing scatter into the other buffer.
int launch_move(pic_ctx* c, int s) {
    Species& sp = c->sp[s];
    if (!sp.sorted) PIC_TRY(launch_bin(c, s));   // tiles need cell-sorted storage
    SpeciesDev d = species_dev(c, sp);
    const int tpr = tiles_per_row(c);
    const unsigned ntiles = (unsigned)(tpr * c->g.gy);
    const bool f64 = c->p.layout == PIC_LAYOUT_F64;
    const int src = sp.cur, dst = sp.cur ^ 1;
    hipEvent_t ev;
    PIC_TRY(timer_begin(c, PIC_STAGE_PUSH, &ev));
    if (sp.n_upper > 0) {
        if (f64 && !d.general_b) tile_kernel<PIC_LAYOUT_F64, 0>(c, sp, d, tpr, ntiles);
        else if (f64) tile_kernel<PIC_LAYOUT_F64, 1>(c, sp, d, tpr, ntiles);
        else if (!d.general_b) tile_kernel<PIC_LAYOUT_CELL32, 0>(c, sp, d, tpr, ntiles);
        else tile_kernel<PIC_LAYOUT_CELL32, 1>(c, sp, d, tpr, ntiles);
        PIC_HIP(hipGetLastError());
    }
    PIC_TRY(timer_end(c, PIC_STAGE_PUSH, ev));
    PIC_TRY(timer_begin(c, PIC_STAGE_BIN, &ev));
    PIC_TRY(scan_counts(c, sp, dst));
    PIC_TRY(timer_end(c, PIC_STAGE_BIN, ev));
    PIC_TRY(timer_begin(c, PIC_STAGE_SCATTER, &ev));
    if (sp.n_upper > 0) {
        if (f64 && !d.general_b) scatter_kernel<PIC_LAYOUT_F64, 0>(c, sp, d, src, dst);
        else if (f64) scatter_kernel<PIC_LAYOUT_F64, 1>(c, sp, d, src, dst);
        else if (!d.general_b) scatter_kernel<PIC_LAYOUT_CELL32, 0>(c, sp, d, src, dst);
        else scatter_kernel<PIC_LAYOUT_CELL32, 1>(c, sp, d, src, dst);
        PIC_HIP(hipGetLastError());
    }
    sp.cur = dst;
    sp.deposit_pending = 0;
    sp.sorted = 1;
    PIC_TRY(timer_end(c, PIC_STAGE_SCATTER, ev));
    return PIC_OK;
}

} // namespace pic
