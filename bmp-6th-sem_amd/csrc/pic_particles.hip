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

// One workgroup per tile of kDepC consecutive cells of one key row; thread
// t owns cell c0 + t — the reference's per-cell reduction and corner meshes
// (:56-148).  The tile's particles (contiguous in cell-sorted storage) are
// staged through LDS in chunks with coalesced loads; each thread then adds
// the fixed-point corner weights of its own cell's particles in registers
// and writes its four corner sums to the node-indexed corner meshes: corner c
// of cell k goes to node k, k+1, k+GX, k+GX+1.  Every cell is written by
// exactly one thread: no atomics; entries no cell writes stay zero.
static constexpr int kDepC = 256;
static constexpr int kDepChunk = 1024;        // particles staged per pass

template <int L> struct DepStage;
template <> struct DepStage<PIC_LAYOUT_F64> {
    double a[kDepChunk / 2], b[kDepChunk / 2], vx[kDepChunk / 2], vy[kDepChunk / 2], vz[kDepChunk / 2];
    static constexpr int n = kDepChunk / 2;
};
template <> struct DepStage<PIC_LAYOUT_CELL32> {
    float a[kDepChunk], b[kDepChunk], vx[kDepChunk], vy[kDepChunk], vz[kDepChunk];
    static constexpr int n = kDepChunk;
};

template <int L>
__global__ __launch_bounds__(kBlock) void k_deposit_tile(
    ParticleBuf b, GridDev g, DepositArgs a, int tiles_per_row, const int32_t* __restrict__ start,
    long long* __restrict__ cq, long long* __restrict__ ce)
{
    using P = typename Lay<L>::P;
    using V = typename Lay<L>::V;
    __shared__ DepStage<L> st;
    constexpr int CH = DepStage<L>::n;
    const int tid = threadIdx.x;
    const int r = blockIdx.x / tiles_per_row;
    const int c0 = (blockIdx.x % tiles_per_row) * kDepC;
    const int cw = min(kDepC, g.gx - c0);
    const int k0 = r * g.gx + c0;
    const int beg = start[k0], end = start[k0 + cw];
    const bool mine = tid < cw;
    const int kx = c0 + tid;
    const bool live = mine && (r + 1 < g.gy) && (kx + 1 < g.gx);
    const int ms = mine ? start[k0 + tid] : 0, me = mine ? start[k0 + tid + 1] : 0;
    long long s0 = 0, s1 = 0, s2 = 0, s3 = 0, e0 = 0, e1 = 0, e2 = 0, e3 = 0;
    for (int cb = beg; cb < end; cb += CH) {
        const int cn = min(CH, end - cb);
        __syncthreads();
        for (int j = tid; j < cn; j += kBlock) {
            st.a[j] = ((const P*)b.pa)[cb + j];
            st.b[j] = ((const P*)b.pb)[cb + j];
            if (a.energy) {
                st.vx[j] = ((const V*)b.vx)[cb + j];
                st.vy[j] = ((const V*)b.vy)[cb + j];
                st.vz[j] = ((const V*)b.vz)[cb + j];
            }
        }
        __syncthreads();
        if (!live) continue;
        const int lo = max(ms, cb) - cb, hi = min(me, cb + cn) - cb;
        for (int j = lo; j < hi; ++j) {
            double w[4];
            if (L == PIC_LAYOUT_F64) {
                double lx = (double)st.a[j] - kx * g.dx;
                double ly = (double)st.b[j] - r * g.dy;
                w[0] = ((g.dx - lx) * (g.dy - ly)) * a.inv_area;
                w[1] = (lx * (g.dy - ly)) * a.inv_area;
                w[2] = ((g.dx - lx) * ly) * a.inv_area;
                w[3] = (lx * ly) * a.inv_area;
            } else {
                double fx = (double)st.a[j], fy = (double)st.b[j];
                w[0] = (1.0 - fx) * (1.0 - fy);
                w[1] = fx * (1.0 - fy);
                w[2] = (1.0 - fx) * fy;
                w[3] = fx * fy;
            }
            s0 += __double2ll_rn(w[0] * a.scale_q);
            s1 += __double2ll_rn(w[1] * a.scale_q);
            s2 += __double2ll_rn(w[2] * a.scale_q);
            s3 += __double2ll_rn(w[3] * a.scale_q);
            if (a.energy) {
                double pvx = (double)st.vx[j], pvy = (double)st.vy[j], pvz = (double)st.vz[j];
                double v2 = pvx * pvx + pvy * pvy + pvz * pvz;
                double en = 0.5 * v2 * a.denom;
                e0 += __double2ll_rn((w[0] * en) * a.scale_e);
                e1 += __double2ll_rn((w[1] * en) * a.scale_e);
                e2 += __double2ll_rn((w[2] * en) * a.scale_e);
                e3 += __double2ll_rn((w[3] * en) * a.scale_e);
            }
        }
    }
    if (!mine) return;
    const int gp = g.gp;
    const int k = k0 + tid;
    cq[k] = s0;
    cq[gp + k + 1] = s1;
    cq[2 * gp + k + g.gx] = s2;
    cq[3 * gp + k + g.gx + 1] = s3;
    if (a.energy) {
        ce[k] = e0;
        ce[gp + k + 1] = e1;
        ce[2 * gp + k + g.gx] = e2;
        ce[3 * gp + k + g.gx + 1] = e3;
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

template <int L, int GENB>
__device__ __forceinline__ Pushed<L> push_one(const ParticleBuf& b, int i, const GridDev& g, const SpeciesDev& s,
                                             const double4* __restrict__ fld) {
    using P = typename Lay<L>::P;
    using V = typename Lay<L>::V;
    double pvx = (double)((const V*)b.vx)[i];
    double pvy = (double)((const V*)b.vy)[i];
    double pvz = (double)((const V*)b.vz)[i];
    Pushed<L> o;
    if (L == PIC_LAYOUT_F64) {
        double px = ((const double*)b.pa)[i];
        double py = ((const double*)b.pb)[i];
        int gx = (int)(px / g.dx);
        int gy = (int)(py / g.dy);
        double lx = px - gx * g.dx;
        double ly = py - gy * g.dy;
        int i00 = gy * g.gx + gx;
        double4 f00 = fld[i00], f10 = fld[i00 + 1], f01 = fld[i00 + g.gx], f11 = fld[i00 + g.gx + 1];
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
        int c = b.cell[i];
        int cgx, cgy;
        split_key(c, g, cgx, cgy);
        double fx = (double)((const float*)b.pa)[i];
        double fy = (double)((const float*)b.pb)[i];
        double w00 = (1.0 - fx) * (1.0 - fy), w10 = fx * (1.0 - fy);
        double w01 = (1.0 - fx) * fy, w11 = fx * fy;
        double4 f00 = fld[c], f10 = fld[c + 1], f01 = fld[c + g.gx], f11 = fld[c + g.gx + 1];
        double Epx = f00.x * w00 + f10.x * w10 + f01.x * w01 + f11.x * w11;
        double Epy = f00.y * w00 + f10.y * w10 + f01.y * w01 + f11.y * w11;
        double Bpz = f00.z * w00 + f10.z * w10 + f01.z * w01 + f11.z * w11;
        boris<GENB>(Epx, Epy, Bpz, s, pvx, pvy, pvz);
        double tx = ((double)cgx + fx) + pvx * s.dtdx;
        double ty = ((double)cgy + fy) + pvy * s.dtdy;
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
// moves less than a cell per step).  Lanes of a wave with the same window bin
// are merged with ballots so each bin takes one LDS atomic per wave; the
// window then reserves each bin's slots with one global atomic, and every
// particle gets its slot inside its new cell.  Particles that land outside
// the window take a slot with a global atomic.  k_scatter_push recomputes the
// identical push and writes each particle to start[key] + slot.
static constexpr int kTileC = 256;
static constexpr int kHW = kTileC + 2, kHR = 3;
static constexpr int kRankBits = 21;            // window bin (< 1024) in bits 21..30
static constexpr int kFallback = (int)0x80000000;

template <int L, int GENB>
__global__ __launch_bounds__(kBlock) void k_push_tile(
    ParticleBuf b, GridDev g, SpeciesDev s, int tiles_per_row, const double4* __restrict__ fld,
    const int32_t* __restrict__ start, int32_t* __restrict__ counts, int32_t* __restrict__ rank,
    unsigned long long* __restrict__ absorbed)
{
    __shared__ int hist[kHR * kHW];
    __shared__ unsigned long long nabs[2];
    const int tid = threadIdx.x, lane = tid & 63;
    const int r = blockIdx.x / tiles_per_row;
    const int c0 = (blockIdx.x % tiles_per_row) * kTileC;
    const int cw = min(kTileC, g.gx - c0);
    const int k0 = r * g.gx + c0;
    for (int j = tid; j < kHR * kHW; j += kBlock) hist[j] = 0;
    if (tid < 2) nabs[tid] = 0ull;
    __syncthreads();
    const int beg = start[k0], end = start[k0 + cw];
    const unsigned long long lt_mask = (1ull << lane) - 1ull;
    // every lane of a wave runs the same number of iterations (ballots below)
    for (int base = beg; base < end; base += kBlock) {
        const int i = base + tid;
        int bin = -1, code = -1;
        if (i < end) {
            Pushed<L> o = push_one<L, GENB>(b, i, g, s, fld);
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
        unsigned long long active = __ballot(bin >= 0);
        while (active) {
            const int leader = __ffsll((unsigned long long)active) - 1;
            const int lb = __shfl(bin, leader, 64);
            const unsigned long long m = __ballot(bin == lb) & active;
            int wbase = 0;
            if (lane == leader) wbase = atomicAdd(&hist[lb], (int)__popcll(m));
            wbase = __shfl(wbase, leader, 64);
            if ((m >> lane) & 1ull) code = (bin << kRankBits) | (wbase + (int)__popcll(m & lt_mask));
            active &= ~m;
        }
        if (i < end) rank[i] = code;
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
        int v = rank[i];
        if (v == -1) continue;
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
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= *nsrc) return;
    Pushed<L> o = push_one<L, GENB>(src, i, g, s, fld);
    if (o.key < 0) return;
    int d = start[o.key] + rank[i];
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
                                   c->stream, sp.buf[sp.cur], c->g, a, tpr, sp.start, (long long*)sp.q_charge,
                                   (long long*)sp.q_energy);
            else
                hipLaunchKernelGGL(k_deposit_tile<PIC_LAYOUT_CELL32>, dim3(ntiles), dim3(kBlock), 0,
                                   c->stream, sp.buf[sp.cur], c->g, a, tpr, sp.start, (long long*)sp.q_charge,
                                   (long long*)sp.q_energy);
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
                       sp.buf[sp.cur], c->g, d, tpr, c->fld, sp.start, sp.counts, sp.rank, sp.absorbed);
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
