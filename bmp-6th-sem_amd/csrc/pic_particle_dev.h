// pic_particle_dev.h — device functions of the particle stages shared by the
// single-run kernels (pic_particles.hip) and the ensemble's persistent kernels
// (pic_ensemble.hip).  Both files are compiled with -ffp-contract=off so every
// FP64 expression rounds exactly as written.
#pragma once

#include "pic_internal.h"

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
// field rows and cell starts a tile touches in that XCD's L2.  With rev the
// XCD walks its band backwards: the particle passes alternate direction, so
// each starts on the data the previous pass touched last, which the
// memory-side cache still holds.
static constexpr int kXcd = 8;
__device__ __forceinline__ int xcd_tile(int b, int ntiles, int rev = 0) {
    const int per = (ntiles + kXcd - 1) / kXcd;
    const int x = b % kXcd, k = b / kXcd;
    const int full = ntiles - (per - 1) * kXcd;     // XCDs that get `per` tiles
    // XCD x owns tiles [first(x), first(x) + count(x)); workgroup b is its k-th
    const int first = x < full ? x * per : full * per + (x - full) * (per - 1);
    const int count = x < full ? per : per - 1;
    return first + (rev ? count - 1 - k : k);
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

// round_fixed for |v| < 2^51, without the range test: one add and one 32-bit
// subtract.  For such v, v + 1.5 * 2^52 lies in [2^52, 2^53), where the
// spacing of doubles is 1, so the add rounds v to the nearest integer (ties
// to even) exactly like __double2ll_rn.
__device__ __forceinline__ long long round_small(double v) {
    return __double_as_longlong(v + 0x1.8p52) - 0x4338000000000000LL;
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
    double e_fast;     // 2^(50 - charge_shift): energy sums take round_small below it
    int    energy;
    int    q_fast;     // charge_shift <= 50: charge weights (< 2 * 2^charge_shift) take round_small
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
// particles is flushed by one to three lanes).  The tile also sums the cell
// left of its first one (its particles are read by both tiles; each uses
// other corners of it), so thread t can form the two node-row sums of node
// n = k0 + t by itself:
//   lower[n]      = corner 0 of cell n + corner 1 of cell n - 1 (node row r),
//   upper[n + GX] = corner 2 of cell n + corner 3 of cell n - 1 (row r + 1),
// stored as mesh 0 and mesh 1 of the corner meshes; a node's deposit is
// lower + upper.  Every node sum is written by exactly one thread, so no
// global atomics; cells on the last node row or column contribute zero (the
// reference skips them, :122), and node row 0's upper sum stays zero.
//
// Weights are formed already scaled by 2^S: for CELL32 (1 - ox) 2^h *
// (1 - oy) 2^(S-h) rounds exactly like ((1 - ox)(1 - oy)) 2^S (scaling by a
// power of two commutes with rounding), and the energy weight
// w 2^S * (en 2^(E-S)) like (w en) 2^E, so the sums equal the oracle's bit
// for bit.
static constexpr int kDepC = 256;
static constexpr int kDepA = kDepC + 1;     // accumulated cells: the tile's and the one left of it
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

// Deposit of one tile (kDepC cells of a key row); acc: 8 x kDepA LDS
// accumulators, local cell l = cell - (k0 - 1); sst: kDepA + 1 starts.
// NQ: consecutive quads per thread (4 NQ particles).  Dense cells hold many
// quads; with NQ > 1 fewer lanes flush into one cell's accumulators, so the
// same-address LDS atomics of a wave serialise less.
template <int L, int NQ = 1>
__device__ __forceinline__ void deposit_tile_body(
    const ParticleBuf& b, const GridDev& g, const DepositArgs& a, int tiles_per_row, int tile,
    const int32_t* __restrict__ start, long long* __restrict__ cq, long long* __restrict__ ce,
    unsigned long long (*acc)[kDepA], int* sst = nullptr)
{
    using P = typename Lay<L>::P;
    using V = typename Lay<L>::V;
    const int tid = threadIdx.x;
    const int r = tile / tiles_per_row;
    const int c0 = (tile % tiles_per_row) * kDepC;
    const int cw = min(kDepC, g.gx - c0);
    const int k0 = r * g.gx + c0;
    const int kl = k0 - 1;                          // local cell 0
    for (int j = tid; j < kDepA; j += kBlock)
#pragma unroll
        for (int c = 0; c < 8; ++c) acc[c][j] = 0ull;
    // the left cell of column 0 is the last cell of the row below: it holds
    // no particles (last node column) and is left out
    const int beg = c0 > 0 ? start[kl] : start[k0];
    const int end = start[k0 + cw];
    // with `sst` the particles' cells come from the cell starts (LDS): the
    // cell array is not read; sst[l] = first particle of local cell l
    if (sst)
        for (int j = tid; j <= cw + 1; j += kBlock) sst[j] = j == 0 ? beg : start[kl + j];
    __syncthreads();
    const int first = beg & ~(kVec - 1);
    constexpr int kSpan = kVec * NQ;
    for (int base0 = first + kSpan * tid; base0 < end; base0 += kSpan * kBlock) {
        int cur = -1, sc = 0;
        long long s0 = 0, s1 = 0, s2 = 0, s3 = 0, e0 = 0, e1 = 0, e2 = 0, e3 = 0;
        auto flush = [&]() {
            const int lc = cur - kl;
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
        for (int qd = 0; qd < NQ; ++qd) {
            const int base = base0 + kVec * qd;
            if (base >= end) break;
            Quad<L> q;
            if (sst) {
                if (qd == 0) {
                    // last local cell l with sst[l] <= the span's first particle of the tile
                    const int p0 = max(base, beg);
                    int lo = 0, hi = cw;
                    while (lo < hi) {
                        const int mid = (lo + hi + 1) >> 1;
                        if (sst[mid] <= p0) lo = mid;
                        else hi = mid - 1;
                    }
                    sc = lo;
                }
            } else {
                load_quad<int>(b.cell, base, q.cell);
            }
            load_quad<P>((const P*)b.pa, base, q.a);
            load_quad<P>((const P*)b.pb, base, q.b);
            if (a.energy) {
                load_quad<V>((const V*)b.vx, base, q.vx);
                load_quad<V>((const V*)b.vy, base, q.vy);
                load_quad<V>((const V*)b.vz, base, q.vz);
            }
#pragma unroll
            for (int i = 0; i < kVec; ++i) {
                const int pi = base + i;
                if (pi < beg || pi >= end) continue;
                int cl;
                if (sst) {
                    while (sst[sc + 1] <= pi) ++sc;
                    cl = kl + sc;
                } else {
                    cl = q.cell[i];
                }
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
                // the range tests of round_fixed are taken once per workgroup
                // (charge) or once per particle (energy): |w| < 2 * 2^S
                if (a.q_fast) {
                    s0 += round_small(w0);
                    s1 += round_small(w1);
                    s2 += round_small(w2);
                    s3 += round_small(w3);
                } else {
                    s0 += round_fixed(w0);
                    s1 += round_fixed(w1);
                    s2 += round_fixed(w2);
                    s3 += round_fixed(w3);
                }
                if (a.energy) {
                    const double pvx = (double)q.vx[i], pvy = (double)q.vy[i], pvz = (double)q.vz[i];
                    const double v2 = pvx * pvx + pvy * pvy + pvz * pvz;
                    const double en = 0.5 * v2 * a.denom;
                    const double es = en * a.scale_eq;    // exact: power of two
                    if (es < a.e_fast) {
                        e0 += round_small(w0 * es);
                        e1 += round_small(w1 * es);
                        e2 += round_small(w2 * es);
                        e3 += round_small(w3 * es);
                    } else {
                        e0 += round_fixed(w0 * es);
                        e1 += round_fixed(w1 * es);
                        e2 += round_fixed(w2 * es);
                        e3 += round_fixed(w3 * es);
                    }
                }
            }
        }
        if (cur >= 0) flush();
    }
    __syncthreads();
    if (tid >= cw) return;
    // node n = k0 + tid: its own cell (local tid + 1) and the cell left of it
    // (local tid); a cell on the last node row or column contributes zero
    const bool row_live = r + 1 < g.gy;
    const bool own = row_live && (c0 + tid + 1 < g.gx);
    const bool left = row_live && c0 + tid > 0;
    const int gp = g.gp;
    const int n = k0 + tid;
    auto pick = [&](int c, int l, bool live) { return live ? (long long)acc[c][l] : 0ll; };
    cq[n] = pick(0, tid + 1, own) + pick(1, tid, left);
    if (row_live) cq[gp + n + g.gx] = pick(2, tid + 1, own) + pick(3, tid, left);
    if (a.energy) {
        ce[n] = pick(4, tid + 1, own) + pick(5, tid, left);
        if (row_live) ce[gp + n + g.gx] = pick(6, tid + 1, own) + pick(7, tid, left);
    }
}

// F64 gather (deterministic_mover, mover.cpp:103-133): gx = (int)(x/dx) — a
// division, unlike the bin key — lx = x - gx*dx, bilinear weights over
// (dx*dy).  The field arrays are padded with zero nodes so the reference's
// flat indexing (gy*GRID_X + gx + 1 etc.) stays in bounds at the last node.
// Boris (mover.cpp:151-166) for Bz; when bx or by is non-zero the same
// rotation is done with the full vector t = (q/m) B dt / 2.
// Boundary (mover.cpp:325,331-352): absorbed when x >= W, x < 0, y >= H or
// y < 0 (CELL32: the same walls in cell units); particles leaving through
// x < 0 are counted separately.  With periodic y the new y is wrapped first.

// Periodic y (mover.cpp:13-22,317-320): modulo(x, y) = fmod for x >= 0; for
// x < 0, 0 when x / y is an integer, else y - fmod(-x, y).  Applied to the new
// y before the wall test (metres; CELL32: cell units, period NY).
__device__ __forceinline__ double modulo_ref(double x, double y) {
    if (x < 0) {
        if ((double)(int)(x / y) == x / y) return 0;
        return y - fmod(-x, y);
    }
    return fmod(x, y);
}

// Particle state after the push, in storage form.
template <int L> struct Pushed {
    using P = typename Lay<L>::P;
    using V = typename Lay<L>::V;
    int key;            // new bin key, -1 when absorbed
    int kx, ky;         // column and row of the new key (0 <= kx < GRID_X)
    bool left;          // absorbed through x < 0
    P a, b;             // F64: x, y   CELL32: offsets
    V vx, vy, vz;
};

// Push modes (template parameter GENB): bit 0 — an in-plane B, the rotation
// about the full vector t = (q/m) B dt / 2; bit 1 — CELL32 with a uniform Bz:
// t and s are the species' constants (SpeciesDev::ut, us), Bz is not
// gathered and the half kick is ce * E (ce = (0.5 q/m) dt).
template <int GENB>
__device__ __forceinline__ void boris(double Epx, double Epy, double Bpz, const SpeciesDev& s,
                                      double& pvx, double& pvy, double& pvz) {
    double hx = s.half_qm * Epx * s.dt;     // 0.5 * pctm * Epx * dt
    double hy = s.half_qm * Epy * s.dt;
    double Vxminus = pvx + hx;
    double Vyminus = pvy + hy;
    if (!(GENB & 1)) {
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

// The same rotation with the species' constant t and s (uniform B).
template <int GB>
__device__ __forceinline__ void boris_const(double hx, double hy, const SpeciesDev& s, double& pvx, double& pvy,
                                            double& pvz) {
    double Vxminus = pvx + hx;
    double Vyminus = pvy + hy;
    if (!GB) {
        const double t = s.ut[2], sf = s.us[2];
        double Vxprime = Vxminus + Vyminus * t;
        double Vyprime = Vyminus - Vxminus * t;
        pvx = Vxminus + Vyprime * sf + hx;
        pvy = Vyminus - Vxprime * sf + hy;
    } else {
        const double tx = s.ut[0], ty = s.ut[1], tz = s.ut[2];
        const double sx = s.us[0], sy = s.us[1], sz = s.us[2];
        double Vzminus = pvz;
        double vpx = Vxminus + (Vyminus * tz - Vzminus * ty);
        double vpy = Vyminus + (Vzminus * tx - Vxminus * tz);
        double vpz = Vzminus + (Vxminus * ty - Vyminus * tx);
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

// With CELL = false the cell array is not read (the caller supplies p.cell).
template <int L, bool CELL = true>
__device__ __forceinline__ PState<L> load_state(const ParticleBuf& b, int i) {
    using P = typename Lay<L>::P;
    using V = typename Lay<L>::V;
    PState<L> p;
    p.cell = (L == PIC_LAYOUT_CELL32 && CELL) ? b.cell[i] : 0;
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

// Field window of a CELL32 tile (cells c0 .. c0+cw-1 of key row r, first key
// k0) in LDS: nodes (r, c0 .. c0+cw) then (r+1, c0 .. c0+cw), as three
// component arrays (Ex, Ey, Bz) of kFW = 2 (TC + 1) doubles each; the four
// corners of every cell of the tile (fld is padded past the last node row, so
// row r+1 exists for the last key row too).
template <int TC>
__device__ __forceinline__ void load_field_window(const double4* __restrict__ fld, const GridDev& g, int k0, int cw,
                                                  double* wf) {
    constexpr int kFW = 2 * (TC + 1);
    for (int j = threadIdx.x; j < kFW; j += kBlock) {
        const int row = j >= TC + 1 ? 1 : 0;
        const int col = j - row * (TC + 1);
        if (col <= cw) {
            const double4 v = fld[k0 + row * g.gx + col];
            wf[j] = v.x;
            wf[kFW + j] = v.y;
            wf[2 * kFW + j] = v.z;
        }
    }
}

// The corners of a particle in cell `cell` of the tile, from the window
// (Bz only with NEEDB; push mode 2/3 does not read it).
template <int TC, bool NEEDB = true>
__device__ __forceinline__ Corners gather_window(int cell, int k0, int c0, int r, const double* wf) {
    constexpr int kFW = 2 * (TC + 1);
    Corners c;
    const int lc = cell - k0;
    c.cgx = c0 + lc;
    c.cgy = r;
    const int u = TC + 1 + lc;
    auto bz = [&](int j) { return NEEDB ? wf[2 * kFW + j] : 0.0; };
    c.f00 = make_double4(wf[lc], wf[kFW + lc], bz(lc), 0.0);
    c.f10 = make_double4(wf[lc + 1], wf[kFW + lc + 1], bz(lc + 1), 0.0);
    c.f01 = make_double4(wf[u], wf[kFW + u], bz(u), 0.0);
    c.f11 = make_double4(wf[u + 1], wf[kFW + u + 1], bz(u + 1), 0.0);
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
        if (g.periodic) py = modulo_ref(py, g.height);
        bool out = (px >= g.width || px < 0.0 || py >= g.height || py < 0);
        o.left = px < 0.0;
        if (out) {
            o.key = -1;
            return o;
        }
        int ngx = (int)(px * g.inv_dx), ngy = (int)(py * g.inv_dy);
        o.key = ngy * g.gx + ngx;
        o.kx = ngx;
        o.ky = ngy;
        o.a = (P)px;
        o.b = (P)py;
    } else {
        double fx = (double)p.a;
        double fy = (double)p.b;
        double w00 = (1.0 - fx) * (1.0 - fy), w10 = fx * (1.0 - fy);
        double w01 = (1.0 - fx) * fy, w11 = fx * fy;
        double Epx = f00.x * w00 + f10.x * w10 + f01.x * w01 + f11.x * w11;
        double Epy = f00.y * w00 + f10.y * w10 + f01.y * w01 + f11.y * w11;
        if constexpr ((GENB & 2) != 0) {
            boris_const<GENB & 1>(s.ce * Epx, s.ce * Epy, s, pvx, pvy, pvz);
        } else {
            double Bpz = f00.z * w00 + f10.z * w10 + f01.z * w01 + f11.z * w11;
            boris<GENB>(Epx, Epy, Bpz, s, pvx, pvy, pvz);
        }
        double tx = ((double)c.cgx + fx) + pvx * s.dtdx;
        double ty = ((double)c.cgy + fy) + pvy * s.dtdy;
        if (g.periodic) ty = modulo_ref(ty, g.ny);
        bool out = (tx >= g.nx || tx < 0.0 || ty >= g.ny || ty < 0.0);
        o.left = tx < 0.0;
        if (out) {
            o.key = -1;
            return o;
        }
        int ngx = (int)tx, ngy = (int)ty;
        o.key = ngy * g.gx + ngx;
        o.kx = ngx;
        o.ky = ngy;
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

// Index of the cell holding particle i among the cells of a sorted range:
// the last j in [0, hi] with st[j] <= i (st[0] <= i < st[hi + 1]; empty
// cells have st[j] == st[j + 1] and are passed over).
__device__ __forceinline__ int cell_of(const int* st, int hi, int i) {
    int lo = 0;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (st[mid] <= i) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// In-place push without re-binning (stage API pic_push(ctx, s, 0)); the
// stored cell becomes the new key (-1 when absorbed).
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
static constexpr int kTileC = 256;              // tile width of the single-run kernels
static constexpr int kHR = 3;
static constexpr int kRankBits = 21;            // window bin (< 1024) in bits 21..30
static constexpr int kFallback = (int)0x80000000;
static constexpr int kCodeCap = 4096;

// Push + slot count of one tile of TC cells; LDS: hist[kHR * (TC + 2)],
// codes[kCodeCap], nabs[2], and with WF (CELL32 only) the tile's field window
// wf[6 * (TC + 1)], from which the gathers read.
//
// SC (CELL32 with WF): the particles' cells come from the tile's cell starts
// instead of the cell array; the tile also copies its starts to start_old
// (the scan that follows overwrites `start`) and records the cell of every
// particle i with i % 256 == 0 in blk_first[i / 256], from which
// k_scatter_push finds its particles' cells.  CAP: entries of `codes`.
template <int L, int GENB, int TC, bool WF = false, bool SC = false, int CAP = kCodeCap>
__device__ __forceinline__ void push_tile_body(
    const ParticleBuf& b, const GridDev& g, const SpeciesDev& s, int tiles_per_row, int tile,
    const double4* __restrict__ fld, const int32_t* __restrict__ start, int32_t* __restrict__ counts,
    int32_t* __restrict__ rank, unsigned long long* __restrict__ absorbed, int* hist, int* codes,
    unsigned long long* nabs, double* wf = nullptr, int32_t* __restrict__ start_old = nullptr,
    int32_t* __restrict__ blk_first = nullptr)
{
    static_assert(!WF || L == PIC_LAYOUT_CELL32, "the field window serves the CELL32 layout");
    static_assert(!SC || WF, "cells from the starts serve the windowed CELL32 tile");
    constexpr int kHW = TC + 2;
    const int tid = threadIdx.x;
    const int r = tile / tiles_per_row;
    const int c0 = (tile % tiles_per_row) * TC;
    const int cw = min(TC, g.gx - c0);
    const int k0 = r * g.gx + c0;
    if constexpr (WF) load_field_window<TC>(fld, g, k0, cw, wf);
    const int beg = start[k0], end = start[k0 + cw];
    // SC: codes[p - beg] first holds particle p's cell in the tile (thread j
    // writes cell j's particles); a thread reads the entry of a particle it
    // owns before replacing it by the particle's code.  Particles past CAP
    // (tiles far above the mean occupancy) search the starts.
    if constexpr (SC)
        for (int j = tid; j < cw; j += kBlock) {
            const int e = min(start[k0 + j + 1] - beg, CAP);
            for (int li = start[k0 + j] - beg; li < e; ++li) codes[li] = j;
        }
    for (int j = tid; j < kHR * kHW; j += kBlock) hist[j] = 0;
    if (tid < 2) nabs[tid] = 0ull;
    __syncthreads();
    // global stores after the barrier, whose fence would wait for them
    if constexpr (SC)
        for (int j = tid; j <= cw; j += kBlock) start_old[k0 + j] = start[k0 + j];
    auto tile_cell = [&](int i) {
        const int li = i - beg;
        return k0 + (li < CAP ? codes[li] : cell_of(start + k0, cw - 1, i));
    };
    PState<L> cur;
    if (beg + tid < end) {
        cur = load_state<L, !SC>(b, beg + tid);
        if constexpr (SC) cur.cell = tile_cell(beg + tid);
    }
    for (int base = beg; base < end; base += kBlock) {
        const int i = base + tid;
        const bool have = i < end;
        Corners cn;
        if (have) {
            if constexpr (WF) cn = gather_window<TC, (GENB & 2) == 0>(cur.cell, k0, c0, r, wf);
            else cn = gather<L>(cur, g, fld);
            if constexpr (SC)
                if ((i & (kBlock - 1)) == 0) blk_first[i / kBlock] = cur.cell;
        }
        PState<L> nxt;
        if (i + kBlock < end) {
            nxt = load_state<L, !SC>(b, i + kBlock);
            if constexpr (SC) nxt.cell = tile_cell(i + kBlock);
        }
        int bin = -1, code = -1;
        if (have) {
            Pushed<L> o = push_state<L, GENB>(cur, cn, g, s);
            if (o.key < 0) {
                if (o.left) atomicAdd(&nabs[0], 1ull);
                atomicAdd(&nabs[1], 1ull);
            } else {
                const int dr = o.ky - (r - 1), dc = o.kx - (c0 - 1);
                if (dr >= 0 && dr < kHR && dc >= 0 && dc < kHW) bin = dr * kHW + dc;
                else code = kFallback | atomicAdd(&counts[o.key], 1);
            }
        }
        if (bin >= 0) code = (bin << kRankBits) | atomicAdd(&hist[bin], 1);
        if (have) {
            if (i - beg < CAP) codes[i - beg] = code;
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
        const int v = li < CAP ? codes[li] : rank[i];
        if (v == -1) continue;                      // absorbed: the scatter skips it
        if (v & kFallback) rank[i] = v & ~kFallback;
        else rank[i] = hist[v >> kRankBits] + (v & ((1 << kRankBits) - 1));
    }
}

// QN consecutive floats starting at a multiple of QN (QN-aligned: one load).
template <int QN>
__device__ __forceinline__ void load_run(const float* __restrict__ p, int i, float out[QN]) {
    if constexpr (QN == 4) {
        const float4 v = *(const float4*)(p + i);
        out[0] = v.x; out[1] = v.y; out[2] = v.z; out[3] = v.w;
    } else if constexpr (QN == 2) {
        const float2 v = *(const float2*)(p + i);
        out[0] = v.x; out[1] = v.y;
    } else {
        out[0] = p[i];
    }
}

// The CELL32 tile push (cells from the starts, field window) with QN
// consecutive particles per thread: the thread loads its particles with one
// QN-wide load per field and pushes them in order; consecutive particles bound
// for the same window bin form a run that takes its slots with one LDS atomic
// (the slots of a run are consecutive).  Dense cells put many particles of
// one cell in a wave, whose same-address atomics serialise; runs cut them by
// up to QN.  Otherwise as push_tile_body<CELL32, GENB, TC, true, true, CAP>.
template <int GENB, int TC, int QN, int CAP = kCodeCap>
__device__ __forceinline__ void push_tile_body_runs(
    const ParticleBuf& b, const GridDev& g, const SpeciesDev& s, int tiles_per_row, int tile,
    const double4* __restrict__ fld, const int32_t* __restrict__ start, int32_t* __restrict__ counts,
    int32_t* __restrict__ rank, unsigned long long* __restrict__ absorbed, int* hist, int* codes,
    unsigned long long* nabs, double* wf, int32_t* __restrict__ start_old, int32_t* __restrict__ blk_first)
{
    constexpr int L = PIC_LAYOUT_CELL32;
    constexpr int kHW = TC + 2;
    const int tid = threadIdx.x;
    const int r = tile / tiles_per_row;
    const int c0 = (tile % tiles_per_row) * TC;
    const int cw = min(TC, g.gx - c0);
    const int k0 = r * g.gx + c0;
    load_field_window<TC>(fld, g, k0, cw, wf);
    const int beg = start[k0], end = start[k0 + cw];
    for (int j = tid; j < cw; j += kBlock) {
        const int e = min(start[k0 + j + 1] - beg, CAP);
        for (int li = start[k0 + j] - beg; li < e; ++li) codes[li] = j;
    }
    for (int j = tid; j < kHR * kHW; j += kBlock) hist[j] = 0;
    if (tid < 2) nabs[tid] = 0ull;
    __syncthreads();
    for (int j = tid; j <= cw; j += kBlock) start_old[k0 + j] = start[k0 + j];
    auto put_code = [&](int i, int v) {
        const int li = i - beg;
        if (li < CAP) codes[li] = v;
        else rank[i] = v;
    };
    const int first = beg & ~(QN - 1);
    for (int q0 = first + QN * tid; q0 < end; q0 += QN * kBlock) {
        float pa[QN], pb[QN], pvx[QN], pvy[QN], pvz[QN];
        load_run<QN>((const float*)b.pa, q0, pa);
        load_run<QN>((const float*)b.pb, q0, pb);
        load_run<QN>((const float*)b.vx, q0, pvx);
        load_run<QN>((const float*)b.vy, q0, pvy);
        load_run<QN>((const float*)b.vz, q0, pvz);
        int cells[QN];
#pragma unroll
        for (int e = 0; e < QN; ++e) {
            const int i = q0 + e, li = i - beg;
            cells[e] = (i >= beg && i < end) ? k0 + (li < CAP ? codes[li] : cell_of(start + k0, cw - 1, i)) : 0;
        }
        int rbin = -1, rn = 0, rfirst = 0;       // pending run: window bin, length, first element
        auto flush = [&]() {
            if (rn > 0) {
                const int s0 = atomicAdd(&hist[rbin], rn);
                for (int k = 0; k < rn; ++k) put_code(q0 + rfirst + k, (rbin << kRankBits) | (s0 + k));
                rn = 0;
            }
        };
#pragma unroll
        for (int e = 0; e < QN; ++e) {
            const int i = q0 + e;
            if (i < beg || i >= end) continue;
            PState<L> cur;
            cur.cell = cells[e];
            cur.a = pa[e];
            cur.b = pb[e];
            cur.vx = pvx[e];
            cur.vy = pvy[e];
            cur.vz = pvz[e];
            if ((i & (kBlock - 1)) == 0) blk_first[i / kBlock] = cur.cell;
            const Corners cn = gather_window<TC, (GENB & 2) == 0>(cur.cell, k0, c0, r, wf);
            const Pushed<L> o = push_state<L, GENB>(cur, cn, g, s);
            int bin = -1, code = -1;
            if (o.key < 0) {
                if (o.left) atomicAdd(&nabs[0], 1ull);
                atomicAdd(&nabs[1], 1ull);
            } else {
                const int dr = o.ky - (r - 1), dc = o.kx - (c0 - 1);
                if (dr >= 0 && dr < kHR && dc >= 0 && dc < kHW) bin = dr * kHW + dc;
                else code = kFallback | atomicAdd(&counts[o.key], 1);
            }
            if (bin >= 0 && bin == rbin && rfirst + rn == e) {
                ++rn;
                continue;
            }
            flush();
            if (bin >= 0) {
                rbin = bin;
                rn = 1;
                rfirst = e;
            } else {
                put_code(i, code);
            }
        }
        flush();
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
        const int v = li < CAP ? codes[li] : rank[i];
        if (v == -1) continue;                      // absorbed: the scatter skips it
        if (v & kFallback) rank[i] = v & ~kFallback;
        else rank[i] = hist[v >> kRankBits] + (v & ((1 << kRankBits) - 1));
    }
}

// host helpers (pic_particles.hip)
DepositArgs deposit_args(const pic_ctx* c, const Species& sp);
SpeciesDev species_dev(const pic_ctx* c, const Species& sp);

} // namespace pic
