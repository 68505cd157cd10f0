// pic_particles.hip — gfx950 kernels for the particle stages of the PIC step:
// encode/bin (counting sort), CIC deposit (int64 fixed point), gather + Boris
// push + wall absorption + re-bin key.
//
// Floating-point contraction is disabled for this file (-ffp-contract=off in
// the build) so that every FP64 expression below rounds exactly as the
// reference's C++ does on the host; the oracle is compiled the same way.
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

// Reference bin key (particles.cpp:229-235, mover.cpp:323-324,356):
// gx = (int)(x * inv_dx), gy = (int)(y * inv_dy), key = gy * GRID_X + gx.
__device__ __forceinline__ int ref_key(double x, double y, const GridDev& g) {
    int gx = (int)(x * g.inv_dx);
    int gy = (int)(y * g.inv_dy);
    return gy * g.gx + gx;
}

// ================================================================ encode
// Upload path: absolute FP64 particles -> storage layout + key.
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
// Counting sort.  count: per-particle slot inside its cell; scan: exclusive
// prefix of the counts (and reset of the counts); scatter: copy to the slot.
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
// CIC weights normalised by dx*dy, converted to int64 at 2^charge_shift and
// added with 64-bit integer atomics: the sum is independent of order.
// Weights follow interpolate (interpolation.cpp:72-98): for the particle's
// key cell (gx, gy) = (key % GRID_X, key / GRID_X), lx = x - gx*dx,
// ly = y - gy*dy; corners (gx,gy) (dx-lx)(dy-ly), (gx+1,gy) lx(dy-ly),
// (gx,gy+1) (dx-lx)ly, (gx+1,gy+1) lx*ly.  Cells on the last node row or
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
__global__ __launch_bounds__(kBlock) void k_deposit(
    const int32_t* __restrict__ nsrc, ParticleBuf b, GridDev g, DepositArgs a,
    unsigned long long* __restrict__ qc, unsigned long long* __restrict__ qe)
{
    using V = typename Lay<L>::V;
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= *nsrc) return;
    int k = b.cell[i];
    int gxk = k % g.gx, gyk = k / g.gx;
    if (gxk + 1 >= g.gx || gyk + 1 >= g.gy) return;
    double w[4];
    cic_weights<L>(b, i, gxk, gyk, g, a, w);
    int node[4] = {k, k + 1, k + g.gx, k + g.gx + 1};
#pragma unroll
    for (int c = 0; c < 4; ++c)
        atomicAdd(qc + node[c], (unsigned long long)__double2ll_rn(w[c] * a.scale_q));
    if (a.energy) {
        double pvx = (double)((const V*)b.vx)[i];
        double pvy = (double)((const V*)b.vy)[i];
        double pvz = (double)((const V*)b.vz)[i];
        double v2 = pvx * pvx + pvy * pvy + pvz * pvz;
        double en = 0.5 * v2 * a.denom;
#pragma unroll
        for (int c = 0; c < 4; ++c)
            atomicAdd(qe + node[c], (unsigned long long)__double2ll_rn((w[c] * en) * a.scale_e));
    }
}

__global__ __launch_bounds__(kBlock) void k_zero_u64(unsigned long long* p, int n) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0ull;
}

__global__ __launch_bounds__(kBlock) void k_zero_f64(double* p, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.0;
}

// final_mesh = W * dx*dy * 2^-S * Q; rho += final_mesh * (-q/eps)  (RhoKernel :165-171)
__global__ __launch_bounds__(kBlock) void k_deposit_finalize(
    int G, const long long* __restrict__ qc, const long long* __restrict__ qe,
    double cq, double ce, double chrg, int energy,
    double* __restrict__ final_mesh, double* __restrict__ energy_mesh, double* __restrict__ rho)
{
    int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= G) return;
    double f = (double)qc[n] * cq;
    final_mesh[n] = f;
    rho[n] += f * chrg;
    if (energy) energy_mesh[n] = (double)qe[n] * ce;
}

// ================================================================= push
// Gather (deterministic_mover, mover.cpp:103-133): gx = (int)(x/dx) — a
// division, unlike the bin key — lx = x - gx*dx, bilinear weights over
// (dx*dy).  The field arrays are padded with zero nodes so the reference's
// flat indexing (gy*GRID_X + gx + 1 etc.) stays in bounds at the last node.
// Boris (mover.cpp:151-166) for Bz; when bx or by is non-zero the same
// rotation is done with the full vector t = (q/m) B dt / 2.
// Boundary (mover.cpp:325,331-352): absorbed when x >= W, x < 0, y >= H or
// y < 0; particles leaving through x < 0 are counted separately.
template <int L, int GENB>
__global__ __launch_bounds__(kBlock) void k_push(
    const int32_t* __restrict__ nsrc, ParticleBuf b, GridDev g, SpeciesDev s,
    const double4* __restrict__ fld, int32_t* __restrict__ key,
    unsigned long long* __restrict__ absorbed)
{
    using P = typename Lay<L>::P;
    using V = typename Lay<L>::V;
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= *nsrc) return;

    double px, py;
    if (L == PIC_LAYOUT_F64) {
        px = ((const double*)b.pa)[i];
        py = ((const double*)b.pb)[i];
    } else {
        int c = b.cell[i];
        int cgx = c % g.gx, cgy = c / g.gx;
        px = ((double)cgx + (double)((const float*)b.pa)[i]) * g.dx;
        py = ((double)cgy + (double)((const float*)b.pb)[i]) * g.dy;
    }
    double pvx = (double)((const V*)b.vx)[i];
    double pvy = (double)((const V*)b.vy)[i];
    double pvz = (double)((const V*)b.vz)[i];

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
        double den = 1 + (tx * tx + ty * ty + tz * tz);
        double sx = 2 * tx / den, sy = 2 * ty / den, sz = 2 * tz / den;
        pvx = Vxminus + (vpy * sz - vpz * sy) + hx;
        pvy = Vyminus + (vpz * sx - vpx * sz) + hy;
        pvz = Vzminus + (vpx * sy - vpy * sx);
    }
    px = px + s.dt * pvx;
    py = py + s.dt * pvy;

    bool out = (px >= g.width || px < 0.0 || py >= g.height || py < 0);
    if (out) {
        key[i] = -1;
        if (px < 0.0) atomicAdd(absorbed, 1ull);
        atomicAdd(absorbed + 1, 1ull);
        return;
    }
    double tx = px * g.inv_dx, ty = py * g.inv_dy;
    int ngx = (int)tx, ngy = (int)ty;
    key[i] = ngy * g.gx + ngx;
    if (L == PIC_LAYOUT_F64) {
        ((P*)b.pa)[i] = (P)px;
        ((P*)b.pb)[i] = (P)py;
    } else {
        ((P*)b.pa)[i] = clamp_offset((float)(tx - (double)ngx));
        ((P*)b.pb)[i] = clamp_offset((float)(ty - (double)ngy));
    }
    ((V*)b.vx)[i] = (V)pvx;
    ((V*)b.vy)[i] = (V)pvy;
    ((V*)b.vz)[i] = (V)pvz;
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
    hipLaunchKernelGGL(k_scan_local, dim3(c->scan_nb), dim3(kBlock), 0, c->stream,
                       sp.counts, c->G, sp.start, c->scan_bsum);
    PIC_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(1024), 0, c->stream,
                       c->scan_bsum, c->scan_nb, sp.start, c->G, sp.ndev + dst);
    PIC_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_scan_add, dim3(blocks_for(c->G)), dim3(kBlock), 0, c->stream,
                       sp.start, c->G, c->scan_bsum);
    PIC_HIP(hipGetLastError());
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
    PIC_TRY(timer_end(c, PIC_STAGE_BIN, ev));
    return PIC_OK;
}

int launch_deposit(pic_ctx* c, int s) {
    Species& sp = c->sp[s];
    hipEvent_t ev;
    PIC_TRY(timer_begin(c, PIC_STAGE_DEPOSIT, &ev));
    const double dx = c->g.dx, dy = c->g.dy;
    DepositArgs a;
    a.inv_area = 1.0 / (dx * dy);
    a.scale_q = std::ldexp(1.0, c->p.charge_shift);
    a.scale_e = std::ldexp(1.0, c->p.energy_shift);
    a.denom = 1 / std::fabs(sp.qm);
    a.energy = c->p.deposit_energy;
    hipLaunchKernelGGL(k_zero_u64, dim3(blocks_for(c->G)), dim3(kBlock), 0, c->stream,
                       (unsigned long long*)sp.q_charge, c->G);
    PIC_HIP(hipGetLastError());
    if (a.energy) {
        hipLaunchKernelGGL(k_zero_u64, dim3(blocks_for(c->G)), dim3(kBlock), 0, c->stream,
                           (unsigned long long*)sp.q_energy, c->G);
        PIC_HIP(hipGetLastError());
    }
    const int64_t n = sp.n_upper;
    if (n > 0) {
        if (c->p.layout == PIC_LAYOUT_F64)
            hipLaunchKernelGGL(k_deposit<PIC_LAYOUT_F64>, dim3(blocks_for(n)), dim3(kBlock), 0, c->stream,
                               sp.ndev + sp.cur, sp.buf[sp.cur], c->g, a,
                               (unsigned long long*)sp.q_charge, (unsigned long long*)sp.q_energy);
        else
            hipLaunchKernelGGL(k_deposit<PIC_LAYOUT_CELL32>, dim3(blocks_for(n)), dim3(kBlock), 0, c->stream,
                               sp.ndev + sp.cur, sp.buf[sp.cur], c->g, a,
                               (unsigned long long*)sp.q_charge, (unsigned long long*)sp.q_energy);
        PIC_HIP(hipGetLastError());
    }
    const double cq = sp.W * (dx * dy) * std::ldexp(1.0, -c->p.charge_shift);
    const double ce = sp.W * (dx * dy) * std::ldexp(1.0, -c->p.energy_shift);
    const double chrg = -sp.charge / c->p.epsilon;
    hipLaunchKernelGGL(k_deposit_finalize, dim3(blocks_for(c->G)), dim3(kBlock), 0, c->stream,
                       c->G, (const long long*)sp.q_charge, (const long long*)sp.q_energy,
                       cq, ce, chrg, a.energy, sp.final_mesh, sp.energy_mesh, c->rho);
    PIC_HIP(hipGetLastError());
    PIC_TRY(timer_end(c, PIC_STAGE_DEPOSIT, ev));
    return PIC_OK;
}

int launch_push(pic_ctx* c, int s) {
    Species& sp = c->sp[s];
    hipEvent_t ev;
    PIC_TRY(timer_begin(c, PIC_STAGE_PUSH, &ev));
    SpeciesDev d;
    d.qm = sp.qm;
    d.half_qm = 0.5 * sp.qm;
    d.denom = 1 / std::fabs(sp.qm);
    d.dt = c->p.dt;
    d.bx = c->p.bx;
    d.by = c->p.by;
    d.general_b = (c->p.bx != 0.0 || c->p.by != 0.0);
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
    PIC_TRY(timer_end(c, PIC_STAGE_PUSH, ev));
    return PIC_OK;
}

} // namespace pic
