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
#include "pic_particle_dev.h"

#include <hip/hip_ext.h>

#include <climits>
#include <cmath>
#include <mutex>
#include <tuple>

namespace pic {

// Launch with optional start / stop events taken by the dispatch itself
// (timing mode 2); the arguments are converted to the kernel's parameter types.
template <typename... KArgs, typename... Args>
static void launch_timed(void (*kernel)(KArgs...), dim3 grid, dim3 block, hipStream_t stream, const hipEvent_t ev[2],
                         Args... args) {
    std::tuple<KArgs...> t{args...};
    void* ptrs[sizeof...(KArgs)];
    std::apply([&](auto&... x) {
        int k = 0;
        ((ptrs[k++] = (void*)&x), ...);
    }, t);
    (void)hipExtLaunchKernel((void*)kernel, grid, block, ptrs, 0, stream, ev[0], ev[1], 0);
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

// The scan of up to two species at once: species blockIdx.y of the set.
struct ScanSet {
    int32_t* counts[2];
    int32_t* start[2];
    int32_t* bsum[2];
    int32_t* nout[2];     // live count after the scan
};

__global__ __launch_bounds__(kBlock) void k_scan_local(ScanSet ss, int G)
{
    int32_t* __restrict__ counts = ss.counts[blockIdx.y];
    int32_t* __restrict__ start = ss.start[blockIdx.y];
    int32_t* __restrict__ bsum = ss.bsum[blockIdx.y];
    __shared__ int wsum[kBlock / 64];
    const int t = threadIdx.x, lane = t & 63, wid = t >> 6;
    const int base = blockIdx.x * kScanChunk + t * kScanPer;
    // a thread's 8 counts as two 16-byte accesses (32-byte aligned: the
    // arrays come from hipMalloc), element by element at the end of the grid
    const bool whole = base + kScanPer <= G;
    int v[kScanPer];
    if (whole) {
        const int4 a = *(const int4*)(counts + base), b = *(const int4*)(counts + base + 4);
        v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = a.w;
        v[4] = b.x; v[5] = b.y; v[6] = b.z; v[7] = b.w;
        const int4 z = make_int4(0, 0, 0, 0);
        *(int4*)(counts + base) = z;
        *(int4*)(counts + base + 4) = z;
    } else {
#pragma unroll
        for (int j = 0; j < kScanPer; ++j) {
            const int idx = base + j;
            v[j] = idx < G ? counts[idx] : 0;
            if (idx < G) counts[idx] = 0;
        }
    }
    int s = 0;
#pragma unroll
    for (int j = 0; j < kScanPer; ++j) s += v[j];
    int incl = wave_incl_scan(s);
    if (lane == 63) wsum[wid] = incl;
    __syncthreads();
    int wpre = 0;
    for (int w = 0; w < wid; ++w) wpre += wsum[w];
    int ex = wpre + incl - s;
    int o[kScanPer];
#pragma unroll
    for (int j = 0; j < kScanPer; ++j) {
        o[j] = ex;
        ex += v[j];
    }
    if (whole) {
        *(int4*)(start + base) = make_int4(o[0], o[1], o[2], o[3]);
        *(int4*)(start + base + 4) = make_int4(o[4], o[5], o[6], o[7]);
    } else {
#pragma unroll
        for (int j = 0; j < kScanPer; ++j)
            if (base + j < G) start[base + j] = o[j];
    }
    if (t == kBlock - 1) bsum[blockIdx.x] = wpre + incl;
}

// Sum of v over the workgroup (every thread gets it); wsum: kBlock / 64 ints.
__device__ __forceinline__ int block_sum(int v, int* wsum) {
    const int t = threadIdx.x;
    const int incl = wave_incl_scan(v);
    if ((t & 63) == 63) wsum[t >> 6] = incl;
    __syncthreads();
    int s = 0;
#pragma unroll
    for (int w = 0; w < kBlock / 64; ++w) s += wsum[w];
    __syncthreads();
    return s;
}

// Grids of at most this many chunks let every workgroup of k_scan_add add up
// the chunk totals before its own (at most 4 loads per thread, from L2);
// larger grids scan the totals once first (k_scan_top), since the summing
// workgroups' reads grow with the square of the chunk count.
static constexpr int kScanDirectChunks = 1024;

// Exclusive scan of the chunk totals in place, with the grand total stored
// after them (bsum[nb]); one workgroup per species.
__global__ __launch_bounds__(kBlock) void k_scan_top(ScanSet ss, int nb)
{
    int32_t* __restrict__ bsum = ss.bsum[blockIdx.y];
    __shared__ int wsum[kBlock / 64];
    const int t = threadIdx.x, lane = t & 63, wid = t >> 6;
    const int per = (nb + kBlock - 1) / kBlock;
    const int b0 = min(t * per, nb), b1 = min(b0 + per, nb);
    int s = 0;
    for (int j = b0; j < b1; ++j) s += bsum[j];
    const int incl = wave_incl_scan(s);
    if (lane == 63) wsum[wid] = incl;
    __syncthreads();
    int ex = incl - s;
    for (int w = 0; w < wid; ++w) ex += wsum[w];
    for (int j = b0; j < b1; ++j) {
        const int v = bsum[j];
        bsum[j] = ex;
        ex += v;
    }
    if (t == kBlock - 1) bsum[nb] = ex;
}

// start += the sum of the chunk totals before this workgroup's chunk.  With
// Scanned false each workgroup adds up the totals it needs; with Scanned true
// k_scan_top has turned them into prefixes.  The last workgroup also writes
// the total to start[G] and the live count.
template <bool Scanned>
__global__ __launch_bounds__(kBlock) void k_scan_add(ScanSet ss, int nb, int G)
{
    int32_t* __restrict__ start = ss.start[blockIdx.y];
    const int32_t* __restrict__ bsum = ss.bsum[blockIdx.y];
    __shared__ int wsum[kBlock / 64];
    const int t = threadIdx.x;
    const int ch = (int)(blockIdx.x * kBlock) / kScanChunk;   // the chunk of this workgroup's counts
    const bool last = blockIdx.x == gridDim.x - 1;
    int pre, total = 0;
    if (Scanned) {
        pre = bsum[ch];
        if (last) total = bsum[nb];
    } else {
        int before = 0, all = 0;
        for (int j = t; j < (last ? nb : ch); j += kBlock) {
            const int v = bsum[j];
            if (j < ch) before += v;
            all += v;
        }
        pre = block_sum(before, wsum);
        if (last) total = block_sum(all, wsum);
    }
    const int i = blockIdx.x * kBlock + t;
    if (i < G) start[i] += pre;
    if (last && t == 0) {
        start[G] = total;
        *ss.nout[blockIdx.y] = total;
    }
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
// Storage not cell-sorted (after an in-place push): per-particle 64-bit
// integer atomics into node-row mesh 0 (each particle's four corners); mesh 1
// (the sums from the row below) was zeroed, and the node pass adds mesh 0 and
// mesh 1 as after a tile deposit.
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
    long long q = cq[n] + cq[gp + n];
    double f = (double)q * cqf;
    final_mesh[n] = f;
    rho[n] += f * chrg;
    if (energy) {
        long long e = ce[n] + ce[gp + n];
        energy_mesh[n] = (double)e * cef;
    }
}

__global__ __launch_bounds__(kBlock) void k_deposit_finalize_all(int G, int gp, FinalizeAll a, int energy,
                                                                 double* __restrict__ rho)
{
    int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= G) return;
    rho[n] = finalize_node(a, n, gp, energy);
}

// ================================================================= push
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

// The particle kernels of the step take one or two species per launch (a
// species set): workgroup t of the launch serves species t % ns, so both
// species' tiles of one key range are dealt to the same XCD (their field rows
// meet in one L2) and the launch drains once for both.
static constexpr int kSetMax = 2;

struct TileSet {
    int ns;
    ParticleBuf b[kSetMax];
    SpeciesDev d[kSetMax];
    const int32_t* start[kSetMax];
    int32_t* counts[kSetMax];
    int32_t* rank[kSetMax];
    unsigned long long* absorbed[kSetMax];
    int32_t* start_old[kSetMax];
    int32_t* blk_first[kSetMax];
};

struct DepositSet {
    int ns;
    ParticleBuf b[kSetMax];
    DepositArgs a[kSetMax];
    const int32_t* start[kSetMax];
    long long* cq[kSetMax];
    long long* ce[kSetMax];
};

struct ScatterSet {
    int ns;
    int nb0;                         // workgroups of species 0
    const int32_t* nsrc[kSetMax];
    ParticleBuf src[kSetMax], dst[kSetMax];
    SpeciesDev d[kSetMax];
    const int32_t* rank[kSetMax];
    const int32_t* start[kSetMax];
    const int32_t* start_old[kSetMax];
    const int32_t* blk_first[kSetMax];
};

// CELL32: the particles' cells come from the cell starts (the cell array is
// neither read here nor written by k_scatter_push); the starts are kept in
// start_old for the scatter, with the cell of every 256th particle in
// blk_first.  kTileCodeCap: the mean tile holds 256 cells x 16 particles =
// 4096, and the workgroup's LDS stays within a fifth of the CU's.
// TC: tile width in cells (256 at the usual occupancy; narrower tiles keep
// dense runs' tiles within the code table).
static constexpr int kTileCodeCap = 4096;
template <int L, int GENB, int TC, int QN = 1>
__global__ __launch_bounds__(kBlock) void k_push_tile(TileSet ts, GridDev g, int tiles_per_row, int ntiles,
                                                      const double4* __restrict__ fld, const int* __restrict__ guard,
                                                      int rev)
{
    // a push enqueued behind a Poisson check that has not passed does nothing
    if (guard && guard[0] != 1) return;
    __shared__ int hist[kHR * (TC + 2)];
    __shared__ int codes[kTileCodeCap];
    __shared__ unsigned long long nabs[2];
    constexpr bool WF = L == PIC_LAYOUT_CELL32;
    __shared__ double wf[WF ? 6 * (TC + 1) : 1];
    const int t = xcd_tile(blockIdx.x, ts.ns * ntiles, rev);
    const int sp = t % ts.ns, tile = t / ts.ns;
    if constexpr (QN > 1) {
        static_assert(L == PIC_LAYOUT_CELL32, "runs of particles per thread serve the CELL32 tile push");
        push_tile_body_runs<GENB, TC, QN, kTileCodeCap>(ts.b[sp], g, ts.d[sp], tiles_per_row, tile, fld,
                                                        ts.start[sp], ts.counts[sp], ts.rank[sp], ts.absorbed[sp],
                                                        hist, codes, nabs, wf, ts.start_old[sp], ts.blk_first[sp]);
        return;
    }
    push_tile_body<L, GENB, TC, WF, WF, kTileCodeCap>(ts.b[sp], g, ts.d[sp], tiles_per_row, tile, fld,
                                                          ts.start[sp], ts.counts[sp], ts.rank[sp], ts.absorbed[sp],
                                                          hist, codes, nabs, wf, ts.start_old[sp], ts.blk_first[sp]);
}

template <int L, int NQ>
__global__ __launch_bounds__(kBlock) void k_deposit_tile(DepositSet ds, GridDev g, int tiles_per_row, int ntiles,
                                                         int rev)
{
    __shared__ unsigned long long acc[8][kDepA];
    __shared__ int sst[kDepA + 1];
    const int t = xcd_tile(blockIdx.x, ds.ns * ntiles, rev);
    const int sp = t % ds.ns, tile = t / ds.ns;
    deposit_tile_body<L, NQ>(ds.b[sp], g, ds.a[sp], tiles_per_row, tile, ds.start[sp], ds.cq[sp], ds.ce[sp], acc,
                             sst);
}

// Recompute the push and write each surviving particle to its sorted slot.
// CELL32: the workgroup's particles are cell-sorted, so their cells are a
// flat key range from kf = blk_first[b] (the cell of its first particle) to
// at most blk_first[b + 1] (the cell of the next workgroup's first particle).
// When that range is short (the usual case: ~16 cells for 256 particles) the
// corner nodes kf .. kl+1 and the same range one node row up are staged in
// LDS with the range's cell starts (start_old: the order before this sort),
// and each particle's cell is found among those starts; otherwise by a search
// of start_old over [kf, G).  The cell array is neither read nor written.
// (Staging the new starts of the keys the workgroup's particles reach in LDS
// too made the kernel slower: those loads precede the barrier, while the
// per-particle start load overlaps the push arithmetic.)
// Workgroups 0 .. nb0-1 (in XCD order) serve species 0, the rest species 1.
static constexpr int kScatterWin = 130;

template <int L, int GENB>
__global__ __launch_bounds__(kBlock) void k_scatter_push(ScatterSet ss, GridDev g, const double4* __restrict__ fld,
                                                         int rev)
{
    constexpr bool C32 = L == PIC_LAYOUT_CELL32;
    __shared__ double4 wf[C32 ? 2 * kScatterWin : 1];
    __shared__ int sst[C32 ? kScatterWin : 1];
    // consecutive particle blocks on one XCD: the partial cache lines two
    // neighbouring blocks write at their boundary meet in one L2
    const int t = xcd_tile(blockIdx.x, gridDim.x, rev);
    const int sp = t < ss.nb0 ? 0 : 1;
    const int b = t - sp * ss.nb0;
    const ParticleBuf src = ss.src[sp];
    const int n = *ss.nsrc[sp];
    const int i0 = b * blockDim.x;
    if (i0 >= n) return;
    const int i = i0 + threadIdx.x;
    const bool have = i < n;
    PState<L> p;
    if (have) p = load_state<L, !C32>(src, i);
    const int rk = have ? ss.rank[sp][i] : 0;
    const int32_t* __restrict__ nstart = ss.start[sp];
    Corners c;
    if constexpr (C32) {
        const int32_t* __restrict__ start_old = ss.start_old[sp];
        const int32_t* __restrict__ blk_first = ss.blk_first[sp];
        const int kf = blk_first[b];
        const int G = g.gx * g.gy;
        const int kl = i0 + kBlock < n ? blk_first[b + 1] : G - 1;
        const int span = kl - kf + 2;                // nodes kf .. kl + 1
        const bool win = span >= 2 && span <= kScatterWin;
        if (win)
            for (int j = threadIdx.x; j < 2 * span; j += kBlock) {
                const int row = j >= span ? 1 : 0;
                const int col = j - row * span;
                wf[j] = fld[kf + row * g.gx + col];
                if (row == 0) sst[col] = start_old[kf + col];
            }
        __syncthreads();
        if (!have) return;
        p.cell = kf + (win ? cell_of(sst, span - 2, i) : cell_of(start_old + kf, G - 1 - kf, i));
        if (win) {
            split_key(p.cell, g, c.cgx, c.cgy);
            const int lc = p.cell - kf;
            c.f00 = wf[lc];
            c.f10 = wf[lc + 1];
            c.f01 = wf[span + lc];
            c.f11 = wf[span + lc + 1];
        } else {
            c = gather<L>(p, g, fld);
        }
    } else {
        if (!have) return;
        c = gather<L>(p, g, fld);
    }
    const Pushed<L> o = push_state<L, GENB>(p, c, g, ss.d[sp]);
    if (o.key < 0) return;
    const int d = nstart[o.key] + rk;
    store_state<L>(ss.dst[sp], d, o);
    if constexpr (!C32) ss.dst[sp].cell[d] = o.key;
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
    PIC_TRY(inject_settle(c));
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
    sp.cells_stale = 0;
    return PIC_OK;
}

// Cell array of the current (sorted) storage from the cell starts: the
// particles start[k] .. start[k+1]-1 are in cell k.
__global__ __launch_bounds__(kBlock) void k_fill_cells(int G, const int32_t* __restrict__ start,
                                                       int32_t* __restrict__ cell) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= G) return;
    const int e = start[k + 1];
    for (int i = start[k]; i < e; ++i) cell[i] = k;
}

int ensure_cells(pic_ctx* c, int s) {
    Species& sp = c->sp[s];
    if (!sp.cells_stale) return PIC_OK;
    hipLaunchKernelGGL(k_fill_cells, dim3(blocks_for(c->G)), dim3(kBlock), 0, c->stream, c->G,
                       (const int32_t*)sp.start, sp.buf[sp.cur].cell);
    PIC_HIP(hipGetLastError());
    sp.cells_stale = 0;
    return PIC_OK;
}

int launch_key_from_cell(pic_ctx* c, int s, int64_t n) {
    PIC_TRY(inject_settle(c));
    PIC_TRY(ensure_cells(c, s));
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

// Exclusive scan of the cell counts into start (counts cleared), the live
// count into ndev[dst]; species s0 and, if s1 >= 0, s1 in the same launches.
static int scan_counts2(pic_ctx* c, int s0, int s1) {
    ScanSet ss{};
    const int n = s1 >= 0 ? 2 : 1;
    const int ids[2] = {s0, s1 >= 0 ? s1 : s0};
    for (int k = 0; k < 2; ++k) {
        Species& sp = c->sp[ids[k]];
        ss.counts[k] = sp.counts;
        ss.start[k] = sp.start;
        ss.bsum[k] = c->scan_bsum + (size_t)k * (c->scan_nb + 1);
        ss.nout[k] = sp.ndev + (sp.cur ^ 1);
    }
    hipLaunchKernelGGL(k_scan_local, dim3(c->scan_nb, n), dim3(kBlock), 0, c->stream, ss, c->G);
    if (c->scan_nb <= kScanDirectChunks) {
        hipLaunchKernelGGL(k_scan_add<false>, dim3(blocks_for(c->G), n), dim3(kBlock), 0, c->stream, ss, c->scan_nb,
                           c->G);
    } else {
        hipLaunchKernelGGL(k_scan_top, dim3(1, n), dim3(kBlock), 0, c->stream, ss, c->scan_nb);
        hipLaunchKernelGGL(k_scan_add<true>, dim3(blocks_for(c->G), n), dim3(kBlock), 0, c->stream, ss, c->scan_nb,
                           c->G);
    }
    PIC_HIP(hipGetLastError());
    return PIC_OK;
}

static int scan_counts(pic_ctx* c, Species& sp, int dst) {
    (void)dst;   // ndev[cur ^ 1]
    return scan_counts2(c, (int)(&sp - c->sp.data()), -1);
}

// Counting sort of the current storage by sp.key (upload / pic_bin / after an
// in-place push).
int launch_bin(pic_ctx* c, int s) {
    PIC_TRY(inject_settle(c));
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
    sp.cells_stale = 0;
    PIC_TRY(timer_end(c, PIC_STAGE_BIN, ev));
    return PIC_OK;
}

DepositArgs deposit_args(const pic_ctx* c, const Species& sp) {
    DepositArgs a;
    const double dx = c->g.dx, dy = c->g.dy;
    a.inv_area = 1.0 / (dx * dy);
    a.scale_q = std::ldexp(1.0, c->p.charge_shift);
    a.scale_e = std::ldexp(1.0, c->p.energy_shift);
    a.scale_eq = std::ldexp(1.0, c->p.energy_shift - c->p.charge_shift);
    a.half_x = std::ldexp(1.0, c->p.charge_shift / 2);
    a.half_y = std::ldexp(1.0, c->p.charge_shift - c->p.charge_shift / 2);
    a.denom = 1 / std::fabs(sp.qm);
    a.e_fast = std::ldexp(1.0, 50 - c->p.charge_shift);
    a.energy = c->p.deposit_energy;
    a.q_fast = c->p.charge_shift <= 50;
    return a;
}

SpeciesDev species_dev(const pic_ctx* c, const Species& sp) {
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
    // constant rotation (CELL32, uniform Bz): t and s as the per-particle
    // push forms them from the gathered Bz, here from the uniform value
    d.ub = c->p.layout == PIC_LAYOUT_CELL32 && c->bz_uniform;
    d.ce = d.half_qm * d.dt;
    const double tx = d.half_qm * d.bx * d.dt, ty = d.half_qm * d.by * d.dt, tz = d.half_qm * c->bz0 * d.dt;
    d.ut[0] = tx;
    d.ut[1] = ty;
    d.ut[2] = tz;
    if (d.general_b) {
        const double rden = 1 / (1 + (tx * tx + ty * ty + tz * tz));
        d.us[0] = (2 * tx) * rden;
        d.us[1] = (2 * ty) * rden;
        d.us[2] = (2 * tz) * rden;
    } else {
        d.us[0] = d.us[1] = 0.0;
        d.us[2] = 2 * tz / (1 + tz * tz);
    }
    return d;
}

// Push mode of the launches: bit 0 in-plane B, bit 1 the constant rotation
// of a uniform Bz (CELL32).
static int push_mode(const pic_ctx* c) {
    const int genb = (c->p.bx != 0.0 || c->p.by != 0.0) ? 1 : 0;
    return genb | ((c->p.layout == PIC_LAYOUT_CELL32 && c->bz_uniform) ? 2 : 0);
}

// Mean particles per cell of the densest species among ids[0 .. n-1].
static double max_occupancy(const pic_ctx* c, const int* ids, int n) {
    const double cells = (double)c->p.NX * (double)c->p.NY;
    double m = 0.0;
    for (int k = 0; k < n; ++k) m = std::max(m, (double)c->sp[ids[k]].n_upper / cells);
    return m;
}

// Tile width of the push: params.tile_cells, else the widest of 256, 128, 64,
// 32 cells whose mean tile holds at most kTileTarget particles, well inside
// the code table (the table's particles find their cells in LDS, the rest
// search the starts).  Measured on MI355X at 16 particles per cell (2 x 2^24
// particles): push 257, 229, 219, 274 us for tiles of 256, 128, 64, 32 cells;
// at 75 (2 x 9.6M): 273, 228, 136, 124, 136 us for 256 .. 16 cells.
static constexpr int kTileTarget = 1024;
static int push_tile_cells(const pic_ctx* c, const int* ids, int n) {
    if (c->p.tile_cells) return c->p.tile_cells;
    const double ppc = max_occupancy(c, ids, n);
    int tc = kTileC;
    while (tc > 32 && tc * ppc > kTileTarget) tc >>= 1;
    return tc;
}

// Consecutive quads per deposit thread: params.deposit_group, else 1 up to
// 32 particles per cell and 2 above (fewer lanes per cell's accumulators;
// measured at 75 per cell: 127 us with 2, 170 with 1, 213 with 4).
static int deposit_quads(const pic_ctx* c, const int* ids, int n) {
    if (c->p.deposit_group) return c->p.deposit_group;
    return max_occupancy(c, ids, n) > 32.0 ? 2 : 1;
}

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
static int deposit_corners(pic_ctx* c, int s);
static int deposit_corners_set(pic_ctx* c, const int* ids, int n);

int launch_deposit(pic_ctx* c, int s) {
    PIC_TRY(deposit_corners(c, s));
    hipEvent_t ev;
    PIC_TRY(timer_begin(c, PIC_STAGE_FINALIZE, &ev));
    PIC_TRY(finalize_deposit(c, c->sp[s]));
    PIC_TRY(timer_end(c, PIC_STAGE_FINALIZE, ev));
    return PIC_OK;
}

// Every species' deposit, and rho written (not accumulated) by one
// node pass over all of them.
int launch_deposit_all(pic_ctx* c, bool defer_finalize) {
    const int ns = (int)c->sp.size();
    if (ns == 0 || ns > kFinMax) {
        PIC_TRY(launch_zero_f64(c, c->rho, c->G));
        for (int s = 0; s < ns; ++s) PIC_TRY(launch_deposit(c, s));
        return PIC_OK;
    }
    for (int s0 = 0; s0 < ns; s0 += kSetMax) {
        const int ids[kSetMax] = {s0, s0 + 1};
        PIC_TRY(deposit_corners_set(c, ids, std::min(kSetMax, ns - s0)));
    }
    hipEvent_t ev;
    PIC_TRY(timer_begin(c, PIC_STAGE_FINALIZE, &ev));
    FinalizeAll a{};
    a.ns = ns;
    const double dx = c->g.dx, dy = c->g.dy;
    for (int s = 0; s < ns; ++s) {
        Species& sp = c->sp[s];
        a.cq[s] = (const long long*)sp.q_charge;
        a.ce[s] = (const long long*)sp.q_energy;
        a.cqf[s] = sp.W * (dx * dy) * std::ldexp(1.0, -c->p.charge_shift);
        a.cef[s] = sp.W * (dx * dy) * std::ldexp(1.0, -c->p.energy_shift);
        a.chrg[s] = -sp.charge / c->p.epsilon;
        a.fm[s] = sp.final_mesh;
        a.em[s] = sp.energy_mesh;
    }
    if (defer_finalize) {
        // the next Poisson solve's first kernel finalizes (and overwrites) rho
        c->fin = a;
        c->fin_pending = 1;
    } else {
        hipLaunchKernelGGL(k_deposit_finalize_all, dim3(blocks_for(c->G)), dim3(kBlock), 0, c->stream, c->G, c->g.gp,
                           a, c->p.deposit_energy, c->rho);
        PIC_HIP(hipGetLastError());
        c->meshes_stale = 0;
    }
    PIC_TRY(timer_end(c, PIC_STAGE_FINALIZE, ev));
    return PIC_OK;
}

// final_mesh / energy_mesh of the deposit in `a` (rho untouched).
__global__ __launch_bounds__(kBlock) void k_finalize_meshes(int G, int gp, FinalizeAll a, int energy) {
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n < G) (void)finalize_node<true>(a, n, gp, energy);
}

int materialize_meshes(pic_ctx* c) {
    if (!c->meshes_stale) return PIC_OK;
    hipLaunchKernelGGL(k_finalize_meshes, dim3(blocks_for(c->G)), dim3(kBlock), 0, c->stream, c->G, c->g.gp, c->fin,
                       c->p.deposit_energy);
    PIC_HIP(hipGetLastError());
    c->meshes_stale = 0;
    return PIC_OK;
}

// Charge (and energy) corner sums of the species in ids[0 .. n-1]: the
// sorted ones by one tile launch for all of them, the others (storage left
// unsorted by an in-place push) by per-particle atomics.
static int deposit_corners_set(pic_ctx* c, const int* ids, int n) {
    hipEvent_t ev;
    PIC_TRY(timer_begin(c, PIC_STAGE_DEPOSIT, &ev));
    DepositSet ds{};
    for (int k = 0; k < n; ++k) {
        Species& sp = c->sp[ids[k]];
        if (sp.deposit_pending) continue;
        DepositArgs a = deposit_args(c, sp);
        if (sp.sorted) {
            const int j = ds.ns++;
            ds.b[j] = sp.buf[sp.cur];
            ds.a[j] = a;
            ds.start[j] = sp.start;
            ds.cq[j] = (long long*)sp.q_charge;
            ds.ce[j] = (long long*)sp.q_energy;
        } else {
            PIC_TRY(inject_settle(c));   // the per-particle grid needs the live-count bound
            const int64_t nacc = 2 * (int64_t)c->g.gp;
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
            PIC_HIP(hipGetLastError());
        }
        sp.deposit_pending = 1;
    }
    if (ds.ns > 0) {
        const int tpr = (c->g.gx + kDepC - 1) / kDepC;
        const int ntiles = tpr * c->g.gy;
        const dim3 grid((unsigned)(ds.ns * ntiles));
        const int nq = deposit_quads(c, ids, n);
        hipEvent_t kt[2];
        PIC_TRY(kt_begin(c, kt, PIC_STAGE_DEPOSIT));
        const int rev = c->sweep;
        c->sweep ^= 1;
#define PIC_DEP(LAY, NQ) launch_timed(k_deposit_tile<LAY, NQ>, grid, dim3(kBlock), c->stream, kt, ds, c->g, tpr, ntiles, rev)
        if (c->p.layout == PIC_LAYOUT_F64) {
            if (nq == 4) PIC_DEP(PIC_LAYOUT_F64, 4);
            else if (nq == 2) PIC_DEP(PIC_LAYOUT_F64, 2);
            else PIC_DEP(PIC_LAYOUT_F64, 1);
        } else {
            if (nq == 4) PIC_DEP(PIC_LAYOUT_CELL32, 4);
            else if (nq == 2) PIC_DEP(PIC_LAYOUT_CELL32, 2);
            else PIC_DEP(PIC_LAYOUT_CELL32, 1);
        }
#undef PIC_DEP
        PIC_HIP(hipGetLastError());
        kt_end(c, PIC_STAGE_DEPOSIT, kt);
    }
    PIC_TRY(timer_end(c, PIC_STAGE_DEPOSIT, ev));
    return PIC_OK;
}

static int deposit_corners(pic_ctx* c, int s) {
    return deposit_corners_set(c, &s, 1);
}

// In-place push, no re-bin (keys left in sp.key).
int launch_push(pic_ctx* c, int s) {
    PIC_TRY(inject_settle(c));
    PIC_TRY(ensure_cells(c, s));
    Species& sp = c->sp[s];
    hipEvent_t ev;
    PIC_TRY(timer_begin(c, PIC_STAGE_PUSH, &ev));
    SpeciesDev d = species_dev(c, sp);
    const int64_t n = sp.n_upper;
    if (n > 0) {
        dim3 grid(blocks_for(n)), blk(kBlock);
        const bool f64 = c->p.layout == PIC_LAYOUT_F64;
#define PIC_PUSH_ARGS grid, blk, 0, c->stream, sp.ndev + sp.cur, sp.buf[sp.cur], c->g, d, c->fld, sp.key, sp.absorbed
        switch (f64 ? (d.general_b ? 1 : 0) : 4 + push_mode(c)) {
        case 0: hipLaunchKernelGGL((k_push<PIC_LAYOUT_F64, 0>), PIC_PUSH_ARGS); break;
        case 1: hipLaunchKernelGGL((k_push<PIC_LAYOUT_F64, 1>), PIC_PUSH_ARGS); break;
        case 4: hipLaunchKernelGGL((k_push<PIC_LAYOUT_CELL32, 0>), PIC_PUSH_ARGS); break;
        case 5: hipLaunchKernelGGL((k_push<PIC_LAYOUT_CELL32, 1>), PIC_PUSH_ARGS); break;
        case 6: hipLaunchKernelGGL((k_push<PIC_LAYOUT_CELL32, 2>), PIC_PUSH_ARGS); break;
        default: hipLaunchKernelGGL((k_push<PIC_LAYOUT_CELL32, 3>), PIC_PUSH_ARGS); break;
        }
#undef PIC_PUSH_ARGS
        PIC_HIP(hipGetLastError());
    }
    sp.deposit_pending = 0;
    sp.sorted = 0;
    PIC_TRY(timer_end(c, PIC_STAGE_PUSH, ev));
    return PIC_OK;
}

template <int L, int GENB, int QN>
static void tile_kernel_q(pic_ctx* c, const TileSet& ts, int tc, const hipEvent_t kt[2], const int* guard) {
    const int tpr = (c->g.gx + tc - 1) / tc, ntiles = tpr * c->g.gy;
    const dim3 grid((unsigned)(ts.ns * ntiles));
    const double4* fld = (const double4*)c->fld;
    const int rev = c->sweep;
    c->sweep ^= 1;
    switch (tc) {
    case 32: launch_timed(k_push_tile<L, GENB, 32, QN>, grid, dim3(kBlock), c->stream, kt, ts, c->g, tpr, ntiles, fld, guard, rev); break;
    case 64: launch_timed(k_push_tile<L, GENB, 64, QN>, grid, dim3(kBlock), c->stream, kt, ts, c->g, tpr, ntiles, fld, guard, rev); break;
    case 128: launch_timed(k_push_tile<L, GENB, 128, QN>, grid, dim3(kBlock), c->stream, kt, ts, c->g, tpr, ntiles, fld, guard, rev); break;
    default: launch_timed(k_push_tile<L, GENB, kTileC, QN>, grid, dim3(kBlock), c->stream, kt, ts, c->g, tpr, ntiles, fld, guard, rev); break;
    }
}

// Consecutive particles per thread of the CELL32 tile push:
// params.push_group, else 1 up to 32 particles per cell and 2 above (measured
// at 75 per cell: push 113 us with 2, 122 with 1, 130 with 4; at 16: 218 with
// 1, 238 with 2, 264 with 4 — one particle per thread keeps the next
// particle's loads in flight behind the current push).
static int push_run(const pic_ctx* c, double ppc) {
    if (c->p.push_group) return c->p.push_group;
    return ppc > 32.0 ? 2 : 1;
}

template <int L, int GENB>
static void tile_kernel(pic_ctx* c, const TileSet& ts, int tc, int qn, const hipEvent_t kt[2], const int* guard) {
    if constexpr (L == PIC_LAYOUT_CELL32) {
        if (qn == 4) return tile_kernel_q<L, GENB, 4>(c, ts, tc, kt, guard);
        if (qn == 2) return tile_kernel_q<L, GENB, 2>(c, ts, tc, kt, guard);
    }
    tile_kernel_q<L, GENB, 1>(c, ts, tc, kt, guard);
}

template <int L, int GENB>
static void scatter_kernel(pic_ctx* c, const ScatterSet& ss, unsigned grid, const hipEvent_t kt[2]) {
    const int rev = c->sweep;
    c->sweep ^= 1;
    launch_timed(k_scatter_push<L, GENB>, dim3(grid), dim3(kBlock), c->stream, kt, ss, c->g, (const double4*)c->fld,
                 rev);
}

// Gather + Boris + walls + re-bin of the species in ids[0 .. n-1]: tile
// push/count in one launch (move_push_set), prefix scan and recomputing
// scatter into the other buffer (move_scatter_set).  Injection runs between
// the two phases: its particles are counted into the same cells before the
// scan.
int move_push_set(pic_ctx* c, const int* ids, int n, const int* guard) {
    PIC_TRY(inject_settle(c));
    TileSet ts{};
    for (int k = 0; k < n; ++k) {
        Species& sp = c->sp[ids[k]];
        if (!sp.sorted) PIC_TRY(launch_bin(c, ids[k]));   // tiles need cell-sorted storage
        if (sp.n_upper == 0) continue;
        const int j = ts.ns++;
        ts.b[j] = sp.buf[sp.cur];
        ts.d[j] = species_dev(c, sp);
        ts.start[j] = sp.start;
        ts.counts[j] = sp.counts;
        ts.rank[j] = sp.rank;
        ts.absorbed[j] = sp.absorbed;
        ts.start_old[j] = sp.start_old;
        ts.blk_first[j] = sp.blk_first;
    }
    const int tc = push_tile_cells(c, ids, n);
    const int qn = push_run(c, max_occupancy(c, ids, n));
    const bool f64 = c->p.layout == PIC_LAYOUT_F64;
    const int mode = push_mode(c);
    hipEvent_t ev;
    PIC_TRY(timer_begin(c, PIC_STAGE_PUSH, &ev));
    if (ts.ns > 0) {
        hipEvent_t kt[2];
        PIC_TRY(kt_begin(c, kt, PIC_STAGE_PUSH));
        switch (f64 ? (mode & 1) : 4 + mode) {
        case 0: tile_kernel<PIC_LAYOUT_F64, 0>(c, ts, tc, qn, kt, guard); break;
        case 1: tile_kernel<PIC_LAYOUT_F64, 1>(c, ts, tc, qn, kt, guard); break;
        case 4: tile_kernel<PIC_LAYOUT_CELL32, 0>(c, ts, tc, qn, kt, guard); break;
        case 5: tile_kernel<PIC_LAYOUT_CELL32, 1>(c, ts, tc, qn, kt, guard); break;
        case 6: tile_kernel<PIC_LAYOUT_CELL32, 2>(c, ts, tc, qn, kt, guard); break;
        default: tile_kernel<PIC_LAYOUT_CELL32, 3>(c, ts, tc, qn, kt, guard); break;
        }
        PIC_HIP(hipGetLastError());
        kt_end(c, PIC_STAGE_PUSH, kt);
    }
    PIC_TRY(timer_end(c, PIC_STAGE_PUSH, ev));
    return PIC_OK;
}

int move_push(pic_ctx* c, int s, const int* guard) {
    return move_push_set(c, &s, 1, guard);
}

int move_push_all(pic_ctx* c, const int* guard) {
    const int ns = (int)c->sp.size();
    for (int s0 = 0; s0 < ns; s0 += kSetMax) {
        const int ids[kSetMax] = {s0, s0 + 1};
        PIC_TRY(move_push_set(c, ids, std::min(kSetMax, ns - s0), guard));
    }
    return PIC_OK;
}

int move_finish(pic_ctx* c, int s) {
    Species& sp = c->sp[s];
    hipEvent_t ev;
    PIC_TRY(timer_begin(c, PIC_STAGE_BIN, &ev));
    PIC_TRY(scan_counts(c, sp, sp.cur ^ 1));
    PIC_TRY(timer_end(c, PIC_STAGE_BIN, ev));
    return move_scatter(c, s);
}

int move_scan_pair(pic_ctx* c) {
    hipEvent_t ev;
    PIC_TRY(timer_begin(c, PIC_STAGE_BIN, &ev));
    PIC_TRY(scan_counts2(c, 0, 1));
    PIC_TRY(timer_end(c, PIC_STAGE_BIN, ev));
    return PIC_OK;
}

int move_scatter_set(pic_ctx* c, const int* ids, int n) {
    ScatterSet ss{};
    unsigned nb[kSetMax] = {0, 0};
    for (int k = 0; k < n; ++k) {
        Species& sp = c->sp[ids[k]];
        const int src = sp.cur, dst = sp.cur ^ 1;
        if (sp.n_upper > 0) {
            const int j = ss.ns++;
            nb[j] = blocks_for(sp.n_upper);
            ss.nsrc[j] = sp.ndev + src;
            ss.src[j] = sp.buf[src];
            ss.dst[j] = sp.buf[dst];
            ss.d[j] = species_dev(c, sp);
            ss.rank[j] = sp.rank;
            ss.start[j] = sp.start;
            ss.start_old[j] = sp.start_old;
            ss.blk_first[j] = sp.blk_first;
        }
        sp.cur = dst;
        sp.deposit_pending = 0;
        sp.sorted = 1;
        sp.cells_stale = c->p.layout == PIC_LAYOUT_CELL32 ? 1 : 0;
    }
    ss.nb0 = (int)nb[0];
    const bool f64 = c->p.layout == PIC_LAYOUT_F64;
    const int mode = push_mode(c);
    hipEvent_t ev;
    PIC_TRY(timer_begin(c, PIC_STAGE_SCATTER, &ev));
    if (ss.ns > 0) {
        const unsigned grid = nb[0] + nb[1];
        hipEvent_t kt[2];
        PIC_TRY(kt_begin(c, kt, PIC_STAGE_SCATTER));
        switch (f64 ? (mode & 1) : 4 + mode) {
        case 0: scatter_kernel<PIC_LAYOUT_F64, 0>(c, ss, grid, kt); break;
        case 1: scatter_kernel<PIC_LAYOUT_F64, 1>(c, ss, grid, kt); break;
        case 4: scatter_kernel<PIC_LAYOUT_CELL32, 0>(c, ss, grid, kt); break;
        case 5: scatter_kernel<PIC_LAYOUT_CELL32, 1>(c, ss, grid, kt); break;
        case 6: scatter_kernel<PIC_LAYOUT_CELL32, 2>(c, ss, grid, kt); break;
        default: scatter_kernel<PIC_LAYOUT_CELL32, 3>(c, ss, grid, kt); break;
        }
        PIC_HIP(hipGetLastError());
        kt_end(c, PIC_STAGE_SCATTER, kt);
    }
    PIC_TRY(timer_end(c, PIC_STAGE_SCATTER, ev));
    return PIC_OK;
}

int move_scatter(pic_ctx* c, int s) {
    return move_scatter_set(c, &s, 1);
}

int move_scatter_pair(pic_ctx* c) {
    const int ids[2] = {0, 1};
    return move_scatter_set(c, ids, 2);
}

int launch_move(pic_ctx* c, int s) {
    PIC_TRY(move_push(c, s));
    return move_finish(c, s);
}

// ====================================================== kinetic energy
// Sum of 0.5 |v|^2 / |q/m| over the live particles (calculate_energy,
// utils.cpp:77-118).  A fixed grid walks the particles grid-stride, each
// workgroup reduces in a fixed tree and writes one partial; one workgroup adds
// the partials in order: the FP64 result does not depend on scheduling.
static constexpr int kKeBlocks = 1024;

template <int L>
__global__ __launch_bounds__(kBlock) void k_kinetic_partial(const int32_t* __restrict__ ndev, ParticleBuf b,
                                                            double denom, double* __restrict__ part) {
    using V = typename Lay<L>::V;
    __shared__ double red[kBlock];
    const int n = *ndev;
    double acc = 0.0;
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += kKeBlocks * kBlock) {
        const double vx = (double)((const V*)b.vx)[i], vy = (double)((const V*)b.vy)[i];
        const double vz = (double)((const V*)b.vz)[i];
        const double v2 = vx * vx + vy * vy + vz * vz;
        acc += 0.5 * v2 / denom;
    }
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int w = kBlock / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

__global__ __launch_bounds__(kKeBlocks) void k_kinetic_total(const double* __restrict__ part, double* out) {
    __shared__ double red[kKeBlocks];
    red[threadIdx.x] = part[threadIdx.x];
    __syncthreads();
    for (int w = kKeBlocks / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = red[0];
}

int launch_kinetic_energy(pic_ctx* c, int s, double* out) {
    PIC_TRY(inject_settle(c));
    Species& sp = c->sp[s];
    static_assert(kKeBlocks + 1 <= PIC_PART_DOUBLES, "partials fit the dot-product scratch");
    double* part = c->part;   // dot-product partials: free between Poisson solves
    const double denom = std::fabs(sp.qm);
    if (c->p.layout == PIC_LAYOUT_F64)
        hipLaunchKernelGGL(k_kinetic_partial<PIC_LAYOUT_F64>, dim3(kKeBlocks), dim3(kBlock), 0, c->stream,
                           sp.ndev + sp.cur, sp.buf[sp.cur], denom, part);
    else
        hipLaunchKernelGGL(k_kinetic_partial<PIC_LAYOUT_CELL32>, dim3(kKeBlocks), dim3(kBlock), 0, c->stream,
                           sp.ndev + sp.cur, sp.buf[sp.cur], denom, part);
    hipLaunchKernelGGL(k_kinetic_total, dim3(1), dim3(kKeBlocks), 0, c->stream, part, part + kKeBlocks);
    PIC_HIP(hipGetLastError());
    PIC_HIP(hipMemcpyAsync(out, part + kKeBlocks, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    PIC_HIP(hipStreamSynchronize(c->stream));
    return PIC_OK;
}

// ======================================== counting sort without allocation
// pic_bin_keys_noalloc sorts cell keys held in caller memory with the bin
// stage's kernels; the scan's chunk totals (and its grand total and live
// count) go to `scratch`, a fixed array of this library on every GPU, so the
// call allocates nothing.  Host code never passes `scratch` itself to a
// kernel: each device's address comes from hipGetSymbolAddress, taken once
// per device on its first call.
static constexpr int kNoAllocMaxKeys = 2048 * 8192;                   // the grid node limit
static constexpr int kNoAllocChunks = kNoAllocMaxKeys / kScanChunk;   // 8192
static __device__ __align__(256) int32_t scratch[kNoAllocChunks + 2];

static constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static int32_t* g_scratch[kMaxDevices];

static int scratch_address(int32_t** out) {
    int dev = 0;
    PIC_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) {
        set_error("pic_bin_keys_noalloc: device %d beyond the %d this library tracks", dev, kMaxDevices);
        return PIC_ERR_RANGE;
    }
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            set_error("pic_bin_keys_noalloc: hipGetSymbolAddress of the scan scratch failed on device %d: %s", dev,
                      hipGetErrorString(e));
            return PIC_ERR_HIP;
        }
        g_scratch[dev] = (int32_t*)p;
    }
    *out = g_scratch[dev];
    return PIC_OK;
}

__global__ __launch_bounds__(kBlock) void k_na_zero(int G, int32_t* __restrict__ counts) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < G) counts[i] = 0;
}

// slot[i] = the key's place among the equal keys so far (-1: key outside [0, G))
__global__ __launch_bounds__(kBlock) void k_na_count(int64_t n, const int32_t* __restrict__ key, int G,
                                                     int32_t* __restrict__ counts, int32_t* __restrict__ slot) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int k = key[i];
    slot[i] = (k >= 0 && k < G) ? atomicAdd(&counts[k], 1) : -1;
}

// counts again from the starts (the scan clears the counts it reads); slots
// become sorted positions
__global__ __launch_bounds__(kBlock) void k_na_finish(int64_t n, int G, const int32_t* __restrict__ key,
                                                      const int32_t* __restrict__ start,
                                                      int32_t* __restrict__ counts, int32_t* __restrict__ slot) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < G) counts[i] = start[i + 1] - start[i];
    if (i < n) {
        const int r = slot[i];
        if (r >= 0) slot[i] = start[key[i]] + r;
    }
}

int bin_keys_noalloc(const int32_t* key, int64_t n, int32_t G, int32_t* counts, int32_t* start, int32_t* slot,
                     hipStream_t s) {
    if (n < 0 || n > (int64_t)INT32_MAX) {
        set_error("pic_bin_keys_noalloc: %lld keys; at most %d", (long long)n, INT32_MAX);
        return PIC_ERR_RANGE;
    }
    if (G < 1 || G > kNoAllocMaxKeys) {
        set_error("pic_bin_keys_noalloc: %d cells; the fixed scan scratch serves 1 to %d", G, kNoAllocMaxKeys);
        return PIC_ERR_RANGE;
    }
    if (!counts || !start || (n > 0 && (!key || !slot))) {
        set_error("pic_bin_keys_noalloc: null pointer");
        return PIC_ERR_ARG;
    }
    if ((((uintptr_t)counts) | ((uintptr_t)start)) & 15) {
        set_error("pic_bin_keys_noalloc: counts and start must be 16-byte aligned (the scan moves 4 ints at once)");
        return PIC_ERR_ARG;
    }
    int32_t* sc = nullptr;
    PIC_TRY(scratch_address(&sc));
    const int nb = (G + kScanChunk - 1) / kScanChunk;
    hipLaunchKernelGGL(k_na_zero, dim3(blocks_for(G)), dim3(kBlock), 0, s, G, counts);
    if (n > 0)
        hipLaunchKernelGGL(k_na_count, dim3(blocks_for(n)), dim3(kBlock), 0, s, n, key, G, counts, slot);
    ScanSet ss{};
    for (int k = 0; k < 2; ++k) {
        ss.counts[k] = counts;
        ss.start[k] = start;
        ss.bsum[k] = sc;
        ss.nout[k] = sc + kNoAllocChunks + 1;
    }
    hipLaunchKernelGGL(k_scan_local, dim3(nb, 1), dim3(kBlock), 0, s, ss, G);
    if (nb <= kScanDirectChunks) {
        hipLaunchKernelGGL(k_scan_add<false>, dim3(blocks_for(G), 1), dim3(kBlock), 0, s, ss, nb, G);
    } else {
        hipLaunchKernelGGL(k_scan_top, dim3(1, 1), dim3(kBlock), 0, s, ss, nb);
        hipLaunchKernelGGL(k_scan_add<true>, dim3(blocks_for(G), 1), dim3(kBlock), 0, s, ss, nb, G);
    }
    hipLaunchKernelGGL(k_na_finish, dim3(blocks_for(std::max<int64_t>(n, G))), dim3(kBlock), 0, s, n, G, key, start,
                       counts, slot);
    PIC_HIP(hipGetLastError());
    return PIC_OK;
}

} // namespace pic

extern "C" int pic_bin_keys_noalloc(const int32_t* key, int64_t n, int32_t G, int32_t* counts, int32_t* start,
                                    int32_t* slot, void* stream) {
    return pic::bin_keys_noalloc(key, n, G, counts, start, slot, (hipStream_t)stream);
}
