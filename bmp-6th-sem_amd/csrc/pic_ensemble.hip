// pic_ensemble.hip — batched execution of an ensemble of independent PIC runs
// of uneven sizes (grids, particle counts, electrode voltage, Bz) in one step.
//
// Every stage of the step is one launch for the whole ensemble:
//   deposit   persistent over (run, tile) items, dealt by workgroup stride
//   finalize  persistent over (run, node chunk): corner sums -> meshes, rho
//   Poisson   one workgroup per run: the run's whole MG-preconditioned PCG
//             (FP64 vectors, FP32 V(2,2) weighted-Jacobi multigrid, direct
//             coarsest solve) with workgroup barriers only
//   E field   persistent over (run, node chunk)
//   push      persistent over (run, species, tile) items: push + LDS slot count
//   scan      one workgroup per (run, species)
//   scatter   persistent over (run, species, particle chunk)
// so the step has a fixed launch sequence with no host decisions and is
// captured once in a hipGraph per buffer parity and replayed.  The tile width
// and the number of persistent workgroups come from pic_ensemble_tune, which
// times the candidates on the ensemble itself.
//
// The runs are ordinary contexts (pic_create, geometry, species, particles);
// the ensemble reads their buffers through device descriptor tables and leaves
// them in the state a serial pic_step would (sorted storage, phi, E, meshes).
#include "pic_particle_dev.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <vector>

namespace pic {

static constexpr int kEnsMaxLev = 12;
static constexpr int kEnsChunk = 8;          // scatter item: kEnsChunk x 256 particles

struct EnsLevel {
    int gx, gy;
    float cx, cy, diag, wd;
    float wd2;                    // second pre-smoothing sweep (first post-smoothing sweep)
    const uint8_t* dir;
    float *e, *b, *t;
};

struct EnsSpecies {
    ParticleBuf src, dst;
    SpeciesDev sd;
    DepositArgs da;
    const int32_t* nsrc;
    int32_t* ndst;
    int32_t *start, *counts, *rank;
    long long *cq, *ce;
    double *final_mesh, *energy_mesh;
    unsigned long long* absorbed;
    double cqf, cef, chrg;
};

struct EnsRun {
    GridDev g;
    int G;
    double *rho, *phi, *ex, *ey;
    const double* bz;
    double4* fld;
    const uint8_t* dir;
    const double* dval;
    EnsSpecies sp[2];
    // Poisson
    int nl;
    EnsLevel lv[kEnsMaxLev];
    double cx, cy, diag;
    double *r, *p, *q;
    const float* inv;
    const int* unk;
    int nu;
    double tol2;
    int max_it, pre, post, sweeps;
    int* warm;                    // 1 once phi holds a solution (device state)
    int direct;                   // phi holds this step's direct solve (the PCG only checks it)
    int* iters;                   // per-run stats of the last solve
    double* relres;
};

// item -> run: largest r with prefix[r] <= item (prefix has R + 1 entries)
__device__ __forceinline__ int find_run(const int* __restrict__ prefix, int R, int item) {
    int lo = 0, hi = R - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (prefix[mid] <= item) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// Items are dealt by workgroup stride: workgroup b of a persistent launch
// takes items b, b + G, b + 2G, ... (G workgroups).  A shared queue counter
// would balance uneven items better, but its same-address atomics serialise
// at about 110 ns per item, which the ~20k items of a kernel cannot afford.

// ----------------------------------------------------------------- deposit
// items: (species, run, deposit tile); prefix over runs of the tiles of one species
__global__ __launch_bounds__(kBlock) void k_ens_deposit(const EnsRun* __restrict__ runs, int R,
                                                        const int* __restrict__ prefix) {
    __shared__ unsigned long long acc[8][kDepA];
    const int per_species = prefix[R];
    for (int it = (int)blockIdx.x; it < 2 * per_species; it += (int)gridDim.x) {
        __syncthreads();   // the previous item's LDS use is over
        const int s = it / per_species, k = it - s * per_species;
        const int r = find_run(prefix, R, k);
        const EnsRun& run = runs[r];
        const EnsSpecies& sp = run.sp[s];
        const int tpr = (run.g.gx + kDepC - 1) / kDepC;
        deposit_tile_body<PIC_LAYOUT_CELL32>(sp.src, run.g, sp.da, tpr, k - prefix[r], sp.start, sp.cq, sp.ce, acc);
    }
}

// corner sums -> final/energy meshes of both species; rho = (0 + f0 c0) + f1 c1
// (the serial step's order); items: (run, chunk of 256 nodes)
__global__ __launch_bounds__(kBlock) void k_ens_finalize(const EnsRun* __restrict__ runs, int R,
                                                         const int* __restrict__ prefix, int energy) {
    for (int it = (int)blockIdx.x; it < prefix[R]; it += (int)gridDim.x) {
        __syncthreads();   // the previous item's LDS use is over
        const int r = find_run(prefix, R, it);
        const EnsRun& run = runs[r];
        const int n = (it - prefix[r]) * kBlock + threadIdx.x;
        if (n >= run.G) continue;
        const int gp = run.g.gp;
        double rho = 0.0;
#pragma unroll
        for (int s = 0; s < 2; ++s) {
            const EnsSpecies& sp = run.sp[s];
            const long long q = sp.cq[n] + sp.cq[gp + n];
            const double f = (double)q * sp.cqf;
            sp.final_mesh[n] = f;
            rho += f * sp.chrg;
            if (energy) {
                const long long e = sp.ce[n] + sp.ce[gp + n];
                sp.energy_mesh[n] = (double)e * sp.cef;
            }
        }
        run.rho[n] = rho;
    }
}

// ----------------------------------------------------------------- Poisson
__device__ __forceinline__ double wg_sum(double v, double* sh) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) sh[wid] = v;
    __syncthreads();
    double s = 0;
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) s += sh[w];
    return s;
}

#define WG_FOR(n, N) for (int n = threadIdx.x; n < (N); n += blockDim.x)

// out = in + w (b - A in) on unknowns (zero_in: out = w b); 0 on Dirichlet
__device__ void wg_jacobi(const EnsLevel& L, const float* in, float* out, bool zero_in, float w) {
    const int G = L.gx * L.gy;
#pragma unroll 4
    WG_FOR(n, G) {
        float v = 0.f;
        if (!L.dir[n]) {
            if (zero_in) {
                v = w * L.b[n];
            } else {
                const float e = in[n];
                const float ax = L.diag * e - L.cx * (in[n - 1] + in[n + 1]) - L.cy * (in[n - L.gx] + in[n + L.gx]);
                v = e + w * (L.b[n] - ax);
            }
        }
        out[n] = v;
    }
    __syncthreads();
}

// M r by one V-cycle; r -> level 0 b; returns the array holding z (level 0)
__device__ float* wg_vcycle(const EnsRun& run, const double* r, float** ecur) {
    const int nl = run.nl;
    {
        const EnsLevel& L = run.lv[0];
#pragma unroll 4
        WG_FOR(n, L.gx * L.gy) L.b[n] = (float)r[n];
        __syncthreads();
    }
    for (int l = 0; l < nl; ++l) {
        const EnsLevel& L = run.lv[l];
        float* e = L.e;
        float* t = L.t;
        if (l == nl - 1) {
            // coarsest: exact solve, or sweeps
            const int G = L.gx * L.gy;
            if (run.nu > 0) {
                WG_FOR(n, G) e[n] = 0.f;
                __syncthreads();
                if ((int)threadIdx.x < run.nu) {
                    float v = 0.f;
                    for (int q = 0; q < run.nu; ++q) v += run.inv[threadIdx.x * run.nu + q] * L.b[run.unk[q]];
                    e[run.unk[threadIdx.x]] = v;
                }
                __syncthreads();
            } else {
                for (int s = 0; s < run.sweeps; ++s) {
                    wg_jacobi(L, e, t, s == 0, L.wd);
                    float* x = e;
                    e = t;
                    t = x;
                }
            }
            ecur[l] = e;
            break;
        }
        // pre-smoothing from zero
        for (int s = 0; s < run.pre; ++s) {
            wg_jacobi(L, e, t, s == 0, s == 1 ? L.wd2 : L.wd);
            float* x = e;
            e = t;
            t = x;
        }
        if (run.pre == 0) {
            WG_FOR(n, L.gx * L.gy) e[n] = 0.f;
            __syncthreads();
        }
        ecur[l] = e;
        // restriction of the residual (full weighting)
        const EnsLevel& C = run.lv[l + 1];
        auto res = [&](int n) -> float {
            if (L.dir[n]) return 0.f;
            const float ax = L.diag * e[n] - L.cx * (e[n - 1] + e[n + 1]) - L.cy * (e[n - L.gx] + e[n + L.gx]);
            return L.b[n] - ax;
        };
#pragma unroll 2
        WG_FOR(n, C.gx * C.gy) {
            float v = 0.f;
            if (!C.dir[n]) {
                const int J = n / C.gx, I = n - J * C.gx;
                const int f = 2 * J * L.gx + 2 * I, gx = L.gx;
                v = 0.0625f * (4.f * res(f) + 2.f * (res(f - 1) + res(f + 1) + res(f - gx) + res(f + gx)) +
                               (res(f - gx - 1) + res(f - gx + 1) + res(f + gx - 1) + res(f + gx + 1)));
            }
            C.b[n] = v;
        }
        __syncthreads();
    }
    // up
    for (int l = nl - 2; l >= 0; --l) {
        const EnsLevel& L = run.lv[l];
        const EnsLevel& C = run.lv[l + 1];
        float* e = ecur[l];
        float* t = (e == L.e) ? L.t : L.e;
        const float* ec = ecur[l + 1];
#pragma unroll 4
        WG_FOR(n, L.gx * L.gy) {
            if (L.dir[n]) continue;
            const int j = n / L.gx, i = n - j * L.gx;
            const int I = i >> 1, J = j >> 1;
            const bool oi = i & 1, oj = j & 1;
            const float* c0 = ec + J * C.gx + I;
            float v;
            if (!oi && !oj) v = c0[0];
            else if (oi && !oj) v = 0.5f * (c0[0] + c0[1]);
            else if (!oi && oj) v = 0.5f * (c0[0] + c0[C.gx]);
            else v = 0.25f * (c0[0] + c0[1] + c0[C.gx] + c0[C.gx + 1]);
            e[n] += v;
        }
        __syncthreads();
        for (int s = 0; s < run.post; ++s) {
            wg_jacobi(L, e, t, false, (s == 0 && run.post == 2) ? L.wd2 : L.wd);
            float* x = e;
            e = t;
            t = x;
        }
        ecur[l] = e;
    }
    return ecur[0];
}

// One workgroup solves run blockIdx.x: overwriteRho, flexible PCG (FP64) with
// the V-cycle, overwritePhi.  Same method as the single-run solver.
__global__ __launch_bounds__(1024) void k_ens_poisson(const EnsRun* __restrict__ runs) {
    __shared__ double sh[16];
    const EnsRun& run = runs[blockIdx.x];
    const int G = run.G, gx = run.g.gx;
    double* x = run.phi;
    double* r = run.r;
    double* p = run.p;
    double* q = run.q;
    const uint8_t* dir = run.dir;
    const double cx = run.cx, cy = run.cy, dg = run.diag;
    const int warm = *run.warm || run.direct;
    // rho[geom] = V; rho = -rho; x0 = warm ? phi : 0 with Dirichlet values
#pragma unroll 4
    WG_FOR(n, G) {
        const double v = dir[n] ? run.dval[n] : run.rho[n];
        run.rho[n] = -v;
        x[n] = dir[n] ? run.dval[n] : (warm ? x[n] : 0.0);
    }
    __syncthreads();
    // ||b||: residual of the Dirichlet-only field; r0 = b - A x0
    double bb = 0.0, rr = 0.0;
    auto D = [&](int m) -> double { return dir[m] ? run.dval[m] : 0.0; };
#pragma unroll 4
    WG_FOR(n, G) {
        double v = 0.0;
        if (!dir[n]) {
            const double b0 = run.rho[n] - (-cx * (D(n - 1) + D(n + 1)) - cy * (D(n - gx) + D(n + gx)));
            bb += b0 * b0;
            const double ax = dg * x[n] - cx * (x[n - 1] + x[n + 1]) - cy * (x[n - gx] + x[n + gx]);
            v = run.rho[n] - ax;
        }
        r[n] = v;
        rr += v * v;
    }
    bb = wg_sum(bb, sh);
    rr = wg_sum(rr, sh);
    const double rr0 = rr;
    int it = 0;
    float* ecur[kEnsMaxLev];
    if (!(rr <= run.tol2 * bb || bb == 0.0)) {
        float* z = wg_vcycle(run, r, ecur);
        double rz = 0.0;
#pragma unroll 4
        WG_FOR(n, G) {
            const double zn = (double)z[n];
            p[n] = zn;
            rz += r[n] * zn;
        }
        rz = wg_sum(rz, sh);
        while (it < run.max_it) {
            ++it;
            double pq = 0.0;
#pragma unroll 4
            WG_FOR(n, G) {
                double v = 0.0;
                if (!dir[n]) v = dg * p[n] - cx * (p[n - 1] + p[n + 1]) - cy * (p[n - gx] + p[n + gx]);
                q[n] = v;
                pq += p[n] * v;
            }
            pq = wg_sum(pq, sh);
            const double alpha = rz / pq;
            rr = 0.0;
            double rzo = 0.0;
#pragma unroll 4
            WG_FOR(n, G) {
                x[n] += alpha * p[n];
                const double rn = r[n] - alpha * q[n];
                r[n] = rn;
                rr += rn * rn;
                rzo += rn * (double)z[n];
            }
            rr = wg_sum(rr, sh);
            rzo = wg_sum(rzo, sh);
            if (!(rr == rr) || rr <= run.tol2 * bb) break;
            z = wg_vcycle(run, r, ecur);
            double rzn = 0.0;
#pragma unroll 4
            WG_FOR(n, G) rzn += r[n] * (double)z[n];
            rzn = wg_sum(rzn, sh);
            const double beta = (rzn - rzo) / rz;
            rz = rzn;
#pragma unroll 4
            WG_FOR(n, G) p[n] = (double)z[n] + beta * p[n];
            __syncthreads();
        }
    }
#pragma unroll 4
    WG_FOR(n, G) if (dir[n]) x[n] = run.dval[n];
    if (threadIdx.x == 0) {
        *run.warm = 1;
        *run.iters = it;
        run.relres[0] = bb > 0 ? sqrt(rr0 / bb) : 0.0;
        run.relres[1] = bb > 0 ? sqrt(rr / bb) : 0.0;
    }
}

// E = -grad phi with the reference's edge copies, packed gather array
__global__ __launch_bounds__(kBlock) void k_ens_efield(const EnsRun* __restrict__ runs, int R,
                                                       const int* __restrict__ prefix) {
    for (int it = (int)blockIdx.x; it < prefix[R]; it += (int)gridDim.x) {
        __syncthreads();   // the previous item's LDS use is over
        const int r = find_run(prefix, R, it);
        const EnsRun& run = runs[r];
        const int n = (it - prefix[r]) * kBlock + threadIdx.x;
        if (n >= run.G) continue;
        const int gx = run.g.gx, gy = run.g.gy;
        const int j = n % gx, i = n / gx;
        const int jj = min(max(j, 1), gx - 2), ii = min(max(i, 1), gy - 2);
        const int m = ii * gx + jj;
        const double* phi = run.phi;
        const double fx = -((phi[m + 1] - phi[m - 1]) / (2 * run.g.dx));
        const double fy = -((phi[m + gx] - phi[m - gx]) / (2 * run.g.dy));
        run.ex[n] = fx;
        run.ey[n] = fy;
        run.fld[n] = make_double4(fx, fy, run.bz[n], 0.0);
    }
}

// ------------------------------------------------------------------- move
// items: (species, run, push tile)
template <int TC>
__global__ __launch_bounds__(kBlock) void k_ens_push(const EnsRun* __restrict__ runs, int R,
                                                     const int* __restrict__ prefix) {
    __shared__ int hist[kHR * (TC + 2)];
    __shared__ int codes[kCodeCap];
    __shared__ unsigned long long nabs[2];
    __shared__ double wf[6 * (TC + 1)];
    const int per_species = prefix[R];
    for (int it = (int)blockIdx.x; it < 2 * per_species; it += (int)gridDim.x) {
        __syncthreads();   // the previous item's LDS use is over
        const int s = it / per_species, k = it - s * per_species;
        const int r = find_run(prefix, R, k);
        const EnsRun& run = runs[r];
        const EnsSpecies& sp = run.sp[s];
        const int tpr = (run.g.gx + TC - 1) / TC;
        // a run's uniform Bz takes the constant rotation, as its serial push does
        if (sp.sd.ub)
            push_tile_body<PIC_LAYOUT_CELL32, 2, TC, true>(sp.src, run.g, sp.sd, tpr, k - prefix[r], run.fld, sp.start,
                                                           sp.counts, sp.rank, sp.absorbed, hist, codes, nabs, wf);
        else
            push_tile_body<PIC_LAYOUT_CELL32, 0, TC, true>(sp.src, run.g, sp.sd, tpr, k - prefix[r], run.fld, sp.start,
                                                           sp.counts, sp.rank, sp.absorbed, hist, codes, nabs, wf);
    }
}

// exclusive scan of one (run, species)'s counts into start (counts reset to
// zero, start[G] and the new live count written); one workgroup each
__global__ __launch_bounds__(1024) void k_ens_scan(const EnsRun* __restrict__ runs) {
    __shared__ int wsum[16];
    __shared__ int carry_s;
    const EnsRun& run = runs[blockIdx.x >> 1];
    const EnsSpecies& sp = run.sp[blockIdx.x & 1];
    const int G = run.G;
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (threadIdx.x == 0) carry_s = 0;
    __syncthreads();
    for (int base = 0; base < G; base += 1024) {
        const int n = base + threadIdx.x;
        const int v = n < G ? sp.counts[n] : 0;
        int incl = v;
        for (int d = 1; d < 64; d <<= 1) {
            const int t = __shfl_up(incl, d, 64);
            if (lane >= d) incl += t;
        }
        if (lane == 63) wsum[wid] = incl;
        __syncthreads();
        int wpre = 0;
        for (int w = 0; w < wid; ++w) wpre += wsum[w];
        const int carry = carry_s;
        if (n < G) {
            sp.start[n] = carry + wpre + incl - v;
            sp.counts[n] = 0;
        }
        __syncthreads();
        if (threadIdx.x == 1023) carry_s = carry + wpre + incl;
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        sp.start[G] = carry_s;
        *sp.ndst = carry_s;
    }
}

// items: (species, run, chunk of kEnsChunk x 256 particles of the source
// storage); each thread takes every 256th particle of the chunk
__global__ __launch_bounds__(kBlock) void k_ens_scatter(const EnsRun* __restrict__ runs, int R,
                                                        const int* __restrict__ prefix) {
    const int per_species = prefix[R];
    for (int it = (int)blockIdx.x; it < 2 * per_species; it += (int)gridDim.x) {
        __syncthreads();   // the previous item's LDS use is over
        const int s = it / per_species, k = it - s * per_species;
        const int r = find_run(prefix, R, k);
        const EnsRun& run = runs[r];
        const EnsSpecies& sp = run.sp[s];
        const int n = *sp.nsrc;
        const int i0 = (k - prefix[r]) * (kBlock * kEnsChunk) + threadIdx.x;
#pragma unroll 2
        for (int j = 0; j < kEnsChunk; ++j) {
            const int i = i0 + j * kBlock;
            if (i >= n) break;
            const PState<PIC_LAYOUT_CELL32> p = load_state<PIC_LAYOUT_CELL32>(sp.src, i);
            const int rk = sp.rank[i];
            const Corners c = gather<PIC_LAYOUT_CELL32>(p, run.g, run.fld);
            const Pushed<PIC_LAYOUT_CELL32> o = sp.sd.ub ? push_state<PIC_LAYOUT_CELL32, 2>(p, c, run.g, sp.sd)
                                                         : push_state<PIC_LAYOUT_CELL32, 0>(p, c, run.g, sp.sd);
            if (o.key < 0) continue;
            const int d = sp.start[o.key] + rk;
            store_state<PIC_LAYOUT_CELL32>(sp.dst, d, o);
            sp.dst.cell[d] = o.key;
        }
    }
}

} // namespace pic

// =============================================================== host side
using namespace pic;

struct pic_ensemble {
    std::vector<pic_ctx*> runs;
    int R = 0;
    hipStream_t stream = nullptr;
    EnsRun* table[2] = {nullptr, nullptr};    // device, per parity
    int *dep_prefix = nullptr, *node_prefix = nullptr, *push_prefix = nullptr, *part_prefix = nullptr;
    // per-run Poisson state (device)
    int* warm = nullptr;
    int* iters = nullptr;
    double* relres = nullptr;
    std::vector<double*> pwork;               // r, p, q per run
    int tile = 256;                           // push tile width (cells)
    int blocks = 0;                           // persistent workgroups per launch
    int parity = 0;                           // 0: the runs' current buffers are their `cur`
    hipGraph_t graph[2] = {nullptr, nullptr};
    hipGraphExec_t exec[2] = {nullptr, nullptr};
    int energy = 1;
    int64_t steps = 0;
    FdBatch* fdb = nullptr;                   // batched direct solves (null: PCG only)
    std::vector<unsigned char> direct;        // per run: solved by fdb
};

namespace {


void destroy_graphs(pic_ensemble* e) {
    for (int k = 0; k < 2; ++k) {
        if (e->exec[k]) (void)hipGraphExecDestroy(e->exec[k]);
        if (e->graph[k]) (void)hipGraphDestroy(e->graph[k]);
        e->exec[k] = nullptr;
        e->graph[k] = nullptr;
    }
}

int build_prefix(pic_ensemble* e, int tile) {
    std::vector<int> pp(e->R + 1, 0);
    for (int r = 0; r < e->R; ++r) {
        const pic_ctx* c = e->runs[r];
        pp[r + 1] = pp[r] + ((c->g.gx + tile - 1) / tile) * c->g.gy;
    }
    PIC_HIP(hipMemcpy(e->push_prefix, pp.data(), pp.size() * sizeof(int), hipMemcpyHostToDevice));
    return PIC_OK;
}

int fill_tables(pic_ensemble* e) {
    for (int par = 0; par < 2; ++par) {
        std::vector<EnsRun> t(e->R);
        for (int r = 0; r < e->R; ++r) {
            pic_ctx* c = e->runs[r];
            EnsRun& d = t[r];
            d.g = c->g;
            d.G = c->G;
            d.rho = c->rho;
            d.phi = c->phi;
            d.ex = c->ex;
            d.ey = c->ey;
            d.bz = c->bz;
            d.fld = c->fld;
            d.dir = c->dir;
            d.dval = c->dval;
            for (int s = 0; s < 2; ++s) {
                Species& sp = c->sp[s];
                const int src = sp.cur ^ par, dst = src ^ 1;
                EnsSpecies& o = d.sp[s];
                o.src = sp.buf[src];
                o.dst = sp.buf[dst];
                o.sd = species_dev(c, sp);
                o.da = deposit_args(c, sp);
                o.nsrc = sp.ndev + src;
                o.ndst = sp.ndev + dst;
                o.start = sp.start;
                o.counts = sp.counts;
                o.rank = sp.rank;
                o.cq = (long long*)sp.q_charge;
                o.ce = (long long*)sp.q_energy;
                o.final_mesh = sp.final_mesh;
                o.energy_mesh = sp.energy_mesh;
                o.absorbed = sp.absorbed;
                const double dx = c->g.dx, dy = c->g.dy;
                o.cqf = sp.W * (dx * dy) * std::ldexp(1.0, -c->p.charge_shift);
                o.cef = sp.W * (dx * dy) * std::ldexp(1.0, -c->p.energy_shift);
                o.chrg = -sp.charge / c->p.epsilon;
            }
            const int L = (int)c->mg.size();
            d.nl = L;
            for (int l = 0; l < L; ++l) {
                const MgLevel& m = c->mg[l];
                EnsLevel& v = d.lv[l];
                v.gx = m.gx;
                v.gy = m.gy;
                v.cx = m.cx;
                v.cy = m.cy;
                v.diag = m.diag;
                mg_sweep_weights(c, m, &v.wd, &v.wd2);
                v.dir = m.dir;
                v.e = m.e;
                v.b = m.b;
                v.t = m.t;
            }
            d.cx = 1.0 / (c->g.dx * c->g.dx);
            d.cy = 1.0 / (c->g.dy * c->g.dy);
            d.diag = 2 * (d.cx + d.cy);
            d.r = e->pwork[3 * r];
            d.p = e->pwork[3 * r + 1];
            d.q = e->pwork[3 * r + 2];
            d.inv = c->mg_inv;
            d.unk = c->mg_unk;
            d.nu = c->mg_nu;
            d.tol2 = c->p.poisson_tol * c->p.poisson_tol;
            d.max_it = c->p.poisson_max_iter;
            d.pre = c->p.mg_pre;
            d.post = c->p.mg_post;
            const MgLevel& last = c->mg.back();
            d.sweeps = std::max(50, 4 * std::max(last.gx, last.gy));
            d.warm = e->warm + r;
            d.direct = e->direct.empty() ? 0 : e->direct[r];
            d.iters = e->iters + r;
            d.relres = e->relres + 2 * r;
        }
        PIC_HIP(hipMemcpy(e->table[par], t.data(), t.size() * sizeof(EnsRun), hipMemcpyHostToDevice));
    }
    return PIC_OK;
}

template <int TC>
void launch_push(pic_ensemble* e, const EnsRun* tab, hipStream_t s) {
    hipLaunchKernelGGL(k_ens_push<TC>, dim3(e->blocks), dim3(kBlock), 0, s, tab, e->R, e->push_prefix);
}

// The whole step for the given parity on stream s.
int enqueue_step(pic_ensemble* e, int par, hipStream_t s) {
    const EnsRun* tab = e->table[par];
    hipLaunchKernelGGL(k_ens_deposit, dim3(e->blocks), dim3(kBlock), 0, s, tab, e->R, e->dep_prefix);
    hipLaunchKernelGGL(k_ens_finalize, dim3(e->blocks), dim3(kBlock), 0, s, tab, e->R, e->node_prefix, e->energy);
    if (e->fdb) PIC_TRY(fast_batch_enqueue(e->fdb, s));
    hipLaunchKernelGGL(k_ens_poisson, dim3(e->R), dim3(1024), 0, s, tab);
    hipLaunchKernelGGL(k_ens_efield, dim3(e->blocks), dim3(kBlock), 0, s, tab, e->R, e->node_prefix);
    switch (e->tile) {
    case 64: launch_push<64>(e, tab, s); break;
    case 128: launch_push<128>(e, tab, s); break;
    default: launch_push<256>(e, tab, s); break;
    }
    hipLaunchKernelGGL(k_ens_scan, dim3(2 * e->R), dim3(1024), 0, s, tab);
    hipLaunchKernelGGL(k_ens_scatter, dim3(e->blocks), dim3(kBlock), 0, s, tab, e->R, e->part_prefix);
    PIC_HIP(hipGetLastError());
    return PIC_OK;
}

int capture(pic_ensemble* e) {
    destroy_graphs(e);
    for (int par = 0; par < 2; ++par) {
        PIC_HIP(hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal));
        const int rc = enqueue_step(e, par, e->stream);
        hipGraph_t g = nullptr;
        const hipError_t ce = hipStreamEndCapture(e->stream, &g);
        PIC_TRY(rc);
        PIC_HIP(ce);
        e->graph[par] = g;
        PIC_HIP(hipGraphInstantiate(&e->exec[par], g, nullptr, nullptr, 0));
    }
    return PIC_OK;
}

// host state of the runs after `n` ensemble steps
void advance_host_state(pic_ensemble* e, int n) {
    if (n & 1) e->parity ^= 1;
    for (pic_ctx* c : e->runs) {
        for (auto& sp : c->sp) {
            if (n & 1) sp.cur ^= 1;
            sp.sorted = 1;
            sp.deposit_pending = 0;
            if (n > 0) sp.cells_stale = 0;   // the ensemble's scatter writes the cell array
        }
        c->have_prev_phi = 1;
        c->guess_set = 0;
        if (n > 0) c->meshes_stale = 0;     // the ensemble's node pass writes the meshes
        if (n > 0) c->e_stale = 0;          // and its field kernel ex, ey with fld
    }
}

} // namespace

extern "C" {

int pic_ensemble_create(pic_ctx** runs, int nruns, pic_ensemble** out) {
    if (!runs || nruns < 1 || !out) {
        set_error("pic_ensemble_create: invalid argument");
        return PIC_ERR_ARG;
    }
    for (int r = 0; r < nruns; ++r) {
        pic_ctx* c = runs[r];
        if (!c || c->sp.size() != 2 || c->p.layout != PIC_LAYOUT_CELL32 || c->p.bx != 0.0 || c->p.by != 0.0 ||
            c->p.poisson_precision != 0 || c->unet || c->inj.on || c->p.periodic ||
            (int)c->mg.size() > kEnsMaxLev) {
            set_error("pic_ensemble_create: run %d must have two species, CELL32 storage, Bz-only field, "
                      "FP64 Poisson, no learned guess, no injection, no periodic y", r);
            return PIC_ERR_UNSUPPORTED;
        }
        if (c->p.deposit_energy != runs[0]->p.deposit_energy) {
            set_error("pic_ensemble_create: runs differ in deposit_energy");
            return PIC_ERR_UNSUPPORTED;
        }
        for (auto& sp : c->sp) {
            if (!sp.sorted || sp.deposit_pending) {
                set_error("pic_ensemble_create: run %d is not in cell-sorted state", r);
                return PIC_ERR_STATE;
            }
        }
        for (int s = 0; s < 2; ++s) PIC_TRY(ensure_cells(c, s));   // the ensemble's kernels read the cells
        PIC_HIP(hipStreamSynchronize(c->stream));
    }
    pic_ensemble* e = new pic_ensemble();
    e->runs.assign(runs, runs + nruns);
    e->R = nruns;
    e->energy = runs[0]->p.deposit_energy;
    auto fail = [&](int rc) {
        pic_ensemble_destroy(e);
        return rc;
    };
    if (hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking) != hipSuccess) {
        set_error("pic_ensemble_create: stream");
        return fail(PIC_ERR_HIP);
    }
    const size_t R = nruns;
    bool ok = hipMalloc(&e->table[0], R * sizeof(EnsRun)) == hipSuccess &&
              hipMalloc(&e->table[1], R * sizeof(EnsRun)) == hipSuccess &&
              hipMalloc(&e->dep_prefix, (R + 1) * sizeof(int)) == hipSuccess &&
              hipMalloc(&e->node_prefix, (R + 1) * sizeof(int)) == hipSuccess &&
              hipMalloc(&e->push_prefix, (R + 1) * sizeof(int)) == hipSuccess &&
              hipMalloc(&e->part_prefix, (R + 1) * sizeof(int)) == hipSuccess &&
              hipMalloc(&e->warm, R * sizeof(int)) == hipSuccess && hipMalloc(&e->iters, R * sizeof(int)) == hipSuccess &&
              hipMalloc(&e->relres, 2 * R * sizeof(double)) == hipSuccess;
    if (!ok) {
        set_error("pic_ensemble_create: allocation failed");
        return fail(PIC_ERR_ALLOC);
    }
    std::vector<int> dp(R + 1, 0), np(R + 1, 0), pp(R + 1, 0), warm(R, 0);
    for (size_t r = 0; r < R; ++r) {
        pic_ctx* c = runs[r];
        dp[r + 1] = dp[r] + ((c->g.gx + kDepC - 1) / kDepC) * c->g.gy;
        np[r + 1] = np[r] + (c->G + kBlock - 1) / kBlock;
        int64_t cap = 0;
        for (auto& sp : c->sp) cap = std::max(cap, sp.n_upper);
        pp[r + 1] = pp[r] + (int)((cap + kBlock * kEnsChunk - 1) / (kBlock * kEnsChunk));
        warm[r] = c->have_prev_phi ? 1 : 0;
        for (int k = 0; k < 3; ++k) {
            double* w = nullptr;
            if (hipMalloc(&w, (size_t)c->G * sizeof(double)) != hipSuccess) {
                set_error("pic_ensemble_create: allocation failed");
                return fail(PIC_ERR_ALLOC);
            }
            e->pwork.push_back(w);
        }
    }
    if (hipMemcpy(e->dep_prefix, dp.data(), dp.size() * sizeof(int), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(e->node_prefix, np.data(), np.size() * sizeof(int), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(e->part_prefix, pp.data(), pp.size() * sizeof(int), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(e->warm, warm.data(), R * sizeof(int), hipMemcpyHostToDevice) != hipSuccess) {
        set_error("pic_ensemble_create: copy failed");
        return fail(PIC_ERR_HIP);
    }
    int dev = 0, ncu = 256;
    if (hipGetDevice(&dev) == hipSuccess) {
        hipDeviceProp_t prop;
        if (hipGetDeviceProperties(&prop, dev) == hipSuccess) ncu = prop.multiProcessorCount;
    }
    e->blocks = 4 * ncu;
    e->tile = 256;
    e->direct.assign(R, 0);
    int rc = fast_batch_create(runs, nruns, &e->fdb, e->direct.data());
    if (rc != PIC_OK) return fail(rc);
    rc = build_prefix(e, e->tile);
    if (rc == PIC_OK) rc = fill_tables(e);
    if (rc == PIC_OK) rc = capture(e);
    if (rc != PIC_OK) return fail(rc);
    *out = e;
    return PIC_OK;
}

void pic_ensemble_destroy(pic_ensemble* e) {
    if (!e) return;
    if (e->stream) (void)hipStreamSynchronize(e->stream);
    destroy_graphs(e);
    fast_batch_free(e->fdb);
    void* ptrs[] = {e->table[0], e->table[1], e->dep_prefix, e->node_prefix, e->push_prefix, e->part_prefix,
                    e->warm, e->iters, e->relres};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    for (double* w : e->pwork)
        if (w) (void)hipFree(w);
    if (e->stream) (void)hipStreamDestroy(e->stream);
    delete e;
}

// nsteps time steps of every run; graph = 1 replays the captured step,
// 0 launches the same kernels directly.
int pic_ensemble_step(pic_ensemble* e, int nsteps, int graph) {
    if (!e || nsteps < 0) {
        set_error("pic_ensemble_step: invalid argument");
        return PIC_ERR_ARG;
    }
    // runs stepped on their own since the last ensemble step: cells from the starts
    for (pic_ctx* c : e->runs) {
        bool stale = false;
        for (int s = 0; s < 2; ++s) {
            stale = stale || c->sp[s].cells_stale;
            PIC_TRY(ensure_cells(c, s));
        }
        if (stale) PIC_HIP(hipStreamSynchronize(c->stream));
    }
    for (int k = 0; k < nsteps; ++k) {
        const int par = (e->parity + k) & 1;
        if (graph) PIC_HIP(hipGraphLaunch(e->exec[par], e->stream));
        else PIC_TRY(enqueue_step(e, par, e->stream));
    }
    PIC_HIP(hipStreamSynchronize(e->stream));
    advance_host_state(e, nsteps);
    e->steps += nsteps;
    return PIC_OK;
}

// Poisson iterations and (initial, final) relative residuals of each run's
// last solve.
int pic_ensemble_poisson_stats(pic_ensemble* e, int* iters, double* relres) {
    if (!e) {
        set_error("pic_ensemble_poisson_stats: null ensemble");
        return PIC_ERR_ARG;
    }
    PIC_HIP(hipStreamSynchronize(e->stream));
    if (iters) PIC_HIP(hipMemcpy(iters, e->iters, e->R * sizeof(int), hipMemcpyDeviceToHost));
    if (relres) PIC_HIP(hipMemcpy(relres, e->relres, 2 * e->R * sizeof(double), hipMemcpyDeviceToHost));
    return PIC_OK;
}

// Autotuner: times `steps` graph-replayed steps for each (tile width,
// persistent workgroups) candidate on the ensemble itself and keeps the
// fastest; the steps advance the runs.  Reports the choice and its ms/step.
int pic_ensemble_tune(pic_ensemble* e, int steps, int* tile, int* blocks, double* ms_per_step) {
    if (!e || steps < 1) {
        set_error("pic_ensemble_tune: invalid argument");
        return PIC_ERR_ARG;
    }
    int ncu = e->blocks / 4;
    const int tiles[] = {64, 128, 256};
    const int mult[] = {2, 4, 8};
    double best = 1e30;
    int bt = e->tile, bb = e->blocks;
    for (int t : tiles)
        for (int m : mult) {
            e->tile = t;
            e->blocks = m * ncu;
            PIC_TRY(build_prefix(e, t));
            PIC_TRY(capture(e));
            PIC_TRY(pic_ensemble_step(e, 1, 1));          // warm-up of this candidate
            const auto t0 = std::chrono::steady_clock::now();
            PIC_TRY(pic_ensemble_step(e, steps, 1));
            const double ms =
                std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count() / steps;
            if (ms < best) {
                best = ms;
                bt = t;
                bb = e->blocks;
            }
        }
    e->tile = bt;
    e->blocks = bb;
    PIC_TRY(build_prefix(e, bt));
    PIC_TRY(capture(e));
    if (tile) *tile = bt;
    if (blocks) *blocks = bb;
    if (ms_per_step) *ms_per_step = best;
    return PIC_OK;
}

int pic_ensemble_steps_done(pic_ensemble* e, int64_t* steps) {
    if (!e || !steps) {
        set_error("pic_ensemble_steps_done: null argument");
        return PIC_ERR_ARG;
    }
    *steps = e->steps;
    return PIC_OK;
}

} // extern "C"
