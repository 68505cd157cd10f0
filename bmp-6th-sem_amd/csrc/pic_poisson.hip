// pic_poisson.hip — Poisson stage on gfx950: the reference's 5-point system
// with Dirichlet rows (poissonSolver.cpp:183-241,384-402,459-544) solved by a
// flexible preconditioned conjugate gradient in FP64 whose preconditioner is
// one FP32 geometric-multigrid V-cycle (weighted Jacobi smoothing, full
// weighting restriction, bilinear prolongation), then E = -grad phi
// (calculateElectricField :433-456).
//
// Unknowns are the non-Dirichlet nodes.  Row i (interior):
//   2(1/dx^2 + 1/dy^2) phi_i - (phi_W + phi_E)/dx^2 - (phi_S + phi_N)/dy^2 = -rho_i
// which is the reference's CSR row times phi = rho after overwriteRho negates
// rho; Dirichlet nodes hold phi = V, which is what the reference's -1 diagonal
// rows give.  Eliminating them leaves a symmetric positive definite system.
//
// One PCG iteration k is, on the fine grid:
//   k_apply_p     p_k = z_{k-1} + beta p_{k-1} (formed at the node and its
//                 neighbours), q = A p_k, partials of p.q
//   k_update_pre2 alpha = (r.z)_{k-1} / (p.q) (each workgroup adds the p.q
//                 partials itself); x += alpha p; r_k = r_{k-1} - alpha q
//                 (node and neighbours);
//                 the V-cycle's two pre-smoothing sweeps from zero on r_k in
//                 the same pass; partials of r.r
//   k_check       convergence flag
//   ...           the rest of the V-cycle (coarser levels, prolongation and
//                 the first post-smoothing sweep fused)
//   k_smooth_dots last post-smoothing sweep -> z_k (FP32), partials of r.z
//                 and r.z_{k-1}
//   k_beta        Polak-Ribiere beta = (r.z_k - r.z_{k-1}) / (r.z)_{k-1}
// r, p and z alternate between two buffers so that no kernel overwrites what a
// neighbouring workgroup still reads.  Dot products leave one partial sum per
// block; a one-block kernel (for p.q, every workgroup of the next kernel)
// re-adds them in a fixed order (deterministic, no atomics).  Convergence is a device flag: once set, every kernel of the
// remaining enqueued iterations returns at once, and the host reads the flag
// only between batches of iterations.  Levels of at most 65 x 65 nodes run
// inside one workgroup's LDS.
//
// Periodic y (generateCoeffsPeriodic, poissonSolver.cpp:82-180): the unknowns
// are rows 0 .. GRID_Y-2 and the south neighbour of row 0 is row GRID_Y-2
// (and the reverse); row GRID_Y-1 is the image of row 0, masked in the solve
// (mask value 2) and copied from row 0 afterwards (overwritePhi :364-382).
// Every level keeps the same form, so the multigrid transfers wrap too.
#include "pic_internal.h"
#include "pic_unet.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>

namespace pic {

static constexpr int kB = 256;
static constexpr int kNP = PIC_PART_BLOCKS;  // blocks of the reduction kernels (partials)

// device scalars
enum Scal { S_RR = 0, S_BB, S_RR0, S_RZ, S_ALPHA, S_BETA, S_G0, S_G1, S_IRR, S_IBB, S_IRR0, S_GPICK, S_N };
static_assert(S_N <= PIC_SCAL_DOUBLES, "the device scalars fit the state block");
static_assert(kNP % kB == 0, "a kB-thread block adds kNP / kB partials per thread");
// partial-sum arrays, kNP doubles each
enum Part { P_RR = 0, P_BB, P_PQ, P_RZ, P_RZO, P_N };

static inline unsigned nblk(int64_t n) { return (unsigned)((n + kB - 1) / kB); }

// ------------------------------------------------------- block reductions
// Sum over the block in a fixed order; the result is returned to every thread.
__device__ __forceinline__ double block_sum_all(double v, double* sh) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) sh[wid] = v;
    __syncthreads();
    double s = 0;
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) s += sh[w];
    __syncthreads();
    return s;
}

__device__ __forceinline__ void write_partial(double v, double* part, double* sh) {
    double s = block_sum_all(v, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// Sum of the kNP partials by a 1024-thread block (two per thread, fixed order).
__device__ __forceinline__ double sum_partials(const double* part, double* sh) {
    double v = part[threadIdx.x] + part[threadIdx.x + 1024];
    return block_sum_all(v, sh);
}
static_assert(kNP == 2048, "sum_partials reads two partials per thread of a 1024-thread block");

// ------------------------------------------------------------ fine kernels
struct Fine {
    int gx, gy;
    double cx, cy, diag;   // 1/dx^2, 1/dy^2, 2(1/dx^2 + 1/dy^2)
    int per;               // periodic y
};

// Offsets of the south / north neighbour of node n (row j) on a gx x gy grid;
// periodic grids wrap rows 0 and gy-2 onto each other (row gy-1 is the image).
__device__ __forceinline__ int south_of(int per, int j, int gx, int gy) {
    return (per && j == 0) ? (gy - 2) * gx : -gx;
}
__device__ __forceinline__ int north_of(int per, int j, int gx, int gy) {
    return (per && j == gy - 2) ? -(gy - 2) * gx : gx;
}
// The same from the flat index (no division): row 0 is n < gx, row gy-2 is
// (gy-2) gx <= n < (gy-1) gx, and the image row is never a stencil centre.
__device__ __forceinline__ int south_n(int per, int n, int gx, int gy) {
    return (per && n < gx) ? (gy - 2) * gx : -gx;
}
__device__ __forceinline__ int north_n(int per, int n, int gx, int gy) {
    return (per && n >= (gy - 2) * gx) ? -(gy - 2) * gx : gx;
}
// row j - 1 / j + 1 with the periodic wrap
__device__ __forceinline__ int row_s(int per, int j, int gy) { return (per && j == 0) ? gy - 2 : j - 1; }
__device__ __forceinline__ int row_n(int per, int j, int gy) { return (per && j == gy - 2) ? 0 : j + 1; }

// rho[geom] = V; rho = -rho  (overwriteRho, poissonSolver.cpp:384-402)
__global__ void k_overwrite_rho(int G, double* rho, const uint8_t* dir, const double* dval) {
    int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= G) return;
    double r = dir[n] == 1 ? dval[n] : rho[n];
    rho[n] = -r;
}

// phi[geom] = V (overwritePhi :364-382).  Mask value 1 marks geometry nodes;
// 2 marks the periodic image row, which takes row 0 first (the reference's
// memcpy) unless a geometry covers it.
__global__ void k_overwrite_phi(int G, int gx, double* phi, const uint8_t* dir, const double* dval) {
    int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= G) return;
    const uint8_t d = dir[n];
    if (d == 1) phi[n] = dval[n];
    else if (d == 2) {
        const int m = n % gx;
        phi[n] = dir[m] == 1 ? dval[m] : phi[m];
    }
}

// x <- guess (or 0) with Dirichlet values imposed; `keep` (optional) receives
// the previous x.  With `prev` the guess is 2 x - prev (x, prev: the last two
// solutions).
__global__ void k_init_guess(int G, double* x, const double* src, const double* prev, double* keep,
                             const uint8_t* dir, const double* dval) {
    int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= G) return;
    const double cur = x[n];
    double v = 0.0;
    if (prev) v = 2.0 * cur - prev[n];
    else if (src) v = src[n];
    if (keep) keep[n] = cur;
    x[n] = dir[n] ? dval[n] : v;
}

// r = rhs - A x on unknown nodes (0 on Dirichlet); partial of r.r
__global__ __launch_bounds__(kB) void k_residual(Fine f, const double* __restrict__ rhs,
                                                 const double* __restrict__ x, const uint8_t* __restrict__ dir,
                                                 double* __restrict__ r, double* part) {
    __shared__ double sh[kB / 64];
    double acc = 0;
    const int G = f.gx * f.gy;
    for (int n = blockIdx.x * blockDim.x + threadIdx.x; n < G; n += gridDim.x * blockDim.x) {
        double v = 0;
        if (!dir[n]) {
            const int so = south_n(f.per, n, f.gx, f.gy), no = north_n(f.per, n, f.gx, f.gy);
            double ax = f.diag * x[n] - f.cx * (x[n - 1] + x[n + 1]) - f.cy * (x[n + so] + x[n + no]);
            v = rhs[n] - ax;
        }
        r[n] = v;
        acc += v * v;
    }
    write_partial(acc, part, sh);
}

// The learned guess and the starting residuals of both candidates in one
// pass over the grid (the grid, loop and partial sums of k_residual): with
// Dirichlet nodes at their values dval, guess = phi + out * (max|r0| / diag)
// on unknown nodes of the N x N image, phi elsewhere; r_phi = rhs - A phi
// into r0 with the partials of |r_phi|^2, r_guess = rhs - A guess into r1
// with those of |r_guess|^2 (the guess at the stencil's nodes formed from phi
// and out on the spot).  Non-periodic grids (the learned guess's).
__global__ __launch_bounds__(kB) void k_guess_residuals(Fine f, const double* __restrict__ rhs,
                                                        const double* __restrict__ phi,
                                                        const uint8_t* __restrict__ dir,
                                                        const double* __restrict__ dval,
                                                        const float* __restrict__ out,
                                                        const unsigned long long* __restrict__ scl, int N,
                                                        double diag, double* __restrict__ guess,
                                                        double* __restrict__ r0, double* __restrict__ r1,
                                                        double* part0, double* part1) {
    __shared__ double sh[kB / 64];
    const double m0 = __longlong_as_double((long long)scl[0]);
    const double sc = (m0 > 0.0 ? m0 : 1.0) / diag;
    auto pv = [&](int n) -> double { return dir[n] ? dval[n] : phi[n]; };
    auto gv = [&](int n, int x, int y) -> double {
        if (dir[n]) return dval[n];
        double v = phi[n];
        if (x < N && y < N) v += (double)out[y * N + x] * sc;
        return v;
    };
    double acc0 = 0, acc1 = 0;
    const int G = f.gx * f.gy;
    for (int n = blockIdx.x * blockDim.x + threadIdx.x; n < G; n += gridDim.x * blockDim.x) {
        const int y = n / f.gx, x = n - y * f.gx;
        double v0 = 0, v1 = 0;
        const double g = gv(n, x, y);
        if (!dir[n]) {
            const double a0 = f.diag * phi[n] - f.cx * (pv(n - 1) + pv(n + 1)) - f.cy * (pv(n - f.gx) + pv(n + f.gx));
            v0 = rhs[n] - a0;
            const double a1 = f.diag * g - f.cx * (gv(n - 1, x - 1, y) + gv(n + 1, x + 1, y))
                            - f.cy * (gv(n - f.gx, x, y - 1) + gv(n + f.gx, x, y + 1));
            v1 = rhs[n] - a1;
        }
        guess[n] = g;
        r0[n] = v0;
        r1[n] = v1;
        acc0 += v0 * v0;
        acc1 += v1 * v1;
    }
    write_partial(acc0, part0, sh);
    write_partial(acc1, part1, sh);
}

// Where the device picked the learned guess (*pick set by k_sum2): phi and
// r0 take the guess and its residual, and its partials replace phi's; else
// phi keeps its values with the Dirichlet nodes at dval (as r0 assumed).
__global__ __launch_bounds__(kB) void k_take_guess(int G, double* __restrict__ phi, double* __restrict__ r0,
                                                   const double* __restrict__ guess, const double* __restrict__ r1,
                                                   double* __restrict__ part0, const double* __restrict__ part1,
                                                   const uint8_t* __restrict__ dir, const double* __restrict__ dval,
                                                   const double* __restrict__ pick) {
    const bool take = *pick > 0.5;
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    for (int n = k; n < G; n += gridDim.x * blockDim.x) {
        if (take) {
            phi[n] = guess[n];
            r0[n] = r1[n];
        } else if (dir[n]) {
            phi[n] = dval[n];
        }
    }
    if (take && k < kNP) part0[k] = part1[k];
}

// k_residual after the direct solve, with overwritePhi in the same pass:
// Dirichlet nodes take their values (the image row of a periodic grid row
// 0's), and the stencil of an unknown node reads a Dirichlet neighbour's
// value from dval, so no node is read while it is written (an unknown node's
// stencil never reaches the image row).  The same residual as k_overwrite_phi
// followed by k_residual.
// With EF (non-periodic grids) the same pass writes E = -grad phi as
// k_efield does (central differences at the clamped interior node, the same
// expressions), reading every phi through the same Dirichlet substitution,
// into the gather array fld only: ex / ey are formed from fld when a caller
// reads them (pic_ctx::e_stale).
struct EfOut {
    double dx, dy;
    const double* bz;
    double *ex, *ey;
    double4* fld;
};

template <bool EF>
__global__ __launch_bounds__(kB) void k_residual_fix(Fine f, const double* __restrict__ rhs, double* __restrict__ x,
                                                     const uint8_t* __restrict__ dir, const double* __restrict__ dval,
                                                     double* __restrict__ r, double* part, EfOut e) {
    __shared__ double sh[kB / 64];
    double acc = 0;
    const int G = f.gx * f.gy;
    auto xe = [&](int m) { return dir[m] == 1 ? dval[m] : x[m]; };
    for (int n = blockIdx.x * blockDim.x + threadIdx.x; n < G; n += gridDim.x * blockDim.x) {
        if constexpr (EF) {
            const int j = n % f.gx, i = n / f.gx;
            const int jj = min(max(j, 1), f.gx - 2), ii = min(max(i, 1), f.gy - 2);
            const int m = ii * f.gx + jj;
            const double fx = -((xe(m + 1) - xe(m - 1)) / (2 * e.dx));
            const double fy = -((xe(m + f.gx) - xe(m - f.gx)) / (2 * e.dy));
            e.fld[n] = make_double4(fx, fy, e.bz[n], 0.0);
        }
        double v = 0;
        const uint8_t d = dir[n];
        if (!d) {
            const int so = south_n(f.per, n, f.gx, f.gy), no = north_n(f.per, n, f.gx, f.gy);
            double ax = f.diag * x[n] - f.cx * (xe(n - 1) + xe(n + 1)) - f.cy * (xe(n + so) + xe(n + no));
            v = rhs[n] - ax;
        } else if (d == 1) {
            x[n] = dval[n];
        } else {
            x[n] = xe(n % f.gx);
        }
        r[n] = v;
        acc += v * v;
    }
    write_partial(acc, part, sh);
}

// The start of every solve in one pass: flags cleared; overwriteRho
// (rho[geom] = V, rho = -rho); the partial of b.b, b = the residual of the
// Dirichlet-only field (unknowns at zero) on the unknown nodes.
__global__ __launch_bounds__(kB) void k_prep_rhs(Fine f, double* __restrict__ rho, const uint8_t* __restrict__ dir,
                                                 const double* __restrict__ dval, double* part_bb, int* flags,
                                                 unsigned long long* ts) {
    __shared__ double sh[kB / 64];
    if (ts && blockIdx.x == 0 && threadIdx.x == 0) ts[0] = wall_clock64();   // the solve's start
    // flags: [0] outer converged (1) / failed (2), [1] outer iterations,
    // [2] inner (mixed precision) converged / failed, [3] inner iterations
    if (blockIdx.x == 0 && threadIdx.x < 4) flags[threadIdx.x] = 0;
    double acc = 0;
    const int G = f.gx * f.gy;
    auto x0 = [&](int m) -> double { return dir[m] ? dval[m] : 0.0; };
    for (int n = blockIdx.x * blockDim.x + threadIdx.x; n < G; n += gridDim.x * blockDim.x) {
        const uint8_t d = dir[n];
        const double r = d == 1 ? dval[n] : rho[n];
        rho[n] = -r;
        if (!d) {
            const int so = south_n(f.per, n, f.gx, f.gy), no = north_n(f.per, n, f.gx, f.gy);
            const double ax = f.diag * 0.0 - f.cx * (x0(n - 1) + x0(n + 1)) - f.cy * (x0(n + so) + x0(n + no));
            const double v = -r - ax;
            acc += v * v;
        }
    }
    write_partial(acc, part_bb, sh);
}

// k_prep_rhs with the deposit's node pass in front: rho[n] is the species
// sum of the corner meshes, then overwritten as above.  One pass instead of
// two over the grid; final_mesh and energy_mesh are left to
// materialize_meshes (read by diagnostics only).
__global__ __launch_bounds__(kB) void k_finalize_prep(FinalizeAll a, int gp, int energy, Fine f,
                                                      double* __restrict__ rho, const uint8_t* __restrict__ dir,
                                                      const double* __restrict__ dval, double* part_bb, int* flags,
                                                      unsigned long long* ts) {
    __shared__ double sh[kB / 64];
    if (blockIdx.x == 0 && threadIdx.x < 4) flags[threadIdx.x] = 0;
    if (ts && blockIdx.x == 0 && threadIdx.x == 0) ts[0] = wall_clock64();   // the solve's start
    double acc = 0;
    const int G = f.gx * f.gy;
    auto x0 = [&](int m) -> double { return dir[m] ? dval[m] : 0.0; };
    for (int n = blockIdx.x * blockDim.x + threadIdx.x; n < G; n += gridDim.x * blockDim.x) {
        const double dep = finalize_node<false>(a, n, gp, energy);
        const uint8_t d = dir[n];
        const double r = d == 1 ? dval[n] : dep;
        rho[n] = -r;
        if (!d) {
            const int so = south_n(f.per, n, f.gx, f.gy), no = north_n(f.per, n, f.gx, f.gy);
            const double ax = f.diag * 0.0 - f.cx * (x0(n - 1) + x0(n + 1)) - f.cy * (x0(n + so) + x0(n + no));
            const double v = -r - ax;
            acc += v * v;
        }
    }
    write_partial(acc, part_bb, sh);
}

// rr = sum of the partials; done when rr <= tol^2 bb (flags[0] = 1), NaN -> 2.
// The first call also records bb (from part_bb, or rr itself when part_bb is
// null) and rr0; later calls count one more iteration unless already done.
// sbb, srr, srr0: the scalar slots (outer solve or inner correction solve).
__device__ __forceinline__ void check_body(const double* part_rr, const double* part_bb, double tol2, int* flags,
                                           double* scal, int first, int sbb, int srr, int srr0, double* sh) {
    if (flags[0]) return;
    double rr = sum_partials(part_rr, sh);
    double bb = first ? (part_bb ? sum_partials(part_bb, sh) : rr) : scal[sbb];
    if (threadIdx.x == 0) {
        if (first) {
            scal[sbb] = bb;
            scal[srr0] = rr;
        } else {
            flags[1] += 1;
        }
        scal[srr] = rr;
        if (!(rr == rr)) flags[0] = 2;
        else if (rr <= tol2 * bb || bb == 0.0) flags[0] = 1;
    }
}

__global__ __launch_bounds__(1024) void k_check(const double* part_rr, const double* part_bb, double tol2,
                                                int* flags, double* scal, int first, int sbb, int srr, int srr0) {
    __shared__ double sh[16];
    check_body(part_rr, part_bb, tol2, flags, scal, first, sbb, srr, srr0, sh);
}

// k_check, then the state block published to the host: copied into the
// mapped host memory `mail` with the wall-clock stamps, a system-scope
// release, then the sequence number the host waits for.
__global__ __launch_bounds__(1024) void k_check_publish(const double* part_rr, const double* part_bb, double tol2,
                                                        int* flags, double* scal, int first, int sbb, int srr,
                                                        int srr0, SolveMail* mail, const unsigned long long* ts,
                                                        unsigned seq) {
    __shared__ double sh[16];
    check_body(part_rr, part_bb, tol2, flags, scal, first, sbb, srr, srr0, sh);
    if (threadIdx.x == 0) {
        for (int k = 0; k < PIC_SCAL_DOUBLES; ++k) mail->scal[k] = scal[k];
        for (int k = 0; k < 4; ++k) mail->flags[k] = flags[k];
        mail->t0 = ts[0];
        mail->t1 = wall_clock64();
        __threadfence_system();
        __hip_atomic_store(&mail->seq, seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
    }
}

// p = z + beta pold (first: p = z) at the node and its neighbours; q = A p;
// partial of p.q.  z, pold and p are zero on Dirichlet nodes.
template <typename T>
__global__ __launch_bounds__(kB) void k_apply_p(Fine f, const float* __restrict__ z, const T* __restrict__ pold,
                                                T* __restrict__ p, const uint8_t* __restrict__ dir,
                                                T* __restrict__ q, double* part_pq, const double* scal,
                                                const int* flags, int first) {
    __shared__ double sh[kB / 64];
    if (flags[0]) return;
    const double beta = first ? 0.0 : scal[S_BETA];
    double acc = 0;
    const int G = f.gx * f.gy;
    auto P = [&](int m) -> double { return first ? (double)z[m] : (double)z[m] + beta * (double)pold[m]; };
    for (int n = blockIdx.x * blockDim.x + threadIdx.x; n < G; n += gridDim.x * blockDim.x) {
        double v = 0, pn = 0;
        if (!dir[n]) {
            pn = P(n);
            const int so = south_n(f.per, n, f.gx, f.gy), no = north_n(f.per, n, f.gx, f.gy);
            v = f.diag * pn - f.cx * (P(n - 1) + P(n + 1)) - f.cy * (P(n + so) + P(n + no));
            acc += pn * v;
        }
        p[n] = (T)pn;
        q[n] = (T)v;
    }
    write_partial(acc, part_pq, sh);
}

// beta = (r.z - r.zold) / (r.z)_prev (Polak-Ribiere; first: 0); keeps r.z.
__global__ __launch_bounds__(1024) void k_beta(const double* part_rz, const double* part_rzo, double* scal,
                                               const int* flags, int first) {
    __shared__ double sh[16];
    if (flags[0]) return;
    const double rz = sum_partials(part_rz, sh);
    const double rzo = first ? 0.0 : sum_partials(part_rzo, sh);
    if (threadIdx.x == 0) {
        scal[S_BETA] = first ? 0.0 : (rz - rzo) / scal[S_RZ];
        scal[S_RZ] = rz;
    }
}

// scal[S_G0], scal[S_G1] = sums of two partial arrays (squared residuals of
// two guesses); scal[S_GPICK] = 1 where the second is lower (else 0)
__global__ __launch_bounds__(1024) void k_sum2(const double* pa, const double* pb, double* scal) {
    __shared__ double sh[16];
    const double a = sum_partials(pa, sh);
    const double b = sum_partials(pb, sh);
    if (threadIdx.x == 0) {
        scal[S_G0] = a;
        scal[S_G1] = b;
        scal[S_GPICK] = b < a ? 1.0 : 0.0;
    }
}

struct Lvl {
    int gx, gy;
    float cx, cy, diag, wd;
    float wd2;             // second sweep of the fused two-sweep kernels
    int per;               // periodic y
};

// x += alpha p; rnew = rold - alpha q (also at the neighbours, for the
// stencil); partial of rnew.rnew.  PRE2: the finest level's two weighted-Jacobi
// sweeps from zero on b = (float) rnew in the same pass (the first sweep is
// pointwise, e1 = wd b, so the second, weight wd2, needs only b on the
// 5-point stencil); b and e are stored for the restriction and the post-smoothing.  Without
// PRE2 only b is stored.
template <int PRE2, typename T>
__global__ __launch_bounds__(kB) void k_update_pre2(Fine f, Lvl L, T* __restrict__ x,
                                                    const T* __restrict__ p, const T* __restrict__ rold,
                                                    T* __restrict__ rnew, const T* __restrict__ q,
                                                    const uint8_t* __restrict__ dir, float* __restrict__ b0,
                                                    float* __restrict__ e0, double* part_rr, double* scal,
                                                    const double* __restrict__ part_pq, const int* flags) {
    __shared__ double sh[kB / 64];
    if (flags[0]) return;
    // alpha = r.z / p.q, every workgroup adding the p.q partials in the same
    // fixed order (workgroup 0 also keeps it in the state block)
    double pq = 0.0;
#pragma unroll
    for (int k = 0; k < kNP / kB; ++k) pq += part_pq[threadIdx.x + k * kB];
    pq = block_sum_all(pq, sh);
    const double alpha = scal[S_RZ] / pq;
    if (blockIdx.x == 0 && threadIdx.x == 0) scal[S_ALPHA] = alpha;
    double acc = 0;
    const int G = f.gx * f.gy;
    auto R = [&](int m) -> double { return (double)rold[m] - alpha * (double)q[m]; };
    for (int n = blockIdx.x * blockDim.x + threadIdx.x; n < G; n += gridDim.x * blockDim.x) {
        if (dir[n]) {
            rnew[n] = (T)0;
            b0[n] = 0.f;
            if (PRE2) e0[n] = 0.f;
            continue;
        }
        x[n] = (T)((double)x[n] + alpha * (double)p[n]);
        const T rt = (T)R(n);
        const double rn = (double)rt;
        rnew[n] = rt;
        acc += rn * rn;
        const float bn = (float)rn;
        b0[n] = bn;
        if (PRE2) {
            // neighbours' residuals rounded like the stored ones
            auto e1 = [&](int m) -> float { return dir[m] ? 0.f : L.wd * (float)(T)R(m); };
            const float c = L.wd * bn;
            const int so = south_n(f.per, n, f.gx, f.gy), no = north_n(f.per, n, f.gx, f.gy);
            const float ax = L.diag * c - L.cx * (e1(n - 1) + e1(n + 1)) - L.cy * (e1(n + so) + e1(n + no));
            e0[n] = c + L.wd2 * (bn - ax);
        }
    }
    write_partial(acc, part_rr, sh);
}

// Last post-smoothing sweep of the finest level: z = ei + wd (b - A ei) (zero
// on Dirichlet), stored in FP32; partials of r.z and r.zold.
template <typename T>
__global__ __launch_bounds__(kB) void k_smooth_dots(Lvl L, const float* __restrict__ ei, float* __restrict__ z,
                                                    const float* __restrict__ b, const uint8_t* __restrict__ dir,
                                                    const T* __restrict__ r, const float* __restrict__ zold,
                                                    double* part_rz, double* part_rzo, const int* flags, int first) {
    __shared__ double sh[kB / 64];
    if (flags[0]) return;
    double a = 0, c = 0;
    const int G = L.gx * L.gy;
    for (int n = blockIdx.x * blockDim.x + threadIdx.x; n < G; n += gridDim.x * blockDim.x) {
        float v = 0.f;
        if (!dir[n]) {
            const float e = ei[n];
            const int so = south_n(L.per, n, L.gx, L.gy), no = north_n(L.per, n, L.gx, L.gy);
            const float ax = L.diag * e - L.cx * (ei[n - 1] + ei[n + 1]) - L.cy * (ei[n + so] + ei[n + no]);
            v = e + L.wd * (b[n] - ax);
        }
        z[n] = v;
        const double rn = (double)r[n];
        a += rn * (double)v;
        if (!first) c += rn * (double)zold[n];
    }
    write_partial(a, part_rz, sh);
    if (!first) write_partial(c, part_rzo, sh);
}

// rf = (float) r; partial of rf.rf (the inner solve's starting residual)
__global__ __launch_bounds__(kB) void k_r_to_f32(int G, const double* __restrict__ r, float* __restrict__ rf,
                                                 double* part) {
    __shared__ double sh[kB / 64];
    double acc = 0;
    for (int n = blockIdx.x * blockDim.x + threadIdx.x; n < G; n += gridDim.x * blockDim.x) {
        const float v = (float)r[n];
        rf[n] = v;
        acc += (double)v * (double)v;
    }
    write_partial(acc, part, sh);
}

__global__ void k_zero_f32v(int G, float* __restrict__ d) {
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n < G) d[n] = 0.f;
}

// iterative-refinement correction in FP64: x += d
__global__ void k_ir_update(int G, double* __restrict__ x, const float* __restrict__ d) {
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n < G) x[n] += (double)d[n];
}

__global__ void k_reset_inner(int* flags) { flags[2] = 0; }

// --------------------------------------------------------------- MG (FP32)
// Level kernels use 2D grids of 64 x 4 node blocks so node indices need no
// division and narrow coarse rows waste few lanes; `wd` is the Jacobi weight
// over the diagonal.
#define MG_IJ(L)                                                              \
    const int i = blockIdx.x * blockDim.x + threadIdx.x;                      \
    const int j = blockIdx.y * blockDim.y + threadIdx.y;                      \
    if (i >= (L).gx || j >= (L).gy) return;                                   \
    const int n = j * (L).gx + i;

template <typename T>
__global__ void k_d2f(int G, const T* __restrict__ r, float* __restrict__ b, const int* flags) {
    if (flags[0]) return;
    int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n < G) b[n] = (float)r[n];
}

// Weighted Jacobi: eo = ei + wd (b - A ei) on unknowns; zero on Dirichlet.
__global__ __launch_bounds__(kB) void k_mg_smooth(Lvl L, const float* __restrict__ ei, float* __restrict__ eo,
                                                  const float* __restrict__ b, const uint8_t* __restrict__ dir,
                                                  int zero_in, const int* flags) {
    if (flags[0]) return;
    MG_IJ(L)
    if (dir[n]) { eo[n] = 0.f; return; }
    if (zero_in) { eo[n] = L.wd * b[n]; return; }
    const int so = south_of(L.per, j, L.gx, L.gy), no = north_of(L.per, j, L.gx, L.gy);
    float e = ei[n];
    float ax = L.diag * e - L.cx * (ei[n - 1] + ei[n + 1]) - L.cy * (ei[n + so] + ei[n + no]);
    eo[n] = e + L.wd * (b[n] - ax);
}

// Two weighted-Jacobi sweeps from zero in one pass: the first sweep is
// pointwise (e1 = wd b), so the second needs only b on the 5-point stencil.
// At the finest level b is the FP64 residual, rounded to FP32 here (and
// stored for the restriction and post-smoothing).
template <typename TB>
__global__ __launch_bounds__(kB) void k_mg_pre2(Lvl L, const TB* __restrict__ bin, float* __restrict__ bout,
                                                float* __restrict__ eo, const uint8_t* __restrict__ dir,
                                                const int* flags) {
    if (flags[0]) return;
    MG_IJ(L)
    float bn = (float)bin[n];
    if (bout) bout[n] = bn;
    if (dir[n]) { eo[n] = 0.f; return; }
    auto e1 = [&](int m) -> float { return dir[m] ? 0.f : L.wd * (float)bin[m]; };
    const int so = south_of(L.per, j, L.gx, L.gy), no = north_of(L.per, j, L.gx, L.gy);
    float c = L.wd * bn;
    float ax = L.diag * c - L.cx * (e1(n - 1) + e1(n + 1)) - L.cy * (e1(n + so) + e1(n + no));
    eo[n] = c + L.wd2 * (bn - ax);
}

// Row of the stored grid that region row rj stands for: periodic grids wrap
// rows modulo gy - 1 (the image row gy - 1 reads as row 0), others give -1
// outside [0, gy).
__device__ __forceinline__ int grid_row(const Lvl& L, int rj) {
    if (L.per) {
        const int p = L.gy - 1;
        return ((rj % p) + p) % p;
    }
    return (rj >= 0 && rj < L.gy) ? rj : -1;
}

// Coarse rhs = full weighting of the fine residual; zero on coarse Dirichlet.
// One workgroup per kRt x kRt coarse nodes: the fine e of the covering region
// (+1 halo) is staged in LDS, each fine residual is computed once into LDS,
// then each thread sums the 3 x 3 weights of its coarse node.
static constexpr int kRt = 16;
static constexpr int kRe = 2 * kRt + 3;     // fine e region edge
static constexpr int kRr = 2 * kRt + 1;     // fine residual region edge
__global__ __launch_bounds__(kRt * kRt) void k_mg_restrict(Lvl F, Lvl C, const float* __restrict__ e,
                                                           const float* __restrict__ b,
                                                           const uint8_t* __restrict__ fdir, float* __restrict__ bc,
                                                           const uint8_t* __restrict__ cdir, const int* flags) {
    __shared__ float se[kRe * kRe];
    __shared__ float sr[kRr * kRr];
    if (flags[0]) return;
    const int tid = threadIdx.x;
    const int I0 = blockIdx.x * kRt, J0 = blockIdx.y * kRt;
    const int fi0 = 2 * I0 - 2, fj0 = 2 * J0 - 2;   // fine origin of the e region
    for (int k = tid; k < kRe * kRe; k += kRt * kRt) {
        const int rj = k / kRe, ri = k - rj * kRe;
        const int gj = grid_row(F, fj0 + rj), fi = fi0 + ri;
        se[k] = (gj >= 0 && fi >= 0 && fi < F.gx) ? e[gj * F.gx + fi] : 0.f;
    }
    __syncthreads();
    for (int k = tid; k < kRr * kRr; k += kRt * kRt) {
        const int rj = k / kRr, ri = k - rj * kRr;
        const int gj = grid_row(F, fj0 + 1 + rj), fi = fi0 + 1 + ri;
        float v = 0.f;
        if (gj >= 0 && fi >= 0 && fi < F.gx) {
            const int n = gj * F.gx + fi;
            if (!fdir[n]) {
                const float* c = &se[(rj + 1) * kRe + ri + 1];
                const float ax = F.diag * c[0] - F.cx * (c[-1] + c[1]) - F.cy * (c[-kRe] + c[kRe]);
                v = b[n] - ax;
            }
        }
        sr[k] = v;
    }
    __syncthreads();
    const int tx = tid % kRt, ty = tid / kRt;
    const int I = I0 + tx, J = J0 + ty;
    if (I >= C.gx || J >= C.gy) return;
    const int n = J * C.gx + I;
    if (cdir[n]) { bc[n] = 0.f; return; }
    const float* c = &sr[(2 * ty + 1) * kRr + 2 * tx + 1];
    const float s = 4.f * c[0] + 2.f * (c[-1] + c[1] + c[-kRr] + c[kRr])
                  + (c[-kRr - 1] + c[-kRr + 1] + c[kRr - 1] + c[kRr + 1]);
    bc[n] = s * 0.0625f;
}

// bilinear interpolation of the coarse correction at fine node (i, j); on a
// periodic grid the coarse row after the last real one is row 0
__device__ __forceinline__ float prolong_at(const Lvl& C, const float* ec, int i, int j) {
    int I = i >> 1, J = j >> 1;
    bool oi = i & 1, oj = j & 1;
    const float* c0 = ec + J * C.gx + I;
    if (!oi && !oj) return c0[0];
    if (oi && !oj) return 0.5f * (c0[0] + c0[1]);
    const int up = north_of(C.per, J, C.gx, C.gy);
    if (!oi && oj) return 0.5f * (c0[0] + c0[up]);
    return 0.25f * (c0[0] + c0[1] + c0[up] + c0[up + 1]);
}

// e += P ec (unknown nodes only) — generic path.
__global__ __launch_bounds__(kB) void k_mg_prolong(Lvl F, Lvl C, float* __restrict__ e,
                                                   const uint8_t* __restrict__ fdir, const float* __restrict__ ec,
                                                   const int* flags) {
    if (flags[0]) return;
    MG_IJ(F)
    if (fdir[n]) return;
    e[n] += prolong_at(C, ec, i, j);
}

// Prolongation fused with the first post-smoothing sweep: e1 = e + P ec on a
// kPx x kPy fine tile and its one-node halo, computed once per node into LDS,
// then eo = one weighted-Jacobi sweep of e1 on the tile.
static constexpr int kPx = 64, kPy = 4;
__global__ __launch_bounds__(kPx * kPy) void k_mg_prolong_sweep(Lvl F, Lvl C, const float* __restrict__ e,
                                                                const float* __restrict__ ec,
                                                                const float* __restrict__ b,
                                                                const uint8_t* __restrict__ dir,
                                                                float* __restrict__ eo, const int* flags) {
    constexpr int W = kPx + 2, H = kPy + 2;
    __shared__ float s1[W * H];
    if (flags[0]) return;
    const int tid = threadIdx.y * kPx + threadIdx.x;
    const int i0 = blockIdx.x * kPx, j0 = blockIdx.y * kPy;
    for (int k = tid; k < W * H; k += kPx * kPy) {
        const int rj = k / W, ri = k - rj * W;
        const int gj = grid_row(F, j0 - 1 + rj), fi = i0 - 1 + ri;
        float v = 0.f;
        if (gj >= 0 && fi >= 0 && fi < F.gx) {
            const int m = gj * F.gx + fi;
            if (!dir[m]) v = e[m] + prolong_at(C, ec, fi, gj);
        }
        s1[k] = v;
    }
    __syncthreads();
    const int i = i0 + (int)threadIdx.x, j = j0 + (int)threadIdx.y;
    if (i >= F.gx || j >= F.gy) return;
    const int n = j * F.gx + i;
    if (dir[n]) { eo[n] = 0.f; return; }
    const float* c = &s1[(threadIdx.y + 1) * W + threadIdx.x + 1];
    const float ax = F.diag * c[0] - F.cx * (c[-1] + c[1]) - F.cy * (c[-W] + c[W]);
    eo[n] = c[0] + F.wd * (b[n] - ax);
}


// ---------------------------------------------- coarse levels in one workgroup
// Levels of at most 65 x 65 nodes run in LDS; the coarsest level is solved
// exactly with its inverse (precomputed on the host, unknowns only).  With
// two pre- and two post-smoothing sweeps each level costs two barriers on the
// way down (fused pre-smoothing, restriction) and two on the way up
// (prolongation fused with the first sweep, second sweep).
static constexpr int kLdsMaxLev = 8;
static constexpr int kLdsNodes = 5800;       // all in-LDS levels together
static constexpr int kLdsPerThread = 5;      // nodes per thread per level (1024 threads)
static constexpr int kDirectMax = 64;        // unknowns of a directly solved coarsest level

struct CoarseLevels {
    int nl, total;
    int gx[kLdsMaxLev], gy[kLdsMaxLev], off[kLdsMaxLev];
    float cx[kLdsMaxLev], cy[kLdsMaxLev], diag[kLdsMaxLev], wd[kLdsMaxLev];
    float wd2[kLdsMaxLev];        // second pre-sweep / first post-sweep weight
    const uint8_t* cdir;          // masks of all levels, concatenated at off[]
    int pre, post, sweeps;
    int per;                      // periodic y
    int nu;                       // unknowns of the coarsest level solved directly (0: use sweeps)
    const float* inv;             // nu x nu inverse
    const int* unk;               // node index of each unknown
};

// south / north offsets of node n (row n / gx) of an in-LDS level
__device__ __forceinline__ void lds_nb(int per, int n, int gx, int gy, int& so, int& no) {
    so = -gx;
    no = gx;
    if (per) {
        const int j = n / gx;
        so = south_of(per, j, gx, gy);
        no = north_of(per, j, gx, gy);
    }
}

// One weighted-Jacobi sweep in place (two barriers: read all, then write).
__device__ __forceinline__ void lds_sweep(float* E, const float* B, const uint8_t* D, int gx, int gy, float cx,
                                          float cy, float diag, float wd, bool zero_in, int per) {
    const int G = gx * gy;
    float nv[kLdsPerThread];
#pragma unroll
    for (int j = 0; j < kLdsPerThread; ++j) {
        int n = threadIdx.x + j * 1024;
        nv[j] = 0.f;
        if (n < G && !D[n]) {
            if (zero_in) {
                nv[j] = wd * B[n];
            } else {
                int so, no;
                lds_nb(per, n, gx, gy, so, no);
                float e = E[n];
                float ax = diag * e - cx * (E[n - 1] + E[n + 1]) - cy * (E[n + so] + E[n + no]);
                nv[j] = e + wd * (B[n] - ax);
            }
        }
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < kLdsPerThread; ++j) {
        int n = threadIdx.x + j * 1024;
        if (n < G) E[n] = nv[j];
    }
    __syncthreads();
}

__device__ __forceinline__ float lds_prolong(const float* ec, int cgx, int cgy, int per, int i, int j) {
    const int I = i >> 1, J = j >> 1;
    const bool oi = i & 1, oj = j & 1;
    const float* c0 = ec + J * cgx + I;
    if (!oi && !oj) return c0[0];
    if (oi && !oj) return 0.5f * (c0[0] + c0[1]);
    const int up = north_of(per, J, cgx, cgy);
    if (!oi && oj) return 0.5f * (c0[0] + c0[up]);
    return 0.25f * (c0[0] + c0[1] + c0[up] + c0[up + 1]);
}

__global__ __launch_bounds__(1024) void k_mg_coarse_cycle(CoarseLevels cl, const float* __restrict__ bglob,
                                                          float* __restrict__ eglob, const int* flags) {
    __shared__ float E[kLdsNodes];
    __shared__ float T[kLdsNodes];
    __shared__ float B[kLdsNodes];
    __shared__ uint8_t D[kLdsNodes];
    __shared__ float INV[kDirectMax * kDirectMax];
    __shared__ int UNK[kDirectMax];
    if (flags[0]) return;
    const int tid = threadIdx.x;
    const int G0 = cl.gx[0] * cl.gy[0];
    const int nn = cl.nu * cl.nu;
    // every global load of the set-up is issued before the first wait
    constexpr int QD = (kLdsNodes + 1023) / 1024;
    constexpr int QI = (kDirectMax * kDirectMax + 1023) / 1024;
    uint8_t dv[QD];
    float bv[kLdsPerThread], iv[QI];
    int uv = 0;
#pragma unroll
    for (int q = 0; q < QD; ++q) {
        const int n = tid + q * 1024;
        dv[q] = n < cl.total ? cl.cdir[n] : (uint8_t)1;
    }
#pragma unroll
    for (int q = 0; q < kLdsPerThread; ++q) {
        const int n = tid + q * 1024;
        bv[q] = n < G0 ? bglob[n] : 0.f;
    }
#pragma unroll
    for (int q = 0; q < QI; ++q) {
        const int n = tid + q * 1024;
        iv[q] = n < nn ? cl.inv[n] : 0.f;
    }
    if (tid < cl.nu) uv = cl.unk[tid];
#pragma unroll
    for (int q = 0; q < QD; ++q) {
        const int n = tid + q * 1024;
        if (n < cl.total) {
            D[n] = dv[q];
            E[n] = 0.f;
        }
    }
#pragma unroll
    for (int q = 0; q < kLdsPerThread; ++q) {
        const int n = tid + q * 1024;
        if (n < G0) B[n] = bv[q];
    }
#pragma unroll
    for (int q = 0; q < QI; ++q) {
        const int n = tid + q * 1024;
        if (n < nn) INV[n] = iv[q];
    }
    if (tid < cl.nu) UNK[tid] = uv;
    __syncthreads();
    // down
    for (int l = 0; l + 1 < cl.nl; ++l) {
        float* e = E + cl.off[l];
        const float* b = B + cl.off[l];
        const uint8_t* d = D + cl.off[l];
        const int gx = cl.gx[l], G = gx * cl.gy[l];
        const float cx = cl.cx[l], cy = cl.cy[l], dg = cl.diag[l], wd = cl.wd[l], wd2 = cl.wd2[l];
        if (cl.pre == 2) {
            // two sweeps from zero: e1 = wd b is pointwise
#pragma unroll
            for (int q = 0; q < kLdsPerThread; ++q) {
                const int n = tid + q * 1024;
                if (n >= G) continue;
                if (d[n]) { e[n] = 0.f; continue; }
                auto e1 = [&](int m) -> float { return d[m] ? 0.f : wd * b[m]; };
                int so, no;
                lds_nb(cl.per, n, gx, cl.gy[l], so, no);
                const float c = wd * b[n];
                const float ax = dg * c - cx * (e1(n - 1) + e1(n + 1)) - cy * (e1(n + so) + e1(n + no));
                e[n] = c + wd2 * (b[n] - ax);
            }
            __syncthreads();
        } else {
            for (int s = 0; s < cl.pre; ++s) lds_sweep(e, b, d, gx, cl.gy[l], cx, cy, dg, wd, s == 0, cl.per);
        }
        // restrict the residual
        const int cgx = cl.gx[l + 1], cgy = cl.gy[l + 1];
        float* bc = B + cl.off[l + 1];
        const uint8_t* dc = D + cl.off[l + 1];
        const int gy = cl.gy[l], per = cl.per;
        auto res = [&](int n, int j) -> float {
            if (d[n]) return 0.f;
            float ax = dg * e[n] - cx * (e[n - 1] + e[n + 1])
                     - cy * (e[n + south_of(per, j, gx, gy)] + e[n + north_of(per, j, gx, gy)]);
            return b[n] - ax;
        };
        for (int n = tid; n < cgx * cgy; n += 1024) {
            if (dc[n]) { bc[n] = 0.f; continue; }
            const int fj = 2 * (n / cgx);
            const int f = fj * gx + 2 * (n % cgx);
            const int js = row_s(per, fj, gy), jn = row_n(per, fj, gy);
            const int fs = f + south_of(per, fj, gx, gy), fn = f + north_of(per, fj, gx, gy);
            float s = 4.f * res(f, fj) + 2.f * (res(f - 1, fj) + res(f + 1, fj) + res(fs, js) + res(fn, jn))
                    + (res(fs - 1, js) + res(fs + 1, js) + res(fn - 1, jn) + res(fn + 1, jn));
            bc[n] = s * 0.0625f;
        }
        __syncthreads();
    }
    // coarsest: exact solve with the precomputed inverse, else sweeps
    {
        const int l = cl.nl - 1;
        float* e = E + cl.off[l];
        const float* b = B + cl.off[l];
        if (cl.nu > 0) {
            if (tid < cl.nu) {
                float v = 0.f;
                for (int q = 0; q < cl.nu; ++q) v += INV[tid * cl.nu + q] * b[UNK[q]];
                e[UNK[tid]] = v;
            }
            __syncthreads();
        } else {
            for (int s = 0; s < cl.sweeps; ++s)
                lds_sweep(e, b, D + cl.off[l], cl.gx[l], cl.gy[l], cl.cx[l], cl.cy[l], cl.diag[l], cl.wd[l], s == 0,
                          cl.per);
        }
    }
    // up
    for (int l = cl.nl - 2; l >= 0; --l) {
        float* e = E + cl.off[l];
        float* t = T + cl.off[l];
        const float* b = B + cl.off[l];
        const uint8_t* d = D + cl.off[l];
        const float* ec = E + cl.off[l + 1];
        const int gx = cl.gx[l], gy = cl.gy[l], G = gx * gy, cgx = cl.gx[l + 1], cgy = cl.gy[l + 1];
        const int per = cl.per;
        const float cx = cl.cx[l], cy = cl.cy[l], dg = cl.diag[l], wd = cl.wd[l], wd2 = cl.wd2[l];
        if (cl.post == 2) {
            // prolongation fused with the first sweep (into t), second sweep back into e
#pragma unroll
            for (int q = 0; q < kLdsPerThread; ++q) {
                const int n = tid + q * 1024;
                if (n >= G) continue;
                if (d[n]) { t[n] = 0.f; continue; }
                const int j = n / gx, i = n - j * gx;
                auto e1 = [&](int m, int mi, int mj) -> float {
                    return d[m] ? 0.f : e[m] + lds_prolong(ec, cgx, cgy, per, mi, mj);
                };
                const float c = e1(n, i, j);
                const float ax = dg * c - cx * (e1(n - 1, i - 1, j) + e1(n + 1, i + 1, j))
                               - cy * (e1(n + south_of(per, j, gx, gy), i, row_s(per, j, gy))
                                     + e1(n + north_of(per, j, gx, gy), i, row_n(per, j, gy)));
                t[n] = c + wd2 * (b[n] - ax);
            }
            __syncthreads();
#pragma unroll
            for (int q = 0; q < kLdsPerThread; ++q) {
                const int n = tid + q * 1024;
                if (n >= G) continue;
                if (d[n]) { e[n] = 0.f; continue; }
                int so, no;
                lds_nb(per, n, gx, gy, so, no);
                const float c = t[n];
                const float ax = dg * c - cx * (t[n - 1] + t[n + 1]) - cy * (t[n + so] + t[n + no]);
                e[n] = c + wd * (b[n] - ax);
            }
            __syncthreads();
        } else {
            for (int n = tid; n < G; n += 1024) {
                if (d[n]) continue;
                const int j = n / gx, i = n - j * gx;
                e[n] += lds_prolong(ec, cgx, cgy, per, i, j);
            }
            __syncthreads();
            for (int s = 0; s < cl.post; ++s) lds_sweep(e, b, d, gx, gy, cx, cy, dg, wd, false, per);
        }
    }
    for (int n = tid; n < G0; n += 1024) eglob[n] = E[n];
}

// ----------------------------------------------------------------- E field
// Interior: Ex = -(phi[j+1]-phi[j-1])/(2dx), Ey = -(phi[i+1]-phi[i-1])/(2dy);
// the reference then copies column 1 -> 0, GX-2 -> GX-1, then row 1 -> 0,
// GY-2 -> GY-1, so every node takes the interior value at the clamped index.
__global__ __launch_bounds__(kB) void k_efield(int gx, int gy, double dx, double dy, const double* __restrict__ phi,
                                               const double* __restrict__ bz, double* __restrict__ ex,
                                               double* __restrict__ ey, double4* __restrict__ fld) {
    int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= gx * gy) return;
    int j = n % gx, i = n / gx;
    int jj = min(max(j, 1), gx - 2), ii = min(max(i, 1), gy - 2);
    int m = ii * gx + jj;
    double fx = -((phi[m + 1] - phi[m - 1]) / (2 * dx));
    double fy = -((phi[m + gx] - phi[m - gx]) / (2 * dy));
    ex[n] = fx;
    ey[n] = fy;
    fld[n] = make_double4(fx, fy, bz[n], 0.0);
}

// calculateElectricFieldPeriodic (poissonSolver.cpp:411-430): Ex one-sided in
// the first and last column, central inside; Ey central with rows taken
// modulo GRID_Y-1 (row GRID_Y-1 is the image of row 0).
__global__ __launch_bounds__(kB) void k_efield_periodic(int gx, int gy, double dx, double dy,
                                                        const double* __restrict__ phi,
                                                        const double* __restrict__ bz, double* __restrict__ ex,
                                                        double* __restrict__ ey, double4* __restrict__ fld) {
    int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= gx * gy) return;
    const int j = n % gx, i = n / gx;
    double fx;
    if (j == 0) fx = -((phi[n + 1] - phi[n]) / dx);
    else if (j == gx - 1) fx = -((phi[n] - phi[n - 1]) / dx);
    else fx = -((phi[n + 1] - phi[n - 1]) / (2 * dx));
    const int p = gy - 1;
    const int up = (i + 1) % p, dn = ((i - 1) + p) % p;
    const double fy = -((phi[up * gx + j] - phi[dn * gx + j]) / (2 * dy));
    ex[n] = fx;
    ey[n] = fy;
    fld[n] = make_double4(fx, fy, bz[n], 0.0);
}


// ============================================================= host side
static Fine fine_of(const pic_ctx* c) {
    Fine f;
    f.gx = c->g.gx;
    f.gy = c->g.gy;
    f.cx = 1.0 / (c->g.dx * c->g.dx);
    f.cy = 1.0 / (c->g.dy * c->g.dy);
    f.diag = 2 * (f.cx + f.cy);
    f.per = c->p.periodic ? 1 : 0;
    return f;
}

static Lvl lvl_of(const MgLevel& m, float w) {
    Lvl L;
    L.gx = m.gx; L.gy = m.gy; L.cx = m.cx; L.cy = m.cy; L.diag = m.diag;
    L.wd = w / m.diag;
    L.wd2 = L.wd;
    L.per = m.per;
    return L;
}

// Weights (over the diagonal) of the two pre-smoothing sweeps of level m; the
// post-smoothing applies them in the reverse order, so the V-cycle stays
// symmetric.  Weighted Jacobi: omega twice.  Chebyshev (two sweeps before and
// after): the steps 1/lambda at the roots of the degree-2 Chebyshev
// polynomial on [a, 2], the part of the spectrum of D^-1 A the coarse grid does
// not correct (a = min(cx, cy) / (cx + cy), 1/2 for square cells); the error
// factor there is 1/T2((2 + a)/(2 - a)) = 0.22 against 0.36 for Jacobi 0.8.
void mg_sweep_weights(const pic_ctx* c, const MgLevel& m, float* w0, float* w1) {
    const double om = c->p.mg_omega;
    if (c->p.mg_smoother == 1 && c->p.mg_pre == 2 && c->p.mg_post == 2) {
        const double a = std::min(m.cx, m.cy) / (double)(m.cx + m.cy), b = 2.0;
        const double mid = 0.5 * (a + b), half = 0.5 * (b - a), r = std::sqrt(0.5);
        *w0 = (float)(1.0 / ((mid + half * r) * m.diag));
        *w1 = (float)(1.0 / ((mid - half * r) * m.diag));
    } else {
        *w0 = *w1 = (float)(om / m.diag);
    }
}

// level operators with the pre-smoothing weights (wd first, wd2 second)
static Lvl lvl_pre(const pic_ctx* c, const MgLevel& m) {
    Lvl L = lvl_of(m, (float)c->p.mg_omega);
    mg_sweep_weights(c, m, &L.wd, &L.wd2);
    return L;
}
// single post-smoothing sweep k (0: first) in wd
static Lvl lvl_post(const pic_ctx* c, const MgLevel& m, int k) {
    Lvl L = lvl_of(m, (float)c->p.mg_omega);
    float w0, w1;
    mg_sweep_weights(c, m, &w0, &w1);
    L.wd = L.wd2 = k == 0 ? w1 : w0;
    return L;
}

static const dim3 kMgBlock(64, 4);
static dim3 grid2d(const MgLevel& m) { return dim3((unsigned)((m.gx + 63) / 64), (unsigned)((m.gy + 3) / 4)); }
static dim3 grid_restrict(const MgLevel& coarse) {
    return dim3((unsigned)((coarse.gx + kRt - 1) / kRt), (unsigned)((coarse.gy + kRt - 1) / kRt));
}

void poisson_free(pic_ctx* c) {
    for (auto& m : c->mg) {
        if (m.dir) (void)hipFree(m.dir);
        if (m.e) (void)hipFree(m.e);
        if (m.b) (void)hipFree(m.b);
        if (m.t) (void)hipFree(m.t);
    }
    c->mg.clear();
    if (c->mg_inv) (void)hipFree(c->mg_inv);
    if (c->mg_unk) (void)hipFree(c->mg_unk);
    if (c->mg_cdir) (void)hipFree(c->mg_cdir);
    c->mg_inv = nullptr;
    c->mg_unk = nullptr;
    c->mg_cdir = nullptr;
    fast_free(c);
    c->mg_nu = 0;
    c->mg_lc = 0;
    c->mg_ctotal = 0;
}

// Inverse of the coarsest level's operator on its unknowns (Gauss-Jordan in
// FP64, stored as FP32) for the exact coarse solve.
static int coarsest_inverse(pic_ctx* c, const std::vector<uint8_t>& dir) {
    const MgLevel& m = c->mg.back();
    std::vector<int> unk;
    for (int n = 0; n < m.gx * m.gy; ++n)
        if (!dir[n]) unk.push_back(n);
    const int nu = (int)unk.size();
    if (nu == 0 || nu > kDirectMax) return PIC_OK;
    std::vector<double> A((size_t)nu * 2 * nu, 0.0);
    const double cx = m.cx, cy = m.cy, dg = m.diag;
    auto col = [&](int node) { for (int q = 0; q < nu; ++q) if (unk[q] == node) return q; return -1; };
    for (int p = 0; p < nu; ++p) {
        double* row = &A[(size_t)p * 2 * nu];
        int n = unk[p];
        row[p] = dg;
        const int j = n / m.gx;
        int so = -m.gx, no = m.gx;
        if (m.per && j == 0) so = (m.gy - 2) * m.gx;
        if (m.per && j == m.gy - 2) no = -(m.gy - 2) * m.gx;
        int nb[4] = {n - 1, n + 1, n + so, n + no};
        double cf[4] = {cx, cx, cy, cy};
        for (int k = 0; k < 4; ++k) {
            int q = col(nb[k]);
            if (q >= 0) row[q] -= cf[k];
        }
        row[nu + p] = 1.0;
    }
    for (int p = 0; p < nu; ++p) {
        int piv = p;
        for (int q = p + 1; q < nu; ++q)
            if (std::fabs(A[(size_t)q * 2 * nu + p]) > std::fabs(A[(size_t)piv * 2 * nu + p])) piv = q;
        if (piv != p)
            for (int k = 0; k < 2 * nu; ++k) std::swap(A[(size_t)p * 2 * nu + k], A[(size_t)piv * 2 * nu + k]);
        double d = A[(size_t)p * 2 * nu + p];
        if (!(std::fabs(d) > 0)) return PIC_OK;
        for (int k = 0; k < 2 * nu; ++k) A[(size_t)p * 2 * nu + k] /= d;
        for (int q = 0; q < nu; ++q) {
            if (q == p) continue;
            double f = A[(size_t)q * 2 * nu + p];
            if (f != 0)
                for (int k = 0; k < 2 * nu; ++k) A[(size_t)q * 2 * nu + k] -= f * A[(size_t)p * 2 * nu + k];
        }
    }
    std::vector<float> inv((size_t)nu * nu);
    for (int p = 0; p < nu; ++p)
        for (int q = 0; q < nu; ++q) inv[(size_t)p * nu + q] = (float)A[(size_t)p * 2 * nu + nu + q];
    PIC_HIP(hipMalloc(&c->mg_inv, inv.size() * sizeof(float)));
    PIC_HIP(hipMalloc(&c->mg_unk, nu * sizeof(int)));
    PIC_HIP(hipMemcpy(c->mg_inv, inv.data(), inv.size() * sizeof(float), hipMemcpyHostToDevice));
    PIC_HIP(hipMemcpy(c->mg_unk, unk.data(), nu * sizeof(int), hipMemcpyHostToDevice));
    c->mg_nu = nu;
    return PIC_OK;
}

// first level whose levels (down to the coarsest) fit the in-LDS cycle
static int lds_first_level(const pic_ctx* c) {
    const int L = (int)c->mg.size();
    for (int l = 0; l < L; ++l) {
        int total = 0;
        bool ok = (L - l) <= kLdsMaxLev;
        for (int m = l; m < L; ++m) {
            int nn = c->mg[m].gx * c->mg[m].gy;
            total += nn;
            if (nn > kLdsPerThread * 1024) ok = false;
        }
        if (ok && total <= kLdsNodes) return l;
    }
    return L - 1;
}

int poisson_setup(pic_ctx* c) {
    const int gx = c->g.gx, gy = c->g.gy, G = c->G;
    // Dirichlet mask and values from the geometry rectangles, in file order
    // (later rectangles overwrite earlier ones, as the reference loops do).
    const int per = c->p.periodic ? 1 : 0;
    std::vector<uint8_t> dir(G, 0);
    std::vector<double> dval(G, 0.0);
    for (const auto& r : c->geom) {
        for (int k = r.y1; k <= r.y2; ++k)
            for (int j = r.x1; j <= r.x2; ++j) {
                if (j < 0 || j >= gx || k < 0 || k >= gy) continue;
                dir[k * gx + j] = 1;
                dval[k * gx + j] = r.potential;
            }
    }
    if (per) {
        // the image row is not an unknown (the reference's system has
        // GRID_X * (GRID_Y-1) rows, poissonSolver.cpp:86)
        for (int j = 0; j < gx; ++j)
            if (!dir[(gy - 1) * gx + j]) dir[(gy - 1) * gx + j] = 2;
    } else {
        // A non-Dirichlet node on the outer frame would make the reference's
        // rows reach outside the grid; require the walls to be covered.
        for (int j = 0; j < gx; ++j)
            if (!dir[j] || !dir[(gy - 1) * gx + j]) {
                set_error("geometry leaves wall node (%d, %d) or (%d, %d) free; the non-periodic solver needs "
                          "Dirichlet walls", j, 0, j, gy - 1);
                return PIC_ERR_FORMAT;
            }
    }
    // x walls: the reference's rows never wrap in x
    for (int k = 0; k < gy; ++k)
        if (!dir[k * gx] || !dir[k * gx + gx - 1]) {
            set_error("geometry leaves wall node (0, %d) or (%d, %d) free; the solver needs Dirichlet walls "
                      "at x = 0 and x = width", k, gx - 1, k);
            return PIC_ERR_FORMAT;
        }
    PIC_HIP(hipMemcpy(c->dir, dir.data(), G, hipMemcpyHostToDevice));
    PIC_HIP(hipMemcpy(c->dval, dval.data(), G * sizeof(double), hipMemcpyHostToDevice));

    // Multigrid hierarchy: vertex-centred coarsening, coarse node (I, J) at fine
    // node (2I, 2J).  A level with an odd cell count in a Dirichlet direction
    // gets one more coarse node than its nodes cover: it lies half a coarse
    // cell beyond the wall and is Dirichlet like the wall (the coarse levels
    // only precondition, so their domain may overhang the fine one).  The
    // transfer kernels never read a fine node of it: they skip Dirichlet
    // coarse nodes, and a fine node at most reaches coarse node (i + 1) / 2.
    // A periodic y direction must keep an even period to coarsen.
    poisson_free(c);
    int lx = gx, ly = gy;
    double hx = c->g.dx, hy = c->g.dy;
    std::vector<uint8_t> cur = dir;
    std::vector<std::vector<uint8_t>> masks;
    while (true) {
        MgLevel m;
        m.gx = lx; m.gy = ly;
        m.cx = (float)(1.0 / (hx * hx));
        m.cy = (float)(1.0 / (hy * hy));
        m.diag = 2.f * (m.cx + m.cy);
        m.per = per;
        size_t nn = (size_t)lx * ly;
        PIC_HIP(hipMalloc(&m.dir, nn));
        PIC_HIP(hipMalloc(&m.e, nn * sizeof(float)));
        PIC_HIP(hipMalloc(&m.b, nn * sizeof(float)));
        PIC_HIP(hipMalloc(&m.t, nn * sizeof(float)));
        PIC_HIP(hipMemcpy(m.dir, cur.data(), nn, hipMemcpyHostToDevice));
        masks.push_back(cur);
        PIC_HIP(hipMemset(m.e, 0, nn * sizeof(float)));
        PIC_HIP(hipMemset(m.b, 0, nn * sizeof(float)));
        PIC_HIP(hipMemset(m.t, 0, nn * sizeof(float)));
        c->mg.push_back(m);
        int cxn = lx - 1, cyn = ly - 1;
        bool can = cxn >= 4 && cyn >= 4 && nn > 25 && (!per || cyn % 2 == 0);
        if (!can) break;
        int nlx = (cxn + 1) / 2 + 1, nly = (cyn + 1) / 2 + 1;
        std::vector<uint8_t> nxt((size_t)nlx * nly, 0);
        for (int J = 0; J < nly; ++J)
            for (int I = 0; I < nlx; ++I) {
                const bool inside = 2 * I < lx && 2 * J < ly;     // else the overhanging node
                uint8_t d = (!inside || cur[(size_t)(2 * J) * lx + 2 * I] != 0) ? 1 : 0;
                if (I == 0 || I == nlx - 1 || (!per && (J == 0 || J == nly - 1))) d = 1;
                if (per && J == nly - 1 && !d) d = 2;
                nxt[(size_t)J * nlx + I] = d;
            }
        cur.swap(nxt);
        lx = nlx; ly = nly;
        hx *= 2; hy *= 2;
    }
    PIC_TRY(coarsest_inverse(c, cur));
    const size_t last = (size_t)c->mg.back().gx * c->mg.back().gy;
    if (last > (size_t)kLdsNodes || last > (size_t)kLdsPerThread * 1024) {
        set_error("coarsest multigrid level %dx%d is too large for the in-LDS coarse solve; "
                  "use NX, NY with more factors of two",
                  c->mg.back().gx, c->mg.back().gy);
        return PIC_ERR_UNSUPPORTED;
    }
    // masks of the in-LDS levels, concatenated in level order
    c->mg_lc = lds_first_level(c);
    std::vector<uint8_t> cd;
    for (size_t l = (size_t)c->mg_lc; l < c->mg.size(); ++l) cd.insert(cd.end(), masks[l].begin(), masks[l].end());
    c->mg_ctotal = (int)cd.size();
    PIC_HIP(hipMalloc(&c->mg_cdir, cd.size()));
    PIC_HIP(hipMemcpy(c->mg_cdir, cd.data(), cd.size(), hipMemcpyHostToDevice));
    PIC_HIP(hipDeviceSynchronize());
    PIC_TRY(fast_setup(c, dir, dval));
    return PIC_OK;
}


int launch_overwrite_rho(pic_ctx* c) {
    hipLaunchKernelGGL(k_overwrite_rho, dim3(nblk(c->G)), dim3(kB), 0, c->stream, c->G, c->rho, c->dir, c->dval);
    PIC_HIP(hipGetLastError());
    return PIC_OK;
}

int launch_overwrite_phi(pic_ctx* c) {
    hipLaunchKernelGGL(k_overwrite_phi, dim3(nblk(c->G)), dim3(kB), 0, c->stream, c->G, c->g.gx, c->phi, c->dir,
                       c->dval);
    PIC_HIP(hipGetLastError());
    return PIC_OK;
}

int launch_efield(pic_ctx* c) {
    c->e_stale = 0;
    hipEvent_t ev;
    PIC_TRY(timer_begin(c, PIC_STAGE_EFIELD, &ev));
    if (c->p.periodic)
        hipLaunchKernelGGL(k_efield_periodic, dim3(nblk(c->G)), dim3(kB), 0, c->stream, c->g.gx, c->g.gy, c->g.dx,
                           c->g.dy, c->phi, c->bz, c->ex, c->ey, c->fld);
    else
        hipLaunchKernelGGL(k_efield, dim3(nblk(c->G)), dim3(kB), 0, c->stream, c->g.gx, c->g.gy, c->g.dx, c->g.dy,
                           c->phi, c->bz, c->ex, c->ey, c->fld);
    PIC_HIP(hipGetLastError());
    PIC_TRY(timer_end(c, PIC_STAGE_EFIELD, ev));
    return PIC_OK;
}

// Pre-smoothing of one level from zero into m.e with the configured sweep count.
static void presmooth(pic_ctx* c, MgLevel& m, const int* flags) {
    const float w = (float)c->p.mg_omega;
    Lvl lv = lvl_of(m, w);
    dim3 g2 = grid2d(m);
    float* a = m.e;
    float* b = m.t;
    for (int s = 0; s < c->p.mg_pre; ++s) {
        hipLaunchKernelGGL(k_mg_smooth, g2, kMgBlock, 0, c->stream, lv, a, b, m.b, m.dir, s == 0 ? 1 : 0, flags);
        std::swap(a, b);
    }
    if (c->p.mg_pre == 0) {
        Lvl zl = lv;
        zl.wd = 0.f;
        hipLaunchKernelGGL(k_mg_smooth, g2, kMgBlock, 0, c->stream, zl, b, a, m.b, m.dir, 1, flags);
    }
    if (a != m.e) std::swap(m.e, m.t);
}

// The V-cycle below the finest level's pre-smoothing (already in m0.e, with
// its right-hand side in m0.b), ending with the last post-smoothing sweep of
// the finest level into z and the partials of r.z and r.zold.
template <typename T>
static int vcycle_rest(pic_ctx* c, const T* r, float* z, const float* zold, int first, const int* flags) {
    const float w = (float)c->p.mg_omega;
    const int L = (int)c->mg.size();
    const int lc = c->mg_lc;
    hipStream_t st = c->stream;
    const bool fpre = c->p.mg_pre == 2, fpost = c->p.mg_post == 2;
    for (int l = 0; l < lc; ++l) {
        MgLevel& m = c->mg[l];
        if (l > 0) {
            if (fpre)
                hipLaunchKernelGGL(k_mg_pre2<float>, grid2d(m), kMgBlock, 0, st, lvl_pre(c, m), m.b, (float*)nullptr,
                                   m.e, m.dir, flags);
            else
                presmooth(c, m, flags);
        }
        MgLevel& n = c->mg[l + 1];
        hipLaunchKernelGGL(k_mg_restrict, grid_restrict(n), dim3(kRt * kRt), 0, st, lvl_of(m, w), lvl_of(n, w), m.e,
                           m.b, m.dir,
                           n.b, n.dir, flags);
    }
    {
        CoarseLevels cl;
        cl.nl = L - lc;
        cl.total = c->mg_ctotal;
        cl.cdir = c->mg_cdir;
        int off = 0;
        for (int l = lc; l < L; ++l) {
            const MgLevel& m = c->mg[l];
            int k = l - lc;
            cl.gx[k] = m.gx; cl.gy[k] = m.gy; cl.off[k] = off;
            cl.cx[k] = m.cx; cl.cy[k] = m.cy; cl.diag[k] = m.diag;
            mg_sweep_weights(c, m, &cl.wd[k], &cl.wd2[k]);
            off += m.gx * m.gy;
        }
        cl.pre = c->p.mg_pre;
        cl.post = c->p.mg_post;
        cl.per = c->p.periodic ? 1 : 0;
        const MgLevel& last = c->mg[L - 1];
        cl.sweeps = std::max(50, 4 * std::max(last.gx, last.gy));
        cl.nu = c->mg_nu;
        cl.inv = c->mg_inv;
        cl.unk = c->mg_unk;
        hipLaunchKernelGGL(k_mg_coarse_cycle, dim3(1), dim3(1024), 0, st, cl, c->mg[lc].b, c->mg[lc].e, flags);
    }
    if (lc == 0) {
        // the whole cycle ran in LDS: z is its output, copied with the dot products
        MgLevel& m = c->mg[0];
        Lvl zl = lvl_of(m, w);
        zl.wd = 0.f;
        hipLaunchKernelGGL(k_smooth_dots<T>, dim3(kNP), dim3(kB), 0, st, zl, m.e, z, m.b, m.dir, r, zold,
                           c->part + P_RZ * kNP, c->part + P_RZO * kNP, flags, first);
    }
    for (int l = lc - 1; l >= 0; --l) {
        MgLevel& m = c->mg[l];
        MgLevel& n = c->mg[l + 1];
        Lvl lv = lvl_of(m, w);
        dim3 g2 = grid2d(m);
        // post-smoothing: all sweeps but the last leave the correction in m.e
        int rest;
        if (fpost) {
            hipLaunchKernelGGL(k_mg_prolong_sweep, g2, dim3(kPx, kPy), 0, st, lvl_post(c, m, 0), lvl_of(n, w), m.e, n.e,
                               m.b, m.dir, m.t, flags);
            std::swap(m.e, m.t);
            rest = 1;
            lv = lvl_post(c, m, 1);
        } else {
            hipLaunchKernelGGL(k_mg_prolong, g2, kMgBlock, 0, st, lv, lvl_of(n, w), m.e, m.dir, n.e, flags);
            rest = c->p.mg_post;
        }
        const int last_here = l == 0 ? 1 : 0;         // the finest level's last sweep is k_smooth_dots
        for (int s = 0; s < rest - last_here; ++s) {
            hipLaunchKernelGGL(k_mg_smooth, g2, kMgBlock, 0, st, lv, m.e, m.t, m.b, m.dir, 0, flags);
            std::swap(m.e, m.t);
        }
        if (l == 0) {
            Lvl zl = lv;
            if (rest == 0) zl.wd = 0.f;                 // no post-smoothing: z = e
            hipLaunchKernelGGL(k_smooth_dots<T>, dim3(kNP), dim3(kB), 0, st, zl, m.e, z, m.b, m.dir, r, zold,
                               c->part + P_RZ * kNP, c->part + P_RZO * kNP, flags, first);
        }
    }
    PIC_HIP(hipGetLastError());
    return PIC_OK;
}

static int read_state(pic_ctx* c) {
    PIC_HIP(hipMemcpyAsync(c->scal_host, c->scal, PIC_STATE_BYTES, hipMemcpyDeviceToHost, c->stream));
    PIC_HIP(hipStreamSynchronize(c->stream));
    return PIC_OK;
}

// Flexible PCG from the residual already in R[0] (type T; its check already
// done): z0 = M r0, then iterations until flags[0] is set or max_it.  x, R, P,
// q are T; z is FP32 (the V-cycle output); dots and scalars are FP64.  The
// host reads the flag between batches (the first as long as *batch_hint); a
// start that already meets the tolerance is found by the device (every kernel
// returns at once once flags[0] is set), so nothing is read before the first
// batch.  count0: the device's iteration counter flags[1] before this run.
template <typename T>
static int run_pcg(pic_ctx* c, T* x, T* const R[2], T* const P[2], T* q, float* const Z[2], int* flags,
                   double tol2, int sbb, int srr, int srr0, int max_it, int* batch_hint, int count0) {
    const Fine f = fine_of(c);
    hipStream_t s = c->stream;
    double* part = c->part;
    MgLevel& m0 = c->mg[0];
    const Lvl l0 = lvl_pre(c, m0);
    const int flag_off = (int)(flags - c->flags);
    if (max_it <= 0) {
        *batch_hint = 0;
        return PIC_OK;
    }
    if (c->p.mg_pre == 2) {
        hipLaunchKernelGGL(k_mg_pre2<T>, grid2d(m0), kMgBlock, 0, s, l0, R[0], m0.b, m0.e, m0.dir, flags);
    } else {
        hipLaunchKernelGGL(k_d2f<T>, dim3(nblk(c->G)), dim3(kB), 0, s, c->G, R[0], m0.b, flags);
        presmooth(c, m0, flags);
    }
    PIC_TRY(vcycle_rest<T>(c, R[0], Z[0], Z[1], 1, flags));
    hipLaunchKernelGGL(k_beta, dim3(1), dim3(1024), 0, s, part + P_RZ * kNP, part + P_RZO * kNP, c->scal, flags, 1);
    PIC_HIP(hipGetLastError());
    int it = 0;
    int batch = *batch_hint > 0 ? *batch_hint : 2;
    bool first = true;
    while (true) {
        if (!first) {
            PIC_TRY(read_state(c));
            if (c->flags_host[flag_off] || it >= max_it) break;
        }
        first = false;
        for (int k = 0; k < batch && it < max_it; ++k) {
            ++it;
            const int cur = it & 1, old = cur ^ 1;
            hipLaunchKernelGGL(k_apply_p<T>, dim3(kNP), dim3(kB), 0, s, f, Z[old], P[old], P[cur], c->dir, q,
                               part + P_PQ * kNP, c->scal, flags, it == 1 ? 1 : 0);
            if (c->p.mg_pre == 2)
                hipLaunchKernelGGL((k_update_pre2<1, T>), dim3(kNP), dim3(kB), 0, s, f, l0, x, P[cur], R[old], R[cur],
                                   q, c->dir, m0.b, m0.e, part + P_RR * kNP, c->scal, part + P_PQ * kNP, flags);
            else
                hipLaunchKernelGGL((k_update_pre2<0, T>), dim3(kNP), dim3(kB), 0, s, f, l0, x, P[cur], R[old], R[cur],
                                   q, c->dir, m0.b, m0.e, part + P_RR * kNP, c->scal, part + P_PQ * kNP, flags);
            hipLaunchKernelGGL(k_check, dim3(1), dim3(1024), 0, s, part + P_RR * kNP, (const double*)nullptr, tol2,
                               flags, c->scal, 0, sbb, srr, srr0);
            if (c->p.mg_pre != 2) presmooth(c, m0, flags);
            PIC_TRY(vcycle_rest<T>(c, R[cur], Z[cur], Z[old], 0, flags));
            hipLaunchKernelGGL(k_beta, dim3(1), dim3(1024), 0, s, part + P_RZ * kNP, part + P_RZO * kNP, c->scal,
                               flags, 0);
            PIC_HIP(hipGetLastError());
        }
        batch = 2;
    }
    *batch_hint = c->flags_host[flag_off + 1] - count0;
    return PIC_OK;
}

// FP32 work vectors of the mixed-precision solve (allocated on first use)
static int ensure_f32(pic_ctx* c) {
    if (c->rf0) return PIC_OK;
    const size_t b = (size_t)c->G * sizeof(float);
    float** bufs[] = {&c->rf0, &c->rf1, &c->pf0, &c->pf1, &c->qf, &c->df};
    for (float** q : bufs) {
        PIC_HIP(hipMalloc((void**)q, b));
        PIC_HIP(hipMemsetAsync(*q, 0, b, c->stream));
    }
    return PIC_OK;
}

static int solve_finish(pic_ctx* c, const SolveTail& t, pic_solve_stats* st, int* redo);

int poisson_solve(pic_ctx* c, int iteration, pic_solve_stats* st, int defer) {
    const Fine f = fine_of(c);
    const int G = c->G;
    hipStream_t s = c->stream;
    // a deposit node pass left to this solve (dropped if the solve fails first)
    const bool fin = c->fin_pending != 0;
    c->fin_pending = 0;
    hipEvent_t ev;
    PIC_TRY(timer_begin(c, PIC_STAGE_POISSON, &ev));
    if (!c->solve_ev[0]) {
        PIC_HIP(hipEventCreate(&c->solve_ev[0]));
        PIC_HIP(hipEventCreate(&c->solve_ev[1]));
    }
    hipEvent_t t0 = c->solve_ev[0];
    double* part = c->part;
    const double tol2 = c->p.poisson_tol * c->p.poisson_tol;
    double* const R[2] = {c->pr, c->pr2};
    double* const P[2] = {c->pp, c->pp2};
    float* const Z[2] = {c->zf0, c->zf1};
    const bool mixed = c->p.poisson_precision == 1;
    if (mixed) PIC_TRY(ensure_f32(c));
    // the path, from host state only: the direct solve unless a form asks for the PCG
    const bool learned = c->unet && !c->guess_set && iteration > 0 && c->have_prev_phi;
    const bool direct = c->fast && !mixed && !learned && !c->guess_set;
    // the direct solve is timed by device wall-clock stamps in its first and
    // last kernels, the PCG paths by events
    if (!direct) PIC_HIP(hipEventRecord(t0, s));
    unsigned long long* ts = direct ? c->tstamp : nullptr;

    // flags cleared, overwriteRho, ||b||^2 (b: residual of the Dirichlet-only field)
    if (fin) {
        // the deposit's node pass, deferred to here, writes rho first
        hipLaunchKernelGGL(k_finalize_prep, dim3(kNP), dim3(kB), 0, s, c->fin, c->g.gp, c->p.deposit_energy, f,
                           c->rho, c->dir, c->dval, part + P_BB * kNP, c->flags, ts);
        c->meshes_stale = 1;
    } else {
        hipLaunchKernelGGL(k_prep_rhs, dim3(kNP), dim3(kB), 0, s, f, c->rho, c->dir, c->dval, part + P_BB * kNP,
                           c->flags, ts);
    }
    // initial guess
    int guess = 0;
    const double* src = nullptr;
    const double* prev = nullptr;
    double* keep = nullptr;
    const bool extrap = c->p.poisson_guess == 1;
    if (c->guess_set) {
        src = c->guess;
        guess = 2;
    } else if (iteration > 0 && c->have_prev_phi) {
        if (extrap && c->have_prev_phi >= 2) {
            prev = c->phi_prev;
            guess = 3;
        } else {
            src = c->phi;
            guess = 1;
        }
        if (extrap) keep = c->phi_prev;
    }
    // learned guess (in place of an extrapolated one): keep phi_{n-1} while
    // pairs are trained on; once the network has trained enough, form the
    // prediction and both starting residuals in one pass, and let the device
    // keep the lower (the host learns the choice with the solve's state)
    bool picked = false;
    if (learned) {
        if (prev) {
            prev = nullptr;
            src = c->phi;
            guess = 1;
        }
        keep = nullptr;
        if (c->learn_trained < c->learn_train)
            PIC_HIP(hipMemcpyAsync(c->phi_prev, c->phi, (size_t)G * sizeof(double), hipMemcpyDeviceToDevice, s));
        if (c->learn_trained >= c->learn_min) {
            UnetPrediction pr{};
            PIC_TRY(unet_predict(c, c->unet, &pr));
            const double diag = 2.0 / (c->g.dx * c->g.dx) + 2.0 / (c->g.dy * c->g.dy);
            hipLaunchKernelGGL(k_guess_residuals, dim3(kNP), dim3(kB), 0, s, f, c->rho, c->phi, c->dir, c->dval,
                               pr.out, pr.scl, pr.N, diag, c->guess, R[0], R[1], part + P_RR * kNP,
                               part + P_RZO * kNP);
            hipLaunchKernelGGL(k_sum2, dim3(1), dim3(1024), 0, s, part + P_RR * kNP, part + P_RZO * kNP, c->scal);
            hipLaunchKernelGGL(k_take_guess, dim3(nblk(G)), dim3(kB), 0, s, G, c->phi, R[0], c->guess, R[1],
                               part + P_RR * kNP, part + P_RZO * kNP, c->dir, c->dval, c->scal + S_GPICK);
            c->learn_tried += 1;
            picked = true;
            guess = 4;
        }
    }
    // direct solve: the PCG below starts from it and only checks (or
    // corrects) the residual
    bool efield_done = false;
    if (direct) {
        if (keep) PIC_HIP(hipMemcpyAsync(keep, c->phi, (size_t)G * sizeof(double), hipMemcpyDeviceToDevice, s));
        PIC_TRY(fast_solve(c, c->phi));   // the interior nodes; the residual imposes the Dirichlet values
        guess = 5;
        if (f.per) {
            hipLaunchKernelGGL(k_residual_fix<false>, dim3(kNP), dim3(kB), 0, s, f, c->rho, c->phi, c->dir, c->dval,
                               R[0], part + P_RR * kNP, EfOut{});
        } else {
            // and E in the same pass
            hipLaunchKernelGGL(k_residual_fix<true>, dim3(kNP), dim3(kB), 0, s, f, c->rho, c->phi, c->dir, c->dval,
                               R[0], part + P_RR * kNP,
                               EfOut{c->g.dx, c->g.dy, c->bz, c->ex, c->ey, (double4*)c->fld});
            efield_done = true;
            c->e_stale = 1;
        }
    } else if (!picked) {
        hipLaunchKernelGGL(k_init_guess, dim3(nblk(G)), dim3(kB), 0, s, G, c->phi, src, prev, keep, c->dir,
                           c->dval);
        hipLaunchKernelGGL(k_residual, dim3(kNP), dim3(kB), 0, s, f, c->rho, c->phi, c->dir, R[0],
                           part + P_RR * kNP);
    }
    if (direct) {
        // the check publishes its state to the host itself (no copy, no event)
        c->mail_seq += 1;
        hipLaunchKernelGGL(k_check_publish, dim3(1), dim3(1024), 0, s, part + P_RR * kNP, part + P_BB * kNP, tol2,
                           c->flags, c->scal, 1, S_BB, S_RR, S_RR0, c->mail_dev, (const unsigned long long*)c->tstamp,
                           c->mail_seq);
    } else {
        hipLaunchKernelGGL(k_check, dim3(1), dim3(1024), 0, s, part + P_RR * kNP, part + P_BB * kNP, tol2, c->flags,
                           c->scal, 1, S_BB, S_RR, S_RR0);
    }
    PIC_HIP(hipGetLastError());
    SolveTail tail;
    tail.direct = direct;
    tail.mixed = mixed;
    tail.learned = learned;
    tail.extrap_keep = extrap && keep;
    tail.guess = guess;
    tail.picked = picked;
    tail.ev = ev;
    if (direct) {
        // The host waits for the check's publication only, not for the work
        // enqueued after it.
        PIC_TRY(timer_end(c, PIC_STAGE_POISSON, ev));
        if (!efield_done) PIC_TRY(launch_efield(c));
        if (defer) {
            // the caller enqueues more work first and calls poisson_finish
            tail.active = 1;
            c->tail = tail;
            return PIC_OK;
        }
    }
    return solve_finish(c, tail, st, nullptr);
}

int poisson_finish(pic_ctx* c, pic_solve_stats* st, int* redo) {
    if (redo) *redo = 0;
    if (!c->tail.active) return PIC_OK;
    const SolveTail t = c->tail;
    c->tail.active = 0;
    return solve_finish(c, t, st, redo);
}

// The rest of a solve after its first check: the wait for the check of a
// direct solve, PCG iterations where needed, the statistics.
// Wait for the check kernel's publication (sequence number c->mail_seq) and
// take its state block; the device time of the solve up to the check from the
// wall-clock stamps.  Should the stream finish without the publication being
// visible, the state block is copied instead.
static int wait_mail(pic_ctx* c, double* ms) {
    const unsigned want = c->mail_seq;
    // the stream is queried only after a while (a query on every spin would
    // disturb the stream's own progress)
    auto next_query = std::chrono::steady_clock::now() + std::chrono::milliseconds(20);
    for (unsigned spin = 1;; ++spin) {
        if (__atomic_load_n(&c->mail->seq, __ATOMIC_ACQUIRE) == want) break;
        if ((spin & 255u) == 0 && std::chrono::steady_clock::now() >= next_query) {
            next_query = std::chrono::steady_clock::now() + std::chrono::milliseconds(20);
            const hipError_t q = hipStreamQuery(c->stream);
            if (q != hipSuccess && q != hipErrorNotReady) {
                set_error("Poisson check: %s", hipGetErrorString(q));
                return PIC_ERR_HIP;
            }
            if (q == hipSuccess && __atomic_load_n(&c->mail->seq, __ATOMIC_ACQUIRE) != want) {
                PIC_HIP(hipMemcpy(c->scal_host, c->scal, PIC_STATE_BYTES, hipMemcpyDeviceToHost));
                *ms = 0.0;
                return PIC_OK;
            }
        }
    }
    std::memcpy(c->scal_host, c->mail->scal, sizeof(c->mail->scal));
    std::memcpy(c->flags_host, c->mail->flags, sizeof(c->mail->flags));
    *ms = (double)(c->mail->t1 - c->mail->t0) / c->wall_khz;
    return PIC_OK;
}

static int solve_finish(pic_ctx* c, const SolveTail& t, pic_solve_stats* st, int* redo) {
    const Fine f = fine_of(c);
    const int G = c->G;
    hipStream_t s = c->stream;
    hipEvent_t ev = t.ev;
    hipEvent_t t0 = c->solve_ev[0], t1 = c->solve_ev[1];
    double* part = c->part;
    const double tol2 = c->p.poisson_tol * c->p.poisson_tol;
    double* const R[2] = {c->pr, c->pr2};
    double* const P[2] = {c->pp, c->pp2};
    float* const Z[2] = {c->zf0, c->zf1};
    const bool mixed = t.mixed, learned = t.learned;
    int guess = t.guess;
    bool fields_done = false;
    double direct_ms = 0.0;
    if (t.direct) {
        PIC_TRY(wait_mail(c, &direct_ms));
        fields_done = c->flags_host[0] == 1;
        if (!fields_done) {
            if (redo) *redo = 1;
            PIC_TRY(timer_begin(c, PIC_STAGE_POISSON, &ev));
            PIC_HIP(hipEventRecord(t0, s));    // the correcting PCG's time is added
        }
    }
    int iters = 0, inner_total = 0;
    if (fields_done) {
        // the direct solution met the tolerance: no PCG iteration
    } else if (!mixed) {
        // the solve's first kernel cleared the iteration counter
        PIC_TRY(run_pcg<double>(c, c->phi, R, P, c->pq, Z, c->flags, tol2, S_BB, S_RR, S_RR0, c->p.poisson_max_iter,
                                &c->last_iters, 0));
        iters = c->flags_host[1];
        inner_total = iters;
    } else {
        // iterative refinement: FP64 residual, FP32 PCG for the correction to
        // the inner relative residual, FP64 update of phi
        float* const RF[2] = {c->rf0, c->rf1};
        float* const PF[2] = {c->pf0, c->pf1};
        const double eta = c->p.poisson_inner_tol > 0 ? c->p.poisson_inner_tol : 1e-4;
        int outer = 0;
        while (true) {
            PIC_TRY(read_state(c));
            if (c->flags_host[0] || outer >= c->p.poisson_max_iter) break;
            hipLaunchKernelGGL(k_r_to_f32, dim3(kNP), dim3(kB), 0, s, G, R[0], RF[0], part + P_RR * kNP);
            hipLaunchKernelGGL(k_reset_inner, dim3(1), dim3(1), 0, s, c->flags);
            hipLaunchKernelGGL(k_check, dim3(1), dim3(1024), 0, s, part + P_RR * kNP, (const double*)nullptr,
                               eta * eta, c->flags + 2, c->scal, 1, S_IBB, S_IRR, S_IRR0);
            hipLaunchKernelGGL(k_zero_f32v, dim3(nblk(G)), dim3(kB), 0, s, G, c->df);
            PIC_HIP(hipGetLastError());
            // the inner counter flags[3] as read at the top of this step
            PIC_TRY(run_pcg<float>(c, c->df, RF, PF, c->qf, Z, c->flags + 2, eta * eta, S_IBB, S_IRR, S_IRR0,
                                   c->p.poisson_max_iter, &c->last_inner, c->flags_host[3]));
            if (c->flags_host[2] == 2) {
                set_error("Poisson inner FP32 PCG diverged (NaN residual)");
                return PIC_ERR_SOLVER;
            }
            hipLaunchKernelGGL(k_ir_update, dim3(nblk(G)), dim3(kB), 0, s, G, c->phi, c->df);
            hipLaunchKernelGGL(k_residual, dim3(kNP), dim3(kB), 0, s, f, c->rho, c->phi, c->dir, R[0],
                               part + P_RR * kNP);
            hipLaunchKernelGGL(k_check, dim3(1), dim3(1024), 0, s, part + P_RR * kNP, (const double*)nullptr, tol2,
                               c->flags, c->scal, 0, S_BB, S_RR, S_RR0);
            PIC_HIP(hipGetLastError());
            ++outer;
        }
        iters = outer;
        inner_total = c->flags_host[3];
    }
    const double bb = c->scal_host[S_BB];
    const double bnorm = std::sqrt(bb);
    const double rel0 = bnorm > 0 ? std::sqrt(c->scal_host[S_RR0]) / bnorm : 0.0;
    const double rel = bnorm > 0 ? std::sqrt(c->scal_host[S_RR]) / bnorm : 0.0;
    if (c->flags_host[0] == 2) {
        set_error("Poisson PCG diverged (NaN residual) after %d iterations", iters);
        return PIC_ERR_SOLVER;
    }
    if (c->flags_host[0] == 0) {
        set_error("Poisson PCG did not reach relative residual %.3e in %d iterations (%.3e)", c->p.poisson_tol,
                  iters, rel);
        return PIC_ERR_SOLVER;
    }
    double ms = direct_ms;
    if (!fields_done) {
        PIC_TRY(launch_overwrite_phi(c));
        PIC_HIP(hipEventRecord(t1, s));
        PIC_HIP(hipEventSynchronize(t1));
        float e = 0;
        PIC_HIP(hipEventElapsedTime(&e, t0, t1));
        ms += e;
    }
    c->guess_set = 0;
    c->have_prev_phi = t.extrap_keep ? 2 : 1;
    if (learned && c->learn_trained < c->learn_train) {
        float loss = 0.f;
        PIC_TRY(unet_train_pair(c, c->unet, c->phi_prev, c->learn_lr, &loss));
        c->learn_trained += 1;
        c->learn_loss = loss;
    }
    if (!fields_done) {
        PIC_TRY(timer_end(c, PIC_STAGE_POISSON, ev));
        PIC_TRY(launch_efield(c));
    }
    if (t.picked) {
        // the device's choice between the prediction and phi_{n-1}
        if (c->scal_host[S_GPICK] > 0.5) c->learn_used += 1;
        else guess = 1;
    }
    if (st) {
        st->iterations = iters;
        st->initial_relres = rel0;
        st->final_relres = rel;
        st->ms = ms;
        st->guess = guess;
        st->inner_iterations = inner_total;
    }
    return PIC_OK;
}

} // namespace pic
