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
#include "pic_internal.h"

#include <algorithm>
#include <cmath>

namespace pic {

static constexpr int kB = 256;
static constexpr int kRedBlocks = 1024;     // fixed grid for reductions (deterministic order)

enum Scal { S_RR = 0, S_BB, S_RZ, S_RZOLD, S_PQ, S_RZPREV, S_ALPHA, S_BETA, S_N };

static inline unsigned nblk(int64_t n) { return (unsigned)((n + kB - 1) / kB); }
static inline unsigned nred(int64_t n) { return (unsigned)std::min<int64_t>(kRedBlocks, (n + kB - 1) / kB); }

// ------------------------------------------------------- block reductions
// Sum of `v` over the block in a fixed order; result valid in thread 0.
__device__ __forceinline__ double block_sum(double v, double* sh) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) sh[wid] = v;
    __syncthreads();
    double s = 0;
    if (threadIdx.x == 0)
        for (int w = 0; w < (int)(blockDim.x >> 6); ++w) s += sh[w];
    __syncthreads();
    return s;
}

// Grid-wide sum of up to two values per thread: block partials go to `part`,
// the last block to finish adds them in block order and writes scal[slot0/1].
__device__ void grid_sum2(double a, double b, double* part, unsigned* count,
                          double* scal, int slot0, int slot1) {
    __shared__ double sh[16];
    __shared__ bool last;
    double sa = block_sum(a, sh);
    double sb = slot1 >= 0 ? block_sum(b, sh) : 0.0;
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = sa;
        part[2 * blockIdx.x + 1] = sb;
        __threadfence();
        unsigned prev = atomicInc(count, 0xffffffffu);
        last = (prev == gridDim.x - 1);
    }
    __syncthreads();
    if (!last) return;
    __threadfence();
    double ta = 0, tb = 0;
    for (unsigned k = threadIdx.x; k < gridDim.x; k += blockDim.x) {
        ta += part[2 * k];
        tb += part[2 * k + 1];
    }
    // combine per-thread partials in a fixed order
    ta = block_sum(ta, sh);
    tb = block_sum(tb, sh);
    if (threadIdx.x == 0) {
        scal[slot0] = ta;
        if (slot1 >= 0) scal[slot1] = tb;
        *count = 0;
    }
}

// ------------------------------------------------------------ fine kernels
struct Fine {
    int gx, gy;
    double cx, cy, diag;   // 1/dx^2, 1/dy^2, 2(1/dx^2 + 1/dy^2)
};

// rho[geom] = V; rho = -rho  (overwriteRho, poissonSolver.cpp:384-402)
__global__ void k_overwrite_rho(int G, double* rho, const uint8_t* dir, const double* dval) {
    int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= G) return;
    double r = dir[n] ? dval[n] : rho[n];
    rho[n] = -r;
}

// phi[geom] = V (overwritePhi :364-382)
__global__ void k_overwrite_phi(int G, double* phi, const uint8_t* dir, const double* dval) {
    int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= G) return;
    if (dir[n]) phi[n] = dval[n];
}

// x <- guess (or 0) with Dirichlet values imposed
__global__ void k_init_guess(int G, double* x, const double* src, const uint8_t* dir, const double* dval) {
    int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= G) return;
    x[n] = dir[n] ? dval[n] : (src ? src[n] : 0.0);
}

// r = rhs - A x on unknown nodes (0 on Dirichlet), rr = r.r.
// rhs = rho after overwriteRho.  x holds V on Dirichlet nodes.
__global__ __launch_bounds__(kB) void k_residual(Fine f, const double* __restrict__ rhs,
                                                 const double* __restrict__ x, const uint8_t* __restrict__ dir,
                                                 double* __restrict__ r, double* part, unsigned* count,
                                                 double* scal, int slot) {
    double acc = 0;
    const int G = f.gx * f.gy;
    for (int n = blockIdx.x * blockDim.x + threadIdx.x; n < G; n += gridDim.x * blockDim.x) {
        double v = 0;
        if (!dir[n]) {
            double ax = f.diag * x[n] - f.cx * (x[n - 1] + x[n + 1]) - f.cy * (x[n - f.gx] + x[n + f.gx]);
            v = rhs[n] - ax;
        }
        r[n] = v;
        acc += v * v;
    }
    grid_sum2(acc, 0, part, count, scal, slot, -1);
}

// q = A p (p = 0 on Dirichlet nodes), pq = p.q
__global__ __launch_bounds__(kB) void k_apply(Fine f, const double* __restrict__ p, const uint8_t* __restrict__ dir,
                                              double* __restrict__ q, double* part, unsigned* count, double* scal) {
    double acc = 0;
    const int G = f.gx * f.gy;
    for (int n = blockIdx.x * blockDim.x + threadIdx.x; n < G; n += gridDim.x * blockDim.x) {
        double v = 0;
        if (!dir[n]) {
            v = f.diag * p[n] - f.cx * (p[n - 1] + p[n + 1]) - f.cy * (p[n - f.gx] + p[n + f.gx]);
            acc += p[n] * v;
        }
        q[n] = v;
    }
    grid_sum2(acc, 0, part, count, scal, S_PQ, -1);
}

// alpha = rz / pq; x += alpha p; r -= alpha q; rr = r.r
__global__ __launch_bounds__(kB) void k_update_xr(int G, double* __restrict__ x, double* __restrict__ r,
                                                  const double* __restrict__ p, const double* __restrict__ q,
                                                  double* part, unsigned* count, double* scal) {
    const double alpha = scal[S_RZ] / scal[S_PQ];
    double acc = 0;
    for (int n = blockIdx.x * blockDim.x + threadIdx.x; n < G; n += gridDim.x * blockDim.x) {
        x[n] += alpha * p[n];
        double rn = r[n] - alpha * q[n];
        r[n] = rn;
        acc += rn * rn;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) scal[S_ALPHA] = alpha;
    grid_sum2(acc, 0, part, count, scal, S_RR, -1);
}

// z = (double) e0; rz = r.z; rzold = r.zold (for the flexible beta)
__global__ __launch_bounds__(kB) void k_precond_out(int G, const float* __restrict__ e0, const double* __restrict__ r,
                                                    double* __restrict__ z, const double* __restrict__ zold,
                                                    double* part, unsigned* count, double* scal, int first) {
    double a = 0, b = 0;
    for (int n = blockIdx.x * blockDim.x + threadIdx.x; n < G; n += gridDim.x * blockDim.x) {
        double zn = (double)e0[n];
        z[n] = zn;
        a += r[n] * zn;
        if (!first) b += r[n] * zold[n];
    }
    grid_sum2(a, b, part, count, scal, S_RZ, first ? -1 : S_RZOLD);
}

// beta = (rz - rzold) / rzprev; p = z + beta p; zold = z; rzprev = rz
__global__ __launch_bounds__(kB) void k_update_p(int G, double* __restrict__ p, const double* __restrict__ z,
                                                 double* __restrict__ zold, const double* scal_in, int first) {
    const double beta = first ? 0.0 : (scal_in[S_RZ] - scal_in[S_RZOLD]) / scal_in[S_RZPREV];
    for (int n = blockIdx.x * blockDim.x + threadIdx.x; n < G; n += gridDim.x * blockDim.x) {
        double zn = z[n];
        p[n] = first ? zn : zn + beta * p[n];
        zold[n] = zn;
    }
}

__global__ void k_roll_rz(double* scal) { scal[S_RZPREV] = scal[S_RZ]; }

// --------------------------------------------------------------- MG (FP32)
struct Lvl {
    int gx, gy;
    float cx, cy, diag;
};

__global__ void k_d2f(int G, const double* __restrict__ r, float* __restrict__ b) {
    int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n < G) b[n] = (float)r[n];
}

// Weighted Jacobi: eo = ei + w/diag (b - A ei) on unknowns; zero on Dirichlet.
// With zero_in the input iterate is taken as 0.
__global__ __launch_bounds__(kB) void k_mg_smooth(Lvl L, float w, const float* __restrict__ ei,
                                                  float* __restrict__ eo, const float* __restrict__ b,
                                                  const uint8_t* __restrict__ dir, int zero_in) {
    int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= L.gx * L.gy) return;
    if (dir[n]) { eo[n] = 0.f; return; }
    if (zero_in) { eo[n] = w * b[n] / L.diag; return; }
    float e = ei[n];
    float ax = L.diag * e - L.cx * (ei[n - 1] + ei[n + 1]) - L.cy * (ei[n - L.gx] + ei[n + L.gx]);
    eo[n] = e + w * (b[n] - ax) / L.diag;
}

__device__ __forceinline__ float mg_res(const Lvl& L, const float* e, const float* b, const uint8_t* dir, int i, int j) {
    int n = j * L.gx + i;
    if (dir[n]) return 0.f;
    float ax = L.diag * e[n] - L.cx * (e[n - 1] + e[n + 1]) - L.cy * (e[n - L.gx] + e[n + L.gx]);
    return b[n] - ax;
}

// Coarse rhs = full weighting of the fine residual; zero on coarse Dirichlet.
__global__ __launch_bounds__(kB) void k_mg_restrict(Lvl F, Lvl C, const float* __restrict__ e,
                                                    const float* __restrict__ b, const uint8_t* __restrict__ fdir,
                                                    float* __restrict__ bc, const uint8_t* __restrict__ cdir) {
    int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= C.gx * C.gy) return;
    if (cdir[n]) { bc[n] = 0.f; return; }
    int I = n % C.gx, J = n / C.gx;
    int i = 2 * I, j = 2 * J;
    float s = 4.f * mg_res(F, e, b, fdir, i, j)
            + 2.f * (mg_res(F, e, b, fdir, i - 1, j) + mg_res(F, e, b, fdir, i + 1, j)
                   + mg_res(F, e, b, fdir, i, j - 1) + mg_res(F, e, b, fdir, i, j + 1))
            + (mg_res(F, e, b, fdir, i - 1, j - 1) + mg_res(F, e, b, fdir, i + 1, j - 1)
             + mg_res(F, e, b, fdir, i - 1, j + 1) + mg_res(F, e, b, fdir, i + 1, j + 1));
    bc[n] = s * 0.0625f;
}

// e += bilinear interpolation of the coarse correction (unknown nodes only).
__global__ __launch_bounds__(kB) void k_mg_prolong(Lvl F, Lvl C, float* __restrict__ e,
                                                   const uint8_t* __restrict__ fdir, const float* __restrict__ ec) {
    int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= F.gx * F.gy) return;
    if (fdir[n]) return;
    int i = n % F.gx, j = n / F.gx;
    int I = i >> 1, J = j >> 1;
    bool oi = i & 1, oj = j & 1;
    float v;
    const float* c0 = ec + J * C.gx + I;
    if (!oi && !oj) v = c0[0];
    else if (oi && !oj) v = 0.5f * (c0[0] + c0[1]);
    else if (!oi && oj) v = 0.5f * (c0[0] + c0[C.gx]);
    else v = 0.25f * (c0[0] + c0[1] + c0[C.gx] + c0[C.gx + 1]);
    e[n] += v;
}

// Coarsest level: weighted-Jacobi sweeps in one workgroup, in LDS.
static constexpr int kCoarseMax = 8192;
__global__ __launch_bounds__(1024) void k_mg_coarse(Lvl L, float w, int sweeps, float* __restrict__ e,
                                                    const float* __restrict__ b, const uint8_t* __restrict__ dir) {
    __shared__ float ea[kCoarseMax / 2];
    __shared__ float eb[kCoarseMax / 2];
    const int G = L.gx * L.gy;
    for (int n = threadIdx.x; n < G; n += blockDim.x) { ea[n] = 0.f; eb[n] = 0.f; }
    __syncthreads();
    float* src = ea;
    float* dst = eb;
    for (int s = 0; s < sweeps; ++s) {
        for (int n = threadIdx.x; n < G; n += blockDim.x) {
            if (dir[n]) { dst[n] = 0.f; continue; }
            float x = src[n];
            float ax = L.diag * x - L.cx * (src[n - 1] + src[n + 1]) - L.cy * (src[n - L.gx] + src[n + L.gx]);
            dst[n] = x + w * (b[n] - ax) / L.diag;
        }
        __syncthreads();
        float* t = src; src = dst; dst = t;
    }
    for (int n = threadIdx.x; n < G; n += blockDim.x) e[n] = src[n];
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

// ============================================================= host side
static Fine fine_of(const pic_ctx* c) {
    Fine f;
    f.gx = c->g.gx;
    f.gy = c->g.gy;
    f.cx = 1.0 / (c->g.dx * c->g.dx);
    f.cy = 1.0 / (c->g.dy * c->g.dy);
    f.diag = 2 * (f.cx + f.cy);
    return f;
}

static Lvl lvl_of(const MgLevel& m) {
    Lvl L;
    L.gx = m.gx; L.gy = m.gy; L.cx = m.cx; L.cy = m.cy; L.diag = m.diag;
    return L;
}

void poisson_free(pic_ctx* c) {
    for (auto& m : c->mg) {
        if (m.dir) (void)hipFree(m.dir);
        if (m.e) (void)hipFree(m.e);
        if (m.b) (void)hipFree(m.b);
        if (m.t) (void)hipFree(m.t);
    }
    c->mg.clear();
}

int poisson_setup(pic_ctx* c) {
    const int gx = c->g.gx, gy = c->g.gy, G = c->G;
    // Dirichlet mask and values from the geometry rectangles, in file order
    // (later rectangles overwrite earlier ones, as the reference loops do).
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
    // A non-Dirichlet node on the outer frame would make the reference's rows
    // reach outside the grid; require the walls to be covered.
    for (int j = 0; j < gx; ++j)
        if (!dir[j] || !dir[(gy - 1) * gx + j]) {
            set_error("geometry leaves wall node (%d, %d) or (%d, %d) free; the non-periodic solver needs Dirichlet walls",
                      j, 0, j, gy - 1);
            return PIC_ERR_FORMAT;
        }
    for (int k = 0; k < gy; ++k)
        if (!dir[k * gx] || !dir[k * gx + gx - 1]) {
            set_error("geometry leaves wall node (0, %d) or (%d, %d) free; the non-periodic solver needs Dirichlet walls",
                      k, gx - 1, k);
            return PIC_ERR_FORMAT;
        }
    PIC_HIP(hipMemcpy(c->dir, dir.data(), G, hipMemcpyHostToDevice));
    PIC_HIP(hipMemcpy(c->dval, dval.data(), G * sizeof(double), hipMemcpyHostToDevice));

    // Multigrid hierarchy: vertex-centred coarsening while both cell counts are even.
    poisson_free(c);
    int lx = gx, ly = gy;
    double hx = c->g.dx, hy = c->g.dy;
    std::vector<uint8_t> cur = dir;
    while (true) {
        MgLevel m;
        m.gx = lx; m.gy = ly;
        m.cx = (float)(1.0 / (hx * hx));
        m.cy = (float)(1.0 / (hy * hy));
        m.diag = 2.f * (m.cx + m.cy);
        size_t nn = (size_t)lx * ly;
        PIC_HIP(hipMalloc(&m.dir, nn));
        PIC_HIP(hipMalloc(&m.e, nn * sizeof(float)));
        PIC_HIP(hipMalloc(&m.b, nn * sizeof(float)));
        PIC_HIP(hipMalloc(&m.t, nn * sizeof(float)));
        PIC_HIP(hipMemcpy(m.dir, cur.data(), nn, hipMemcpyHostToDevice));
        PIC_HIP(hipMemset(m.e, 0, nn * sizeof(float)));
        PIC_HIP(hipMemset(m.b, 0, nn * sizeof(float)));
        PIC_HIP(hipMemset(m.t, 0, nn * sizeof(float)));
        c->mg.push_back(m);
        int cxn = lx - 1, cyn = ly - 1;
        bool can = (cxn % 2 == 0) && (cyn % 2 == 0) && cxn >= 4 && cyn >= 4 && nn > 25;
        if (!can) break;
        int nlx = cxn / 2 + 1, nly = cyn / 2 + 1;
        std::vector<uint8_t> nxt((size_t)nlx * nly, 0);
        for (int J = 0; J < nly; ++J)
            for (int I = 0; I < nlx; ++I) {
                bool d = cur[(size_t)(2 * J) * lx + 2 * I] != 0;
                if (I == 0 || J == 0 || I == nlx - 1 || J == nly - 1) d = true;
                nxt[(size_t)J * nlx + I] = d;
            }
        cur.swap(nxt);
        lx = nlx; ly = nly;
        hx *= 2; hy *= 2;
    }
    PIC_HIP(hipDeviceSynchronize());
    if ((size_t)c->mg.back().gx * c->mg.back().gy > (size_t)kCoarseMax / 2) {
        set_error("coarsest multigrid level %dx%d is too large for the in-LDS coarse solve; "
                  "use NX, NY with more factors of two",
                  c->mg.back().gx, c->mg.back().gy);
        return PIC_ERR_UNSUPPORTED;
    }
    return PIC_OK;
}

int launch_overwrite_rho(pic_ctx* c) {
    hipLaunchKernelGGL(k_overwrite_rho, dim3(nblk(c->G)), dim3(kB), 0, c->stream, c->G, c->rho, c->dir, c->dval);
    PIC_HIP(hipGetLastError());
    return PIC_OK;
}

int launch_overwrite_phi(pic_ctx* c) {
    hipLaunchKernelGGL(k_overwrite_phi, dim3(nblk(c->G)), dim3(kB), 0, c->stream, c->G, c->phi, c->dir, c->dval);
    PIC_HIP(hipGetLastError());
    return PIC_OK;
}

int launch_efield(pic_ctx* c) {
    hipEvent_t ev;
    PIC_TRY(timer_begin(c, PIC_STAGE_EFIELD, &ev));
    hipLaunchKernelGGL(k_efield, dim3(nblk(c->G)), dim3(kB), 0, c->stream, c->g.gx, c->g.gy, c->g.dx, c->g.dy,
                       c->phi, c->bz, c->ex, c->ey, c->fld);
    PIC_HIP(hipGetLastError());
    PIC_TRY(timer_end(c, PIC_STAGE_EFIELD, ev));
    return PIC_OK;
}

// One V-cycle: c->mg[0].b holds the rhs, result in c->mg[0].e.
static int vcycle(pic_ctx* c) {
    const float w = (float)c->p.mg_omega;
    const int L = (int)c->mg.size();
    hipStream_t st = c->stream;
    for (int l = 0; l < L - 1; ++l) {
        MgLevel& m = c->mg[l];
        Lvl lv = lvl_of(m);
        unsigned nb = nblk((int64_t)m.gx * m.gy);
        // pre-smoothing, starting from zero
        float* a = m.e;
        float* b = m.t;
        for (int s = 0; s < c->p.mg_pre; ++s) {
            hipLaunchKernelGGL(k_mg_smooth, dim3(nb), dim3(kB), 0, st, lv, w, a, b, m.b, m.dir, s == 0 ? 1 : 0);
            std::swap(a, b);
        }
        if (c->p.mg_pre == 0)
            hipLaunchKernelGGL(k_mg_smooth, dim3(nb), dim3(kB), 0, st, lv, 0.f, b, a, m.b, m.dir, 1);
        if (a != m.e) std::swap(m.e, m.t);   // keep the iterate in m.e
        MgLevel& n = c->mg[l + 1];
        hipLaunchKernelGGL(k_mg_restrict, dim3(nblk((int64_t)n.gx * n.gy)), dim3(kB), 0, st, lv, lvl_of(n),
                           m.e, m.b, m.dir, n.b, n.dir);
    }
    {
        MgLevel& m = c->mg[L - 1];
        int sweeps = std::max(50, 4 * std::max(m.gx, m.gy));
        hipLaunchKernelGGL(k_mg_coarse, dim3(1), dim3(1024), 0, st, lvl_of(m), w, sweeps, m.e, m.b, m.dir);
    }
    for (int l = L - 2; l >= 0; --l) {
        MgLevel& m = c->mg[l];
        MgLevel& n = c->mg[l + 1];
        Lvl lv = lvl_of(m);
        unsigned nb = nblk((int64_t)m.gx * m.gy);
        hipLaunchKernelGGL(k_mg_prolong, dim3(nb), dim3(kB), 0, st, lv, lvl_of(n), m.e, m.dir, n.e);
        float* a = m.e;
        float* b = m.t;
        for (int s = 0; s < c->p.mg_post; ++s) {
            hipLaunchKernelGGL(k_mg_smooth, dim3(nb), dim3(kB), 0, st, lv, w, a, b, m.b, m.dir, 0);
            std::swap(a, b);
        }
        if (a != m.e) std::swap(m.e, m.t);
    }
    PIC_HIP(hipGetLastError());
    return PIC_OK;
}

static int precondition(pic_ctx* c, int first) {
    MgLevel& m0 = c->mg[0];
    hipLaunchKernelGGL(k_d2f, dim3(nblk(c->G)), dim3(kB), 0, c->stream, c->G, c->pr, m0.b);
    PIC_TRY(vcycle(c));
    hipLaunchKernelGGL(k_precond_out, dim3(nred(c->G)), dim3(kB), 0, c->stream, c->G, c->mg[0].e, c->pr, c->pz,
                       c->pzold, c->red, c->red_count, c->scal, first);
    PIC_HIP(hipGetLastError());
    return PIC_OK;
}

static int read_scal(pic_ctx* c) {
    PIC_HIP(hipMemcpyAsync(c->scal_host, c->scal, S_N * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    PIC_HIP(hipStreamSynchronize(c->stream));
    return PIC_OK;
}

int poisson_solve(pic_ctx* c, int iteration, pic_solve_stats* st) {
    const Fine f = fine_of(c);
    const int G = c->G;
    hipStream_t s = c->stream;
    hipEvent_t ev;
    PIC_TRY(timer_begin(c, PIC_STAGE_POISSON, &ev));
    hipEvent_t t0, t1;
    PIC_HIP(hipEventCreate(&t0));
    PIC_HIP(hipEventCreate(&t1));
    PIC_HIP(hipEventRecord(t0, s));

    PIC_TRY(launch_overwrite_rho(c));
    // ||b||: residual of the Dirichlet-only field (unknowns at zero) -> pb
    hipLaunchKernelGGL(k_init_guess, dim3(nblk(G)), dim3(kB), 0, s, G, c->pq, (const double*)nullptr, c->dir, c->dval);
    hipLaunchKernelGGL(k_residual, dim3(nred(G)), dim3(kB), 0, s, f, c->rho, c->pq, c->dir, c->pb, c->red,
                       c->red_count, c->scal, (int)S_BB);
    int guess = 0;
    const double* src = nullptr;
    if (c->guess_set) { src = c->guess; guess = 2; }
    else if (iteration > 0 && c->have_prev_phi) { src = c->phi; guess = 1; }
    hipLaunchKernelGGL(k_init_guess, dim3(nblk(G)), dim3(kB), 0, s, G, c->phi, src, c->dir, c->dval);
    hipLaunchKernelGGL(k_residual, dim3(nred(G)), dim3(kB), 0, s, f, c->rho, c->phi, c->dir, c->pr, c->red,
                       c->red_count, c->scal, (int)S_RR);
    PIC_HIP(hipGetLastError());
    PIC_TRY(read_scal(c));
    const double bb = c->scal_host[S_BB];
    const double tol = c->p.poisson_tol;
    double rr = c->scal_host[S_RR];
    const double bnorm = std::sqrt(bb);
    double rel0 = bnorm > 0 ? std::sqrt(rr) / bnorm : 0.0;
    double rel = rel0;
    int it = 0;
    if (bnorm > 0 && rel > tol) {
        PIC_TRY(precondition(c, 1));
        hipLaunchKernelGGL(k_update_p, dim3(nred(G)), dim3(kB), 0, s, G, c->pp, c->pz, c->pzold, c->scal, 1);
        hipLaunchKernelGGL(k_roll_rz, dim3(1), dim3(1), 0, s, c->scal);
        for (it = 1; it <= c->p.poisson_max_iter; ++it) {
            hipLaunchKernelGGL(k_apply, dim3(nred(G)), dim3(kB), 0, s, f, c->pp, c->dir, c->pq, c->red,
                               c->red_count, c->scal);
            hipLaunchKernelGGL(k_update_xr, dim3(nred(G)), dim3(kB), 0, s, G, c->phi, c->pr, c->pp, c->pq, c->red,
                               c->red_count, c->scal);
            PIC_HIP(hipGetLastError());
            PIC_TRY(read_scal(c));
            rr = c->scal_host[S_RR];
            rel = std::sqrt(rr) / bnorm;
            if (!(rel == rel)) {
                set_error("Poisson PCG diverged (NaN residual) at iteration %d", it);
                return PIC_ERR_SOLVER;
            }
            if (rel <= tol) break;
            PIC_TRY(precondition(c, 0));
            hipLaunchKernelGGL(k_update_p, dim3(nred(G)), dim3(kB), 0, s, G, c->pp, c->pz, c->pzold, c->scal, 0);
            hipLaunchKernelGGL(k_roll_rz, dim3(1), dim3(1), 0, s, c->scal);
        }
        if (it > c->p.poisson_max_iter) {
            set_error("Poisson PCG did not reach relative residual %.3e in %d iterations (%.3e)", tol,
                      c->p.poisson_max_iter, rel);
            return PIC_ERR_SOLVER;
        }
    }
    PIC_TRY(launch_overwrite_phi(c));
    PIC_HIP(hipEventRecord(t1, s));
    PIC_HIP(hipEventSynchronize(t1));
    float ms = 0;
    PIC_HIP(hipEventElapsedTime(&ms, t0, t1));
    (void)hipEventDestroy(t0);
    (void)hipEventDestroy(t1);
    c->guess_set = 0;
    c->have_prev_phi = 1;
    PIC_TRY(timer_end(c, PIC_STAGE_POISSON, ev));
    if (st) {
        st->iterations = it;
        st->initial_relres = rel0;
        st->final_relres = rel;
        st->ms = ms;
        st->guess = guess;
    }
    return PIC_OK;
}

} // namespace pic
