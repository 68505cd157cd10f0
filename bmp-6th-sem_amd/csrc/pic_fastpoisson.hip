// pic_fastpoisson.hip — direct solve of the reference's Poisson system
// (poissonSolver.cpp:183-241,459-544: the 5-point rows of the unknown nodes,
// Dirichlet rows phi = V) for walls held at fixed potentials plus internal
// Dirichlet nodes (electrodes).  The reference factors the system once with
// PARDISO; here the same system is solved exactly by
//
//   * the discrete sine transform: on the interior of the rectangle (m x n
//     nodes, walls eliminated into the right-hand side) the operator
//     T = cx L_m (x) I + cy I (x) L_n, L = tridiag(-1, 2, -1), is diagonal in
//     the orthonormal DST-I basis S[j][k] = sqrt(2/(N+1)) sin(pi (j+1)(k+1)/(N+1)):
//     T^-1 f = S_n ((S_n F S_m) ./ Lambda) S_m, Lambda[k][l] =
//     cy (2 - 2 cos(pi (k+1)/(n+1))) + cx (2 - 2 cos(pi (l+1)/(m+1)));
//     the products with S run as FFTs (one LDS-resident complex FFT of length
//     2(N+1) per pair of rows) when N + 1 is a power of two, else as FP64
//     GEMMs on the matrix cores (k_dgemm);
//   * periodic y (the reference's generateCoeffsPeriodic, :82-180): the
//     system's rows are grid rows 0 .. gy-2 with the y neighbours wrapped, so
//     the y factor is the periodic second difference, diagonal in the real
//     Fourier basis Y (cos / sin columns, eigenvalues 2 - 2 cos(2 pi p / N));
//     T^-1 f = Y ((Y^T F S_m) ./ Lambda) S_m by GEMMs, every x eigenvalue is
//     positive so Lambda has no zero;
//   * a capacitance correction for the internal Dirichlet nodes: with Q the
//     internal Dirichlet nodes next to an unknown (P of them), sources c on Q
//     are chosen so that u = T^-1 (f + B c) equals V on Q:
//     c = C^-1 (V_Q - (T^-1 f)_Q), C = (T^-1)_QQ (P x P, symmetric positive
//     definite, inverted once at set-up).  The unknown rows only see Q among
//     the internal Dirichlet nodes, so u solves the reference's system there.
//
// Per solve: four transform passes (FFT passes, or GEMMs of n x n x m /
// m x m x n), the correction (a few line vectors folded into the third FFT
// pass, or an n x nq x m product), a few small kernels.  The PCG of
// pic_poisson.hip then starts from this solution and only checks (or, if
// rounding left the residual above the tolerance, corrects) it.
#include "pic_internal.h"

#include <algorithm>
#include <array>
#include <cmath>
#include <vector>

namespace pic {

static constexpr int kFdMaxP = 1024;      // internal Dirichlet nodes next to unknowns
static constexpr int kFdMaxN = 2047;      // interior nodes per direction
static constexpr int kFdB = 256;
static constexpr int kLines = 16;         // covering lines handled by the line kernels (more: GEMMs)
static constexpr int kSeg = 64;           // row segments of the line products

struct FdRun;

struct FastPoisson {
    int m = 0, n = 0, P = 0, nr = 0, nc = 0;
    int per = 0;                             // periodic y: Sn is the real Fourier basis, rows 0 .. n-1 of the grid
    double *Sm = nullptr, *Sn = nullptr;     // m x m, n x n; row = node, column = mode
    double* Dinv = nullptr;                  // n x m
    double *F = nullptr, *W = nullptr, *U = nullptr, *G = nullptr;   // n x m work
    double* SnR = nullptr;                   // nr x n: rows of S_n at the covering rows
    double* V1 = nullptr;                    // nr x m: S_n[rows] Uhat
    double* SmC = nullptr;                   // nc x m: rows of S_m at the covering columns
    double* V2 = nullptr;                    // n x nc (column-major): Uhat S_m[:, cols]
    double* part = nullptr;                  // kSeg x nr x m partial row-line products
    double* part2 = nullptr;                 // ceil(m/64) x nc x n partial column-line products
    int nq = 0;                              // distinct rows of Q
    double* SnQ = nullptr;                   // n x nq: S_n[k][row a]
    double* Z = nullptr;                     // nq x m: sum over q on row a of c_q S_m[col(q)][l]
    double* cq = nullptr;                    // P: the sources c = C^-1 (V_Q - u0_Q), in row order
    int *rq_ptr = nullptr, *rq_idx = nullptr;   // the q of distinct row a: rq_idx[rq_ptr[a] .. rq_ptr[a+1])
    int* rq_col = nullptr;                   // P: column of rq_idx[t]
    double* Cinv = nullptr;                  // P x P
    double* VQ = nullptr;                    // P: potentials on Q
    double* dq = nullptr;                    // P: V_Q - u0_Q
    int *qri = nullptr, *qcol = nullptr;     // P: covering line of q (row a >= 0, column j as -1 - j); column of q
    int* qrow = nullptr;                     // P: row of q
    // sine transforms by FFT when m + 1 and n + 1 are powers of two (else GEMMs)
    int fft = 0, lm = 0, ln = 0;             // log2(2 (m + 1)), log2(2 (n + 1))
    double2 *twm = nullptr, *twn = nullptr;  // FFT twiddles of both lengths, in the order the stages read them
    double* Fb = nullptr;                    // n x m: wall values moved to the right-hand side
    FdRun* self = nullptr;                   // device: this run's descriptor (a table of one)
    FdRun* hself = nullptr;                  // host copy
    // line form of the correction (FFT transforms, one-pass covering lines):
    // the sources of covering line v are lt[lptr[v] .. lptr[v+1]) (indices
    // into cq), at row-of-S index lidx (the column for a row line, the row
    // for a column line); LV holds one vector per line (lr x m, then lc x n)
    int line_form = 0;
    int *lptr = nullptr, *lt = nullptr, *lidx = nullptr;
    double* LV = nullptr;
};

static void* dmalloc(size_t bytes, bool& ok) {
    void* p = nullptr;
    if (ok && hipMalloc(&p, std::max<size_t>(bytes, 8)) != hipSuccess) {
        ok = false;
        p = nullptr;
    }
    return p;
}

void fast_free(pic_ctx* c) {
    FastPoisson* f = c->fast;
    if (!f) return;
    void* ptrs[] = {f->Sm, f->Sn, f->Dinv, f->F, f->W, f->U, f->G, f->SnR, f->V1, f->SmC, f->V2, f->SnQ, f->cq,
                    f->rq_ptr, f->rq_idx, f->rq_col, f->Z, f->Cinv, f->VQ, f->dq, f->qri, f->qcol, f->qrow, f->part, f->part2, f->twm, f->twn, f->Fb, f->self,
                    f->lptr, f->lt, f->lidx, f->LV};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    delete f->hself;
    delete f;
    c->fast = nullptr;
}

// ------------------------------------------------------------------ kernels
// F (n x m, system row r at grid node (r + per ? 0 : 1, c + 1)): rhs on
// unknowns plus the wall values moved to the right-hand side (the x walls, and
// the y walls unless y is periodic); 0 on internal Dirichlet nodes.
__global__ __launch_bounds__(kFdB) void k_fd_rhs(int gx, int n, int per, double cx, double cy,
                                                 const double* __restrict__ rhs, const uint8_t* __restrict__ dir,
                                                 const double* __restrict__ dval, double* __restrict__ F) {
    const int m = gx - 2;
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= m * n) return;
    const int r = k / m, q = k - r * m;
    const int node = (r + (per ? 0 : 1)) * gx + q + 1;
    double v = 0.0;
    if (dir[node] == 0) {
        v = rhs[node];
        if (q == 0) v += cx * dval[node - 1];
        if (q == m - 1) v += cx * dval[node + 1];
        if (!per && r == 0) v += cy * dval[node - gx];
        if (!per && r == n - 1) v += cy * dval[node + gx];
    }
    F[k] = v;
}

// X += G .* Dinv
__global__ __launch_bounds__(kFdB) void k_fd_addscaled(int N, double* __restrict__ X, const double* __restrict__ G,
                                                       const double* __restrict__ Dinv) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < N) X[k] += G[k] * Dinv[k];
}

// ------------------------------------------------------------ FP64 GEMM
// C (M x N, row-major, ldc) = op(A) op(B) on the matrix cores
// (v_mfma_f64_16x16x4_f64), with op(A)[i][k] = TA ? A[k lda + i] : A[i lda + k]
// and op(B)[k][j] = TB ? B[j ldb + k] : B[k ldb + j].  EP: 0 store, 1 store
// times S[i ldc + j] (elementwise), 2 add to C.  A workgroup computes a 32 x 32
// tile, one 16 x 16 quadrant per wave, over k chunks of 32 staged in LDS (the
// next chunk is loaded into registers while the current one is multiplied);
// every output is summed in the same order, so results are reproducible.
// Lane l of an MFMA holds A[l % 16][l / 16], B[l / 16][l % 16] and receives
// D[l / 16 + 4 e][l % 16], e < 4 (rows interleaved over the lane groups).
typedef double d4v __attribute__((ext_vector_type(4)));
static constexpr int kGT = 32, kGK = 64;
static constexpr int kGE = kGT * kGK / 256;   // elements of each operand chunk per thread

template <bool TA, bool TB, int EP>
__global__ __launch_bounds__(256) void k_dgemm(int M, int N, int K, const double* __restrict__ A, int lda,
                                               const double* __restrict__ B, int ldb, double* __restrict__ C,
                                               int ldc, const double* __restrict__ S) {
    __shared__ double As[kGK][kGT + 1];     // As[k][i]
    __shared__ double Bs[kGK][kGT + 1];     // Bs[k][j]
    const int i0 = blockIdx.y * kGT, j0 = blockIdx.x * kGT;
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const int wi = (w >> 1) * 16, wj = (w & 1) * 16;
    double ra[kGE], rb[kGE];
    // element e of thread t: (row, k) of the A chunk and (k, column) of the B
    // chunk, the fast index along the stored rows so that loads coalesce
    auto a_pos = [&](int e, int& ii, int& kk) {
        const int idx = t + 256 * e;
        if (TA) { kk = idx / kGT; ii = idx % kGT; }
        else { ii = idx / kGK; kk = idx % kGK; }
    };
    auto b_pos = [&](int e, int& kk, int& jj) {
        const int idx = t + 256 * e;
        if (TB) { jj = idx / kGK; kk = idx % kGK; }
        else { kk = idx / kGT; jj = idx % kGT; }
    };
    auto load = [&](int k0) {
#pragma unroll
        for (int e = 0; e < kGE; ++e) {
            int ii, kk, jj, kb;
            a_pos(e, ii, kk);
            const int gi = i0 + ii, gk = k0 + kk;
            ra[e] = (gi < M && gk < K) ? (TA ? A[(size_t)gk * lda + gi] : A[(size_t)gi * lda + gk]) : 0.0;
            b_pos(e, kb, jj);
            const int gj = j0 + jj, gk2 = k0 + kb;
            rb[e] = (gj < N && gk2 < K) ? (TB ? B[(size_t)gj * ldb + gk2] : B[(size_t)gk2 * ldb + gj]) : 0.0;
        }
    };
    // two accumulators (even and odd k steps) so that consecutive MFMAs do not
    // wait for each other; added at the end
    d4v acc = {0.0, 0.0, 0.0, 0.0}, acc2 = {0.0, 0.0, 0.0, 0.0};
    load(0);
    for (int k0 = 0; k0 < K; k0 += kGK) {
        __syncthreads();                    // the previous chunk is consumed
#pragma unroll
        for (int e = 0; e < kGE; ++e) {
            int ii, kk, jj, kb;
            a_pos(e, ii, kk);
            As[kk][ii] = ra[e];
            b_pos(e, kb, jj);
            Bs[kb][jj] = rb[e];
        }
        __syncthreads();
        if (k0 + kGK < K) load(k0 + kGK);
#pragma unroll
        for (int kk = 0; kk < kGK; kk += 8) {
            const double a = As[kk + (lane >> 4)][wi + (lane & 15)];
            const double b = Bs[kk + (lane >> 4)][wj + (lane & 15)];
            const double a2 = As[kk + 4 + (lane >> 4)][wi + (lane & 15)];
            const double b2 = Bs[kk + 4 + (lane >> 4)][wj + (lane & 15)];
            acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc, 0, 0, 0);
            acc2 = __builtin_amdgcn_mfma_f64_16x16x4f64(a2, b2, acc2, 0, 0, 0);
        }
    }
    acc += acc2;
    const int j = j0 + wj + (lane & 15);
    if (j >= N) return;
#pragma unroll
    for (int e = 0; e < 4; ++e) {
        const int i = i0 + wi + (lane >> 4) + 4 * e;
        if (i >= M) continue;
        const size_t o = (size_t)i * ldc + j;
        if (EP == 0) C[o] = acc[e];
        else if (EP == 1) C[o] = acc[e] * S[o];
        else C[o] += acc[e];
    }
}

template <bool TA, bool TB, int EP>
static void dgemm(hipStream_t s, int M, int N, int K, const double* A, int lda, const double* B, int ldb, double* C,
                  int ldc, const double* S = nullptr) {
    hipLaunchKernelGGL((k_dgemm<TA, TB, EP>), dim3((unsigned)((N + kGT - 1) / kGT), (unsigned)((M + kGT - 1) / kGT)),
                       dim3(256), 0, s, M, N, K, A, lda, B, ldb, C, ldc, S);
}

__device__ __forceinline__ double wave_sum(double v) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    return v;
}

// ------------------------------------------------------ per-run descriptor
// Everything the correction kernels and the FFT passes read of one run's
// direct solve.  Kernels take a table of these and pick their run by the last
// grid dimension, so one launch serves a single run (a table of one) or a
// batch of runs with the same grid (an ensemble).
struct FdRun {
    int gx, m, n, P, nq;
    int lr, lc;                              // covering rows / columns taken by k_fd_lines (0: GEMM or none)
    double sgn;                              // +1: rho holds overwriteRho's result; -1: the raw deposit
    const double* rho;
    const uint8_t* dir;
    double* phi;                             // output: the interior nodes
    const double *Fb, *Dinv;
    double *W, *U, *G;
    const double *SnR, *SmC;
    double *V1, *V2, *part, *part2;
    const double *Sn, *Sm;
    const int *qri, *qrow, *qcol;
    const double* VQ;
    double* dq;
    const double* Cinv;
    const int *rq_idx, *rq_col, *rq_ptr;
    double *cq, *Z;
    const double* SnQ;
    const double2 *twm, *twn;
    int cmode;                               // correction: 0 none, 1 the spectrum G, 2 line form (LV)
    int rl2;                                 // the row lines (1 .. kPassLines of them) come from the second FFT pass
    const int *lptr, *lt, *lidx;
    double* LV;
};

// d_q = V_q - u0(r, c) at q = (r, c), u0 = S_n Uhat S_m.  A q covered by row a
// of the line cover: u0 = V1[a] . S_m[c]; by column j: u0 = S_n[r] . V2[:, j].
// One wave per q.
template <bool B>
__global__ __launch_bounds__(kFdB) void k_fd_u0q(const FdRun one, const FdRun* __restrict__ runs) {
    __shared__ double part[4];
    const FdRun& f = B ? runs[blockIdx.z] : one;
    const int q = blockIdx.x, lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    if (q >= f.P) return;                    // uniform over the workgroup
    const int line = f.qri[q];
    const double *a, *b;
    int len;
    if (line >= 0 && f.rl2) {
        // the second FFT pass left four wave partials per value
        const double4* __restrict__ a4 = (const double4*)(f.V1 + (size_t)line * f.m * 4);
        b = f.Sm + (size_t)f.qcol[q] * f.m;
        double s = 0.0;
#pragma unroll 4
        for (int l = threadIdx.x; l < f.m; l += kFdB) {
            const double4 p = a4[l];
            s += ((p.x + p.y) + (p.z + p.w)) * b[l];
        }
        s = wave_sum(s);
        if (lane == 0) part[wv] = s;
        __syncthreads();
        if (threadIdx.x == 0) f.dq[q] = f.VQ[q] - ((part[0] + part[1]) + (part[2] + part[3]));
        return;
    }
    if (line >= 0) {
        a = f.V1 + (size_t)line * f.m;
        b = f.Sm + (size_t)f.qcol[q] * f.m;
        len = f.m;
    } else {
        a = f.V2 + (size_t)(-1 - line) * f.n;
        b = f.Sn + (size_t)f.qrow[q] * f.n;
        len = f.n;
    }
    double s = 0.0;
#pragma unroll 4
    for (int l = threadIdx.x; l < len; l += kFdB) s += a[l] * b[l];
    s = wave_sum(s);
    if (lane == 0) part[wv] = s;
    __syncthreads();
    if (threadIdx.x == 0) f.dq[q] = f.VQ[q] - ((part[0] + part[1]) + (part[2] + part[3]));
}

// Covering lines in one pass over Uhat (lr, lc <= kLines; either may be 0):
//   rows:    V1[a][l] = sum_k S_n[row a][k] Uhat[k][l]
//   columns: V2[j][k] = sum_l Uhat[k][l] S_m[col j][l]
// Workgroup (64 columns, one of kSeg row segments), four waves over
// interleaved rows.  Row lines: per-segment partials of each column (part);
// column lines: per-column-block partials of each row, one wave sum per row
// (part2).  k_fd_linesum adds both in a fixed order.
template <bool B>
__global__ __launch_bounds__(kFdB) void k_fd_lines(const FdRun one, const FdRun* __restrict__ runs) {
    __shared__ double sh[4][kLines][64];
    const FdRun& f = B ? runs[blockIdx.z] : one;
    const int nr = f.lr, nc = f.lc, n = f.n, m = f.m;
    const int lx = threadIdx.x & 63, ph = threadIdx.x >> 6;
    const int l = blockIdx.x * 64 + lx;
    const int per = (n + kSeg - 1) / kSeg;
    const int k0 = blockIdx.y * per, k1 = min(n, k0 + per);
    double acc[kLines], smc[kLines];
#pragma unroll
    for (int a = 0; a < kLines; ++a) {
        acc[a] = 0.0;
        smc[a] = (a < nc && l < m) ? f.SmC[(size_t)a * m + l] : 0.0;
    }
    for (int k = k0 + ph; k < k1; k += 4) {
        const double u = l < m ? f.U[(size_t)k * m + l] : 0.0;
#pragma unroll
        for (int a = 0; a < kLines; ++a)
            if (a < nr) acc[a] += f.SnR[(size_t)a * n + k] * u;
#pragma unroll
        for (int j = 0; j < kLines; ++j)
            if (j < nc) {
                const double v = wave_sum(u * smc[j]);
                if (lx == 0) f.part2[((size_t)blockIdx.x * nc + j) * n + k] = v;
            }
    }
    if (nr == 0) return;
#pragma unroll
    for (int a = 0; a < kLines; ++a) sh[ph][a][lx] = acc[a];
    __syncthreads();
    if (ph == 0 && l < m)
        for (int a = 0; a < nr; ++a)
            f.part[((size_t)blockIdx.y * nr + a) * m + l] = (sh[0][a][lx] + sh[1][a][lx]) + (sh[2][a][lx] + sh[3][a][lx]);
}

// V1 (lr x m) from the kSeg segment partials; V2 (lc x n) from the
// ceil(m/64) column-block partials
template <bool B>
__global__ __launch_bounds__(kFdB) void k_fd_linesum(const FdRun one, const FdRun* __restrict__ runs) {
    const FdRun& f = B ? runs[blockIdx.z] : one;
    const int nr = f.lr, nc = f.lc, n = f.n, m = f.m, ncb = (m + 63) / 64;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < nr * m) {
        double s = 0.0;
        for (int seg = 0; seg < kSeg; ++seg) s += f.part[(size_t)seg * nr * m + t];
        f.V1[t] = s;
    } else if (t < nr * m + nc * n) {
        const int u = t - nr * m;
        const int j = u / n, k = u - j * n;
        double s = 0.0;
        for (int cb = 0; cb < ncb; ++cb) s += f.part2[((size_t)cb * nc + j) * n + k];
        f.V2[(size_t)j * n + k] = s;
    }
}

// The sources in row order: c_t = (C^-1 d)_q for q = rq_idx[t]; one wave per t
template <bool B>
__global__ __launch_bounds__(kFdB) void k_fd_sources(const FdRun one, const FdRun* __restrict__ runs) {
    __shared__ double part[4];
    const FdRun& f = B ? runs[blockIdx.z] : one;
    const int P = f.P;
    const int t = blockIdx.x, lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    if (t >= P) return;                      // uniform over the workgroup
    const int q = f.rq_idx[t];
    double s = 0.0;
#pragma unroll 4
    for (int p = threadIdx.x; p < P; p += kFdB) s += f.Cinv[(size_t)q * P + p] * f.dq[p];
    s = wave_sum(s);
    if (lane == 0) part[wv] = s;
    __syncthreads();
    if (threadIdx.x == 0) f.cq[t] = (part[0] + part[1]) + (part[2] + part[3]);
}

// Column lines alone (lc <= kLines), one workgroup per row k of Uhat (its
// four waves split the row; every load of a thread issues at once):
//   V2[j][k] = sum_l Uhat[k][l] S_m[col j][l]
template <bool B>
__global__ __launch_bounds__(kFdB) void k_fd_collines(const FdRun one, const FdRun* __restrict__ runs) {
    __shared__ double part[4][kLines];
    const FdRun& f = B ? runs[blockIdx.z] : one;
    const int k = blockIdx.x, lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int nc = f.lc, m = f.m;
    if (k >= f.n) return;                    // uniform over the workgroup
    const double* __restrict__ u = f.U + (size_t)k * m;
    constexpr int kIt = (kFdMaxN + kFdB - 1) / kFdB;
    double x[kIt];
#pragma unroll
    for (int i = 0; i < kIt; ++i) {
        const int l = threadIdx.x + kFdB * i;
        x[i] = l < m ? u[l] : 0.0;
    }
#pragma unroll
    for (int jj = 0; jj < kLines; ++jj) {
        if (jj >= nc) break;
        const double* __restrict__ sm = f.SmC + (size_t)jj * m;
        double a = 0.0;
#pragma unroll
        for (int i = 0; i < kIt; ++i) {
            const int l = threadIdx.x + kFdB * i;
            if (l < m) a += x[i] * sm[l];
        }
        a = wave_sum(a);
        if (lane == 0) part[wv][jj] = a;
    }
    __syncthreads();
    if (threadIdx.x < nc)
        f.V2[(size_t)threadIdx.x * f.n + k] =
            (part[0][threadIdx.x] + part[1][threadIdx.x]) + (part[2][threadIdx.x] + part[3][threadIdx.x]);
}

// Z[a][l] = sum over the sources t of distinct row a of c_t S_m[col t][l].
// Workgroup (64 columns, row a): the row's (c_t, col t) staged in LDS, wave w
// sums t = w, w + 4, ..., the four wave sums added in a fixed order.
template <bool B>
__global__ __launch_bounds__(kFdB) void k_fd_zrows(const FdRun one, const FdRun* __restrict__ runs) {
    __shared__ double cs[kFdMaxP];
    __shared__ int cl[kFdMaxP];
    __shared__ double part[4][64];
    const FdRun& f = B ? runs[blockIdx.z] : one;
    const int m = f.m;
    const int a = blockIdx.y;
    const int t0 = f.rq_ptr[a], nt = f.rq_ptr[a + 1] - t0;
    for (int t = threadIdx.x; t < nt; t += kFdB) {
        cs[t] = f.cq[t0 + t];
        cl[t] = f.rq_col[t0 + t];
    }
    __syncthreads();
    const int lx = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int l = min((int)blockIdx.x * 64 + lx, m - 1);
    double sum = 0.0;
#pragma unroll 8
    for (int t = w; t < nt; t += 4) sum += cs[t] * f.Sm[(size_t)cl[t] * m + l];
    part[w][lx] = sum;
    __syncthreads();
    if (w == 0 && (int)blockIdx.x * 64 + lx < m)
        f.Z[(size_t)a * m + l] = (part[0][lx] + part[1][lx]) + (part[2][lx] + part[3][lx]);
}

// The correction's spectrum G = SnQ Z (n x nq times nq x m, FP64): 64 x 64
// output tiles, k in chunks of 16 staged in LDS, 4 x 4 outputs per thread
// (rows ty + 16 i, columns tx + 16 j).
static constexpr int kCT = 64, kCK = 16;
template <bool B>
__global__ __launch_bounds__(kFdB) void k_fd_corr(const FdRun one, const FdRun* __restrict__ runs) {
    __shared__ double As[kCK][kCT + 1];
    __shared__ double Bs[kCK][kCT];
    const FdRun& f = B ? runs[blockIdx.z] : one;
    const int n = f.n, m = f.m, K = f.nq;
    const int row0 = blockIdx.y * kCT, col0 = blockIdx.x * kCT;
    const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
    double acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = 0.0;
    for (int k0 = 0; k0 < K; k0 += kCK) {
        // A tile: rows row0 .. +63, k k0 .. +15 (transposed into As[k][row]);
        // B tile: k k0 .. +15, columns col0 .. +63
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const int idx = threadIdx.x + kFdB * e;
            const int ar = idx >> 4, ak = idx & 15;
            const int gr = row0 + ar, gk = k0 + ak;
            As[ak][ar] = (gr < n && gk < K) ? f.SnQ[(size_t)gr * K + gk] : 0.0;
            const int bk = idx >> 6, bc = idx & 63;
            const int gk2 = k0 + bk, gc = col0 + bc;
            Bs[bk][bc] = (gk2 < K && gc < m) ? f.Z[(size_t)gk2 * m + gc] : 0.0;
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < kCK; ++kk) {
            double av[4], bv[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) av[i] = As[kk][ty + 16 * i];
#pragma unroll
            for (int j = 0; j < 4; ++j) bv[j] = Bs[kk][tx + 16 * j];
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j) acc[i][j] += av[i] * bv[j];
        }
        __syncthreads();
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int r = row0 + ty + 16 * i;
        if (r >= n) continue;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int cidx = col0 + tx + 16 * j;
            if (cidx < m) f.G[(size_t)r * m + cidx] = acc[i][j];
        }
    }
}

// Line form of the correction's spectrum.  Every q of Q lies on exactly one
// covering line, so G = sum over row lines a of S_n[row a] (x) Zr_a plus sum
// over column lines j of Vc_j (x) S_m[col j], with
//   Zr_a[l] = sum_{q on row line a} c_q S_m[col q][l]   (length m),
//   Vc_j[k] = sum_{q on column line j} c_q S_n[row q][k] (length n):
// a handful of rank-1 terms (four for a rectangle) that the third transform
// pass adds while it loads the spectrum, instead of a dense n x nq x m
// product.  Workgroup (64 outputs, line): the line's sources and their S
// rows are staged in LDS, wave w sums sources w, w + 4, ..., and the four
// wave sums are added in a fixed order.
template <bool B>
__global__ __launch_bounds__(kFdB) void k_fd_linevec(const FdRun one, const FdRun* __restrict__ runs) {
    __shared__ double cs[kFdMaxP];
    __shared__ int li[kFdMaxP];
    __shared__ double part[4][64];
    const FdRun& f = B ? runs[blockIdx.z] : one;
    const int line = blockIdx.y;
    const bool row = line < f.lr;
    const int len = row ? f.m : f.n;
    if (line >= f.lr + f.lc || (int)blockIdx.x * 64 >= len) return;   // uniform over the workgroup
    const double* S = row ? f.Sm : f.Sn;
    const int t0 = f.lptr[line], nt = f.lptr[line + 1] - t0;
    for (int u = threadIdx.x; u < nt; u += kFdB) {
        cs[u] = f.cq[f.lt[t0 + u]];
        li[u] = f.lidx[t0 + u];
    }
    __syncthreads();
    const int lx = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int o = blockIdx.x * 64 + lx, oc = min(o, len - 1);
    double sum = 0.0;
#pragma unroll 8
    for (int u = w; u < nt; u += 4) sum += cs[u] * S[(size_t)li[u] * len + oc];
    part[w][lx] = sum;
    __syncthreads();
    if (w == 0 && o < len) {
        const size_t base = row ? (size_t)line * f.m : (size_t)f.lr * f.m + (size_t)(line - f.lr) * f.n;
        f.LV[base + o] = (part[0][lx] + part[1][lx]) + (part[2][lx] + part[3][lx]);
    }
}

// full-grid solution: the system's unknowns (rows j0 .. j0 + n - 1) from U,
// every other node its Dirichlet value (the periodic image row is copied from
// row 0 afterwards, with the Dirichlet values)
__global__ __launch_bounds__(kFdB) void k_fd_out(int gx, int gy, int j0, int n, const double* __restrict__ U,
                                                 const uint8_t* __restrict__ dir, const double* __restrict__ dval,
                                                 double* __restrict__ out) {
    const int node = blockIdx.x * blockDim.x + threadIdx.x;
    if (node >= gx * gy) return;
    const int j = node / gx, i = node - j * gx;
    double v = dval[node];
    if (dir[node] == 0 && i > 0 && i < gx - 1 && j >= j0 && j < j0 + n) v = U[(size_t)(j - j0) * (gx - 2) + (i - 1)];
    out[node] = v;
}

// Rows [k0, k0 + kc) of the factor A with C = A A^T:
// A[q][(k - k0) m + l] = S_n[row q][k] S_m[col q][l] sqrt(Dinv[k][l]).
__global__ __launch_bounds__(kFdB) void k_fd_amat(int P, int n, int m, int k0, int kc, const double* __restrict__ Sn,
                                                  const double* __restrict__ Sm, const double* __restrict__ Dinv,
                                                  const int* __restrict__ qrow, const int* __restrict__ qcol,
                                                  double* __restrict__ A) {
    const size_t K = (size_t)kc * m;
    const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (size_t)P * K) return;
    const int q = (int)(idx / K);
    const size_t rem = idx - (size_t)q * K;
    const int kk = (int)(rem / m), l = (int)(rem - (size_t)kk * m);
    const int k = k0 + kk;
    double v = 0.0;
    if (k < n) v = Sn[(size_t)qrow[q] * n + k] * Sm[(size_t)qcol[q] * m + l] * sqrt(Dinv[(size_t)k * m + l]);
    A[idx] = v;
}

// ------------------------------------------------------- sine transform by FFT
// DST-I of length N = M/2 - 1 (M = 2^L) through a complex FFT of length M:
// the odd extension z = (0, x_0 .. x_{N-1}, 0, -x_{N-1} .. -x_0) has
// Z[k+1] = -2i sum_j x_j sin(pi (j+1)(k+1)/(N+1)).  Two real rows a, b go into
// one FFT as z_a + i z_b; both transforms are odd and real, so
// (S a)_k = -Im Z[k+1] s/2 and (S b)_k = Re Z[k+1] s/2, s = sqrt(2/(N+1)).
// The FFT is a Stockham radix-4 sequence (a final radix-2 stage for odd L) in
// LDS: every thread reads its butterflies' inputs, the workgroup synchronises,
// every thread writes its outputs.  Each pass reads rows and writes the result
// transposed, so two passes are the 2-D transform back in the input layout;
// the loaders and storers fold the right-hand side, the spectral scaling, the
// correction and the full-grid output into the passes.  Every global load of
// a pass is issued before the first wait (rows, twiddles, then the gathers of
// the storer), so a workgroup waits for memory twice.
static constexpr int kDstT = 256;          // threads per workgroup
static constexpr int kDstPts = 2048;       // complex points per workgroup (at least)

// LDS slot of point x: within each row of 16 slots (one pass over the 64
// banks by 16 lanes of 16 bytes) the low four bits are XORed with 5 (row mod 4).
// Contiguous runs of 16 points stay in one row; the radix-4 stores of the
// first two stages (lane stride 4 slots, and runs of 4 at a row stride of 16)
// land on 16 distinct slots per 16 lanes instead of 4.
__device__ __forceinline__ int zsw(int x) { return x ^ (((x >> 4) & 3) * 5); }

__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// plain rows: X[row][col], row length ld
struct DstLoadRows {
    const double* X;
    int ld;
    __device__ double operator()(int row, int col) const { return X[(size_t)row * ld + col]; }
};

// rows of U + G .* Dinv (the corrected spectrum; U alone without sources);
// G the stored spectrum (mode 1) or its line form (mode 2: the rank-1 terms
// of the covering lines, k_fd_linevec)
struct DstLoadCorrected {
    const double *U, *G, *Dinv;
    int ld, mode, n, lr, lc;
    const double *SnR, *SmC, *LV;
    __device__ double operator()(int row, int col) const {
        const size_t i = (size_t)row * ld + col;
        if (mode == 0) return U[i];
        if (mode == 1) return U[i] + G[i] * Dinv[i];
        // the first two lines of each kind unrolled (a rectangle's cover has
        // two rows and two columns), so their loads issue together
        double g = 0.0;
        const double* V = LV + (size_t)lr * ld;
#pragma unroll
        for (int a = 0; a < 2; ++a)
            if (a < lr) g += SnR[(size_t)a * n + row] * LV[(size_t)a * ld + col];
#pragma unroll
        for (int j = 0; j < 2; ++j)
            if (j < lc) g += V[(size_t)j * n + row] * SmC[(size_t)j * ld + col];
        for (int a = 2; a < lr; ++a) g += SnR[(size_t)a * n + row] * LV[(size_t)a * ld + col];
        for (int j = 2; j < lc; ++j) g += V[(size_t)j * n + row] * SmC[(size_t)j * ld + col];
        return U[i] + g * Dinv[i];
    }
};

// The line form of the correction on the GEMM path: Uhat += G ./ Lambda with
// G the covering lines' rank-1 terms (the loader of the third FFT pass,
// applied in place)
template <bool B>
__global__ __launch_bounds__(kFdB) void k_fd_addlines(const FdRun one, const FdRun* __restrict__ runs) {
    const FdRun& f = B ? runs[blockIdx.z] : one;
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= f.n * f.m) return;
    const int row = k / f.m, col = k - row * f.m;
    const DstLoadCorrected ld{f.U, f.G, f.Dinv, f.m, 2, f.n, f.lr, f.lc, f.SnR, f.SmC, f.LV};
    f.U[k] = ld(row, col);
}

// the right-hand side F at interior node (r + 1, c + 1): rhs on unknowns plus
// the wall terms Fb (fixed by the geometry), 0 on internal Dirichlet nodes.
// rhs = sgn * rho: rho after overwriteRho (sgn 1) or the raw deposit (sgn -1;
// overwriteRho negates the unknown nodes).
struct DstLoadRhs {
    int gx, m;
    double sgn;
    const double* rho;
    const uint8_t* dir;
    const double* Fb;
    __device__ double operator()(int r, int q) const {
        const int node = (r + 1) * gx + q + 1;
        const double v = sgn * rho[node] + Fb[(size_t)r * m + q];
        return dir[node] ? 0.0 : v;
    }
};

// transposed: Y[t][row], row length R
struct DstStoreT {
    double* Y;
    int R;
    struct Aux {};
    __device__ Aux fetch(int, int) const { return {}; }
    __device__ double put(int t, int row, double v, Aux) const {
        Y[(size_t)t * R + row] = v;
        return v;
    }
};

// transposed and scaled by Dinv (the same layout as Y)
struct DstStoreTScaled {
    double* Y;
    const double* Dinv;
    int R;
    using Aux = double;
    __device__ Aux fetch(int t, int row) const { return Dinv[(size_t)t * R + row]; }
    __device__ double put(int t, int row, double v, Aux d) const {
        const double y = v * d;
        Y[(size_t)t * R + row] = y;
        return y;
    }
};

// the grid's interior node (t + 1, row + 1) (the caller imposes the
// Dirichlet values afterwards)
struct DstStoreGrid {
    double* out;
    int gx;
    struct Aux {};
    __device__ Aux fetch(int, int) const { return {}; }
    __device__ double put(int t, int row, double v, Aux) const {
        out[(t + 1) * gx + row + 1] = v;
        return v;
    }
};

// Row lines of the covering (at most kPassLines) taken by the second pass
// while it stores the spectrum: V1[a][row] = sum_t S_n[row a][t] Uhat[t][row]
// for the pass's two rows (columns of Uhat).  Each of the workgroup's four
// waves leaves its partial sum (V1 as lr x R x 4); their reader adds the four
// in a fixed order, so the pass needs no extra barrier.
static constexpr int kPassLines = 2;
struct PassLines {
    int lr;                                  // 0: none
    const double* SnR;                       // lr x N
    double* V1;                              // lr x R x 4 wave partials
};

// One pass: R rows of N = 2^(L-1) - 1 values in, N rows of R values out.
// Workgroup w (after the XCD-aware remap: the workgroups of one XCD take
// consecutive row pairs, so the transposed stores of one cache line meet in
// one L2) transforms NB = max(1, 2048 / M) row pairs.
// LDS of a pass: the workgroup's points and, up to L = 11, the twiddles
// (64 KB in all); declared once by the kernel, whatever bodies it inlines.
template <int L>
struct DstLds {
    static constexpr int M = 1 << L;
    static constexpr int PTS = (M >= kDstPts ? 1 : kDstPts / M) * M;
    static constexpr int size = PTS + (L <= 11 ? M : 0);
};

template <int L, class Load, class Store, bool LINES = false>
__device__ __forceinline__ void dst_body(double2* lds, int R, int nwg, const Load& ld, const Store& st,
                                         const double2* __restrict__ tw, const PassLines pl = PassLines{}) {
    constexpr int M = 1 << L;
    constexpr int N = M / 2 - 1;
    constexpr int NB = M >= kDstPts ? 1 : kDstPts / M;
    constexpr int PTS = NB * M;
    constexpr int NQ = PTS / 4 / kDstT;                  // radix-4 butterflies per thread
    constexpr int NE = NB * N;                           // values per row of the workgroup's pairs
    constexpr int IT = (NE + kDstT - 1) / kDstT;
    constexpr bool kTwLds = L <= 11;                     // twiddles staged in LDS (64 KB in all)
    static_assert(DstLds<L>::size == PTS + (kTwLds ? M : 0), "LDS layout of a pass");
    double2* z = lds;
    double2* tws = lds + PTS;
    const double2* twp = kTwLds ? tws : tw;
    const int per = (int)(gridDim.x >> 3);
    const int w = (int)(blockIdx.x & 7) * per + (int)(blockIdx.x >> 3);
    if (w >= nwg) return;
    const int tid = threadIdx.x;
    __builtin_assume(tid < kDstT);
    const int r0 = w * 2 * NB;
    // loads: indices clamped into the matrix so that none waits behind a branch
    // (with LINES the S_n rows of the line sums too, so that they arrive during
    // the FFT)
    double va[IT], vb[IT];
    double sn[LINES ? kPassLines : 1][IT];
    if constexpr (LINES) {
#pragma unroll
        for (int it = 0; it < IT; ++it) {
            const int t = min(tid + kDstT * it, NE - 1);
#pragma unroll
            for (int a = 0; a < kPassLines; ++a) sn[a][it] = a < pl.lr ? pl.SnR[(size_t)a * N + t] : 0.0;
        }
    }
#pragma unroll
    for (int it = 0; it < IT; ++it) {
        const int e = min(tid + kDstT * it, NE - 1);
        const int b = e / N, j = e - b * N;
        const int ra = r0 + 2 * b;
        va[it] = ld(min(ra, R - 1), j);
        vb[it] = ld(min(ra + 1, R - 1), j);
    }
    if constexpr (kTwLds) {
        constexpr int TI = (M + kDstT - 1) / kDstT;
        double2 t[TI];
#pragma unroll
        for (int it = 0; it < TI; ++it) t[it] = tw[min(tid + kDstT * it, M - 1)];
#pragma unroll
        for (int it = 0; it < TI; ++it)
            if (M % kDstT == 0 || tid + kDstT * it < M) tws[tid + kDstT * it] = t[it];
    }
#pragma unroll
    for (int it = 0; it < IT; ++it) {
        const int e = tid + kDstT * it;
        if (e < NE) {
            const int b = e / N, j = e - b * N;
            const int ra = r0 + 2 * b;
            const double a = ra < R ? va[it] : 0.0;
            const double c = ra + 1 < R ? vb[it] : 0.0;
            z[zsw(b * M + j + 1)] = make_double2(a, c);
            z[zsw(b * M + M - 1 - j)] = make_double2(-a, -c);
        }
    }
    if (tid < NB) {
        z[zsw(tid * M)] = make_double2(0.0, 0.0);
        z[zsw(tid * M + M / 2)] = make_double2(0.0, 0.0);
    }
    __syncthreads();
#pragma unroll
    for (int s = 0; s < L / 2; ++s) {
        const int p = 1 << (2 * s);
        double2 u[NQ][4];
        int dst[NQ];
#pragma unroll
        for (int t = 0; t < NQ; ++t) {
            const int ig = tid + kDstT * t;
            const int base = (ig >> (L - 2)) << L;
            const int i = ig & (M / 4 - 1);
            const int k = i & (p - 1);
#pragma unroll
            for (int r = 0; r < 4; ++r) u[t][r] = z[zsw(base + i + r * (M / 4))];
            if (s > 0) {
#pragma unroll
                for (int r = 1; r < 4; ++r) u[t][r] = cmul(u[t][r], twp[r * p + k - 4]);
            }
            const double2 v0 = make_double2(u[t][0].x + u[t][2].x, u[t][0].y + u[t][2].y);
            const double2 v1 = make_double2(u[t][0].x - u[t][2].x, u[t][0].y - u[t][2].y);
            const double2 v2 = make_double2(u[t][1].x + u[t][3].x, u[t][1].y + u[t][3].y);
            const double2 v3 = make_double2(u[t][1].y - u[t][3].y, u[t][3].x - u[t][1].x);   // (u1 - u3)(-i)
            u[t][0] = make_double2(v0.x + v2.x, v0.y + v2.y);
            u[t][1] = make_double2(v1.x + v3.x, v1.y + v3.y);
            u[t][2] = make_double2(v0.x - v2.x, v0.y - v2.y);
            u[t][3] = make_double2(v1.x - v3.x, v1.y - v3.y);
            dst[t] = base + ((i - k) << 2) + k;
        }
        __syncthreads();
#pragma unroll
        for (int t = 0; t < NQ; ++t)
#pragma unroll
            for (int r = 0; r < 4; ++r) z[zsw(dst[t] + r * p)] = u[t][r];
        __syncthreads();
    }
    if (L & 1) {
        // radix 2 with p = M/2: butterfly i pairs i and i + M/2 in place
#pragma unroll
        for (int t = 0; t < 2 * NQ; ++t) {
            const int ig = tid + kDstT * t;
            const int base = (ig >> (L - 1)) << L;
            const int i = ig & (M / 2 - 1);
            const double2 a = z[zsw(base + i)];
            const double2 b = cmul(z[zsw(base + i + M / 2)], twp[M / 2 - 4 + i]);
            z[zsw(base + i)] = make_double2(a.x + b.x, a.y + b.y);
            z[zsw(base + i + M / 2)] = make_double2(a.x - b.x, a.y - b.y);
        }
        __syncthreads();
    }
    // the storer's gathers first, then the stores
    typename Store::Aux xa[IT], xb[IT];
#pragma unroll
    for (int it = 0; it < IT; ++it) {
        const int e = min(tid + kDstT * it, NE - 1);
        const int b = e / N, t = e - b * N;
        const int ra = r0 + 2 * b;
        xa[it] = st.fetch(t, min(ra, R - 1));
        xb[it] = st.fetch(t, min(ra + 1, R - 1));
    }
    const double sc = 0.5 * sqrt(2.0 / (double)(N + 1));
    double acc[LINES ? kPassLines : 1][2] = {};
#pragma unroll
    for (int it = 0; it < IT; ++it) {
        const int e = tid + kDstT * it;
        if (e < NE) {
            const int b = e / N, t = e - b * N;
            const int ra = r0 + 2 * b;
            const double2 v = z[zsw(b * M + t + 1)];
            double ya = 0.0, yb = 0.0;
            if (ra < R) ya = st.put(t, ra, -v.y * sc, xa[it]);
            if (ra + 1 < R) yb = st.put(t, ra + 1, v.x * sc, xb[it]);
            if constexpr (LINES) {
#pragma unroll
                for (int a = 0; a < kPassLines; ++a) {
                    acc[a][0] += sn[a][it] * ya;
                    acc[a][1] += sn[a][it] * yb;
                }
            }
        }
    }
    if constexpr (LINES) {
        static_assert(NB == 1 && kDstT == 256, "the line sums assume one row pair per workgroup of four waves");
        const int lane = tid & 63, wv = tid >> 6;
#pragma unroll
        for (int a = 0; a < kPassLines; ++a)
#pragma unroll
            for (int c = 0; c < 2; ++c) {
                const double w = wave_sum(acc[a][c]);
                if (lane == 0 && a < pl.lr && r0 + c < R) pl.V1[((size_t)a * R + r0 + c) * 4 + wv] = w;
            }
    }
}

// The four passes of a direct solve, run blockIdx.y of the table:
//   1: rows of F (rhs) -> W (m x n)          2: rows of W -> Uhat = . / Lambda
//   3: rows of Uhat (+ G ./ Lambda) -> W      4: rows of W -> the grid's interior
template <int L, int KIND, bool B>
__global__ __launch_bounds__(kDstT) void k_fd_dst(const FdRun one, const FdRun* __restrict__ runs, int R, int nwg) {
    __shared__ double2 lds[DstLds<L>::size];
    const FdRun& f = B ? runs[blockIdx.y] : one;
    if constexpr (KIND == 1)
        dst_body<L>(lds, R, nwg, DstLoadRhs{f.gx, f.m, f.sgn, f.rho, f.dir, f.Fb}, DstStoreT{f.W, f.n}, f.twm);
    else if constexpr (KIND == 2) {
        // the row lines ride along where a workgroup holds one row pair (L >= 11)
        if constexpr (L >= 11) {
            if (f.rl2) {
                dst_body<L, DstLoadRows, DstStoreTScaled, true>(lds, R, nwg, DstLoadRows{f.W, f.n},
                                                                DstStoreTScaled{f.U, f.Dinv, f.m}, f.twn,
                                                                PassLines{f.lr, f.SnR, f.V1});
                return;
            }
        }
        dst_body<L>(lds, R, nwg, DstLoadRows{f.W, f.n}, DstStoreTScaled{f.U, f.Dinv, f.m}, f.twn);
    }
    else if constexpr (KIND == 3)
        dst_body<L>(lds, R, nwg,
                    DstLoadCorrected{f.U, f.G, f.Dinv, f.m, f.cmode, f.n, f.lr, f.lc, f.SnR, f.SmC, f.LV},
                    DstStoreT{f.W, f.n}, f.twm);
    else
        dst_body<L>(lds, R, nwg, DstLoadRows{f.W, f.n}, DstStoreGrid{f.phi, f.gx}, f.twn);
}

// Runs of a launch: one run's descriptor by value (kernel argument), or a
// device table of `n` runs indexed by the last grid dimension.
struct FdRuns {
    const FdRun* one;
    const FdRun* tab;
    int n;
};

template <int L, int KIND>
static void dst_pass_l(const FdRuns& r, int R, hipStream_t s) {
    constexpr int M = 1 << L;
    constexpr int NB = M >= kDstPts ? 1 : kDstPts / M;
    const int nwg = (R + 2 * NB - 1) / (2 * NB);
    const int grid = (nwg + 7) / 8 * 8;
    if (r.tab)
        hipLaunchKernelGGL((k_fd_dst<L, KIND, true>), dim3((unsigned)grid, (unsigned)r.n), dim3(kDstT), 0, s, FdRun{},
                           r.tab, R, nwg);
    else
        hipLaunchKernelGGL((k_fd_dst<L, KIND, false>), dim3((unsigned)grid), dim3(kDstT), 0, s, *r.one,
                           (const FdRun*)nullptr, R, nwg);
}

// One pass over the runs (same grid): DST-I of R rows of length 2^(L-1) - 1 each
template <int KIND>
static void dst_pass(int L, const FdRuns& r, int R, hipStream_t s) {
    switch (L) {
    case 3: dst_pass_l<3, KIND>(r, R, s); break;
    case 4: dst_pass_l<4, KIND>(r, R, s); break;
    case 5: dst_pass_l<5, KIND>(r, R, s); break;
    case 6: dst_pass_l<6, KIND>(r, R, s); break;
    case 7: dst_pass_l<7, KIND>(r, R, s); break;
    case 8: dst_pass_l<8, KIND>(r, R, s); break;
    case 9: dst_pass_l<9, KIND>(r, R, s); break;
    case 10: dst_pass_l<10, KIND>(r, R, s); break;
    case 11: dst_pass_l<11, KIND>(r, R, s); break;
    case 12: dst_pass_l<12, KIND>(r, R, s); break;
    default: break;   // fast_setup admits 3 <= L <= 12 only
    }
}

// ------------------------------------------------------------------- host
static std::vector<double> dst_matrix(int N) {
    std::vector<double> S((size_t)N * N);
    const double a = std::sqrt(2.0 / (N + 1));
    const double w = M_PI / (N + 1);
    for (int j = 0; j < N; ++j)
        for (int k = 0; k < N; ++k) {
            // reduce (j+1)(k+1) modulo 2(N+1) so the sine argument stays below 2 pi
            const long long t = ((long long)(j + 1) * (k + 1)) % (2LL * (N + 1));
            S[(size_t)j * N + k] = a * std::sin(w * (double)t);
        }
    return S;
}

// Real orthonormal eigenbasis of the periodic second difference of period N
// (row j = node, column k = mode): 1/sqrt(N); sqrt(2/N) cos(2 pi p j / N) and
// sqrt(2/N) sin(2 pi p j / N) for 0 < p < N/2; (-1)^j / sqrt(N) for even N.
// lam[k] = 2 - 2 cos(2 pi p / N) for column k's frequency p.
static std::vector<double> fourier_matrix(int N, std::vector<double>& lam) {
    std::vector<double> Y((size_t)N * N);
    lam.assign(N, 0.0);
    const long double pi = 3.141592653589793238462643383279502884L;
    const double a0 = 1.0 / std::sqrt((double)N), a = std::sqrt(2.0 / N);
    for (int j = 0; j < N; ++j) Y[(size_t)j * N] = a0;
    int k = 1;
    for (int p = 1; 2 * p < N; ++p, k += 2) {
        lam[k] = lam[k + 1] = (double)(2.0L - 2.0L * cosl(2.0L * pi * (long double)p / (long double)N));
        for (int j = 0; j < N; ++j) {
            // reduce p j modulo N so the argument stays below 2 pi
            const long long t = ((long long)p * j) % N;
            const long double ang = 2.0L * pi * (long double)t / (long double)N;
            Y[(size_t)j * N + k] = a * (double)cosl(ang);
            Y[(size_t)j * N + k + 1] = a * (double)sinl(ang);
        }
    }
    if (N % 2 == 0) {
        lam[N - 1] = 4.0;
        for (int j = 0; j < N; ++j) Y[(size_t)j * N + N - 1] = (j % 2) ? -a0 : a0;
    }
    return Y;
}

// Explicit inverse of the SPD matrix C (P x P, row-major) by Cholesky.
static bool spd_inverse(std::vector<double>& C, int P) {
    std::vector<double> L((size_t)P * P, 0.0);
    for (int j = 0; j < P; ++j) {
        double d = C[(size_t)j * P + j];
        for (int k = 0; k < j; ++k) d -= L[(size_t)j * P + k] * L[(size_t)j * P + k];
        if (!(d > 0)) return false;
        const double ljj = std::sqrt(d);
        L[(size_t)j * P + j] = ljj;
        for (int i = j + 1; i < P; ++i) {
            double s = C[(size_t)i * P + j];
            for (int k = 0; k < j; ++k) s -= L[(size_t)i * P + k] * L[(size_t)j * P + k];
            L[(size_t)i * P + j] = s / ljj;
        }
    }
    // Linv (lower), then C^-1 = Linv^T Linv
    std::vector<double> Li((size_t)P * P, 0.0);
    for (int i = 0; i < P; ++i) {
        Li[(size_t)i * P + i] = 1.0 / L[(size_t)i * P + i];
        for (int j = 0; j < i; ++j) {
            double s = 0.0;
            for (int k = j; k < i; ++k) s += L[(size_t)i * P + k] * Li[(size_t)k * P + j];
            Li[(size_t)i * P + j] = -s / L[(size_t)i * P + i];
        }
    }
    for (int i = 0; i < P; ++i)
        for (int j = 0; j <= i; ++j) {
            double s = 0.0;
            for (int k = i; k < P; ++k) s += Li[(size_t)k * P + i] * Li[(size_t)k * P + j];
            C[(size_t)i * P + j] = s;
            C[(size_t)j * P + i] = s;
        }
    return true;
}

// The descriptor of context c's direct solve; sgn as FdRun::sgn.
static FdRun run_of(const pic_ctx* c, double sgn) {
    const FastPoisson* f = c->fast;
    FdRun r{};
    r.gx = c->g.gx;
    r.m = f->m;
    r.n = f->n;
    r.P = f->P;
    r.nq = f->nq;
    r.lr = f->nr <= kLines ? f->nr : 0;
    r.lc = f->nc <= kLines ? f->nc : 0;
    r.sgn = sgn;
    r.rho = c->rho;
    r.dir = c->dir;
    r.phi = c->phi;
    r.Fb = f->Fb;
    r.Dinv = f->Dinv;
    r.W = f->W;
    r.U = f->U;
    r.G = f->G;
    r.SnR = f->SnR;
    r.SmC = f->SmC;
    r.V1 = f->V1;
    r.V2 = f->V2;
    r.part = f->part;
    r.part2 = f->part2;
    r.Sn = f->Sn;
    r.Sm = f->Sm;
    r.qri = f->qri;
    r.qrow = f->qrow;
    r.qcol = f->qcol;
    r.VQ = f->VQ;
    r.dq = f->dq;
    r.Cinv = f->Cinv;
    r.rq_idx = f->rq_idx;
    r.rq_col = f->rq_col;
    r.rq_ptr = f->rq_ptr;
    r.cq = f->cq;
    r.Z = f->Z;
    r.SnQ = f->SnQ;
    r.twm = f->twm;
    r.twn = f->twn;
    r.cmode = f->P == 0 ? 0 : (f->line_form ? 2 : 1);
    r.rl2 = (f->fft && f->ln >= 11 && f->nr >= 1 && f->nr <= kPassLines) ? 1 : 0;
    r.lptr = f->lptr;
    r.lt = f->lt;
    r.lidx = f->lidx;
    r.LV = f->LV;
    return r;
}

// The electrode correction of the runs (dimensions of f; V1/V2 lines beyond
// kLines are left to the caller): covering lines, u0 on Q, sources, their
// rows and the spectrum G by the FP64 tile product k_fd_corr (one run or a
// batch, so a captured graph holds kernels only).
template <bool B>
static void launch_fd(void (*k)(const FdRun, const FdRun*), dim3 g, const FdRuns& r, hipStream_t s) {
    if (B) hipLaunchKernelGGL(k, dim3(g.x, g.y, (unsigned)r.n), dim3(kFdB), 0, s, FdRun{}, r.tab);
    else hipLaunchKernelGGL(k, g, dim3(kFdB), 0, s, *r.one, (const FdRun*)nullptr);
}

template <bool B>
static int enqueue_correction(FastPoisson* f, const FdRuns& r, hipStream_t s) {
    const int m = f->m, n = f->n, P = f->P;
    const int lr = f->nr <= kLines ? f->nr : 0, lc = f->nc <= kLines ? f->nc : 0;
    const bool rows_in_pass = f->fft && f->ln >= 11 && f->nr >= 1 && f->nr <= kPassLines;   // FdRun::rl2
    if (rows_in_pass || lr == 0) {
        // the row lines came from the second FFT pass (or there are none)
        if (lc > 0) launch_fd<B>(k_fd_collines<B>, dim3((unsigned)n), r, s);
    } else {
        launch_fd<B>(k_fd_lines<B>, dim3((unsigned)((m + 63) / 64), kSeg), r, s);
        launch_fd<B>(k_fd_linesum<B>, dim3((unsigned)((lr * m + lc * n + kFdB - 1) / kFdB)), r, s);
    }
    launch_fd<B>(k_fd_u0q<B>, dim3((unsigned)P), r, s);
    launch_fd<B>(k_fd_sources<B>, dim3((unsigned)P), r, s);
    if (f->line_form) {
        // one vector per covering line; the third transform pass adds the terms
        launch_fd<B>(k_fd_linevec<B>, dim3((unsigned)((std::max(m, n) + 63) / 64), (unsigned)(f->nr + f->nc)), r, s);
        return PIC_OK;
    }
    launch_fd<B>(k_fd_zrows<B>, dim3((unsigned)((m + 63) / 64), (unsigned)f->nq), r, s);
    // G = S_n[:, rows of Q] Z (n x nq times nq x m)
    launch_fd<B>(k_fd_corr<B>, dim3((unsigned)((m + kCT - 1) / kCT), (unsigned)((n + kCT - 1) / kCT)), r, s);
    return PIC_OK;
}

// Set up the direct solve for the current geometry (`dir`: 1 = Dirichlet,
// the walls included).  Leaves c->fast null when the configuration is not
// covered (periodic y, too many internal Dirichlet nodes next to unknowns,
// large grids, or turned off); the PCG then solves alone.
int fast_setup(pic_ctx* c, const std::vector<uint8_t>& dir, const std::vector<double>& dval) {
    fast_free(c);
    const int gx = c->g.gx, gy = c->g.gy;
    // Periodic y: the system's rows are grid rows 0 .. gy-2 (row gy-1 is the
    // image of row 0), their y neighbours wrap, and only the x walls move to
    // the right-hand side.  Otherwise the rows are 1 .. gy-2.
    const int per = c->p.periodic ? 1 : 0;
    const int m = gx - 2, n = per ? gy - 1 : gy - 2, j0 = per ? 0 : 1;
    if (!c->p.poisson_direct || m < 1 || n < 1 + per || m > kFdMaxN || n > kFdMaxN) return PIC_OK;
    std::vector<int> qrow, qcol;
    std::vector<double> vq;
    for (int r = 0; r < n; ++r)
        for (int q = 1; q < gx - 1; ++q) {
            const int node = (r + j0) * gx + q;
            if (dir[node] != 1) continue;
            const int south = (r == 0 && per) ? (n - 1) * gx + q : node - gx;
            const int north = (r == n - 1 && per) ? q : node + gx;
            if (!dir[node - 1] || !dir[node + 1] || !dir[south] || !dir[north]) {
                qrow.push_back(r);
                qcol.push_back(q - 1);
                vq.push_back(dval[node]);
            }
        }
    const int P = (int)qrow.size();
    if (P > kFdMaxP) return PIC_OK;
    FastPoisson* f = new FastPoisson();
    c->fast = f;
    f->m = m;
    f->n = n;
    f->P = P;
    f->per = per;
    // Cover Q greedily with rows and columns (the line holding most uncovered
    // nodes first): u0 on Q then needs one product per covering line, e.g. two
    // rows and two columns for a rectangular electrode.
    std::vector<int> rows, cols, qri(P, 0);
    {
        std::vector<char> done(P, 0);
        int left = P;
        while (left > 0) {
            std::vector<int> rc(n, 0), cc(m, 0);
            for (int q = 0; q < P; ++q)
                if (!done[q]) {
                    ++rc[qrow[q]];
                    ++cc[qcol[q]];
                }
            const int br = (int)(std::max_element(rc.begin(), rc.end()) - rc.begin());
            const int bc = (int)(std::max_element(cc.begin(), cc.end()) - cc.begin());
            const bool by_row = rc[br] >= cc[bc];
            for (int q = 0; q < P; ++q) {
                if (done[q] || (by_row ? qrow[q] != br : qcol[q] != bc)) continue;
                done[q] = 1;
                --left;
                qri[q] = by_row ? (int)rows.size() : -1 - (int)cols.size();
            }
            if (by_row) rows.push_back(br);
            else cols.push_back(bc);
        }
    }
    f->nr = (int)rows.size();
    f->nc = (int)cols.size();
    const int nr = f->nr, nc = f->nc;

    const std::vector<double> Sm = dst_matrix(m);
    std::vector<double> lamy(n);
    std::vector<double> Sn;
    if (per) Sn = fourier_matrix(n, lamy);
    else {
        Sn = (n == m) ? Sm : dst_matrix(n);
        for (int k = 0; k < n; ++k) lamy[k] = 2.0 - 2.0 * std::cos(M_PI * (k + 1) / (n + 1));
    }
    const double cx = 1.0 / (c->g.dx * c->g.dx), cy = 1.0 / (c->g.dy * c->g.dy);
    std::vector<double> Dinv((size_t)n * m);
    for (int k = 0; k < n; ++k) {
        const double ly = cy * lamy[k];
        for (int l = 0; l < m; ++l) {
            const double lx = cx * (2.0 - 2.0 * std::cos(M_PI * (l + 1) / (m + 1)));
            Dinv[(size_t)k * m + l] = 1.0 / (ly + lx);
        }
    }
    std::vector<double> SnR((size_t)std::max(nr, 1) * n), SmC((size_t)std::max(nc, 1) * m);
    // distinct rows of Q and the q on each (ascending q)
    std::vector<int> qr_list, rq_ptr(1, 0), rq_idx;
    {
        std::vector<std::vector<int>> by_row(n);
        for (int q = 0; q < P; ++q) by_row[qrow[q]].push_back(q);
        for (int r = 0; r < n; ++r)
            if (!by_row[r].empty()) {
                qr_list.push_back(r);
                rq_idx.insert(rq_idx.end(), by_row[r].begin(), by_row[r].end());
                rq_ptr.push_back((int)rq_idx.size());
            }
    }
    const int nq = (int)qr_list.size();
    f->nq = nq;
    std::vector<double> SnQ((size_t)n * std::max(nq, 1));
    for (int a = 0; a < nr; ++a)
        for (int k = 0; k < n; ++k) SnR[(size_t)a * n + k] = Sn[(size_t)rows[a] * n + k];
    for (int j = 0; j < nc; ++j)
        for (int l = 0; l < m; ++l) SmC[(size_t)j * m + l] = Sm[(size_t)cols[j] * m + l];
    for (int k = 0; k < n; ++k)
        for (int a = 0; a < nq; ++a) SnQ[(size_t)k * nq + a] = Sn[(size_t)qr_list[a] * n + k];

    bool ok = true;
    const size_t nm = (size_t)n * m;
    f->Sm = (double*)dmalloc((size_t)m * m * 8, ok);
    f->Sn = (double*)dmalloc((size_t)n * n * 8, ok);
    f->Dinv = (double*)dmalloc(nm * 8, ok);
    f->F = (double*)dmalloc(nm * 8, ok);
    f->W = (double*)dmalloc(nm * 8, ok);
    f->U = (double*)dmalloc(nm * 8, ok);
    f->G = (double*)dmalloc(nm * 8, ok);
    f->SnR = (double*)dmalloc(SnR.size() * 8, ok);
    f->V1 = (double*)dmalloc((size_t)std::max(nr, 1) * m * 4 * 8, ok);   // room for the pass's wave partials
    f->SmC = (double*)dmalloc(SmC.size() * 8, ok);
    f->V2 = (double*)dmalloc((size_t)std::max(nc, 1) * n * 8, ok);
    f->qrow = (int*)dmalloc((size_t)std::max(P, 1) * 4, ok);
    // FFT transforms for N + 1 = 2^(L-1), 3 <= L <= 12 (N = m, n)
    auto fft_log = [](int N) {
        int L = 0;
        while ((2 << L) < 2 * (N + 1)) ++L;
        L += 1;
        return ((1 << L) == 2 * (N + 1) && L >= 3 && L <= 12) ? L : 0;
    };
    f->lm = fft_log(m);
    f->ln = fft_log(n);
    f->fft = !per && c->p.poisson_direct == 1 && f->lm > 0 && f->ln > 0;
    std::vector<double2> twm, twn;
    if (f->fft) {
        // w^e, w = exp(-2 pi i / M), in the order the FFT reads them: radix-4
        // stage s (p = 4^s, s >= 1) takes w^((r k) 2^(L-2-2s)) for r = 1..3,
        // k < p at r p + k - 4 (consecutive lanes read consecutive entries);
        // a final radix-2 stage (odd L) takes w^i, i < M/2, at M/2 - 4 + i
        auto table = [](int L) {
            const int M = 1 << L;
            const long double pi = 3.141592653589793238462643383279502884L;
            auto w = [&](int e) {
                const long double a = 2.0L * pi * (long double)e / (long double)M;
                return make_double2((double)cosl(a), (double)-sinl(a));
            };
            std::vector<double2> t(M, make_double2(0.0, 0.0));
            for (int st = 1; st < L / 2; ++st) {
                const int p = 1 << (2 * st);
                for (int r = 1; r < 4; ++r)
                    for (int k = 0; k < p; ++k) t[r * p + k - 4] = w((r * k) << (L - 2 - 2 * st));
            }
            if (L & 1)
                for (int i = 0; i < M / 2; ++i) t[M / 2 - 4 + i] = w(i);
            return t;
        };
        twm = table(f->lm);
        twn = table(f->ln);
        f->twm = (double2*)dmalloc(twm.size() * sizeof(double2), ok);
        f->twn = (double2*)dmalloc(twn.size() * sizeof(double2), ok);
        f->Fb = (double*)dmalloc(nm * 8, ok);
    }
    if (nr > 0 && nr <= kLines) f->part = (double*)dmalloc((size_t)kSeg * nr * m * 8, ok);
    if (nc > 0 && nc <= kLines) f->part2 = (double*)dmalloc((size_t)((m + 63) / 64) * nc * n * 8, ok);
    f->SnQ = (double*)dmalloc(SnQ.size() * 8, ok);
    f->Z = (double*)dmalloc((size_t)std::max(nq, 1) * m * 8, ok);
    f->cq = (double*)dmalloc((size_t)std::max(P, 1) * 8, ok);
    f->rq_ptr = (int*)dmalloc(rq_ptr.size() * 4, ok);
    f->rq_idx = (int*)dmalloc((size_t)std::max(P, 1) * 4, ok);
    f->rq_col = (int*)dmalloc((size_t)std::max(P, 1) * 4, ok);
    f->Cinv = (double*)dmalloc((size_t)std::max(P, 1) * std::max(P, 1) * 8, ok);
    f->VQ = (double*)dmalloc((size_t)std::max(P, 1) * 8, ok);
    f->dq = (double*)dmalloc((size_t)std::max(P, 1) * 8, ok);
    f->qri = (int*)dmalloc((size_t)std::max(P, 1) * 4, ok);
    f->qcol = (int*)dmalloc((size_t)std::max(P, 1) * 4, ok);
    int* qrow_d = (int*)dmalloc((size_t)std::max(P, 1) * 4, ok);
    if (!ok) {
        if (qrow_d) (void)hipFree(qrow_d);
        fast_free(c);
        set_error("direct Poisson solve: device allocation failed");
        return PIC_ERR_ALLOC;
    }
    PIC_HIP(hipMemcpy(f->Sm, Sm.data(), (size_t)m * m * 8, hipMemcpyHostToDevice));
    PIC_HIP(hipMemcpy(f->Sn, Sn.data(), (size_t)n * n * 8, hipMemcpyHostToDevice));
    PIC_HIP(hipMemcpy(f->Dinv, Dinv.data(), nm * 8, hipMemcpyHostToDevice));
    PIC_HIP(hipMemcpy(f->SnR, SnR.data(), SnR.size() * 8, hipMemcpyHostToDevice));
    PIC_HIP(hipMemcpy(f->SmC, SmC.data(), SmC.size() * 8, hipMemcpyHostToDevice));
    PIC_HIP(hipMemcpy(f->SnQ, SnQ.data(), SnQ.size() * 8, hipMemcpyHostToDevice));
    PIC_HIP(hipMemcpy(f->rq_ptr, rq_ptr.data(), rq_ptr.size() * 4, hipMemcpyHostToDevice));
    if (P > 0) {
        std::vector<int> rq_col(P);
        for (int t = 0; t < P; ++t) rq_col[t] = qcol[rq_idx[t]];
        PIC_HIP(hipMemcpy(f->rq_idx, rq_idx.data(), (size_t)P * 4, hipMemcpyHostToDevice));
        PIC_HIP(hipMemcpy(f->rq_col, rq_col.data(), (size_t)P * 4, hipMemcpyHostToDevice));
    }
    if (f->fft) {
        PIC_HIP(hipMemcpy(f->twm, twm.data(), twm.size() * sizeof(double2), hipMemcpyHostToDevice));
        PIC_HIP(hipMemcpy(f->twn, twn.data(), twn.size() * sizeof(double2), hipMemcpyHostToDevice));
        // wall terms of F: the Dirichlet values of the wall nodes next to the interior
        std::vector<double> Fb(nm, 0.0);
        for (int r = 0; r < n; ++r)
            for (int q = 0; q < m; ++q) {
                const int node = (r + 1) * gx + q + 1;
                double v = 0.0;
                if (q == 0) v += cx * dval[node - 1];
                if (q == m - 1) v += cx * dval[node + 1];
                if (r == 0) v += cy * dval[node - gx];
                if (r == n - 1) v += cy * dval[node + gx];
                Fb[(size_t)r * m + q] = v;
            }
        PIC_HIP(hipMemcpy(f->Fb, Fb.data(), nm * 8, hipMemcpyHostToDevice));
    }
    if (P > 0) {
        PIC_HIP(hipMemcpy(f->VQ, vq.data(), (size_t)P * 8, hipMemcpyHostToDevice));
        PIC_HIP(hipMemcpy(f->qri, qri.data(), (size_t)P * 4, hipMemcpyHostToDevice));
        PIC_HIP(hipMemcpy(f->qcol, qcol.data(), (size_t)P * 4, hipMemcpyHostToDevice));
        PIC_HIP(hipMemcpy(qrow_d, qrow.data(), (size_t)P * 4, hipMemcpyHostToDevice));
        PIC_HIP(hipMemcpy(f->qrow, qrow.data(), (size_t)P * 4, hipMemcpyHostToDevice));
        // C = A A^T over the spectral index, in chunks of kc rows of k
        const int kc = std::max(1, std::min(n, (int)((size_t)(256u << 20) / ((size_t)P * m * 8))));
        double* A = nullptr;
        double* Cd = nullptr;
        PIC_HIP(hipMalloc((void**)&A, (size_t)P * kc * m * 8));
        PIC_HIP(hipMalloc((void**)&Cd, (size_t)P * P * 8));
        int rc = PIC_OK;
        for (int k0 = 0; k0 < n && rc == PIC_OK; k0 += kc) {
            const size_t tot = (size_t)P * kc * m;
            hipLaunchKernelGGL(k_fd_amat, dim3((unsigned)((tot + kFdB - 1) / kFdB)), dim3(kFdB), 0, c->stream, P, n, m,
                               k0, kc, f->Sn, f->Sm, f->Dinv, qrow_d, f->qcol, A);
            // C (+)= A A^T over this chunk of the spectral index
            if (k0 == 0) dgemm<false, true, 0>(c->stream, P, P, kc * m, A, kc * m, A, kc * m, Cd, P);
            else dgemm<false, true, 2>(c->stream, P, P, kc * m, A, kc * m, A, kc * m, Cd, P);
            if (hipGetLastError() != hipSuccess) rc = PIC_ERR_HIP;
        }
        std::vector<double> C((size_t)P * P);
        if (rc == PIC_OK && hipMemcpyAsync(C.data(), Cd, (size_t)P * P * 8, hipMemcpyDeviceToHost, c->stream) != hipSuccess)
            rc = PIC_ERR_HIP;
        if (rc == PIC_OK && hipStreamSynchronize(c->stream) != hipSuccess) rc = PIC_ERR_HIP;
        (void)hipFree(A);
        (void)hipFree(Cd);
        (void)hipFree(qrow_d);
        qrow_d = nullptr;
        if (rc != PIC_OK) {
            fast_free(c);
            set_error("direct Poisson solve: capacitance matrix assembly failed");
            return rc;
        }
        if (!spd_inverse(C, P)) {
            // not positive definite in floating point: leave the PCG alone
            fast_free(c);
            return PIC_OK;
        }
        PIC_HIP(hipMemcpy(f->Cinv, C.data(), (size_t)P * P * 8, hipMemcpyHostToDevice));
    }
    if (qrow_d) (void)hipFree(qrow_d);
    if (P > 0 && nr <= kLines && nc <= kLines) {
        // line form: the sources of each covering line (rows first, then
        // columns), q ascending; lt indexes the row-ordered sources cq
        std::vector<int> tpos(P);
        for (int t = 0; t < P; ++t) tpos[rq_idx[t]] = t;
        std::vector<std::vector<int>> members(nr + nc);
        for (int q = 0; q < P; ++q) members[qri[q] >= 0 ? qri[q] : nr + (-1 - qri[q])].push_back(q);
        std::vector<int> lptr(1, 0), lt, lidx;
        for (int v = 0; v < nr + nc; ++v) {
            for (int q : members[v]) {
                lt.push_back(tpos[q]);
                lidx.push_back(v < nr ? qcol[q] : qrow[q]);
            }
            lptr.push_back((int)lt.size());
        }
        bool okl = true;
        f->lptr = (int*)dmalloc(lptr.size() * 4, okl);
        f->lt = (int*)dmalloc(lt.size() * 4, okl);
        f->lidx = (int*)dmalloc(lidx.size() * 4, okl);
        f->LV = (double*)dmalloc(((size_t)nr * m + (size_t)nc * n) * 8, okl);
        if (!okl) {
            fast_free(c);
            set_error("direct Poisson solve: device allocation failed");
            return PIC_ERR_ALLOC;
        }
        PIC_HIP(hipMemcpy(f->lptr, lptr.data(), lptr.size() * 4, hipMemcpyHostToDevice));
        PIC_HIP(hipMemcpy(f->lt, lt.data(), lt.size() * 4, hipMemcpyHostToDevice));
        PIC_HIP(hipMemcpy(f->lidx, lidx.data(), lidx.size() * 4, hipMemcpyHostToDevice));
        f->line_form = 1;
    }
    const FdRun self = run_of(c, 1.0);
    if (hipMalloc((void**)&f->self, sizeof(FdRun)) != hipSuccess) {
        fast_free(c);
        set_error("direct Poisson solve: device allocation failed");
        return PIC_ERR_ALLOC;
    }
    PIC_HIP(hipMemcpy(f->self, &self, sizeof(FdRun), hipMemcpyHostToDevice));
    f->hself = new FdRun(self);
    PIC_HIP(hipDeviceSynchronize());
    return PIC_OK;
}

// The solution of the current right-hand side (c->rho after overwriteRho)
// into the interior nodes of `out`; the caller imposes the Dirichlet values
// (internal Dirichlet nodes hold values of the unconstrained solve until then;
// the GEMM path writes them already).
int fast_solve(pic_ctx* c, double* out) {
    FastPoisson* f = c->fast;
    const int gx = c->g.gx, gy = c->g.gy, m = f->m, n = f->n, P = f->P, nr = f->nr, nc = f->nc;
    const int nm = m * n;
    const double cx = 1.0 / (c->g.dx * c->g.dx), cy = 1.0 / (c->g.dy * c->g.dy);
    hipStream_t s = c->stream;
    const unsigned nb = (unsigned)((nm + kFdB - 1) / kFdB);
    if (out != c->phi) {
        set_error("direct Poisson solve: the output is the context's phi");
        return PIC_ERR_STATE;
    }
    const FdRuns one{f->hself, nullptr, 1};
    if (f->fft) {
        // rows of F (n x m) -> W (m x n) -> Uhat = (S_n F S_m) ./ Lambda (n x m)
        dst_pass<1>(f->lm, one, n, s);
        dst_pass<2>(f->ln, one, m, s);
    } else {
        hipLaunchKernelGGL(k_fd_rhs, dim3(nb), dim3(kFdB), 0, s, gx, n, f->per, cx, cy, c->rho, c->dir, c->dval,
                           f->F);
        // W = F S_m, Uhat = (Y^T W) ./ Lambda (Y = S_n, or the Fourier basis)
        dgemm<false, false, 0>(s, n, m, m, f->F, m, f->Sm, m, f->W, m);
        dgemm<true, false, 1>(s, n, m, n, f->Sn, n, f->W, m, f->U, m, f->Dinv);
    }
    if (P > 0) {
        // u0 on Q through the covering lines: V1 = S_n[rows] Uhat (nr x m),
        // V2 = S_m[cols] Uhat^T (nc x n)
        // (a few lines: one pass over Uhat; many lines: GEMMs)
        if (nr > kLines) dgemm<false, false, 0>(s, nr, m, n, f->SnR, n, f->U, m, f->V1, m);
        if (nc > kLines) dgemm<false, true, 0>(s, nc, n, m, f->SmC, m, f->U, m, f->V2, n);
        PIC_TRY(enqueue_correction<false>(f, one, s));
        if (!f->fft) {
            if (f->line_form) launch_fd<false>(k_fd_addlines<false>, dim3(nb), one, s);
            else hipLaunchKernelGGL(k_fd_addscaled, dim3(nb), dim3(kFdB), 0, s, nm, f->U, f->G, f->Dinv);
        }
    }
    if (f->fft) {
        // rows of Uhat (+ G ./ Lambda) -> W (m x n) -> the grid's interior nodes
        dst_pass<3>(f->lm, one, n, s);
        dst_pass<4>(f->ln, one, m, s);
        PIC_HIP(hipGetLastError());
        return PIC_OK;
    }
    // U = Y Uhat S_m
    dgemm<false, false, 0>(s, n, m, n, f->Sn, n, f->U, m, f->W, m);
    dgemm<false, false, 0>(s, n, m, m, f->W, m, f->Sm, m, f->G, m);
    hipLaunchKernelGGL(k_fd_out, dim3((unsigned)((gx * gy + kFdB - 1) / kFdB)), dim3(kFdB), 0, s, gx, gy,
                       f->per ? 0 : 1, n, f->G, c->dir, c->dval, out);
    PIC_HIP(hipGetLastError());
    return PIC_OK;
}

// ------------------------------------------------------------ batched solves
// Direct solves of many runs at once (an ensemble): runs whose solves have the
// same shape (grid, |Q|, covering lines, rows of Q) form a group with one
// device table; each pass is one launch per group.  The runs read their raw
// deposit (sgn -1), so the batch goes before the ensemble's overwriteRho.
struct FdBatch {
    struct Group {
        FastPoisson* f;                      // dimensions (any member)
        FdRun* tab;                          // device
        int nruns;
    };
    std::vector<Group> groups;
};

void fast_batch_free(FdBatch* b) {
    if (!b) return;
    for (auto& g : b->groups)
        if (g.tab) (void)hipFree(g.tab);
    delete b;
}

int fast_batch_create(pic_ctx* const* runs, int n, FdBatch** out, unsigned char* used) {
    *out = nullptr;
    FdBatch* b = new FdBatch();
    std::vector<std::vector<int>> members;
    std::vector<std::array<int, 10>> keys;
    for (int r = 0; r < n; ++r) {
        used[r] = 0;
        const pic_ctx* c = runs[r];
        const FastPoisson* f = c->fast;
        if (!f || !f->fft || f->nr > kLines || f->nc > kLines) continue;
        const std::array<int, 10> k{c->g.gx, c->g.gy, f->m, f->n, f->P, f->nq, f->nr, f->nc, f->lm, f->ln};
        size_t gi = 0;
        while (gi < keys.size() && keys[gi] != k) ++gi;
        if (gi == keys.size()) {
            keys.push_back(k);
            members.emplace_back();
        }
        members[gi].push_back(r);
        used[r] = 1;
    }
    for (size_t gi = 0; gi < keys.size(); ++gi) {
        std::vector<FdRun> t;
        for (int r : members[gi]) t.push_back(run_of(runs[r], -1.0));
        FdBatch::Group g{runs[members[gi][0]]->fast, nullptr, (int)t.size()};
        if (hipMalloc((void**)&g.tab, t.size() * sizeof(FdRun)) != hipSuccess ||
            hipMemcpy(g.tab, t.data(), t.size() * sizeof(FdRun), hipMemcpyHostToDevice) != hipSuccess) {
            if (g.tab) (void)hipFree(g.tab);
            fast_batch_free(b);
            set_error("batched direct Poisson solve: table allocation failed");
            return PIC_ERR_ALLOC;
        }
        b->groups.push_back(g);
    }
    *out = b;
    return PIC_OK;
}

// Every run of the batch: phi's interior nodes from its raw deposit.
int fast_batch_enqueue(const FdBatch* b, hipStream_t s) {
    for (const auto& g : b->groups) {
        FastPoisson* f = g.f;
        const FdRuns r{nullptr, g.tab, g.nruns};
        dst_pass<1>(f->lm, r, f->n, s);
        dst_pass<2>(f->ln, r, f->m, s);
        if (f->P > 0) PIC_TRY(enqueue_correction<true>(f, r, s));
        dst_pass<3>(f->lm, r, f->n, s);
        dst_pass<4>(f->ln, r, f->m, s);
    }
    PIC_HIP(hipGetLastError());
    return PIC_OK;
}

} // namespace pic

using namespace pic;

extern "C" int pic_dgemm_check(int M, int N, int K, int ta, int tb, int ep, const double* a, const double* b,
                               const double* sc, double* cmat) {
    if (M < 1 || N < 1 || K < 1 || M > 4096 || N > 4096 || K > 65536 || ep < 0 || ep > 2 || !a || !b || !cmat ||
        (ep == 1 && !sc)) {
        set_error("pic_dgemm_check: invalid arguments (M %d, N %d, K %d, ep %d)", M, N, K, ep);
        return PIC_ERR_ARG;
    }
    const size_t na = (size_t)M * K, nb_ = (size_t)K * N, nc = (size_t)M * N;
    double *da = nullptr, *db = nullptr, *dc = nullptr, *ds = nullptr;
    PIC_HIP(hipMalloc(&da, na * 8));
    PIC_HIP(hipMalloc(&db, nb_ * 8));
    PIC_HIP(hipMalloc(&dc, nc * 8));
    PIC_HIP(hipMalloc(&ds, nc * 8));
    PIC_HIP(hipMemcpy(da, a, na * 8, hipMemcpyHostToDevice));
    PIC_HIP(hipMemcpy(db, b, nb_ * 8, hipMemcpyHostToDevice));
    PIC_HIP(hipMemcpy(dc, cmat, nc * 8, hipMemcpyHostToDevice));
    if (sc) PIC_HIP(hipMemcpy(ds, sc, nc * 8, hipMemcpyHostToDevice));
    const int lda = ta ? M : K, ldb = tb ? K : N;
    hipStream_t s = nullptr;
#define PIC_DG(TA, TB)                                                                   \
    do {                                                                                 \
        if (ep == 0) dgemm<TA, TB, 0>(s, M, N, K, da, lda, db, ldb, dc, N);              \
        else if (ep == 1) dgemm<TA, TB, 1>(s, M, N, K, da, lda, db, ldb, dc, N, ds);     \
        else dgemm<TA, TB, 2>(s, M, N, K, da, lda, db, ldb, dc, N);                      \
    } while (0)
    if (ta && tb) PIC_DG(true, true);
    else if (ta) PIC_DG(true, false);
    else if (tb) PIC_DG(false, true);
    else PIC_DG(false, false);
#undef PIC_DG
    PIC_HIP(hipGetLastError());
    PIC_HIP(hipMemcpy(cmat, dc, nc * 8, hipMemcpyDeviceToHost));
    (void)hipFree(da);
    (void)hipFree(db);
    (void)hipFree(dc);
    (void)hipFree(ds);
    return PIC_OK;
}
