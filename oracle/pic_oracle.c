/*
 * pic_oracle.c — CPU restatement of the reference's 2D3V electrostatic PIC
 * time step.  TEST INFRASTRUCTURE ONLY: loaded by tests/, by the smoke check
 * of the graft entry point and by bench.py's CPU-baseline leg, always as the
 * checker / the timed CPU baseline, never by the product library.
 *
 * Every function restates the reference serially (or with OpenMP over
 * independent work items, never changing a floating-point summation order)
 * and cites the reference lines it follows.  Built with -ffp-contract=off so
 * each FP64 expression rounds exactly as written.
 *
 * Reference paths are relative to SYCL-PIC-poissonsolver/.
 */
#include "pic_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------ binning */
/* Reference key: gx = (int)(x * inv_dx), gy = (int)(y * inv_dy) with
 * inv_dx = 1.0 / dx; key = gy * GRID_X + gx
 * (src/particles.cpp:229-235; src/mover.cpp:309-324,356). */
int32_t orc_key(double x, double y, int gx_nodes, double dx, double dy) {
    const double inv_dx = 1.0 / dx;
    const double inv_dy = 1.0 / dy;
    int gx = (int)(x * inv_dx);
    int gy = (int)(y * inv_dy);
    return gy * gx_nodes + gx;
}

void orc_keys(int64_t n, const double* x, const double* y, int gx_nodes, double dx, double dy, int32_t* key) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) key[i] = orc_key(x[i], y[i], gx_nodes, dx, dy);
}

/* Membership of the reference hashmap after insert + cleanup
 * (include/sycl_hashmap.hpp:131-231; src/utils.cpp:30-44), restated
 * race-free: the particles of key k are exactly those with key == k.
 * counts[k] = |members|, start = exclusive prefix (start[G] = live total),
 * order = particle indices grouped by key, ascending index within a key.
 * Negative keys (absorbed particles) are dropped. */
void orc_bin(int64_t n, const int32_t* key, int G, int32_t* counts, int32_t* start, int64_t* order) {
    memset(counts, 0, (size_t)G * sizeof(int32_t));
    for (int64_t i = 0; i < n; ++i)
        if (key[i] >= 0) counts[key[i]]++;
    int64_t s = 0;
    for (int k = 0; k < G; ++k) {
        start[k] = (int32_t)s;
        s += counts[k];
    }
    start[G] = (int32_t)s;
    if (!order) return;
    int32_t* cur = (int32_t*)malloc((size_t)G * sizeof(int32_t));
    memcpy(cur, start, (size_t)G * sizeof(int32_t));
    for (int64_t i = 0; i < n; ++i)
        if (key[i] >= 0) order[cur[key[i]]++] = i;
    free(cur);
}

/* ------------------------------------------------------------ deposit */
/* interpolate (src/interpolation.cpp:13-176), FP64, for the particles of
 * each key in `order` (index order within a key):
 *   gx = key % GRID_X, gy = key / GRID_X (:72-73);
 *   lx = x - gx*dx, ly = y - gy*dy (:86-87);
 *   s1 += (dx-lx)(dy-ly), s2 += lx(dy-ly), s3 += (dx-lx)ly, s4 += lx*ly (:90-93);
 *   enrgy = 0.5 * v2 * (1/|q/m|); sKe += (corner weight) * enrgy (:88-98);
 *   keys on the last node column/row are skipped (:122);
 *   corner meshes = sums * W (:139-146); final = corner0+corner1+corner2+corner3
 *   at each node (:150-163); rho += final * (-q/eps) (:164-169). */
void orc_deposit_ref(int64_t n, const double* x, const double* y, const double* vx, const double* vy,
                     const double* vz, const int32_t* start, const int64_t* order, int gxn, int gyn, double dx,
                     double dy, double W, double qm, double charge, double eps, double* final_mesh,
                     double* energy_mesh, double* rho) {
    (void)n;
    const int G = gxn * gyn;
    double* corner = (double*)calloc((size_t)4 * G, sizeof(double));
    double* corner_e = (double*)calloc((size_t)4 * G, sizeof(double));
    const double denom = 1 / fabs(qm);
#pragma omp parallel for schedule(dynamic, 256)
    for (int k = 0; k < G; ++k) {
        const int gx = k % gxn;
        const int gy = k / gxn;
        double s1 = 0, s2 = 0, s3 = 0, s4 = 0;
        double s1e = 0, s2e = 0, s3e = 0, s4e = 0;
        for (int32_t m = start[k]; m < start[k + 1]; ++m) {
            int64_t i = order[m];
            double lx = x[i] - gx * dx;
            double ly = y[i] - gy * dy;
            double v2 = vx[i] * vx[i] + vy[i] * vy[i] + vz[i] * vz[i];
            double enrgy = 0.5 * v2 * denom;
            s1 += (dx - lx) * (dy - ly);
            s2 += lx * (dy - ly);
            s3 += (dx - lx) * ly;
            s4 += lx * ly;
            s1e += ((dx - lx) * (dy - ly)) * enrgy;
            s2e += (lx * (dy - ly)) * enrgy;
            s3e += (dx - lx) * ly * enrgy;
            s4e += lx * ly * enrgy;
        }
        if ((gx + 1) >= gxn || (gy + 1) >= gyn) continue;
        const int p1 = gy * gxn + gx, p2 = gy * gxn + (gx + 1);
        const int p3 = (gy + 1) * gxn + gx, p4 = (gy + 1) * gxn + (gx + 1);
        corner[0 * (size_t)G + p1] = s1 * W;
        corner[1 * (size_t)G + p2] = s2 * W;
        corner[2 * (size_t)G + p3] = s3 * W;
        corner[3 * (size_t)G + p4] = s4 * W;
        corner_e[0 * (size_t)G + p1] = s1e * W;
        corner_e[1 * (size_t)G + p2] = s2e * W;
        corner_e[2 * (size_t)G + p3] = s3e * W;
        corner_e[3 * (size_t)G + p4] = s4e * W;
    }
    const double chrg = -charge / eps;
#pragma omp parallel for schedule(static)
    for (int p = 0; p < G; ++p) {
        double sum = 0.0, sum_e = 0.0;
        for (int kb = 0; kb < 4; ++kb) {
            sum += corner[kb * (size_t)G + p];
            sum_e += corner_e[kb * (size_t)G + p];
        }
        final_mesh[p] = sum;
        if (energy_mesh) energy_mesh[p] = sum_e;
        rho[p] += final_mesh[p] * chrg;
    }
    free(corner);
    free(corner_e);
}

/* Fixed-point form of the same weights, as the GPU computes them: each
 * corner weight normalised by dx*dy and rounded to nearest at 2^cs (energy
 * weights at 2^es eV), summed as integers — order independent. */
void orc_deposit_fixed_f64(int64_t n, const double* x, const double* y, const double* vx, const double* vy,
                           const double* vz, const int32_t* key, int gxn, int gyn, double dx, double dy, int cs,
                           int es, double qm, int64_t* qc, int64_t* qe) {
    const int G = gxn * gyn;
    memset(qc, 0, (size_t)G * sizeof(int64_t));
    if (qe) memset(qe, 0, (size_t)G * sizeof(int64_t));
    const double inv_area = 1.0 / (dx * dy);
    const double sq = ldexp(1.0, cs), se = ldexp(1.0, es);
    const double denom = 1 / fabs(qm);
    for (int64_t i = 0; i < n; ++i) {
        int k = key[i];
        if (k < 0) continue;
        int gx = k % gxn, gy = k / gxn;
        if (gx + 1 >= gxn || gy + 1 >= gyn) continue;
        double lx = x[i] - gx * dx;
        double ly = y[i] - gy * dy;
        double w[4];
        w[0] = ((dx - lx) * (dy - ly)) * inv_area;
        w[1] = (lx * (dy - ly)) * inv_area;
        w[2] = ((dx - lx) * ly) * inv_area;
        w[3] = (lx * ly) * inv_area;
        int node[4] = {k, k + 1, k + gxn, k + gxn + 1};
        for (int c = 0; c < 4; ++c) qc[node[c]] += llrint(w[c] * sq);
        if (qe) {
            double v2 = vx[i] * vx[i] + vy[i] * vy[i] + vz[i] * vz[i];
            double en = 0.5 * v2 * denom;
            for (int c = 0; c < 4; ++c) qe[node[c]] += llrint((w[c] * en) * se);
        }
    }
}

void orc_deposit_fixed_cell32(int64_t n, const int32_t* cell, const float* ox, const float* oy, const float* vx,
                              const float* vy, const float* vz, int gxn, int gyn, int cs, int es, double qm,
                              int64_t* qc, int64_t* qe) {
    const int G = gxn * gyn;
    memset(qc, 0, (size_t)G * sizeof(int64_t));
    if (qe) memset(qe, 0, (size_t)G * sizeof(int64_t));
    const double sq = ldexp(1.0, cs), se = ldexp(1.0, es);
    const double denom = 1 / fabs(qm);
    for (int64_t i = 0; i < n; ++i) {
        int k = cell[i];
        if (k < 0) continue;
        int gx = k % gxn, gy = k / gxn;
        if (gx + 1 >= gxn || gy + 1 >= gyn) continue;
        double fx = (double)ox[i], fy = (double)oy[i];
        double w[4];
        w[0] = (1.0 - fx) * (1.0 - fy);
        w[1] = fx * (1.0 - fy);
        w[2] = (1.0 - fx) * fy;
        w[3] = fx * fy;
        int node[4] = {k, k + 1, k + gxn, k + gxn + 1};
        for (int c = 0; c < 4; ++c) qc[node[c]] += llrint(w[c] * sq);
        if (qe) {
            double pvx = vx[i], pvy = vy[i], pvz = vz[i];
            double v2 = pvx * pvx + pvy * pvy + pvz * pvz;
            double en = 0.5 * v2 * denom;
            for (int c = 0; c < 4; ++c) qe[node[c]] += llrint((w[c] * en) * se);
        }
    }
}

/* ------------------------------------------------------------- E field */
/* calculateElectricField (src/poissonSolver.cpp:433-456), literally: interior
 * central differences, then columns 0 / GX-1 copy their neighbours for every
 * row, then rows 0 / GY-1 copy theirs. */
void orc_efield(int gxn, int gyn, double dx, double dy, const double* phi, double* ex, double* ey) {
    for (int i = 1; i < gyn - 1; i++)
        for (int j = 1; j < gxn - 1; j++) {
            ex[i * gxn + j] = -((phi[i * gxn + (j + 1)] - phi[i * gxn + (j - 1)]) / (2 * dx));
            ey[i * gxn + j] = -((phi[(i + 1) * gxn + j] - phi[(i - 1) * gxn + j]) / (2 * dy));
        }
    for (int i = 0; i < gyn; i++) {
        ex[i * gxn] = ex[i * gxn + 1];
        ey[i * gxn] = ey[i * gxn + 1];
        ex[(i + 1) * gxn - 1] = ex[(i + 1) * gxn - 2];
        ey[(i + 1) * gxn - 1] = ey[(i + 1) * gxn - 2];
    }
    for (int i = 0; i < gxn; i++) {
        ex[i] = ex[gxn + i];
        ey[i] = ey[gxn + i];
        ex[gxn * (gyn - 1) + i] = ex[gxn * (gyn - 2) + i];
        ey[gxn * (gyn - 1) + i] = ey[gxn * (gyn - 2) + i];
    }
}

/* calculateElectricFieldPeriodic (src/poissonSolver.cpp:411-430): Ex one-sided
 * at the first and last column, central inside; Ey central with row indices
 * taken modulo GRID_Y-1, for every row including the image row. */
void orc_efield_periodic(int gxn, int gyn, double dx, double dy, const double* phi, double* ex, double* ey) {
    for (int i = 0; i < gyn; i++) {
        ex[i * gxn] = -((phi[i * gxn + 1] - phi[i * gxn]) / dx);
        for (int j = 1; j < gxn - 1; j++)
            ex[i * gxn + j] = -((phi[i * gxn + (j + 1)] - phi[i * gxn + (j - 1)]) / (2 * dx));
        ex[(i + 1) * gxn - 1] = -((phi[(i + 1) * gxn - 1] - phi[(i + 1) * gxn - 2]) / dx);
    }
    for (int j = 0; j < gxn; j++)
        for (int i = 0; i < gyn; i++)
            ey[i * gxn + j] = -((phi[((i + 1) % (gyn - 1)) * gxn + j] -
                                 phi[(((i - 1) + gyn - 1) % (gyn - 1)) * gxn + j]) / (2 * dy));
}

/* modulo (src/mover.cpp:13-22): the periodic-y wrap of the mover (:317-320). */
double orc_modulo(double x, double y) {
    if (x < 0) {
        if ((int)(x / y) == x / y) return 0;
        return y - fmod(-x, y);
    }
    return fmod(x, y);
}

/* ---------------------------------------------------------------- push */
static inline double fld_at(const double* f, int64_t G, int64_t idx) { return idx < G ? f[idx] : 0.0; }

/* Boris velocity update of deterministic_mover (src/mover.cpp:151-162) for
 * Bz; with bx or by non-zero the same rotation with the full t = (q/m) B dt/2
 * vector: v' = v- + v- x t, v+ = v- + v' x s, s = 2t * (1 / (1 + |t|^2)). */
static inline void orc_boris(double Epx, double Epy, double Bpz, double pctm, double dt, double bx, double by,
                             double* pvx_, double* pvy_, double* pvz_) {
    double pvx = *pvx_, pvy = *pvy_, pvz = *pvz_;
    double Vxminus = pvx + 0.5 * pctm * Epx * dt;
    double Vyminus = pvy + 0.5 * pctm * Epy * dt;
    if (bx == 0.0 && by == 0.0) {
        double t = 0.5 * pctm * Bpz * dt;
        double Vxprime = Vxminus + Vyminus * t;
        double Vyprime = Vyminus - Vxminus * t;
        double s = 2 * t / (1 + t * t);
        pvx = Vxminus + Vyprime * s + 0.5 * pctm * Epx * dt;
        pvy = Vyminus - Vxprime * s + 0.5 * pctm * Epy * dt;
    } else {
        double tx = 0.5 * pctm * bx * dt;
        double ty = 0.5 * pctm * by * dt;
        double tz = 0.5 * pctm * Bpz * dt;
        double Vzminus = pvz;
        double vpx = Vxminus + (Vyminus * tz - Vzminus * ty);
        double vpy = Vyminus + (Vzminus * tx - Vxminus * tz);
        double vpz = Vzminus + (Vxminus * ty - Vyminus * tx);
        double rden = 1 / (1 + (tx * tx + ty * ty + tz * tz));
        double sx = (2 * tx) * rden, sy = (2 * ty) * rden, sz = (2 * tz) * rden;
        pvx = Vxminus + (vpy * sz - vpz * sy) + 0.5 * pctm * Epx * dt;
        pvy = Vyminus + (vpz * sx - vpx * sz) + 0.5 * pctm * Epy * dt;
        pvz = Vzminus + (vpx * sy - vpy * sx);
    }
    *pvx_ = pvx;
    *pvy_ = pvy;
    *pvz_ = pvz;
}

/* deterministic_mover (src/mover.cpp:103-176) + the boundary test and new key
 * of MoverKernel (:304-364).  Gather: gx = (int)(x/dx) (division), lx = x - gx*dx,
 * bilinear sum over (dx*dy) at flat nodes gy*GX+gx, +1, +GX, +GX+1 (reads past
 * the last node see 0).  Absorbed (returns -1) when x >= W, x < 0, y >= H or
 * y < 0 (:325); absorbed[0] counts x < 0 (:344-352), absorbed[1] all. */
static int32_t push_one(double* px_, double* py_, double* pvx_, double* pvy_, double* pvz_, const double* ex,
                        const double* ey, const double* bz, int gxn, int gyn, double dx, double dy, double width,
                        double height, double qm, double dt, double bx, double by, int periodic,
                        int64_t* absorbed) {
    const int64_t G = (int64_t)gxn * gyn;
    double px = *px_, py = *py_;
    int gx = (int)(px / dx);
    int gy = (int)(py / dy);
    double lx = px - gx * dx;
    double ly = py - gy * dy;
    int64_t i00 = (int64_t)gy * gxn + gx;
    int64_t i10 = i00 + 1, i01 = i00 + gxn, i11 = i00 + gxn + 1;
    double Epx = (fld_at(ex, G, i00) * (dx - lx) * (dy - ly) + fld_at(ex, G, i10) * lx * (dy - ly) +
                  fld_at(ex, G, i01) * (dx - lx) * ly + fld_at(ex, G, i11) * lx * ly) / (dx * dy);
    double Epy = (fld_at(ey, G, i00) * (dx - lx) * (dy - ly) + fld_at(ey, G, i10) * lx * (dy - ly) +
                  fld_at(ey, G, i01) * (dx - lx) * ly + fld_at(ey, G, i11) * lx * ly) / (dx * dy);
    double Bpz = (fld_at(bz, G, i00) * (dx - lx) * (dy - ly) + fld_at(bz, G, i10) * lx * (dy - ly) +
                  fld_at(bz, G, i01) * (dx - lx) * ly + fld_at(bz, G, i11) * lx * ly) / (dx * dy);
    orc_boris(Epx, Epy, Bpz, qm, dt, bx, by, pvx_, pvy_, pvz_);
    px = px + dt * (*pvx_);
    py = py + dt * (*pvy_);
    if (periodic) py = orc_modulo(py, height);   /* mover.cpp:317-320 */
    *px_ = px;
    *py_ = py;
    int out = (px >= width || px < 0.0 || py >= height || py < 0);
    if (out) {
        if (absorbed) {
            if (px < 0.0) absorbed[0]++;
            absorbed[1]++;
        }
        return -1;
    }
    (void)gyn;
    return orc_key(px, py, gxn, dx, dy);
}

int32_t orc_push_one(double* px_, double* py_, double* pvx_, double* pvy_, double* pvz_, const double* ex,
                     const double* ey, const double* bz, int gxn, int gyn, double dx, double dy, double width,
                     double height, double qm, double dt, double bx, double by, int64_t* absorbed) {
    return push_one(px_, py_, pvx_, pvy_, pvz_, ex, ey, bz, gxn, gyn, dx, dy, width, height, qm, dt, bx, by, 0,
                    absorbed);
}

void orc_push_f64_p(int64_t n, double* x, double* y, double* vx, double* vy, double* vz, int32_t* key,
                    const double* ex, const double* ey, const double* bz, int gxn, int gyn, double dx, double dy,
                    double width, double height, double qm, double dt, double bx, double by, int periodic,
                    int64_t* absorbed) {
    int64_t a0 = 0, a1 = 0;
#pragma omp parallel for schedule(static) reduction(+ : a0, a1)
    for (int64_t i = 0; i < n; ++i) {
        int64_t ab[2] = {0, 0};
        key[i] = push_one(&x[i], &y[i], &vx[i], &vy[i], &vz[i], ex, ey, bz, gxn, gyn, dx, dy, width, height,
                          qm, dt, bx, by, periodic, ab);
        a0 += ab[0];
        a1 += ab[1];
    }
    if (absorbed) {
        absorbed[0] += a0;
        absorbed[1] += a1;
    }
}

void orc_push_f64(int64_t n, double* x, double* y, double* vx, double* vy, double* vz, int32_t* key,
                  const double* ex, const double* ey, const double* bz, int gxn, int gyn, double dx, double dy,
                  double width, double height, double qm, double dt, double bx, double by, int64_t* absorbed) {
    orc_push_f64_p(n, x, y, vx, vy, vz, key, ex, ey, bz, gxn, gyn, dx, dy, width, height, qm, dt, bx, by, 0,
                   absorbed);
}

/* CELL32 storage (the build's compact layout): key cell (cgx, cgy), float
 * offsets (fx, fy) in cell units, float velocities.  Push in FP64 in cell
 * units: bilinear weights (1-fx)(1-fy), fx(1-fy), (1-fx)fy, fx*fy on nodes
 * cell, +1, +GX, +GX+1; the reference's Boris step; new position
 * t = (cgx + fx) + v * (dt * (1.0/dx)); absorbed when t >= NX or t < 0 (same
 * for y); new cell (int)t and offset (float)(t - (int)t) kept below 1.
 * When every node holds the same Bz (the reference's initMagneticField,
 * src/particles.cpp:709-719), the rotation is the same for every particle:
 * t = (0.5 q/m) B dt and s = 2t / (1 + |t|^2) are formed once from that
 * value (instead of from the gathered Bz, which can differ from it in the
 * last bit), and the half kick is ce * E with ce = (0.5 q/m) dt. */
static inline float clamp_offset(float o) { return o < 1.0f ? o : 0x1.fffffep-1f; }

/* The Boris rotation of orc_boris with constant t and s (the uniform-B form). */
static inline void orc_boris_const(double hx, double hy, int genb, const double* t, const double* s,
                                   double* pvx_, double* pvy_, double* pvz_) {
    double Vxminus = *pvx_ + hx;
    double Vyminus = *pvy_ + hy;
    if (!genb) {
        double Vxprime = Vxminus + Vyminus * t[2];
        double Vyprime = Vyminus - Vxminus * t[2];
        *pvx_ = Vxminus + Vyprime * s[2] + hx;
        *pvy_ = Vyminus - Vxprime * s[2] + hy;
    } else {
        double Vzminus = *pvz_;
        double vpx = Vxminus + (Vyminus * t[2] - Vzminus * t[1]);
        double vpy = Vyminus + (Vzminus * t[0] - Vxminus * t[2]);
        double vpz = Vzminus + (Vxminus * t[1] - Vyminus * t[0]);
        *pvx_ = Vxminus + (vpy * s[2] - vpz * s[1]) + hx;
        *pvy_ = Vyminus + (vpz * s[0] - vpx * s[2]) + hy;
        *pvz_ = Vzminus + (vpx * s[1] - vpy * s[0]);
    }
}

void orc_push_cell32_p(int64_t n, int32_t* cell, float* ox, float* oy, float* vx, float* vy, float* vz,
                       int32_t* key, const double* ex, const double* ey, const double* bz, int gxn, int gyn,
                       double dx, double dy, double width, double height, double qm, double dt, double bx,
                       double by, int periodic, int64_t* absorbed) {
    int64_t a0 = 0, a1 = 0;
    const int64_t G = (int64_t)gxn * gyn;
    const double dtdx = dt * (1.0 / dx), dtdy = dt * (1.0 / dy);
    const double NX = gxn - 1, NY = gyn - 1;
    (void)width;
    (void)height;
    int uniform = 1;
    for (int64_t k = 1; k < G && uniform; ++k)
        if (bz[k] != bz[0]) uniform = 0;
    const int genb = bx != 0.0 || by != 0.0;
    const double half_qm = 0.5 * qm, ce = half_qm * dt;
    double ut[3] = {half_qm * bx * dt, half_qm * by * dt, half_qm * bz[0] * dt}, us[3] = {0.0, 0.0, 0.0};
    if (genb) {
        double rden = 1 / (1 + (ut[0] * ut[0] + ut[1] * ut[1] + ut[2] * ut[2]));
        for (int k = 0; k < 3; ++k) us[k] = (2 * ut[k]) * rden;
    } else {
        us[2] = 2 * ut[2] / (1 + ut[2] * ut[2]);
    }
#pragma omp parallel for schedule(static) reduction(+ : a0, a1)
    for (int64_t i = 0; i < n; ++i) {
        int c = cell[i];
        int cgx = c % gxn, cgy = c / gxn;
        double fx = (double)ox[i], fy = (double)oy[i];
        double w00 = (1.0 - fx) * (1.0 - fy), w10 = fx * (1.0 - fy), w01 = (1.0 - fx) * fy, w11 = fx * fy;
        int64_t n00 = c, n10 = c + 1, n01 = c + gxn, n11 = c + gxn + 1;
        double Epx = fld_at(ex, G, n00) * w00 + fld_at(ex, G, n10) * w10 + fld_at(ex, G, n01) * w01 +
                     fld_at(ex, G, n11) * w11;
        double Epy = fld_at(ey, G, n00) * w00 + fld_at(ey, G, n10) * w10 + fld_at(ey, G, n01) * w01 +
                     fld_at(ey, G, n11) * w11;
        double pvx = vx[i], pvy = vy[i], pvz = vz[i];
        if (uniform) {
            orc_boris_const(ce * Epx, ce * Epy, genb, ut, us, &pvx, &pvy, &pvz);
        } else {
            double Bpz = fld_at(bz, G, n00) * w00 + fld_at(bz, G, n10) * w10 + fld_at(bz, G, n01) * w01 +
                         fld_at(bz, G, n11) * w11;
            orc_boris(Epx, Epy, Bpz, qm, dt, bx, by, &pvx, &pvy, &pvz);
        }
        double tx = ((double)cgx + fx) + pvx * dtdx;
        double ty = ((double)cgy + fy) + pvy * dtdy;
        if (periodic) ty = orc_modulo(ty, NY);   /* the wrap in cell units, period NY */
        if (tx >= NX || tx < 0.0 || ty >= NY || ty < 0.0) {
            if (tx < 0.0) a0++;
            a1++;
            key[i] = -1;
            continue;
        }
        int ngx = (int)tx, ngy = (int)ty;
        int32_t k = ngy * gxn + ngx;
        key[i] = k;
        cell[i] = k;
        ox[i] = clamp_offset((float)(tx - (double)ngx));
        oy[i] = clamp_offset((float)(ty - (double)ngy));
        vx[i] = (float)pvx;
        vy[i] = (float)pvy;
        vz[i] = (float)pvz;
    }
    if (absorbed) {
        absorbed[0] += a0;
        absorbed[1] += a1;
    }
}

void orc_push_cell32(int64_t n, int32_t* cell, float* ox, float* oy, float* vx, float* vy, float* vz, int32_t* key,
                     const double* ex, const double* ey, const double* bz, int gxn, int gyn, double dx, double dy,
                     double width, double height, double qm, double dt, double bx, double by, int64_t* absorbed) {
    orc_push_cell32_p(n, cell, ox, oy, vx, vy, vz, key, ex, ey, bz, gxn, gyn, dx, dy, width, height, qm, dt, bx,
                      by, 0, absorbed);
}

/* Encode absolute positions into CELL32 storage (the GPU upload path). */
void orc_encode_cell32(int64_t n, const double* x, const double* y, const double* vx, const double* vy,
                       const double* vz, int gxn, double dx, double dy, int32_t* cell, float* ox, float* oy,
                       float* fvx, float* fvy, float* fvz) {
    const double inv_dx = 1.0 / dx, inv_dy = 1.0 / dy;
    for (int64_t i = 0; i < n; ++i) {
        double tx = x[i] * inv_dx, ty = y[i] * inv_dy;
        int gx = (int)tx, gy = (int)ty;
        cell[i] = gy * gxn + gx;
        ox[i] = clamp_offset((float)(tx - (double)gx));
        oy[i] = clamp_offset((float)(ty - (double)gy));
        fvx[i] = (float)vx[i];
        fvy[i] = (float)vy[i];
        fvz[i] = (float)vz[i];
    }
}

/* ------------------------------------------------------------- Poisson */
/* readGeometry node rule (src/particles.cpp:677-681): (int)((GRID-1) * frac). */
void orc_geometry_nodes(int gxn, int gyn, int n, const double* rects, int* nodes, double* pot) {
    for (int i = 0; i < n; ++i) {
        nodes[4 * i + 0] = (int)((gxn - 1) * rects[5 * i + 0]);
        nodes[4 * i + 1] = (int)((gyn - 1) * rects[5 * i + 1]);
        nodes[4 * i + 2] = (int)((gxn - 1) * rects[5 * i + 2]);
        nodes[4 * i + 3] = (int)((gyn - 1) * rects[5 * i + 3]);
        pot[i] = rects[5 * i + 4];
    }
}

/* setPotentialBoundary (src/poissonSolver.cpp:64-79) plus the value each
 * Dirichlet node gets from overwriteRho / overwritePhi (:364-402): inclusive
 * rectangles in order, the last one covering a node wins. */
void orc_dirichlet(int gxn, int gyn, int ngeom, const int* nodes, const double* pot, uint8_t* mask, double* dval) {
    const int G = gxn * gyn;
    memset(mask, 0, (size_t)G);
    memset(dval, 0, (size_t)G * sizeof(double));
    for (int i = 0; i < ngeom; ++i)
        for (int k = nodes[4 * i + 1]; k <= nodes[4 * i + 3]; k++)
            for (int j = nodes[4 * i + 0]; j <= nodes[4 * i + 2]; j++) {
                mask[k * gxn + j] = 1;
                dval[k * gxn + j] = pot[i];
            }
}

/* overwriteRho (:384-402): rho[geometry] = V, then rho = -rho. */
void orc_overwrite_rho(int G, const uint8_t* mask, const double* dval, double* rho) {
    for (int i = 0; i < G; ++i)
        if (mask[i]) rho[i] = dval[i];
    for (int i = 0; i < G; ++i) rho[i] = -rho[i];
}

/* The reference's CSR matrix (generateCoeffs :183-241), 1-based: Dirichlet
 * row -1 on the diagonal; other rows [-1/dy^2, -1/dx^2, 2(1/dx^2+1/dy^2),
 * -1/dx^2, -1/dy^2] at columns i-GX, i-1, i, i+1, i+GX.  Returns nnz. */
int64_t orc_assemble(int gxn, int gyn, double dx, double dy, const uint8_t* mask, int* ia, int* ja, double* a) {
    const int G = gxn * gyn;
    const double inv_dx2 = 1.0 / (dx * dx), inv_dy2 = 1.0 / (dy * dy);
    int64_t index = 0;
    ia[0] = 1;
    for (int i = 0; i < G; i++) {
        if (mask[i] == 1) {
            a[index] = -1;
            ja[index] = i + 1;
            index++;
            ia[i + 1] = ia[i] + 1;
        } else {
            a[index] = -inv_dy2;
            a[index + 1] = -inv_dx2;
            a[index + 2] = 2 * (inv_dx2 + inv_dy2);
            a[index + 3] = -inv_dx2;
            a[index + 4] = -inv_dy2;
            ja[index] = i - (gxn - 1) - 1 + 1;
            ja[index + 1] = i - 1 + 1;
            ja[index + 2] = i + 1;
            ja[index + 3] = i + 1 + 1;
            ja[index + 4] = i + (gxn - 1) + 1 + 1;
            ia[i + 1] = ia[i] + 5;
            index += 5;
        }
    }
    return index;
}

/* Residual of the reference system: ||rhs - A phi||_2 / ||rhs||_2 over
 * non-Dirichlet rows, where rhs = rho after overwriteRho. */
double orc_relres(int gxn, int gyn, double dx, double dy, const uint8_t* mask, const double* rhs, const double* phi,
                  const double* dval) {
    const int G = gxn * gyn;
    const double cx = 1.0 / (dx * dx), cy = 1.0 / (dy * dy), diag = 2 * (cx + cy);
    double rr = 0, bb = 0;
    for (int n = 0; n < G; ++n) {
        if (mask[n]) continue;
        double ax = diag * phi[n] - cx * (phi[n - 1] + phi[n + 1]) - cy * (phi[n - gxn] + phi[n + gxn]);
        double r = rhs[n] - ax;
        /* right-hand side of the reduced system: rhs + Dirichlet neighbours */
        double b = rhs[n];
        if (mask[n - 1]) b += cx * dval[n - 1];
        if (mask[n + 1]) b += cx * dval[n + 1];
        if (mask[n - gxn]) b += cy * dval[n - gxn];
        if (mask[n + gxn]) b += cy * dval[n + gxn];
        rr += r * r;
        bb += b * b;
    }
    return bb > 0 ? sqrt(rr / bb) : sqrt(rr);
}

/* Direct FP64 solve of the reference system (the role PARDISO plays at
 * :505-535), by band Cholesky of the symmetric form: Dirichlet rows/columns
 * become identity and their values move to the right-hand side.  Band width
 * GRID_X; intended for grids up to ~257 x 257.  phi gets V on Dirichlet nodes
 * (overwritePhi).  Returns 0, or -1 if the matrix is not positive definite. */
int orc_poisson_direct(int gxn, int gyn, double dx, double dy, const uint8_t* mask, const double* dval,
                       const double* rhs, double* phi) {
    const int64_t G = (int64_t)gxn * gyn;
    const int bw = gxn;                      /* lower band width */
    const double cx = 1.0 / (dx * dx), cy = 1.0 / (dy * dy), diag = 2 * (cx + cy);
    /* band storage: L[i][d] = A[i][i-d], d = 0..bw */
    double* L = (double*)calloc((size_t)G * (bw + 1), sizeof(double));
    double* b = (double*)malloc((size_t)G * sizeof(double));
    if (!L || !b) {
        free(L);
        free(b);
        return -2;
    }
#define BAND(i, d) L[(size_t)(i) * (bw + 1) + (d)]
    for (int64_t n = 0; n < G; ++n) {
        if (mask[n]) {
            BAND(n, 0) = 1.0;
            b[n] = dval[n];
            continue;
        }
        BAND(n, 0) = diag;
        b[n] = rhs[n];
        int64_t nb[4] = {n - 1, n + 1, n - gxn, n + gxn};
        double cf[4] = {cx, cx, cy, cy};
        for (int q = 0; q < 4; ++q) {
            int64_t m = nb[q];
            if (mask[m]) {
                b[n] += cf[q] * dval[m];
            } else if (m < n) {
                BAND(n, n - m) = -cf[q];
            }
        }
    }
    /* in-place band Cholesky: A = L L^T */
    for (int64_t j = 0; j < G; ++j) {
        double s = BAND(j, 0);
        int64_t k0 = j - bw < 0 ? 0 : j - bw;
        for (int64_t k = k0; k < j; ++k) {
            double l = BAND(j, j - k);
            s -= l * l;
        }
        if (!(s > 0)) {
            free(L);
            free(b);
            return -1;
        }
        double d = sqrt(s);
        BAND(j, 0) = d;
        int64_t iend = j + bw >= G ? G - 1 : j + bw;
        for (int64_t i = j + 1; i <= iend; ++i) {
            double t = BAND(i, i - j);
            int64_t kk0 = i - bw < 0 ? 0 : i - bw;
            if (kk0 < k0) kk0 = k0;
            for (int64_t k = kk0; k < j; ++k) t -= BAND(i, i - k) * BAND(j, j - k);
            BAND(i, i - j) = t / d;
        }
    }
    /* forward: L y = b */
    for (int64_t i = 0; i < G; ++i) {
        double t = b[i];
        int64_t k0 = i - bw < 0 ? 0 : i - bw;
        for (int64_t k = k0; k < i; ++k) t -= BAND(i, i - k) * b[k];
        b[i] = t / BAND(i, 0);
    }
    /* backward: L^T x = y */
    for (int64_t i = G - 1; i >= 0; --i) {
        double t = b[i];
        int64_t kend = i + bw >= G ? G - 1 : i + bw;
        for (int64_t k = i + 1; k <= kend; ++k) t -= BAND(k, k - i) * b[k];
        b[i] = t / BAND(i, 0);
    }
#undef BAND
    for (int64_t n = 0; n < G; ++n) phi[n] = mask[n] ? dval[n] : b[n];
    free(L);
    free(b);
    return 0;
}

/* --------------------------------------------------------------- PARDISO */
#ifdef ORC_HAVE_MKL
/* Public LP64 interface of Intel oneMKL PARDISO (MKL_INT = int). */
extern void pardisoinit(void* pt, const int* mtype, int* iparm);
extern void pardiso(void* pt, const int* maxfct, const int* mnum, const int* mtype, const int* phase, const int* n,
                    const double* a, const int* ia, const int* ja, int* perm, const int* nrhs, int* iparm,
                    const int* msglvl, double* b, double* x, int* error);
#endif

typedef struct {
    void* pt[64];
    int iparm[64];
    int n, mtype;
    int* ia;
    int* ja;
    double* a;
} OrcPardiso;

int orc_have_pardiso(void) {
#ifdef ORC_HAVE_MKL
    return 1;
#else
    return 0;
#endif
}

/* initPardiso + phases 11 and 22 (:256-318, :505-525): mtype 11, defaults
 * from pardisoinit. */
void* orc_pardiso_create(int gxn, int gyn, double dx, double dy, const uint8_t* mask) {
#ifdef ORC_HAVE_MKL
    /* GNU OpenMP threading layer: the oracle itself uses libgomp */
    setenv("MKL_THREADING_LAYER", "GNU", 0);
    OrcPardiso* h = (OrcPardiso*)calloc(1, sizeof(OrcPardiso));
    const int G = gxn * gyn;
    h->n = G;
    h->mtype = 11;
    h->ia = (int*)malloc((size_t)(G + 1) * sizeof(int));
    h->ja = (int*)malloc((size_t)5 * G * sizeof(int));
    h->a = (double*)malloc((size_t)5 * G * sizeof(double));
    orc_assemble(gxn, gyn, dx, dy, mask, h->ia, h->ja, h->a);
    pardisoinit(h->pt, &h->mtype, h->iparm);
    int maxfct = 1, mnum = 1, nrhs = 1, msglvl = 0, error = 0, phase = 11;
    double dd = 0;
    pardiso(h->pt, &maxfct, &mnum, &h->mtype, &phase, &h->n, h->a, h->ia, h->ja, NULL, &nrhs, h->iparm, &msglvl, &dd,
            &dd, &error);
    if (error != 0) {
        fprintf(stderr, "PARDISO phase 11 error: %d\n", error);
        return NULL;
    }
    phase = 22;
    pardiso(h->pt, &maxfct, &mnum, &h->mtype, &phase, &h->n, h->a, h->ia, h->ja, NULL, &nrhs, h->iparm, &msglvl, &dd,
            &dd, &error);
    if (error != 0) {
        fprintf(stderr, "PARDISO phase 22 error: %d\n", error);
        return NULL;
    }
    return h;
#else
    (void)gxn; (void)gyn; (void)dx; (void)dy; (void)mask;
    return NULL;
#endif
}

/* phase 33 (:527-535): solve A phi = rhs (rhs = rho after overwriteRho). */
int orc_pardiso_solve(void* hv, double* rhs, double* phi) {
#ifdef ORC_HAVE_MKL
    OrcPardiso* h = (OrcPardiso*)hv;
    int maxfct = 1, mnum = 1, nrhs = 1, msglvl = 0, error = 0, phase = 33;
    pardiso(h->pt, &maxfct, &mnum, &h->mtype, &phase, &h->n, h->a, h->ia, h->ja, NULL, &nrhs, h->iparm, &msglvl, rhs,
            phi, &error);
    return error;
#else
    (void)hv; (void)rhs; (void)phi;
    return -99;
#endif
}

void orc_pardiso_destroy(void* hv) {
#ifdef ORC_HAVE_MKL
    OrcPardiso* h = (OrcPardiso*)hv;
    if (!h) return;
    int maxfct = 1, mnum = 1, nrhs = 1, msglvl = 0, error = 0, phase = -1;
    double dd = 0;
    pardiso(h->pt, &maxfct, &mnum, &h->mtype, &phase, &h->n, h->a, h->ia, h->ja, NULL, &nrhs, h->iparm, &msglvl, &dd,
            &dd, &error);
    free(h->ia);
    free(h->ja);
    free(h->a);
    free(h);
#else
    (void)hv;
#endif
}

int orc_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
