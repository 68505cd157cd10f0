/*
 * pic_baseline.c — the CPU baseline: the reference's time step run on the
 * host cores with the oracle's restated stages.  TEST INFRASTRUCTURE ONLY
 * (timed by bench.py's CPU-baseline leg and used by the multi-step checks).
 *
 * Particles are FP64 array-of-structs records like the reference's Point
 * (include/point.hpp:11-14).  One step follows src/main.cpp:130-257:
 * zero rho; deposit each species (interpolate); Poisson (overwriteRho,
 * PARDISO phase 33 or the oracle's direct solve, overwritePhi); E field;
 * push each species; re-bin (the reference's hashmap insert/delete + cleanup,
 * here a stable counting sort by the reference key).  OpenMP parallelises
 * over particles and over cells; no floating-point sum changes order.
 */
#include "pic_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef struct {
    double x, y, z, vx, vy, vz;
    int32_t cell;
} OrcPoint;

typedef struct {
    OrcPoint* p;
    OrcPoint* q;
    int64_t n, cap;
    double qm, charge, W;
    double* final_mesh;
    double* energy_mesh;
    int32_t* start;
    int64_t absorbed[2];
} OrcSpecies;

struct OrcSim {
    int gxn, gyn, G;
    double dx, dy, width, height, dt, eps, bx, by;
    int ns;
    OrcSpecies sp[4];
    double *rho, *phi, *ex, *ey, *bz, *rhs;
    uint8_t* mask;
    double* dval;
    void* pardiso;
    int deposit_energy;
    double t[5];
    int32_t* hist;      /* threads x G */
    int nthreads;
};

static double now_s(void) {
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec + 1e-9 * ts.tv_nsec;
}

OrcSim* orc_sim_create(int NX, int NY, double width, double height, double dt, double eps, double b0, double bx,
                       double by, int ngeom, const double* rects, int deposit_energy, int use_pardiso) {
    OrcSim* s = (OrcSim*)calloc(1, sizeof(OrcSim));
    s->gxn = NX + 1;
    s->gyn = NY + 1;
    s->G = s->gxn * s->gyn;
    s->width = width;
    s->height = height;
    s->dx = width / NX;
    s->dy = height / NY;
    s->dt = dt;
    s->eps = eps;
    s->bx = bx;
    s->by = by;
    s->deposit_energy = deposit_energy;
    const size_t G = (size_t)s->G;
    s->rho = (double*)calloc(G, sizeof(double));
    s->rhs = (double*)calloc(G, sizeof(double));
    s->phi = (double*)calloc(G, sizeof(double));
    s->ex = (double*)calloc(G, sizeof(double));
    s->ey = (double*)calloc(G, sizeof(double));
    s->bz = (double*)malloc(G * sizeof(double));
    s->mask = (uint8_t*)calloc(G, 1);
    s->dval = (double*)calloc(G, sizeof(double));
    for (size_t i = 0; i < G; ++i) s->bz[i] = b0;
    int* nodes = (int*)malloc((size_t)4 * (ngeom > 0 ? ngeom : 1) * sizeof(int));
    double* pot = (double*)malloc((size_t)(ngeom > 0 ? ngeom : 1) * sizeof(double));
    orc_geometry_nodes(s->gxn, s->gyn, ngeom, rects, nodes, pot);
    orc_dirichlet(s->gxn, s->gyn, ngeom, nodes, pot, s->mask, s->dval);
    free(nodes);
    free(pot);
    for (size_t i = 0; i < G; ++i)
        if (s->mask[i]) s->phi[i] = s->dval[i];
    if (use_pardiso && orc_have_pardiso()) s->pardiso = orc_pardiso_create(s->gxn, s->gyn, s->dx, s->dy, s->mask);
#ifdef _OPENMP
    s->nthreads = omp_get_max_threads();
#else
    s->nthreads = 1;
#endif
    s->hist = (int32_t*)malloc((size_t)s->nthreads * G * sizeof(int32_t));
    return s;
}

int orc_sim_uses_pardiso(OrcSim* s) { return s->pardiso != NULL; }

/* stable counting sort of sp->p[0..n) by .cell into sp->q, dropping cell < 0 */
static void bin_species(OrcSim* s, OrcSpecies* sp) {
    const int T = s->nthreads, G = s->G;
    const int64_t n = sp->n;
    const int64_t chunk = (n + T - 1) / T;
    int32_t* hist = s->hist;
#pragma omp parallel num_threads(T)
    {
#ifdef _OPENMP
        int t = omp_get_thread_num();
#else
        int t = 0;
#endif
        int32_t* h = hist + (size_t)t * G;
        memset(h, 0, (size_t)G * sizeof(int32_t));
        int64_t a = t * chunk, b = a + chunk < n ? a + chunk : n;
        for (int64_t i = a; i < b; ++i)
            if (sp->p[i].cell >= 0) h[sp->p[i].cell]++;
    }
    /* offsets: key-major, thread-minor (keeps index order within a key) */
    int64_t run = 0;
    for (int k = 0; k < G; ++k) {
        sp->start[k] = (int32_t)run;
        for (int t = 0; t < T; ++t) {
            int32_t c = hist[(size_t)t * G + k];
            hist[(size_t)t * G + k] = (int32_t)run;
            run += c;
        }
    }
    sp->start[G] = (int32_t)run;
#pragma omp parallel num_threads(T)
    {
#ifdef _OPENMP
        int t = omp_get_thread_num();
#else
        int t = 0;
#endif
        int32_t* h = hist + (size_t)t * G;
        int64_t a = t * chunk, b = a + chunk < n ? a + chunk : n;
        for (int64_t i = a; i < b; ++i) {
            int k = sp->p[i].cell;
            if (k >= 0) sp->q[h[k]++] = sp->p[i];
        }
    }
    OrcPoint* tmp = sp->p;
    sp->p = sp->q;
    sp->q = tmp;
    sp->n = run;
}

int orc_sim_add_species(OrcSim* s, double charge, double mass, double sim_count, double actual, int64_t n,
                        const double* x, const double* y, const double* vx, const double* vy, const double* vz) {
    if (s->ns >= 4) return -1;
    OrcSpecies* sp = &s->sp[s->ns];
    sp->qm = charge / mass;
    sp->charge = charge;
    /* updateWeighingFactor (src/particles.cpp:303-305), surfArea = W*H (:431) */
    sp->W = actual * (s->width * s->height) / (sim_count * s->dx * s->dy * s->dx * s->dy);
    sp->cap = n;
    sp->n = n;
    sp->p = (OrcPoint*)malloc((size_t)(n > 0 ? n : 1) * sizeof(OrcPoint));
    sp->q = (OrcPoint*)malloc((size_t)(n > 0 ? n : 1) * sizeof(OrcPoint));
    sp->final_mesh = (double*)calloc((size_t)s->G, sizeof(double));
    sp->energy_mesh = (double*)calloc((size_t)s->G, sizeof(double));
    sp->start = (int32_t*)calloc((size_t)s->G + 1, sizeof(int32_t));
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) {
        sp->p[i].x = x[i];
        sp->p[i].y = y[i];
        sp->p[i].z = 0;
        sp->p[i].vx = vx[i];
        sp->p[i].vy = vy[i];
        sp->p[i].vz = vz[i];
        sp->p[i].cell = orc_key(x[i], y[i], s->gxn, s->dx, s->dy);
    }
    bin_species(s, sp);
    s->ns++;
    return s->ns - 1;
}

/* interpolate over the binned AoS (src/interpolation.cpp:13-176) */
static void deposit_species(OrcSim* s, OrcSpecies* sp) {
    const int G = s->G, gxn = s->gxn, gyn = s->gyn;
    const double dx = s->dx, dy = s->dy, W = sp->W;
    const double denom = 1 / fabs(sp->qm);
    const int energy = s->deposit_energy;
    double* fm = sp->final_mesh;
    double* em = sp->energy_mesh;
    /* each node sums its four corner contributions in corner order 0..3 */
#pragma omp parallel for schedule(static)
    for (int p = 0; p < G; ++p) {
        fm[p] = 0;
        em[p] = 0;
    }
    double* corner = (double*)malloc((size_t)8 * G * sizeof(double));
#pragma omp parallel for schedule(static)
    for (size_t i = 0; i < (size_t)8 * G; ++i) corner[i] = 0;
#pragma omp parallel for schedule(dynamic, 512)
    for (int k = 0; k < G; ++k) {
        const int gx = k % gxn, gy = k / gxn;
        if ((gx + 1) >= gxn || (gy + 1) >= gyn) continue;
        double s1 = 0, s2 = 0, s3 = 0, s4 = 0, s1e = 0, s2e = 0, s3e = 0, s4e = 0;
        for (int32_t m = sp->start[k]; m < sp->start[k + 1]; ++m) {
            const OrcPoint* q = &sp->p[m];
            double lx = q->x - gx * dx;
            double ly = q->y - gy * dy;
            s1 += (dx - lx) * (dy - ly);
            s2 += lx * (dy - ly);
            s3 += (dx - lx) * ly;
            s4 += lx * ly;
            if (energy) {
                double v2 = q->vx * q->vx + q->vy * q->vy + q->vz * q->vz;
                double enrgy = 0.5 * v2 * denom;
                s1e += ((dx - lx) * (dy - ly)) * enrgy;
                s2e += (lx * (dy - ly)) * enrgy;
                s3e += (dx - lx) * ly * enrgy;
                s4e += lx * ly * enrgy;
            }
        }
        corner[0 * (size_t)G + k] = s1 * W;
        corner[1 * (size_t)G + k + 1] = s2 * W;
        corner[2 * (size_t)G + k + gxn] = s3 * W;
        corner[3 * (size_t)G + k + gxn + 1] = s4 * W;
        corner[4 * (size_t)G + k] = s1e * W;
        corner[5 * (size_t)G + k + 1] = s2e * W;
        corner[6 * (size_t)G + k + gxn] = s3e * W;
        corner[7 * (size_t)G + k + gxn + 1] = s4e * W;
    }
    const double chrg = -sp->charge / s->eps;
#pragma omp parallel for schedule(static)
    for (int p = 0; p < G; ++p) {
        double sum = 0, sum_e = 0;
        for (int kb = 0; kb < 4; ++kb) {
            sum += corner[kb * (size_t)G + p];
            sum_e += corner[(4 + kb) * (size_t)G + p];
        }
        fm[p] = sum;
        em[p] = sum_e;
        s->rho[p] += sum * chrg;
    }
    free(corner);
}

static void push_species(OrcSim* s, OrcSpecies* sp) {
    int64_t a0 = 0, a1 = 0;
#pragma omp parallel for schedule(static) reduction(+ : a0, a1)
    for (int64_t i = 0; i < sp->n; ++i) {
        OrcPoint* q = &sp->p[i];
        int64_t ab[2] = {0, 0};
        double pz = q->z;
        q->cell = orc_push_one(&q->x, &q->y, &q->vx, &q->vy, &q->vz, s->ex, s->ey, s->bz, s->gxn, s->gyn, s->dx,
                               s->dy, s->width, s->height, sp->qm, s->dt, s->bx, s->by, ab);
        q->z = pz + s->dt * q->vz;   /* pz = pz + dt * pvz (mover.cpp:166) */
        a0 += ab[0];
        a1 += ab[1];
    }
    sp->absorbed[0] += a0;
    sp->absorbed[1] += a1;
}

/* Solve with PARDISO when available, else the oracle's direct solve. */
static int solve(OrcSim* s) {
    const int G = s->G;
    orc_overwrite_rho(G, s->mask, s->dval, s->rho);   /* rho is rewritten in place, as the reference does */
    memcpy(s->rhs, s->rho, (size_t)G * sizeof(double));
    int rc;
    if (s->pardiso)
        rc = orc_pardiso_solve(s->pardiso, s->rhs, s->phi);
    else
        rc = orc_poisson_direct(s->gxn, s->gyn, s->dx, s->dy, s->mask, s->dval, s->rhs, s->phi);
    for (int i = 0; i < G; ++i)
        if (s->mask[i]) s->phi[i] = s->dval[i];   /* overwritePhi */
    return rc;
}

int orc_sim_step(OrcSim* s) {
    double t0 = now_s();
    memset(s->rho, 0, (size_t)s->G * sizeof(double));
    for (int k = 0; k < s->ns; ++k) deposit_species(s, &s->sp[k]);
    double t1 = now_s();
    int rc = solve(s);
    double t2 = now_s();
    orc_efield(s->gxn, s->gyn, s->dx, s->dy, s->phi, s->ex, s->ey);
    double t3 = now_s();
    for (int k = 0; k < s->ns; ++k) push_species(s, &s->sp[k]);
    double t4 = now_s();
    for (int k = 0; k < s->ns; ++k) bin_species(s, &s->sp[k]);
    double t5 = now_s();
    s->t[0] += t1 - t0;
    s->t[1] += t2 - t1;
    s->t[2] += t3 - t2;
    s->t[3] += t4 - t3;
    s->t[4] += t5 - t4;
    return rc;
}

void orc_sim_times(OrcSim* s, double* t5) { memcpy(t5, s->t, sizeof(s->t)); memset(s->t, 0, sizeof(s->t)); }

int64_t orc_sim_count(OrcSim* s, int k) { return s->sp[k].n; }

void orc_sim_get_grid(OrcSim* s, int which, double* out) {
    const double* src = NULL;
    switch (which) {
    case 0: src = s->rho; break;
    case 1: src = s->phi; break;
    case 2: src = s->ex; break;
    case 3: src = s->ey; break;
    case 4: src = s->bz; break;
    default:
        if (which >= 8 && which < 8 + s->ns) src = s->sp[which - 8].final_mesh;
        if (which >= 16 && which < 16 + s->ns) src = s->sp[which - 16].energy_mesh;
    }
    if (src) memcpy(out, src, (size_t)s->G * sizeof(double));
}

void orc_sim_get_particles(OrcSim* s, int k, double* x, double* y, double* vx, double* vy, double* vz,
                           int32_t* cell) {
    OrcSpecies* sp = &s->sp[k];
    for (int64_t i = 0; i < sp->n; ++i) {
        x[i] = sp->p[i].x;
        y[i] = sp->p[i].y;
        vx[i] = sp->p[i].vx;
        vy[i] = sp->p[i].vy;
        vz[i] = sp->p[i].vz;
        cell[i] = sp->p[i].cell;
    }
}

void orc_sim_destroy(OrcSim* s) {
    if (!s) return;
    for (int k = 0; k < s->ns; ++k) {
        free(s->sp[k].p);
        free(s->sp[k].q);
        free(s->sp[k].final_mesh);
        free(s->sp[k].energy_mesh);
        free(s->sp[k].start);
    }
    if (s->pardiso) orc_pardiso_destroy(s->pardiso);
    free(s->rho);
    free(s->rhs);
    free(s->phi);
    free(s->ex);
    free(s->ey);
    free(s->bz);
    free(s->mask);
    free(s->dval);
    free(s->hist);
    free(s);
}
