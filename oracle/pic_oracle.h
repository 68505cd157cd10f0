/*
 * pic_oracle.h — CPU restatement of the reference PIC time step.
 * TEST INFRASTRUCTURE ONLY: used by tests/, the smoke check and the CPU
 * baseline of bench.py as the checker; never linked into the product.
 */
#ifndef PIC_ORACLE_H
#define PIC_ORACLE_H
#define _GNU_SOURCE
#include <stdint.h>

int32_t orc_key(double x, double y, int gx_nodes, double dx, double dy);
void orc_keys(int64_t n, const double* x, const double* y, int gx_nodes, double dx, double dy, int32_t* key);
void orc_bin(int64_t n, const int32_t* key, int G, int32_t* counts, int32_t* start, int64_t* order);
void orc_deposit_ref(int64_t n, const double* x, const double* y, const double* vx, const double* vy,
                     const double* vz, const int32_t* start, const int64_t* order, int gxn, int gyn, double dx,
                     double dy, double W, double qm, double charge, double eps, double* final_mesh,
                     double* energy_mesh, double* rho);
void orc_deposit_fixed_f64(int64_t n, const double* x, const double* y, const double* vx, const double* vy,
                           const double* vz, const int32_t* key, int gxn, int gyn, double dx, double dy, int cs,
                           int es, double qm, int64_t* qc, int64_t* qe);
void orc_deposit_fixed_cell32(int64_t n, const int32_t* cell, const float* ox, const float* oy, const float* vx,
                              const float* vy, const float* vz, int gxn, int gyn, int cs, int es, double qm,
                              int64_t* qc, int64_t* qe);
void orc_efield(int gxn, int gyn, double dx, double dy, const double* phi, double* ex, double* ey);
int32_t orc_push_one(double* px, double* py, double* pvx, double* pvy, double* pvz, const double* ex,
                     const double* ey, const double* bz, int gxn, int gyn, double dx, double dy, double width,
                     double height, double qm, double dt, double bx, double by, int64_t* absorbed);
void orc_push_f64(int64_t n, double* x, double* y, double* vx, double* vy, double* vz, int32_t* key,
                  const double* ex, const double* ey, const double* bz, int gxn, int gyn, double dx, double dy,
                  double width, double height, double qm, double dt, double bx, double by, int64_t* absorbed);
void orc_push_cell32(int64_t n, int32_t* cell, float* ox, float* oy, float* vx, float* vy, float* vz, int32_t* key,
                     const double* ex, const double* ey, const double* bz, int gxn, int gyn, double dx, double dy,
                     double width, double height, double qm, double dt, double bx, double by, int64_t* absorbed);
void orc_push_f64_p(int64_t n, double* x, double* y, double* vx, double* vy, double* vz, int32_t* key,
                    const double* ex, const double* ey, const double* bz, int gxn, int gyn, double dx, double dy,
                    double width, double height, double qm, double dt, double bx, double by, int periodic,
                    int64_t* absorbed);
void orc_push_cell32_p(int64_t n, int32_t* cell, float* ox, float* oy, float* vx, float* vy, float* vz,
                       int32_t* key, const double* ex, const double* ey, const double* bz, int gxn, int gyn,
                       double dx, double dy, double width, double height, double qm, double dt, double bx,
                       double by, int periodic, int64_t* absorbed);
void orc_efield_periodic(int gxn, int gyn, double dx, double dy, const double* phi, double* ex, double* ey);
double orc_modulo(double x, double y);
void orc_encode_cell32(int64_t n, const double* x, const double* y, const double* vx, const double* vy,
                       const double* vz, int gxn, double dx, double dy, int32_t* cell, float* ox, float* oy,
                       float* fvx, float* fvy, float* fvz);
void orc_geometry_nodes(int gxn, int gyn, int n, const double* rects, int* nodes, double* pot);
void orc_dirichlet(int gxn, int gyn, int ngeom, const int* nodes, const double* pot, uint8_t* mask, double* dval);
void orc_overwrite_rho(int G, const uint8_t* mask, const double* dval, double* rho);
int64_t orc_assemble(int gxn, int gyn, double dx, double dy, const uint8_t* mask, int* ia, int* ja, double* a);
double orc_relres(int gxn, int gyn, double dx, double dy, const uint8_t* mask, const double* rhs, const double* phi,
                  const double* dval);
int orc_poisson_direct(int gxn, int gyn, double dx, double dy, const uint8_t* mask, const double* dval,
                       const double* rhs, double* phi);
int orc_have_pardiso(void);
void* orc_pardiso_create(int gxn, int gyn, double dx, double dy, const uint8_t* mask);
int orc_pardiso_solve(void* h, double* rhs, double* phi);
void orc_pardiso_destroy(void* h);
int orc_num_threads(void);

typedef struct OrcSim OrcSim;
OrcSim* orc_sim_create(int NX, int NY, double width, double height, double dt, double eps, double b0, double bx,
                       double by, int ngeom, const double* rects, int deposit_energy, int use_pardiso);
int orc_sim_uses_pardiso(OrcSim* s);
int orc_sim_add_species(OrcSim* s, double charge, double mass, double sim_count, double actual, int64_t n,
                        const double* x, const double* y, const double* vx, const double* vy, const double* vz);
int orc_sim_step(OrcSim* s);
void orc_sim_times(OrcSim* s, double* t5);
int64_t orc_sim_count(OrcSim* s, int k);
void orc_sim_get_grid(OrcSim* s, int which, double* out);
void orc_sim_get_particles(OrcSim* s, int k, double* x, double* y, double* vx, double* vy, double* vz,
                           int32_t* cell);
void orc_sim_destroy(OrcSim* s);
#endif
