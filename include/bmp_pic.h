/*
 * bmp_pic.h — C ABI of the MI355X (gfx950) 2D3V electrostatic PIC time step.
 *
 * Two groups of entry points:
 *
 *  1. The drop-in group `sycl_pic_*`: the same symbol names, signatures and
 *     argument meaning as the reference's plain C wrapper
 *     (SYCL-PIC-poissonsolver/src/sycl_pic_ctypes.cpp:46-272) plus the two
 *     extra calls of its CPython module (update_fields_from_phi, grid x/y;
 *     SYCL-PIC-poissonsolver/src/sycl_pic_module.cpp:110-223).  A ctypes caller
 *     of the reference library binds these unchanged.  One simulation per
 *     process, global state, every call synchronous — as in the reference.
 *
 *  2. The stage group `pic_*`: the reference's per-stage C++ entry points
 *     (bin, deposit, Poisson, gather + Boris push, geometry) over an opaque
 *     context handle instead of `sycl::queue&` / `DeviceView&` / globals.
 *
 * Errors: every `pic_*` call returns a status code (PIC_OK = 0, negative on
 * error) and `pic_last_error()` describes the last failure.  Where the
 * reference calls `std::exit(code)`, the matching status below is returned
 * instead; the library never exits the process.
 *
 * All pointers in this header are HOST pointers unless stated otherwise.
 * Grids are row-major, node (gx, gy) at index gy * GRID_X + gx with
 * GRID_X = NX + 1, GRID_Y = NY + 1 (SYCL-PIC-poissonsolver/src/particles.cpp:356-368).
 */
#ifndef BMP_PIC_H
#define BMP_PIC_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- status */
#define PIC_OK                 0
#define PIC_ERR_ARG           -1   /* bad argument; also bad device string (sycl_pic_ctypes.cpp:57-59 returns -1) */
#define PIC_ERR_FILE          -2   /* input file missing (reference exit(1): particles.cpp:326-329, :639-642) */
#define PIC_ERR_FORMAT        -3   /* malformed M_PICDAT.DAT / geometry.conf (reference exit(2)) */
#define PIC_ERR_ALLOC         -4   /* allocation failure (reference exit(3)) */
#define PIC_ERR_SOLVER        -5   /* Poisson solve failed (reference exit(5)/exit(6), poissonSolver.cpp:291-318,511-534) */
#define PIC_ERR_PERIODIC_GEOM -9   /* periodic flag with > 3 geometries (reference exit(9), particles.cpp:651-654) */
#define PIC_ERR_HIP          -10   /* HIP runtime error */
#define PIC_ERR_RANGE        -11   /* capacity / fixed-point headroom exceeded (reference exit(1) on map overflow, mover.cpp:412-419) */
#define PIC_ERR_UNSUPPORTED  -12   /* configuration this build does not implement (e.g. a learned guess with periodic y) */
#define PIC_ERR_STATE        -13   /* call made in the wrong order (e.g. before init) */

/* ------------------------------------------------------- storage layouts */
/* PIC_LAYOUT_F64: per particle x, y, vx, vy, vz as double plus an int32 cell
 *   key — the reference's Point precision (point.hpp:11-14) without the unused
 *   z coordinate and hashmap slot pointer.
 * PIC_LAYOUT_CELL32: per particle int32 cell key, float in-cell offsets
 *   ox, oy in [0, 1) (units of dx, dy) and float vx, vy, vz: 24 bytes. */
#define PIC_LAYOUT_F64     0
#define PIC_LAYOUT_CELL32  1

/* ----------------------------------------------------------- grid names */
#define PIC_GRID_RHO        0   /* rho (reference `rho`, types.cpp:33)            */
#define PIC_GRID_PHI        1   /* phi (reference `phi`)                            */
#define PIC_GRID_EX         2   /* electricField[].fx                               */
#define PIC_GRID_EY         3   /* electricField[].fy                               */
#define PIC_GRID_BZ         4   /* magneticField (Bz on nodes)                      */
#define PIC_GRID_FINAL0     8   /* + species: Particle::final_mesh (species_ds.hpp:26) */
#define PIC_GRID_ENERGY0   16   /* + species: Particle::energy_mesh                 */

/* -------------------------------------------------------------- stages */
#define PIC_STAGE_DEPOSIT   0   /* per-cell sums of the CIC deposit (node-row sums)      */
#define PIC_STAGE_POISSON   1   /* whole Poisson solve                                   */
#define PIC_STAGE_EFIELD    2   /* E = -grad phi                                         */
#define PIC_STAGE_PUSH      3   /* gather + Boris + walls + slot count (move)            */
#define PIC_STAGE_BIN       4   /* prefix scan of the counts (+ count/scatter of a re-bin) */
#define PIC_STAGE_SCATTER   5   /* recomputed push written to the sorted slots (move)    */
#define PIC_STAGE_FINALIZE  6   /* node sums into final_mesh, energy, rho (pic_step: in the
                                   Poisson stage, the solve's first kernel)              */
#define PIC_NSTAGES         7

typedef struct pic_ctx pic_ctx;

/* Simulation parameters.  Field names follow the reference globals
 * (types.hpp:8-133); fill with pic_params_default() then override. */
typedef struct pic_params {
    int    NX, NY;            /* cells in x / y (M_PICDAT.DAT "NX NY")                       */
    double width, height;     /* GRID_WIDTH, GRID_HEIGHT in m                                 */
    double dt;                /* time step, s                                                 */
    double epsilon;           /* permittivity, F/m                                            */
    double b0;                /* uniform Bz filled into magneticField, T (initMagneticField) */
    double bx, by;            /* uniform in-plane B, T; 0 = the reference's Bz-only Boris     */
    double surf_area;         /* surfArea for the weighting factor (particles.cpp:431-432)   */
    int    periodic;          /* periodicBoundary: 1 = periodic y (poissonSolver.cpp:82-180,411-430,
                                 mover.cpp:317-320); needs Dirichlet walls at x = 0 and x = width  */
    int    layout;            /* PIC_LAYOUT_*                                                 */
    int    charge_shift;      /* fixed-point scale 2^-charge_shift of normalised CIC weights  */
    int    energy_shift;      /* fixed-point scale 2^-energy_shift eV of energy weights        */
    int    deposit_energy;    /* 1: also build energy_mesh each deposit (reference behaviour) */
    double poisson_tol;       /* relative residual ||b - A phi|| / ||b|| to stop at            */
    int    poisson_max_iter;  /* cap on outer (FP64) iterations                                */
    int    mg_pre, mg_post;   /* weighted-Jacobi sweeps before / after coarse correction      */
    double mg_omega;          /* Jacobi weight                                                 */
    int    poisson_guess;     /* initial guess: 0 previous phi (the reference's warm start),
                                 1 linear extrapolation 2 phi_n-1 - phi_n-2 of the last two solves */
    int    poisson_precision; /* 0: FP64 PCG (FP32 multigrid preconditioner);
                                 1: iterative refinement — FP64 residual, dot products and update
                                 of phi, FP32 PCG vectors for each correction solve            */
    double poisson_inner_tol; /* relative residual of each FP32 correction solve (precision 1)  */
    int    mg_smoother;       /* 0: weighted Jacobi (mg_omega); 1: Chebyshev steps on the upper
                                 spectrum of D^-1 A (with mg_pre = mg_post = 2)                */
    int    poisson_direct;    /* 1: solve exactly (FP64 sine transforms in x and y, or a real
                                 Fourier basis in a periodic y, + capacitance correction for
                                 internal Dirichlet nodes) and let the PCG only check the
                                 result; used for grids of at most 2047 interior nodes (or
                                 period) per direction and 1024 internal Dirichlet nodes next to
                                 unknowns, with precision 0 and no learned guess.  The
                                 transforms run as FFTs when NX and NY are powers of two and y
                                 is not periodic, else as FP64 matrix-core GEMMs; 2: GEMMs
                                 always.  0: PCG alone                                         */
    int    tile_cells;        /* cells per tile of the push: 32, 64, 128 or 256; 0 = the widest
                                 whose mean tile holds at most 1024 particles (at least 32)    */
    int    deposit_group;     /* consecutive groups of 4 particles per deposit thread: 1, 2 or 4;
                                 0 = 1 up to 32 particles per cell, 2 above                    */
    int    push_group;        /* consecutive particles per thread of the CELL32 tile push: 1, 2 or 4;
                                 0 = 1 up to 32 particles per cell, 2 above                    */
} pic_params;

/* Per-solve statistics of pic_poisson_solve. */
typedef struct pic_solve_stats {
    int    iterations;        /* PCG iterations (precision 0) or refinement steps (precision 1) */
    double initial_relres;    /* ||b - A phi0|| / ||b|| of the initial guess       */
    double final_relres;      /* same after the solve                              */
    double ms;                /* device time of the solve (HIP events)             */
    int    guess;             /* 0 zero guess, 1 previous phi, 2 caller-set guess,
                                 3 extrapolated from the last two solves,
                                 4 learned (U-Net) prediction, 5 direct solve      */
    int    inner_iterations;  /* PCG iterations in all (precision 1: FP32 ones)   */
} pic_solve_stats;

/* ====================================================== stage group */

/* Fill defaults (dx = dy = 1e-4 m, 128 x 128 cells, ...). */
void pic_params_default(pic_params* p);

/* Create a context on HIP device 0.  Allocates grids and the multigrid
 * hierarchy; geometry defaults to four grounded walls. */
int  pic_create(const pic_params* p, pic_ctx** out);
void pic_destroy(pic_ctx* ctx);

/* Geometry (replaces readGeometry, particles.cpp:636-686, and
 * setPotentialBoundary, poissonSolver.cpp:64-79).  `rects` holds n rows of
 * (x1, y1, x2, y2, V) as fractions of the domain; node index =
 * (int)((GRID-1) * fraction); rectangles are inclusive; later rows win. */
int  pic_set_geometry(pic_ctx* ctx, const double* rects, int n);
/* Parse a geometry.conf file (same format as the reference). */
int  pic_read_geometry(pic_ctx* ctx, const char* path);
/* Node rectangles actually in use: out_nodes n x (x1, y1, x2, y2), out_pot n. */
int  pic_get_geometry(pic_ctx* ctx, int* out_nodes, double* out_pot, int* n);

/* Register a species (PICDAT species table row, particles.cpp:590-621):
 * q/m = charge / mass (initCtoM :298-300); weighting factor
 * W = actual * surf_area / (sim * dx*dy*dx*dy) (updateWeighingFactor :303-305).
 * `capacity` = max particles ever held. */
int  pic_add_species(pic_ctx* ctx, double charge, double mass, double sim_count,
                     double actual_count, int64_t capacity, int* species_out);

/* Project particle generator (host only; replaces initParticlesRandomly,
 * particles.cpp:127-173, which is time-seeded).  Counter-based: particle i of
 * (seed, stream) is independent of n and thread count.  Positions uniform in
 * [0, width) x [0, height), rejecting points inside the node rectangle
 * excl = (x1, y1, x2, y2) (meters = node * dx) when excl != NULL.
 * Velocities Maxwellian, sigma = sqrt(|q/m| T) per component. */
int  pic_generate_particles(uint64_t seed, int stream, int64_t n,
                            double width, double height, double dx, double dy,
                            const int* excl, double qm_abs, double temp_eV,
                            double* x, double* y, double* vx, double* vy, double* vz);

/* Upload n particles (absolute FP64 state) into species s, encode them into
 * the context layout and bin them (counting sort by the reference key
 * (int)(x * (1.0/dx)), particles.cpp:229-235).  Replaces the InsertKernel. */
int  pic_upload_particles(pic_ctx* ctx, int s, int64_t n, const double* x, const double* y,
                          const double* vx, const double* vy, const double* vz);
/* Upload raw CELL32 state (cell, ox, oy, vx, vy, vz) and bin it. */
int  pic_upload_cell32(pic_ctx* ctx, int s, int64_t n, const int32_t* cell,
                       const float* ox, const float* oy,
                       const float* vx, const float* vy, const float* vz);
/* Live particle count of species s (after the last bin). */
int64_t pic_particle_count(pic_ctx* ctx, int s);
/* Download the live particles in storage (cell-sorted) order, decoded to
 * FP64: F64 layout returns the stored values, CELL32 returns
 * x = (gx + ox) * dx.  `cell` gets the bin key.  Any pointer may be NULL. */
int  pic_download_particles(pic_ctx* ctx, int s, double* x, double* y,
                            double* vx, double* vy, double* vz, int32_t* cell);
/* Download raw CELL32 storage.  Any pointer may be NULL. */
int  pic_download_cell32(pic_ctx* ctx, int s, int32_t* cell, float* ox, float* oy,
                         float* vx, float* vy, float* vz);

/* Bin species s: counting sort of the live particles by cell key, dropping
 * dead ones (replaces hashmap insert/delete + cleanup_map + compaction,
 * sycl_hashmap.hpp:131-231, utils.cpp:30-44,249-372). */
int  pic_bin(pic_ctx* ctx, int s);
/* Per-key particle counts, GRID_X * GRID_Y ints (key_idx after cleanup). */
int  pic_get_cell_counts(pic_ctx* ctx, int s, int32_t* counts);
/* Exclusive prefix of the counts, GRID_X * GRID_Y + 1 ints. */
int  pic_get_cell_start(pic_ctx* ctx, int s, int32_t* start);

/* rho = 0 (sycl_pic_zero_rho, sycl_pic_ctypes.cpp:121-124). */
int  pic_zero_rho(pic_ctx* ctx);
/* CIC charge (and energy) deposit of species s (interpolate,
 * interpolation.cpp:13-176): final_mesh[s], energy_mesh[s], and
 * rho += final_mesh[s] * (-q_s / epsilon). */
int  pic_deposit(pic_ctx* ctx, int s);
/* Raw fixed-point accumulators of the last deposit of species s, summed per
 * node (GRID_X*GRID_Y int64 each; either pointer may be NULL). */
int  pic_get_deposit_fixed(pic_ctx* ctx, int s, int64_t* q_charge, int64_t* q_energy);

/* Poisson solve (poissonSolver, poissonSolver.cpp:459-544): rho is rewritten
 * in place as the reference does (overwriteRho: rho[geometry] = V, then
 * rho = -rho), phi solves the 5-point system with Dirichlet nodes held at V,
 * then E = -grad phi (calculateElectricField :433-456).  iteration == 0 starts
 * from a zero guess; later iterations warm-start from the previous phi, or
 * from the guess set with pic_set_guess().  `st` may be NULL. */
int  pic_poisson_solve(pic_ctx* ctx, int iteration, pic_solve_stats* st);
/* Use `phi0` (GRID_X*GRID_Y doubles) as the next solve's initial guess. */
int  pic_set_guess(pic_ctx* ctx, const double* phi0);
/* overwritePhi + calculateElectricField on the current phi
 * (py_update_fields_from_phi, sycl_pic_module.cpp:216-223). */
int  pic_update_fields_from_phi(pic_ctx* ctx);

/* Gather E and Bz at each particle, Boris push, wall absorption, new cell key
 * (mover / deterministic_mover, mover.cpp:103-176,304-367).  With rebin != 0
 * the species is re-binned afterwards (the reference re-inserts into its
 * hashmap inside the mover). */
int  pic_push(pic_ctx* ctx, int s, int rebin);

/* One full time step in the reference order (main.cpp:130-257):
 * zero rho; deposit every species; Poisson + E; push + re-bin every species. */
int  pic_step(pic_ctx* ctx, int iteration, pic_solve_stats* st);

/* Copy a grid (PIC_GRID_*) to / from the host, GRID_X * GRID_Y doubles. */
int  pic_get_grid(pic_ctx* ctx, int which, double* out);
int  pic_set_grid(pic_ctx* ctx, int which, const double* in);

/* Absorption counters of species s since creation: particles that left
 * through x < 0 (the reference `new_additions`, mover.cpp:344-352) and all
 * absorbed particles.  Either pointer may be NULL. */
int  pic_get_absorbed(pic_ctx* ctx, int s, int64_t* left, int64_t* total);

/* Total kinetic energy of species s, sum of 0.5 |v|^2 / |q/m| over the live
 * particles (calculate_energy, utils.cpp:77-118); summed in a fixed order, so
 * repeated calls on the same state give the same double. */
int  pic_kinetic_energy(pic_ctx* ctx, int s, double* energy);

/* ------------------------------------------------------------ injection */
/* New particles after the push of every species (main.cpp:232-245), species 0
 * the electrons and species 1 the ions:
 *   re-injection (generate_new_electrons, utils.cpp:374-499): as many
 *     electrons as left through x < 0 minus the ions that did (never
 *     negative), at x = x_inject, y uniform in [0, height), Maxwellian at
 *     te_reinject_eV (the reference passes its electron temperature as an int,
 *     utils.cpp:381);
 *   pairs (generate_pairs, utils.cpp:610-703): target = ceil(height dt (2/PI)
 *     (x22 - x11) ionimax / 1.67e6) per step with PI = 3.141592654; attempt a
 *     (a < max(10 target, 100)) draws x = xm + asin(2u - 1)(x22 - x11)/PI and is
 *     accepted when 0 < x < width; the first `target` accepted attempts give an
 *     electron (te_eV) and an ion (ti_eV) at the same (x, y).
 * Each velocity component is normal with variance |q/m| T.  Draws come from a
 * counter-based generator keyed by (seed, injection count, stream, index)
 * instead of the reference's time-seeded std::rand / ran2: reruns give the
 * same particles.  The reference exits when the storage would overflow
 * (utils.cpp:388-393,685-688); here the excess is not added and is counted
 * in `dropped`, as are re-injected electrons beyond max_reinject per step. */
typedef struct pic_injection {
    int      reinject;        /* 1: re-inject electrons that left through x < 0  */
    int      pairs;           /* 1: generate ionization pairs                    */
    double   x_inject;        /* re-injection plane, m (utils.cpp:400: 0.024)    */
    double   te_reinject_eV;  /* temperature of re-injected electrons            */
    double   te_eV, ti_eV;    /* pair temperatures (electronTemp, ionTemp)       */
    double   x11, x22, xm;    /* ionization zone, m (types.cpp:103-105)          */
    double   ionimax;         /* peak ionization rate (utils.cpp:619: 5.23e23)   */
    int64_t  max_reinject;    /* re-injected electrons staged per step           */
    uint64_t seed;
} pic_injection;

/* Reference values: x_inject 0.024, te 10, ti 0.026, x11 0.0025, x22 0.01,
 * xm 0.00625, ionimax 5.23e23, both kinds on, max_reinject 1 << 20, seed 1. */
void pic_injection_default(pic_injection* cfg);
/* Enable injection in pic_step (cfg != NULL) or turn it off (NULL).  Needs two
 * species.  Counters of exits start from the current absorption counters. */
int  pic_set_injection(pic_ctx* ctx, const pic_injection* cfg);
/* One injection now into the binned storage of both species (the drop-in's
 * generate_new_particles, sycl_pic_ctypes.cpp:212-236): the new particles are
 * appended and both species re-binned. */
int  pic_inject(pic_ctx* ctx);
/* Pairs per step of the configuration (the ceil above) and the attempts cap. */
int  pic_injection_target(pic_ctx* ctx, int64_t* target, int64_t* attempts);
/* Injections so far, re-injected electrons, pairs, dropped particles in all,
 * and the particles added by the last injection to species 0 and 1.  Any
 * pointer may be NULL. */
int  pic_injection_stats(pic_ctx* ctx, int64_t* injections, int64_t* reinjected, int64_t* pairs,
                         int64_t* dropped, int64_t* last_electrons, int64_t* last_ions);

/* Counting sort of n cell keys in device memory that allocates nothing.  No
 * counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops).  It is the key
 * half of pic_bin (the same count and scan kernels): counts[k] (G ints) = the
 * keys equal to k; start (G + 1 ints) = their exclusive prefix, start[G] = the
 * keys inside [0, G); slot[i] (n ints) = the sorted position of key i
 * (start[key[i]] plus its place among the equal keys, in no fixed order), or
 * -1 for a key outside [0, G) (a dead particle).  All pointers are DEVICE
 * pointers; counts and start 16-byte aligned.  The work is enqueued on
 * `stream` (a hipStream_t; NULL = the null stream) and the call does not wait.
 * The scan's chunk totals live in a fixed array of this library on each GPU
 * (33 KB, kept until the process ends): G is limited to 16,777,216 keys and n
 * to 2^31 - 1 (PIC_ERR_RANGE beyond), and calls that overlap on one GPU race
 * on that array.  Every call clears what it needs; the array keeps the last
 * call's totals. */
int  pic_bin_keys_noalloc(const int32_t* key, int64_t n, int32_t G, int32_t* counts, int32_t* start, int32_t* slot,
                          void* stream);

/* Block until all work of the context is done. */
int  pic_sync(pic_ctx* ctx);
/* Timing: 0 off; 1 per stage (HIP event markers around each stage's
 * kernels); 2 per kernel for the three particle kernels only (the tile push,
 * the scatter and the tile deposit, in the PUSH, SCATTER and DEPOSIT slots),
 * timestamps taken by the kernel dispatches themselves, so a measured loop
 * carries no markers; 3 as 2 for the scatter kernel alone.  pic_stage_times reads the accumulated device
 * milliseconds and launch counts per stage (PIC_NSTAGES entries each; either
 * pointer may be NULL) and resets them. */
int  pic_set_timing(pic_ctx* ctx, int on);
int  pic_stage_times(pic_ctx* ctx, double* ms, int64_t* launches);

/* Grid shape of the context: GRID_X, GRID_Y. */
int  pic_grid_shape(pic_ctx* ctx, int* grid_x, int* grid_y);
/* Text of the last error (thread-local). */
const char* pic_last_error(void);

/* ============================================ learned initial guess */
/* A 4-level U-Net (3x3 convolutions, 16-64 channels, BF16 weights and
 * activations, FP32 accumulation on MFMA) trained online with Adam on the
 * solver's own (rho_n, phi_{n-1}) -> phi_n pairs; its prediction is tried as
 * the PCG initial guess and kept only when its starting residual is lower
 * than the previous solution's.  No counterpart in the reference, whose ML
 * path replaces the solver by a surrogate (poisson_ml_bridge.py:55-106); the
 * PCG here always runs to the stated residual. */
typedef struct pic_unet pic_unet;

/* Network for an n x n node image (n a multiple of 128); weights from `seed`. */
int     pic_unet_create(int n, uint64_t seed, pic_unet** out);
void    pic_unet_destroy(pic_unet* u);
int64_t pic_unet_num_params(pic_unet* u);
/* FP32 master parameters, flattened: per convolution W[9][cout][cin] then
 * b[cout] (14 convolutions), then the head's 16 weights and bias. */
int     pic_unet_get_params(pic_unet* u, float* w);
int     pic_unet_set_params(pic_unet* u, const float* w);
int     pic_unet_get_grads(pic_unet* u, float* g);
/* x: [n][n][2] scaled inputs; out: [n][n] predicted scaled change. */
int     pic_unet_forward(pic_unet* u, const float* x, float* out);
/* Forward, backward and (lr > 0) one Adam step; *loss = masked mean squared
 * error before the step; gradients stay readable by pic_unet_get_grads. */
int     pic_unet_train_step(pic_unet* u, const float* x, const float* target, const uint8_t* mask, float lr,
                            float* loss);
/* The same with the loss used online: mean over mask of
 * (r - (A/diag) d)^2, d = output on masked nodes (0 elsewhere), A/diag with unit
 * diagonal and off-diagonals -cxd (x neighbours), -cyd (y neighbours). */
int     pic_unet_train_step_residual(pic_unet* u, const float* x, const double* r, const uint8_t* mask, float cxd,
                                     float cyd, float lr, float* loss);
/* Average device ms of a forward pass and of forward + backward over `reps`
 * runs on the data currently on the device (no weight update). */
int     pic_unet_benchmark(pic_unet* u, int reps, float* fwd_ms, float* train_ms);
/* 16x16x16 BF16 MFMA on row-major 16x16 inputs (layout check). */
int     pic_unet_mfma_probe(const float* a, const float* b, float* d);
/* 16x16x32 BF16 MFMA (the forward convolutions' instruction) on row-major
 * a (16 x 32) and b (32 x 16): d (16 x 16) = a b (layout check). */
int     pic_unet_mfma32_probe(const float* a, const float* b, float* d);
/* The direct Poisson solve's FP64 matrix-core GEMM on host arrays (check of
 * the kernel): C (M x N, row-major) = op(A) op(B) with op(A) = A (M x K) or,
 * ta != 0, A^T of an array stored K x M; op(B) = B (K x N) or, tb != 0, B^T of
 * an array stored N x K.  ep 0: C = product; 1: C = product .* S; 2: C +=
 * product (C read first).  M, N <= 4096, K <= 65536. */
int     pic_dgemm_check(int M, int N, int K, int ta, int tb, int ep, const double* a, const double* b,
                        const double* s, double* c);

/* Use a learned guess in pic_poisson_solve: train on the first `train_steps`
 * solves (Adam, learning rate lr), offer the prediction from solve
 * `min_trained` on.  The image is the first n x n nodes (n = NX rounded down
 * to a multiple of 128); the rest of the grid keeps the previous solution. */
int     pic_learned_enable(pic_ctx* ctx, int train_steps, int min_trained, float lr, uint64_t seed);
/* The context's network, null when the learned guess is off. */
pic_unet* pic_learned_unet(pic_ctx* ctx);
/* Counters: pairs trained on, solves where the prediction was tried, where it
 * was used (lower starting residual), and the last training loss. */
int     pic_learned_stats(pic_ctx* ctx, int* trained, int* tried, int* used, float* last_loss);

/* ======================================================= ensembles */
/* Batched execution of independent runs of uneven sizes: every stage is one
 * launch for all runs (persistent work-queue kernels for deposit, E, push and
 * scatter; direct Poisson solves batched over the runs of one shape, checked
 * by one workgroup per run, which runs the PCG for the runs without a direct
 * solve; one workgroup per run and species for the scan), the step is
 * captured once in a HIP graph and replayed.  The runs are contexts already
 * set up (geometry, two species, particles uploaded), CELL32 storage, Bz-only
 * field, FP64 Poisson; their geometry must not change while the ensemble
 * exists (it holds their solver buffers).  After the ensemble's steps they are
 * in the state a serial pic_step would leave.  No counterpart in the
 * reference (one simulation per process, sycl_pic_ctypes.cpp:39-41). */
typedef struct pic_ensemble pic_ensemble;

int  pic_ensemble_create(pic_ctx** runs, int nruns, pic_ensemble** out);
void pic_ensemble_destroy(pic_ensemble* e);
/* nsteps steps of every run; graph = 1 replays the captured step. */
int  pic_ensemble_step(pic_ensemble* e, int nsteps, int graph);
/* Iterations and (initial, final) relative residual of each run's last solve. */
int  pic_ensemble_poisson_stats(pic_ensemble* e, int* iters, double* relres);
/* Times `steps` steps per candidate (push tile width 64/128/256 cells x
 * persistent workgroups 2/4/8 per CU) and keeps the fastest; advances the runs. */
int  pic_ensemble_tune(pic_ensemble* e, int steps, int* tile, int* blocks, double* ms_per_step);
int  pic_ensemble_steps_done(pic_ensemble* e, int64_t* steps);

/* =================================================== drop-in group */
/* Same names/signatures/semantics as SYCL-PIC-poissonsolver/src/sycl_pic_ctypes.cpp.
 * init reads ./M_PICDAT.DAT and ./geometry.conf from the working directory;
 * dev_type "gpu" or "cpu" (sycl_pic_ctypes.cpp:53-60): both run on the GPU,
 * the only device this library drives; any other value returns -1;
 * wg and ps are accepted and ignored (the counting sort has no probing
 * width); input1/input2 are ignored as in the reference (particles.cpp:174-176).
 * Particles are drawn by pic_generate_particles with the seed in the
 * environment variable BMP_PIC_SEED (default 1). */
int     sycl_pic_init(const char* dev_type, int wg, int ps, const char* input1, const char* input2); /* sycl_pic_ctypes.cpp:49-88 */
int     sycl_pic_get_grid_size(void);               /* :91-93   */
int     sycl_pic_get_grid_x(void);                  /* sycl_pic_module.cpp:110-115 */
int     sycl_pic_get_grid_y(void);                  /* sycl_pic_module.cpp:117-122 */
int     sycl_pic_get_timesteps(void);               /* :95-97   */
int     sycl_pic_get_print_interval(void);          /* :99-101  */
int     sycl_pic_get_num_species(void);             /* :103-105 */
int     sycl_pic_get_particle_count(int species);   /* :107-109 */
double* sycl_pic_get_rho_ptr(void);                 /* :112-114 host mirror, refreshed after each deposit / solve */
double* sycl_pic_get_phi_ptr(void);                 /* :116-118 host mirror; caller writes are read back by update_fields_from_phi */
void    sycl_pic_zero_rho(void);                    /* :121-124 */
double  sycl_pic_charge_deposition(int species);    /* :127-150 */
double  sycl_pic_poisson_solve(int iteration);      /* :153-158 */
void    sycl_pic_update_fields_from_phi(void);      /* sycl_pic_module.cpp:216-223 */
double  sycl_pic_move_particles(int species, int iteration);                          /* :161-176 */
double  sycl_pic_compact_particles(int species, double* out_alloc_t, double* out_scan_t); /* :179-198 */
double  sycl_pic_cleanup_map(int species);          /* :201-209 */
double  sycl_pic_generate_new_particles(void);      /* :212-236 */
void    sycl_pic_print_diagnostics(const char* dir, int iteration); /* :239-244 */
void    sycl_pic_save_meshes(void);                 /* :247-264 */
void    sycl_pic_finalize(void);                    /* :267-270 */

#ifdef __cplusplus
}
#endif
#endif /* BMP_PIC_H */
