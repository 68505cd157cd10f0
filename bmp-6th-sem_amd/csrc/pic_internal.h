// pic_internal.h — context layout and helpers shared by the host runtime and
// the gfx950 kernels of the PIC time step.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "../../include/bmp_pic.h"

// per-block partial sums of the Poisson dot products: 8 arrays of 2048
#define PIC_PART_BLOCKS 2048
#define PIC_PART_DOUBLES (8 * PIC_PART_BLOCKS)
// solver state block: 16 double scalars, then 4 int flags
#define PIC_SCAL_DOUBLES 16
#define PIC_STATE_BYTES (PIC_SCAL_DOUBLES * 8 + 4 * 4)

// the state block as the check kernel publishes it to the host
struct SolveMail {
    double scal[PIC_SCAL_DOUBLES];
    int flags[4];
    unsigned long long t0, t1;     // device wall clock at the solve's start and at the check
    unsigned seq;
    unsigned pad;
};

namespace pic {

// ------------------------------------------------------------------ errors
void set_error(const char* fmt, ...);

#define PIC_HIP(call)                                                          \
    do {                                                                       \
        hipError_t e_ = (call);                                                \
        if (e_ != hipSuccess) {                                                \
            ::pic::set_error("%s:%d %s: %s", __FILE__, __LINE__, #call,       \
                             hipGetErrorString(e_));                           \
            return PIC_ERR_HIP;                                                \
        }                                                                      \
    } while (0)

#define PIC_TRY(expr)                                                          \
    do {                                                                       \
        int rc_ = (expr);                                                      \
        if (rc_ != PIC_OK) return rc_;                                         \
    } while (0)

// -------------------------------------------------------- device geometry
// Everything a particle kernel needs about the grid, passed by value.
struct GridDev {
    int    gx, gy;          // GRID_X, GRID_Y (nodes)
    int    gp;              // stride of one corner mesh: G + GRID_X + 2
    float  rcp_gx;          // 1 / GRID_X for the key -> (gx, gy) split
    double dx, dy;
    double width, height;   // GRID_WIDTH, GRID_HEIGHT
    double inv_dx, inv_dy;  // 1.0 / dx, 1.0 / dy as the reference computes them
    double nx, ny;          // NX, NY (cells): the walls in cell units
    int    periodic;        // periodic y (mover.cpp:317-320)
};

// Per-species constants for the particle kernels.
struct SpeciesDev {
    double qm;        // q/m (CtoM)
    double half_qm;   // 0.5 * q/m as the reference groups it: (0.5 * pctm)
    double denom;     // 1 / |q/m| for the energy weight
    double dt;
    double dtdx, dtdy; // dt / dx, dt / dy (cell-unit position update)
    double bx, by;    // uniform in-plane B (0 = reference Bz-only push)
    int    general_b; // 1 when bx or by is non-zero
    // CELL32 with a uniform Bz (push modes 2, 3): the rotation is the same for
    // every particle, so its factors are formed once here, and the electric
    // half kick is ce * E with ce = (0.5 q/m) dt
    int    ub;        // 1 when these constants apply
    double ce;
    double ut[3];     // t = (0.5 q/m) B dt  (Bz-only: ut[2] is t)
    double us[3];     // s = 2 t / (1 + |t|^2)
};

// Particle storage: structure of arrays, double-buffered for the sort.
struct ParticleBuf {
    void*    pa = nullptr;   // F64: x (double)   CELL32: ox (float)
    void*    pb = nullptr;   // F64: y (double)   CELL32: oy (float)
    void*    vx = nullptr;   // F64: double       CELL32: float
    void*    vy = nullptr;
    void*    vz = nullptr;
    int32_t* cell = nullptr; // bin key gy*GRID_X + gx
};

// Node pass of the deposit for all species at once (at most kFinMax):
// final_s = corner sums x cqf_s, energy_s likewise, and
// rho = sum_s final_s * chrg_s in species order (the sum a zeroed rho
// accumulates species by species).
static constexpr int kFinMax = 4;
struct FinalizeAll {
    int ns;
    const long long* cq[kFinMax];
    const long long* ce[kFinMax];
    double cqf[kFinMax], cef[kFinMax], chrg[kFinMax];
    double* fm[kFinMax];
    double* em[kFinMax];
};

#ifdef __HIPCC__
// rho at node n returned and, with MESH, final_mesh / energy_mesh of node n
// written; a node's fixed-point deposit is the sum of corner meshes 0 and 1
// (node-row sums of the cells above and below it).  The products are rounded
// before the sum (no fused multiply-add) in every file that includes this.
template <bool MESH = true>
__device__ __forceinline__ double finalize_node(const FinalizeAll& a, int n, int gp, int energy) {
#pragma clang fp contract(off)
    double r = 0.0;
#pragma unroll
    for (int s = 0; s < kFinMax; ++s) {
        if (s >= a.ns) break;
        const long long* cq = a.cq[s];
        long long q = cq[n] + cq[gp + n];
        double f = (double)q * a.cqf[s];
        if (MESH) a.fm[s][n] = f;
        r += f * a.chrg[s];
        if (MESH && energy) {
            const long long* ce = a.ce[s];
            long long e = ce[n] + ce[gp + n];
            a.em[s][n] = (double)e * a.cef[s];
        }
    }
    return r;
}
#endif

struct Species {
    double charge = 0, mass = 0, qm = 0, W = 0, sim = 0, actual = 0;
    int64_t capacity = 0;
    int64_t n_upper = 0;      // host upper bound on live particles (grid size of particle launches)
    ParticleBuf buf[2];
    int cur = 0;              // buffer holding the sorted live particles
    int32_t* ndev = nullptr;  // device: live particles held in buf[0], buf[1]
    int32_t* key = nullptr;   // per-particle new key (-1 = absorbed)
    int32_t* rank = nullptr;  // per-particle slot inside its cell
    int32_t* counts = nullptr;// GRID_X*GRID_Y
    int32_t* start = nullptr; // GRID_X*GRID_Y + 1 (start[G] = live count)
    int32_t* start_old = nullptr; // CELL32 step: `start` of the order being pushed (the scan overwrites start)
    int32_t* blk_first = nullptr; // CELL32 step: cell of particle 256 b of the order being pushed
    int64_t* q_charge = nullptr; // fixed-point corner meshes: 2 x (G + GRID_X + 2) int64 (node-row sums)
    int64_t* q_energy = nullptr;
    double*  final_mesh = nullptr;
    double*  energy_mesh = nullptr;
    unsigned long long* absorbed = nullptr; // [0] left through x<0, [1] all absorbed
    int deposit_pending = 0;  // q_charge/q_energy hold the deposit of the stored particles
    int sorted = 0;           // storage is cell-sorted and `start` describes it
    int cells_stale = 0;      // CELL32: buf[cur].cell not written (the step derives cells from `start`)
};

// One multigrid level (FP32 storage).
struct MgLevel {
    int gx = 0, gy = 0;
    float cx = 0, cy = 0, diag = 0;   // 1/hx^2, 1/hy^2, 2(cx+cy)
    int per = 0;                      // periodic y: rows 0 .. gy-2 wrap, row gy-1 is the image
    uint8_t* dir = nullptr;           // 1 = Dirichlet node, 2 = periodic image row (not an unknown)
    float* e = nullptr;               // correction
    float* b = nullptr;               // right-hand side
    float* t = nullptr;               // scratch
};

struct StageTimer {
    std::vector<std::pair<hipEvent_t, hipEvent_t>> pending;
    double ms = 0;
    int64_t launches = 0;
};

struct Geometry {
    int x1, y1, x2, y2;
    double potential;
};

struct FastPoisson;   // direct Poisson solve (pic_fastpoisson.hip)

// A Poisson solve whose end (the wait for its check, statistics) is deferred
// so the caller can enqueue work behind it first.
struct SolveTail {
    int active = 0;
    bool direct = false, mixed = false, learned = false, extrap_keep = false;
    int guess = 0;
    bool picked = false;            // the device chose between a learned guess and phi_{n-1}
    hipEvent_t ev = nullptr;        // stage timer start
};

// Particle injection state (pic_inject.hip).
struct Injection {
    int on = 0;
    pic_injection cfg{};
    int target = 0, attempts = 0;   // pairs per step; attempts cap
    int ecap = 0;                   // re-injected electrons staged per step
    ParticleBuf stage[2];           // staged new particles per species (storage layout)
    int32_t* key[2] = {nullptr, nullptr};
    int32_t* rank[2] = {nullptr, nullptr};
    long long* state = nullptr;     // device counters
    long long* host = nullptr;      // pinned copy of the counters and live counts
    hipEvent_t ev = nullptr;        // recorded after the copy
    int pending = 0;                // host copy in flight: n_upper not yet raised
};

} // namespace pic

struct pic_ctx {
    pic_params p{};
    pic::GridDev g{};
    int G = 0;                 // GRID_X * GRID_Y
    hipStream_t stream = nullptr;
    std::vector<pic::Species> sp;
    std::vector<pic::Geometry> geom;
    // fields (device)
    double* rho = nullptr;
    double* phi = nullptr;
    double* ex = nullptr;      // separate components for the API / diagnostics
    double* ey = nullptr;
    double* bz = nullptr;
    int bz_uniform = 1;        // every node of bz holds the same value, bz0
    int sweep = 0;             // direction of the next particle pass (xcd_tile's rev); alternates
    double bz0 = 0.0;
    double4* fld = nullptr;    // packed (Ex, Ey, Bz, 0) per node, padded with zeros
    uint8_t* dir = nullptr;    // Dirichlet mask (fine grid)
    double* dval = nullptr;    // Dirichlet value per node (0 elsewhere)
    // Poisson work vectors (FP64, fine grid)
    double *pr = nullptr, *pr2 = nullptr;     // residual, alternating between iterations
    double *pp = nullptr, *pp2 = nullptr;     // search direction, alternating
    double* pq = nullptr;                     // A p
    float *zf0 = nullptr, *zf1 = nullptr;     // preconditioned residual z (FP32), alternating
    // FP32 correction solve of the mixed-precision path (allocated on first use)
    float *rf0 = nullptr, *rf1 = nullptr, *pf0 = nullptr, *pf1 = nullptr, *qf = nullptr, *df = nullptr;
    int last_inner = 0;
    double* guess = nullptr;
    double* phi_prev = nullptr;               // phi of the solve before the last (extrapolated guess)
    int guess_set = 0;
    int have_prev_phi = 0;                    // solutions kept: 0, 1 (phi) or 2 (phi, phi_prev)
    int last_iters = 0;
    // learned initial guess
    pic_unet* unet = nullptr;
    int learn_train = 0, learn_min = 0;       // pairs to train on; pairs before the guess is offered
    float learn_lr = 1e-3f;
    int learn_trained = 0, learn_tried = 0, learn_used = 0;
    float learn_loss = 0.f;
    std::vector<pic::MgLevel> mg;
    float* mg_inv = nullptr;        // inverse of the coarsest level's operator (unknowns only)
    int* mg_unk = nullptr;          // node index of each coarsest-level unknown
    int mg_nu = 0;
    uint8_t* mg_cdir = nullptr;     // Dirichlet masks of the in-LDS levels, concatenated
    int mg_lc = 0;                  // first level of the in-LDS coarse cycle
    int mg_ctotal = 0;              // nodes of the in-LDS levels
    double* part = nullptr;         // per-block partial sums of the dot products
    double* scal = nullptr;         // device scalars (the state block)
    int* flags = nullptr;           // in the state block: [0] converged / failed, [1] iterations, [2], [3] inner
                                    // solve
    double* scal_host = nullptr;    // pinned copy of the state block
    int* flags_host = nullptr;      // in it
    int scan_nb = 0;
    int32_t* scan_bsum = nullptr;   // chunk totals of two scans (2 x (scan_nb + 1))
    pic::Injection inj;
    pic::FastPoisson* fast = nullptr;   // direct solve of the current geometry (null: PCG only)
    // timing
    int timing = 0;
    pic::StageTimer timers[PIC_NSTAGES];
    std::vector<hipEvent_t> event_pool;   // recorded-and-read stage events, reused
    hipEvent_t solve_ev[2] = {nullptr, nullptr};   // start / end of a Poisson solve
    // The direct solve's check is published by the check kernel itself into
    // mapped, coherent host memory (state block, device wall-clock stamps,
    // then a sequence number after a system-scope fence): no copy and no
    // event in the step; the host waits for the sequence number.
    SolveMail* mail = nullptr;                     // host view
    SolveMail* mail_dev = nullptr;                 // device view of the same memory
    unsigned mail_seq = 0;                         // sequence number of the last publication
    unsigned long long* tstamp = nullptr;          // device: wall clock at the solve's start
    double wall_khz = 0;                           // wall-clock rate
    pic::SolveTail tail;                             // a deferred solve end
    pic::FinalizeAll fin{};                          // a deposit node pass left to the next solve
    int fin_pending = 0;
    // final_mesh / energy_mesh lag the fixed-point sums of the deposit in
    // `fin` (the step's node pass writes rho only); materialize_meshes writes
    // them when a caller reads or writes one
    int meshes_stale = 0;
    // ex / ey lag the gather array fld (the direct solve's residual pass
    // writes E into fld only); materialize_e forms them when a caller reads
    // or writes a field grid
    int e_stale = 0;
};

namespace pic {

// kernels / launchers (pic_particles.hip)
int launch_encode_f64(pic_ctx* c, int s, int64_t n, const double* x, const double* y,
                      const double* vx, const double* vy, const double* vz);
int launch_bin(pic_ctx* c, int s);
int launch_deposit(pic_ctx* c, int s);            // adds species s to rho
// every species; writes rho, or with defer_finalize leaves the node pass to
// the next Poisson solve (c->fin)
int launch_deposit_all(pic_ctx* c, bool defer_finalize = false);
int materialize_meshes(pic_ctx* c);        // final_mesh / energy_mesh of the deferred node pass
int materialize_e(pic_ctx* c);             // ex / ey from fld
int launch_push(pic_ctx* c, int s);
int launch_move(pic_ctx* c, int s);
// launch_move in phases, for injection between them; with `guard`, the push
// does nothing unless guard[0] == 1 when it runs (a Poisson check passed)
int move_push(pic_ctx* c, int s, const int* guard = nullptr);
int move_push_all(pic_ctx* c, const int* guard = nullptr);   // every species, two per launch
int move_finish(pic_ctx* c, int s);              // scan + scatter of species s
int move_scan_pair(pic_ctx* c);                   // the scans of species 0 and 1 in one set of launches
int move_scatter(pic_ctx* c, int s);              // the scatter after the scan
int move_scatter_pair(pic_ctx* c);                // the scatters of species 0 and 1 in one launch
int launch_zero_f64(pic_ctx* c, double* p, int64_t n);
// CELL32: write buf[cur].cell from `start` when the step left it unwritten
int ensure_cells(pic_ctx* c, int s);

// injection (pic_inject.hip)
int inject_setup(pic_ctx* c, const pic_injection* cfg);
void inject_free(pic_ctx* c);
int inject_settle(pic_ctx* c);                    // raise n_upper by the last injection
int inject_generate(pic_ctx* c);                  // staged particles of this step
int inject_count(pic_ctx* c, int s);              // their slots in the cell counts
int inject_scatter(pic_ctx* c, int s, int dst);   // and into the sorted buffer
int inject_publish(pic_ctx* c);                   // counters to the host, asynchronously
int inject_append_rebin(pic_ctx* c);              // standalone injection into binned storage
int inject_stats(pic_ctx* c, int64_t* out);       // injections, reinjected, pairs, dropped, last e, last i
int launch_kinetic_energy(pic_ctx* c, int s, double* out);   // host result, synchronous

// Poisson (pic_poisson.hip)
int poisson_setup(pic_ctx* c);
void poisson_free(pic_ctx* c);
// phi and E.  defer: a direct solve returns with its check in flight (E
// enqueued); poisson_finish waits for it, runs PCG iterations if the check
// failed (*redo = 1: phi and E were recomputed) and fills st.
int poisson_solve(pic_ctx* c, int iteration, pic_solve_stats* st, int defer = 0);
int poisson_finish(pic_ctx* c, pic_solve_stats* st, int* redo);
// direct solve (pic_fastpoisson.hip): set-up for the geometry masks, solve of
// the current right-hand side into the interior nodes of a grid (Dirichlet
// values imposed by the caller)
int fast_setup(pic_ctx* c, const std::vector<uint8_t>& dir, const std::vector<double>& dval);
void fast_free(pic_ctx* c);
int fast_solve(pic_ctx* c, double* out);
// direct solves of many contexts at once (from their raw deposits); used[r] = 1
// for the contexts the batch solves
struct FdBatch;
int fast_batch_create(pic_ctx* const* runs, int n, FdBatch** out, unsigned char* used);
int fast_batch_enqueue(const FdBatch* b, hipStream_t s);
void fast_batch_free(FdBatch* b);
// weights over the diagonal of the two pre-smoothing sweeps of a level (the
// post-smoothing uses them in reverse order)
void mg_sweep_weights(const pic_ctx* c, const MgLevel& m, float* w0, float* w1);
int launch_overwrite_phi(pic_ctx* c);
int launch_overwrite_rho(pic_ctx* c);
int launch_efield(pic_ctx* c);

// timing helpers (pic_host.cpp)
int timer_begin(pic_ctx* c, int stage, hipEvent_t* ev);
int timer_end(pic_ctx* c, int stage, hipEvent_t ev0);
int kt_begin(pic_ctx* c, hipEvent_t ev[2], int stage);   // timing modes 2, 3: events of one kernel dispatch
void kt_end(pic_ctx* c, int stage, const hipEvent_t ev[2]);

} // namespace pic
