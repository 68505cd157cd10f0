// pic_host.cpp — host runtime of the stage API: context lifetime, species
// storage, upload / download, stage sequencing and timing.  Every device
// operation is issued on the context's own stream.
#include "pic_internal.h"

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstring>

namespace pic {

static thread_local char g_err[1024] = "";

void set_error(const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
}

// Stage events come from a pool (returned by pic_stage_times) so the timed
// loop does not create events.
static int timer_event(pic_ctx* c, hipEvent_t* ev) {
    if (!c->event_pool.empty()) {
        *ev = c->event_pool.back();
        c->event_pool.pop_back();
        return PIC_OK;
    }
    PIC_HIP(hipEventCreate(ev));
    return PIC_OK;
}

int timer_begin(pic_ctx* c, int stage, hipEvent_t* ev) {
    *ev = nullptr;
    if (c->timing != 1) return PIC_OK;
    PIC_TRY(timer_event(c, ev));
    PIC_HIP(hipEventRecord(*ev, c->stream));
    return PIC_OK;
}

int timer_end(pic_ctx* c, int stage, hipEvent_t ev0) {
    if (c->timing != 1 || !ev0) return PIC_OK;
    hipEvent_t ev1;
    PIC_TRY(timer_event(c, &ev1));
    PIC_HIP(hipEventRecord(ev1, c->stream));
    c->timers[stage].pending.emplace_back(ev0, ev1);
    c->timers[stage].launches++;
    return PIC_OK;
}

// Kernel events (timing mode 2: every particle kernel; mode 3: the scatter
// kernel only): start / stop timestamps written by the dispatch of one kernel,
// no marker in the stream.
int kt_begin(pic_ctx* c, hipEvent_t ev[2], int stage) {
    ev[0] = ev[1] = nullptr;
    if (c->timing != 2 && !(c->timing == 3 && stage == PIC_STAGE_SCATTER)) return PIC_OK;
    PIC_TRY(timer_event(c, &ev[0]));
    PIC_TRY(timer_event(c, &ev[1]));
    return PIC_OK;
}

void kt_end(pic_ctx* c, int stage, const hipEvent_t ev[2]) {
    if (!ev[0]) return;
    c->timers[stage].pending.emplace_back(ev[0], ev[1]);
    c->timers[stage].launches++;
}

int launch_key_from_cell(pic_ctx* c, int s, int64_t n);   // pic_particles.hip

} // namespace pic

using namespace pic;

extern "C" const char* pic_last_error(void) { return g_err; }

extern "C" void pic_params_default(pic_params* p) {
    std::memset(p, 0, sizeof(*p));
    p->NX = 128;
    p->NY = 128;
    p->width = 128 * 1e-4;
    p->height = 128 * 1e-4;
    p->dt = 1e-11;
    p->epsilon = 8.854e-12;
    p->b0 = 0.0;
    p->bx = 0.0;
    p->by = 0.0;
    p->surf_area = 0.0;        // 0 = width * height
    p->periodic = 0;
    p->layout = PIC_LAYOUT_F64;
    p->charge_shift = 40;
    p->energy_shift = 32;
    p->deposit_energy = 1;
    p->poisson_tol = 1e-10;
    p->poisson_max_iter = 200;
    p->mg_pre = 2;
    p->mg_post = 2;
    p->mg_omega = 0.8;
    p->mg_smoother = 1;
    p->poisson_direct = 1;
    p->poisson_guess = 0;
    p->poisson_precision = 0;
    p->poisson_inner_tol = 1e-4;
    p->tile_cells = 0;
    p->deposit_group = 0;
    p->push_group = 0;
}

static void free_species(pic::Species& s) {
    for (auto& b : s.buf) {
        if (b.pa) (void)hipFree(b.pa);
        if (b.pb) (void)hipFree(b.pb);
        if (b.vx) (void)hipFree(b.vx);
        if (b.vy) (void)hipFree(b.vy);
        if (b.vz) (void)hipFree(b.vz);
        if (b.cell) (void)hipFree(b.cell);
        b = pic::ParticleBuf{};
    }
    void* ptrs[] = {s.ndev, s.key, s.rank, s.counts, s.start, s.start_old, s.blk_first, s.q_charge, s.q_energy,
                    s.final_mesh, s.energy_mesh, s.absorbed};
    for (void* q : ptrs)
        if (q) (void)hipFree(q);
}

extern "C" void pic_destroy(pic_ctx* c) {
    if (!c) return;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    inject_free(c);
    for (auto& s : c->sp) free_species(s);
    poisson_free(c);
    if (c->unet) pic_unet_destroy(c->unet);
    void* ptrs[] = {c->rho, c->phi, c->ex, c->ey, c->bz, c->fld, c->dir, c->dval, c->pr, c->pr2, c->pp, c->pp2,
                    c->pq, c->zf0, c->zf1, c->guess, c->phi_prev, c->part, c->scal, c->scan_bsum,
                    c->rf0, c->rf1, c->pf0, c->pf1, c->qf, c->df};
    for (void* q : ptrs)
        if (q) (void)hipFree(q);
    if (c->scal_host) (void)hipHostFree(c->scal_host);
    if (c->mail) (void)hipHostFree(c->mail);
    if (c->tstamp) (void)hipFree(c->tstamp);
    for (auto& t : c->timers)
        for (auto& e : t.pending) {
            (void)hipEventDestroy(e.first);
            (void)hipEventDestroy(e.second);
        }
    for (hipEvent_t e : c->event_pool) (void)hipEventDestroy(e);
    for (hipEvent_t e : c->solve_ev)
        if (e) (void)hipEventDestroy(e);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

static int node_of_fraction(int gridn, double frac) {
    // (int)(GRID - 1) * frac, assigned to an int: truncation (particles.cpp:677-680)
    return (int)((gridn - 1) * frac);
}

static int apply_geometry(pic_ctx* c, const double* rects, int n) {
    if (n < 0 || n > 4096) {
        set_error("geometry count %d out of range", n);
        return PIC_ERR_ARG;
    }
    if (c->p.periodic == 1 && n > 3) {
        set_error("Inconsistant input files: Periodic boundary expects geometry count = 3");
        return PIC_ERR_PERIODIC_GEOM;
    }
    std::vector<pic::Geometry> g;
    for (int i = 0; i < n; ++i) {
        const double* r = rects + 5 * i;
        pic::Geometry q;
        q.x1 = node_of_fraction(c->g.gx, r[0]);
        q.y1 = node_of_fraction(c->g.gy, r[1]);
        q.x2 = node_of_fraction(c->g.gx, r[2]);
        q.y2 = node_of_fraction(c->g.gy, r[3]);
        q.potential = r[4];
        g.push_back(q);
    }
    auto old = c->geom;
    c->geom = g;
    int rc = poisson_setup(c);
    if (rc != PIC_OK) {
        c->geom = old;
        if (!old.empty()) (void)poisson_setup(c);
        return rc;
    }
    c->have_prev_phi = 0;
    PIC_TRY(launch_overwrite_phi(c));
    PIC_HIP(hipStreamSynchronize(c->stream));
    return PIC_OK;
}

extern "C" int pic_create(const pic_params* p, pic_ctx** out) {
    if (!p || !out) {
        set_error("pic_create: null argument");
        return PIC_ERR_ARG;
    }
    *out = nullptr;
    if (p->NX < 2 || p->NY < 2 || !(p->width > 0) || !(p->height > 0) || !(p->dt > 0) || !(p->epsilon > 0)) {
        set_error("pic_create: invalid grid / time step / epsilon");
        return PIC_ERR_ARG;
    }
    if ((int64_t)(p->NX + 1) * (p->NY + 1) > (int64_t)2048 * 8192) {
        set_error("pic_create: grid of %d x %d cells exceeds the 16.7M-node limit", p->NX, p->NY);
        return PIC_ERR_ARG;
    }
    if (p->periodic != 0 && p->periodic != 1) {
        set_error("pic_create: periodic must be 0 or 1");
        return PIC_ERR_ARG;
    }
    if (p->tile_cells != 0 && p->tile_cells != 32 && p->tile_cells != 64 && p->tile_cells != 128 &&
        p->tile_cells != 256) {
        set_error("pic_create: tile_cells must be 0 (automatic), 32, 64, 128 or 256 (got %d)", p->tile_cells);
        return PIC_ERR_ARG;
    }
    if (p->push_group != 0 && p->push_group != 1 && p->push_group != 2 && p->push_group != 4) {
        set_error("pic_create: push_group must be 0 (automatic), 1, 2 or 4 (got %d)", p->push_group);
        return PIC_ERR_ARG;
    }
    if (p->deposit_group != 0 && p->deposit_group != 1 && p->deposit_group != 2 && p->deposit_group != 4) {
        set_error("pic_create: deposit_group must be 0 (automatic), 1, 2 or 4 (got %d)", p->deposit_group);
        return PIC_ERR_ARG;
    }
    if (p->poisson_direct < 0 || p->poisson_direct > 2) {
        set_error("pic_create: poisson_direct must be 0, 1 or 2");
        return PIC_ERR_ARG;
    }
    if (p->layout != PIC_LAYOUT_F64 && p->layout != PIC_LAYOUT_CELL32) {
        set_error("pic_create: unknown layout %d", p->layout);
        return PIC_ERR_ARG;
    }
    if (p->charge_shift < 8 || p->charge_shift > 52 || p->energy_shift < 0 || p->energy_shift > 52) {
        set_error("pic_create: fixed-point shifts out of range");
        return PIC_ERR_ARG;
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
        set_error("pic_create: no HIP device visible (this library runs on the GPU only)");
        return PIC_ERR_HIP;
    }
    pic_ctx* c = new pic_ctx();
    c->p = *p;
    if (c->p.surf_area <= 0) c->p.surf_area = p->width * p->height;
    c->g.gx = p->NX + 1;
    c->g.gy = p->NY + 1;
    c->g.width = p->width;
    c->g.height = p->height;
    c->g.dx = p->width / p->NX;     // dx = GRID_WIDTH / NX (particles.cpp:360)
    c->g.dy = p->height / p->NY;
    c->g.inv_dx = 1.0 / c->g.dx;
    c->g.inv_dy = 1.0 / c->g.dy;
    c->g.nx = p->NX;
    c->g.ny = p->NY;
    c->g.periodic = p->periodic;
    c->g.rcp_gx = 1.0f / (float)c->g.gx;
    c->G = c->g.gx * c->g.gy;
    c->g.gp = c->G + c->g.gx + 2;
    const size_t G = c->G;
    auto fail = [&](int rc) { pic_destroy(c); return rc; };
#define ALLOC(ptr, bytes)                                                      \
    do {                                                                       \
        if (hipMalloc((void**)&(ptr), (bytes)) != hipSuccess) {                \
            set_error("pic_create: hipMalloc of %zu bytes failed", (size_t)(bytes)); \
            return fail(PIC_ERR_ALLOC);                                        \
        }                                                                      \
    } while (0)
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        set_error("pic_create: hipStreamCreate failed");
        return fail(PIC_ERR_HIP);
    }
    const size_t pad = G + (size_t)c->g.gx + 2;   // gather reads up to gy*GX+gx+GX+1
    ALLOC(c->rho, G * sizeof(double));
    ALLOC(c->phi, G * sizeof(double));
    ALLOC(c->ex, G * sizeof(double));
    ALLOC(c->ey, G * sizeof(double));
    ALLOC(c->bz, G * sizeof(double));
    ALLOC(c->fld, pad * sizeof(double4));
    ALLOC(c->dir, G);
    ALLOC(c->dval, G * sizeof(double));
    ALLOC(c->pr, G * sizeof(double));
    ALLOC(c->pr2, G * sizeof(double));
    ALLOC(c->pp, G * sizeof(double));
    ALLOC(c->pp2, G * sizeof(double));
    ALLOC(c->pq, G * sizeof(double));
    ALLOC(c->zf0, G * sizeof(float));
    ALLOC(c->zf1, G * sizeof(float));
    ALLOC(c->guess, G * sizeof(double));
    ALLOC(c->phi_prev, G * sizeof(double));
    ALLOC(c->part, PIC_PART_DOUBLES * sizeof(double));
    // solver scalars and flags in one block (one copy brings both to the host)
    ALLOC(c->scal, PIC_STATE_BYTES);
    c->flags = (int*)(c->scal + PIC_SCAL_DOUBLES);
    c->scan_nb = (int)((G + 2047) / 2048);   // 2048-count chunks of the cell scan (<= 8192 by the node limit)
    ALLOC(c->scan_bsum, 2 * ((size_t)c->scan_nb + 1) * sizeof(int32_t));
#undef ALLOC
    if (hipHostMalloc((void**)&c->scal_host, PIC_STATE_BYTES, hipHostMallocDefault) != hipSuccess) {
        set_error("pic_create: hipHostMalloc failed");
        return fail(PIC_ERR_ALLOC);
    }
    c->flags_host = (int*)(c->scal_host + PIC_SCAL_DOUBLES);
    if (hipHostMalloc((void**)&c->mail, sizeof(SolveMail), hipHostMallocMapped | hipHostMallocCoherent) !=
            hipSuccess ||
        hipHostGetDevicePointer((void**)&c->mail_dev, c->mail, 0) != hipSuccess) {
        set_error("pic_create: mapped host memory for the solve check failed");
        return fail(PIC_ERR_ALLOC);
    }
    std::memset(c->mail, 0, sizeof(SolveMail));
    if (hipMalloc((void**)&c->tstamp, 2 * sizeof(unsigned long long)) != hipSuccess) {
        set_error("pic_create: device allocation failed");
        return fail(PIC_ERR_ALLOC);
    }
    {
        int dev = 0, khz = 0;
        PIC_HIP(hipGetDevice(&dev));
        PIC_HIP(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev));
        c->wall_khz = khz > 0 ? (double)khz : 100000.0;
    }
    (void)hipMemset(c->rho, 0, G * sizeof(double));
    (void)hipMemset(c->phi, 0, G * sizeof(double));
    (void)hipMemset(c->ex, 0, G * sizeof(double));
    (void)hipMemset(c->ey, 0, G * sizeof(double));
    (void)hipMemset(c->fld, 0, pad * sizeof(double4));
    (void)hipMemset(c->scal, 0, PIC_STATE_BYTES);
    (void)hipMemset(c->part, 0, PIC_PART_DOUBLES * sizeof(double));
    (void)hipMemset(c->pp, 0, G * sizeof(double));
    (void)hipMemset(c->pp2, 0, G * sizeof(double));
    (void)hipMemset(c->scal, 0, 16 * sizeof(double));
    {
        std::vector<double> b(G, p->b0);   // initMagneticField: uniform B0 (particles.cpp:709-719)
        c->bz_uniform = 1;
        c->bz0 = p->b0;
        if (hipMemcpy(c->bz, b.data(), G * sizeof(double), hipMemcpyHostToDevice) != hipSuccess) {
            set_error("pic_create: copy of Bz failed");
            return fail(PIC_ERR_HIP);
        }
    }
    // default geometry: four grounded walls (M_PICDAT/geometry.conf default);
    // periodic y keeps the two x walls only
    const double walls[20] = {0, 0, 0, 1, 0, 1, 0, 1, 1, 0, 0, 1, 1, 1, 0, 0, 0, 1, 0, 0};
    int rc = apply_geometry(c, walls, p->periodic ? 2 : 4);
    if (rc != PIC_OK) return fail(rc);
    // E = 0 initially (the reference computes E only after the first solve)
    std::vector<double4> f(G);
    {
        std::vector<double> b(G, p->b0);
        for (size_t i = 0; i < G; ++i) f[i] = make_double4(0.0, 0.0, b[i], 0.0);
        if (hipMemcpy(c->fld, f.data(), G * sizeof(double4), hipMemcpyHostToDevice) != hipSuccess) {
            set_error("pic_create: copy of the field array failed");
            return fail(PIC_ERR_HIP);
        }
    }
    if (hipDeviceSynchronize() != hipSuccess) {
        set_error("pic_create: device synchronisation failed");
        return fail(PIC_ERR_HIP);
    }
    *out = c;
    return PIC_OK;
}

extern "C" int pic_set_geometry(pic_ctx* c, const double* rects, int n) {
    if (!c || (!rects && n > 0)) {
        set_error("pic_set_geometry: null argument");
        return PIC_ERR_ARG;
    }
    return apply_geometry(c, rects, n);
}

extern "C" int pic_get_geometry(pic_ctx* c, int* out_nodes, double* out_pot, int* n) {
    if (!c || !n) {
        set_error("pic_get_geometry: null argument");
        return PIC_ERR_ARG;
    }
    int cap = *n;
    *n = (int)c->geom.size();
    for (int i = 0; i < (int)c->geom.size() && i < cap; ++i) {
        if (out_nodes) {
            out_nodes[4 * i + 0] = c->geom[i].x1;
            out_nodes[4 * i + 1] = c->geom[i].y1;
            out_nodes[4 * i + 2] = c->geom[i].x2;
            out_nodes[4 * i + 3] = c->geom[i].y2;
        }
        if (out_pot) out_pot[i] = c->geom[i].potential;
    }
    return PIC_OK;
}

// Parse geometry.conf with the reference's scanf formats (particles.cpp:647-684).
extern "C" int pic_read_geometry(pic_ctx* c, const char* path) {
    if (!c || !path) {
        set_error("pic_read_geometry: null argument");
        return PIC_ERR_ARG;
    }
    FILE* fd = fopen(path, "r");
    if (!fd) {
        set_error("File not found: %s", path);
        return PIC_ERR_FILE;
    }
    char u1[150], u2[150], tmp;
    int count = 0;
    if (fscanf(fd, "%149[^\n] %149[^\n] %d %c", u1, u2, &count, &tmp) <= 3) {
        fclose(fd);
        set_error("Inappropriate file Format: %s", path);
        return PIC_ERR_FORMAT;
    }
    if (c->p.periodic == 1 && count > 3) {
        fclose(fd);
        set_error("Inconsistant input files: Periodic boundary expects geometry count = 3");
        return PIC_ERR_PERIODIC_GEOM;
    }
    if (fscanf(fd, "%149[^\n] %149[^\n] ", u1, u2) <= 1) {
        fclose(fd);
        set_error("Inappropriate file Format: %s", path);
        return PIC_ERR_FORMAT;
    }
    if (count < 0 || count > 4096) {
        fclose(fd);
        set_error("Inappropriate file Format: %s (geometry count %d)", path, count);
        return PIC_ERR_FORMAT;
    }
    std::vector<double> rects(5 * (size_t)count);
    for (int i = 0; i < count; ++i) {
        double* r = &rects[5 * (size_t)i];
        if (fscanf(fd, "%lf %lf %lf %lf %lf", r, r + 1, r + 2, r + 3, r + 4) <= 4) {
            fclose(fd);
            set_error("Inappropriate file Format: %s (row %d)", path, i);
            return PIC_ERR_FORMAT;
        }
    }
    fclose(fd);
    return apply_geometry(c, rects.data(), count);
}

extern "C" int pic_add_species(pic_ctx* c, double charge, double mass, double sim_count, double actual_count,
                               int64_t capacity, int* species_out) {
    if (!c || !(mass > 0) || charge == 0 || !(sim_count > 0) || capacity < 0 || capacity > ((int64_t)1 << 31) - 1) {
        set_error("pic_add_species: invalid argument");
        return PIC_ERR_ARG;
    }
    if (c->sp.size() >= 8) {
        set_error("pic_add_species: at most 8 species");
        return PIC_ERR_ARG;
    }
    pic::Species s;
    s.charge = charge;
    s.mass = mass;
    s.qm = charge / mass;                        // initCtoM
    s.sim = sim_count;
    s.actual = actual_count;
    const double dx = c->g.dx, dy = c->g.dy;
    s.W = actual_count * c->p.surf_area / (sim_count * dx * dy * dx * dy);   // updateWeighingFactor
    s.capacity = capacity;
    // particle arrays rounded up to 64 entries: kernels load 4 consecutive
    // particles at once from 16-byte aligned groups
    const size_t n = ((size_t)std::max<int64_t>(capacity, 1) + 63) & ~(size_t)63;
    const size_t G = c->G;
    const bool f64 = c->p.layout == PIC_LAYOUT_F64;
    const size_t ps = f64 ? sizeof(double) : sizeof(float);
    bool ok = true;
    for (auto& b : s.buf) {
        ok = ok && hipMalloc(&b.pa, n * ps) == hipSuccess;
        ok = ok && hipMalloc(&b.pb, n * ps) == hipSuccess;
        ok = ok && hipMalloc(&b.vx, n * ps) == hipSuccess;
        ok = ok && hipMalloc(&b.vy, n * ps) == hipSuccess;
        ok = ok && hipMalloc(&b.vz, n * ps) == hipSuccess;
        ok = ok && hipMalloc((void**)&b.cell, n * sizeof(int32_t)) == hipSuccess;
    }
    ok = ok && hipMalloc((void**)&s.ndev, 2 * sizeof(int32_t)) == hipSuccess;
    ok = ok && hipMalloc((void**)&s.key, n * sizeof(int32_t)) == hipSuccess;
    ok = ok && hipMalloc((void**)&s.rank, n * sizeof(int32_t)) == hipSuccess;
    ok = ok && hipMalloc((void**)&s.counts, G * sizeof(int32_t)) == hipSuccess;
    ok = ok && hipMalloc((void**)&s.start, (G + 1) * sizeof(int32_t)) == hipSuccess;
    ok = ok && hipMalloc((void**)&s.start_old, (G + 1) * sizeof(int32_t)) == hipSuccess;
    ok = ok && hipMalloc((void**)&s.blk_first, (n / 256 + 2) * sizeof(int32_t)) == hipSuccess;
    const size_t nacc = 2 * (size_t)c->g.gp;   // two node-indexed meshes: node-row sums from below and above
    ok = ok && hipMalloc((void**)&s.q_charge, nacc * sizeof(int64_t)) == hipSuccess;
    ok = ok && hipMalloc((void**)&s.q_energy, nacc * sizeof(int64_t)) == hipSuccess;
    ok = ok && hipMalloc((void**)&s.final_mesh, G * sizeof(double)) == hipSuccess;
    ok = ok && hipMalloc((void**)&s.energy_mesh, G * sizeof(double)) == hipSuccess;
    ok = ok && hipMalloc((void**)&s.absorbed, 2 * sizeof(unsigned long long)) == hipSuccess;
    if (!ok) {
        free_species(s);
        set_error("pic_add_species: device allocation for %lld particles failed", (long long)capacity);
        return PIC_ERR_ALLOC;
    }
    (void)hipMemset(s.ndev, 0, 2 * sizeof(int32_t));
    (void)hipMemset(s.counts, 0, G * sizeof(int32_t));
    (void)hipMemset(s.start, 0, (G + 1) * sizeof(int32_t));
    (void)hipMemset(s.start_old, 0, (G + 1) * sizeof(int32_t));
    (void)hipMemset(s.blk_first, 0, (n / 256 + 2) * sizeof(int32_t));
    (void)hipMemset(s.q_charge, 0, nacc * sizeof(int64_t));
    (void)hipMemset(s.q_energy, 0, nacc * sizeof(int64_t));
    (void)hipMemset(s.final_mesh, 0, G * sizeof(double));
    (void)hipMemset(s.energy_mesh, 0, G * sizeof(double));
    (void)hipMemset(s.absorbed, 0, 2 * sizeof(unsigned long long));
    PIC_HIP(hipDeviceSynchronize());   // null-stream memsets done before the context stream uses them
    c->sp.push_back(s);
    if (species_out) *species_out = (int)c->sp.size() - 1;
    return PIC_OK;
}

static int check_species(pic_ctx* c, int s, const char* fn) {
    if (!c) {
        set_error("%s: null context", fn);
        return PIC_ERR_ARG;
    }
    if (s < 0 || s >= (int)c->sp.size()) {
        set_error("%s: species %d out of range", fn, s);
        return PIC_ERR_ARG;
    }
    return PIC_OK;
}

extern "C" int pic_upload_particles(pic_ctx* c, int s, int64_t n, const double* x, const double* y,
                                    const double* vx, const double* vy, const double* vz) {
    PIC_TRY(check_species(c, s, "pic_upload_particles"));
    pic::Species& sp = c->sp[s];
    if (n < 0 || n > sp.capacity || (n > 0 && (!x || !y || !vx || !vy || !vz))) {
        set_error("pic_upload_particles: n = %lld exceeds capacity %lld or null array", (long long)n,
                  (long long)sp.capacity);
        return PIC_ERR_RANGE;
    }
    for (int64_t i = 0; i < n; ++i)
        if (!(x[i] >= 0.0 && x[i] < c->g.width && y[i] >= 0.0 && y[i] < c->g.height)) {
            set_error("pic_upload_particles: particle %lld at (%g, %g) is outside the domain", (long long)i, x[i],
                      y[i]);
            return PIC_ERR_ARG;
        }
    double* d = nullptr;
    if (n > 0) {
        PIC_HIP(hipMalloc(&d, 5 * (size_t)n * sizeof(double)));
        const double* src[5] = {x, y, vx, vy, vz};
        for (int k = 0; k < 5; ++k)
            PIC_HIP(hipMemcpyAsync(d + k * n, src[k], n * sizeof(double), hipMemcpyHostToDevice, c->stream));
    }
    int rc = launch_encode_f64(c, s, n, d, d + n, d + 2 * n, d + 3 * n, d + 4 * n);
    if (rc == PIC_OK) rc = launch_bin(c, s);
    if (rc == PIC_OK && hipStreamSynchronize(c->stream) != hipSuccess) {
        set_error("pic_upload_particles: stream error");
        rc = PIC_ERR_HIP;
    }
    if (d) (void)hipFree(d);
    return rc;
}

extern "C" int pic_upload_cell32(pic_ctx* c, int s, int64_t n, const int32_t* cell, const float* ox,
                                 const float* oy, const float* vx, const float* vy, const float* vz) {
    PIC_TRY(check_species(c, s, "pic_upload_cell32"));
    pic::Species& sp = c->sp[s];
    if (c->p.layout != PIC_LAYOUT_CELL32) {
        set_error("pic_upload_cell32: context layout is not CELL32");
        return PIC_ERR_STATE;
    }
    if (n < 0 || n > sp.capacity) {
        set_error("pic_upload_cell32: n exceeds capacity");
        return PIC_ERR_RANGE;
    }
    for (int64_t i = 0; i < n; ++i) {
        int k = cell[i];
        if (k < 0 || k >= c->G || !(ox[i] >= 0.f && ox[i] < 1.f) || !(oy[i] >= 0.f && oy[i] < 1.f)) {
            set_error("pic_upload_cell32: particle %lld has an invalid cell or offset", (long long)i);
            return PIC_ERR_ARG;
        }
    }
    pic::ParticleBuf& b = sp.buf[sp.cur];
    if (n > 0) {
        PIC_HIP(hipMemcpyAsync(b.cell, cell, n * sizeof(int32_t), hipMemcpyHostToDevice, c->stream));
        PIC_HIP(hipMemcpyAsync(b.pa, ox, n * sizeof(float), hipMemcpyHostToDevice, c->stream));
        PIC_HIP(hipMemcpyAsync(b.pb, oy, n * sizeof(float), hipMemcpyHostToDevice, c->stream));
        PIC_HIP(hipMemcpyAsync(b.vx, vx, n * sizeof(float), hipMemcpyHostToDevice, c->stream));
        PIC_HIP(hipMemcpyAsync(b.vy, vy, n * sizeof(float), hipMemcpyHostToDevice, c->stream));
        PIC_HIP(hipMemcpyAsync(b.vz, vz, n * sizeof(float), hipMemcpyHostToDevice, c->stream));
    }
    sp.cells_stale = 0;   // the uploaded cells are the storage's
    PIC_TRY(launch_key_from_cell(c, s, n));
    sp.deposit_pending = 0;
    PIC_TRY(launch_bin(c, s));
    PIC_HIP(hipStreamSynchronize(c->stream));
    return PIC_OK;
}

static int live_count(pic_ctx* c, int s, int32_t* n) {
    pic::Species& sp = c->sp[s];
    PIC_HIP(hipMemcpyAsync(n, sp.ndev + sp.cur, sizeof(int32_t), hipMemcpyDeviceToHost, c->stream));
    PIC_HIP(hipStreamSynchronize(c->stream));
    return PIC_OK;
}

extern "C" int64_t pic_particle_count(pic_ctx* c, int s) {
    if (check_species(c, s, "pic_particle_count") != PIC_OK) return -1;
    int32_t n = 0;
    if (live_count(c, s, &n) != PIC_OK) return -1;
    return n;
}

extern "C" int pic_download_cell32(pic_ctx* c, int s, int32_t* cell, float* ox, float* oy, float* vx, float* vy,
                                   float* vz) {
    PIC_TRY(check_species(c, s, "pic_download_cell32"));
    if (c->p.layout != PIC_LAYOUT_CELL32) {
        set_error("pic_download_cell32: context layout is not CELL32");
        return PIC_ERR_STATE;
    }
    if (cell) PIC_TRY(pic::ensure_cells(c, s));
    int32_t n = 0;
    PIC_TRY(live_count(c, s, &n));
    pic::ParticleBuf& b = c->sp[s].buf[c->sp[s].cur];
    if (n > 0) {
        if (cell) PIC_HIP(hipMemcpy(cell, b.cell, n * sizeof(int32_t), hipMemcpyDeviceToHost));
        if (ox) PIC_HIP(hipMemcpy(ox, b.pa, n * sizeof(float), hipMemcpyDeviceToHost));
        if (oy) PIC_HIP(hipMemcpy(oy, b.pb, n * sizeof(float), hipMemcpyDeviceToHost));
        if (vx) PIC_HIP(hipMemcpy(vx, b.vx, n * sizeof(float), hipMemcpyDeviceToHost));
        if (vy) PIC_HIP(hipMemcpy(vy, b.vy, n * sizeof(float), hipMemcpyDeviceToHost));
        if (vz) PIC_HIP(hipMemcpy(vz, b.vz, n * sizeof(float), hipMemcpyDeviceToHost));
    }
    return PIC_OK;
}

extern "C" int pic_download_particles(pic_ctx* c, int s, double* x, double* y, double* vx, double* vy, double* vz,
                                      int32_t* cell) {
    PIC_TRY(check_species(c, s, "pic_download_particles"));
    int32_t n = 0;
    PIC_TRY(live_count(c, s, &n));
    if (n == 0) return PIC_OK;
    pic::ParticleBuf& b = c->sp[s].buf[c->sp[s].cur];
    if (c->p.layout == PIC_LAYOUT_F64) {
        if (x) PIC_HIP(hipMemcpy(x, b.pa, n * sizeof(double), hipMemcpyDeviceToHost));
        if (y) PIC_HIP(hipMemcpy(y, b.pb, n * sizeof(double), hipMemcpyDeviceToHost));
        if (vx) PIC_HIP(hipMemcpy(vx, b.vx, n * sizeof(double), hipMemcpyDeviceToHost));
        if (vy) PIC_HIP(hipMemcpy(vy, b.vy, n * sizeof(double), hipMemcpyDeviceToHost));
        if (vz) PIC_HIP(hipMemcpy(vz, b.vz, n * sizeof(double), hipMemcpyDeviceToHost));
        if (cell) PIC_HIP(hipMemcpy(cell, b.cell, n * sizeof(int32_t), hipMemcpyDeviceToHost));
        return PIC_OK;
    }
    std::vector<int32_t> k(n);
    std::vector<float> a(n), bb(n), u(n), v(n), w(n);
    PIC_TRY(pic_download_cell32(c, s, k.data(), a.data(), bb.data(), u.data(), v.data(), w.data()));
    const int gx = c->g.gx;
    for (int32_t i = 0; i < n; ++i) {
        int cgx = k[i] % gx, cgy = k[i] / gx;
        if (x) x[i] = ((double)cgx + (double)a[i]) * c->g.dx;
        if (y) y[i] = ((double)cgy + (double)bb[i]) * c->g.dy;
        if (vx) vx[i] = u[i];
        if (vy) vy[i] = v[i];
        if (vz) vz[i] = w[i];
        if (cell) cell[i] = k[i];
    }
    return PIC_OK;
}

extern "C" int pic_bin(pic_ctx* c, int s) {
    PIC_TRY(check_species(c, s, "pic_bin"));
    // re-bin the current storage: keys from the stored cells
    pic::Species& sp = c->sp[s];
    int32_t n = 0;
    PIC_TRY(live_count(c, s, &n));
    PIC_TRY(launch_key_from_cell(c, s, n));
    PIC_TRY(launch_bin(c, s));
    PIC_HIP(hipStreamSynchronize(c->stream));
    (void)sp;
    return PIC_OK;
}

extern "C" int pic_get_cell_start(pic_ctx* c, int s, int32_t* start) {
    PIC_TRY(check_species(c, s, "pic_get_cell_start"));
    if (!start) {
        set_error("pic_get_cell_start: null output");
        return PIC_ERR_ARG;
    }
    PIC_HIP(hipStreamSynchronize(c->stream));
    PIC_HIP(hipMemcpy(start, c->sp[s].start, (c->G + 1) * sizeof(int32_t), hipMemcpyDeviceToHost));
    return PIC_OK;
}

extern "C" int pic_get_cell_counts(pic_ctx* c, int s, int32_t* counts) {
    PIC_TRY(check_species(c, s, "pic_get_cell_counts"));
    if (!counts) {
        set_error("pic_get_cell_counts: null output");
        return PIC_ERR_ARG;
    }
    std::vector<int32_t> st(c->G + 1);
    PIC_TRY(pic_get_cell_start(c, s, st.data()));
    for (int i = 0; i < c->G; ++i) counts[i] = st[i + 1] - st[i];
    return PIC_OK;
}

extern "C" int pic_zero_rho(pic_ctx* c) {
    if (!c) {
        set_error("pic_zero_rho: null context");
        return PIC_ERR_ARG;
    }
    PIC_TRY(launch_zero_f64(c, c->rho, c->G));
    PIC_HIP(hipStreamSynchronize(c->stream));
    return PIC_OK;
}

extern "C" int pic_deposit(pic_ctx* c, int s) {
    PIC_TRY(check_species(c, s, "pic_deposit"));
    PIC_TRY(launch_deposit(c, s));
    PIC_HIP(hipStreamSynchronize(c->stream));
    return PIC_OK;
}

extern "C" int pic_get_deposit_fixed(pic_ctx* c, int s, int64_t* qc, int64_t* qe) {
    PIC_TRY(check_species(c, s, "pic_get_deposit_fixed"));
    PIC_HIP(hipStreamSynchronize(c->stream));
    const size_t gp = c->g.gp;
    std::vector<int64_t> tmp(2 * gp);
    int64_t* outs[2] = {qc, qe};
    int64_t* srcs[2] = {c->sp[s].q_charge, c->sp[s].q_energy};
    for (int a = 0; a < 2; ++a) {
        if (!outs[a]) continue;
        PIC_HIP(hipMemcpy(tmp.data(), srcs[a], 2 * gp * sizeof(int64_t), hipMemcpyDeviceToHost));
        for (int n = 0; n < c->G; ++n) outs[a][n] = tmp[n] + tmp[gp + n];
    }
    return PIC_OK;
}

extern "C" int pic_poisson_solve(pic_ctx* c, int iteration, pic_solve_stats* st) {
    if (!c) {
        set_error("pic_poisson_solve: null context");
        return PIC_ERR_ARG;
    }
    PIC_TRY(poisson_solve(c, iteration, st));   // also computes E
    PIC_HIP(hipStreamSynchronize(c->stream));
    return PIC_OK;
}

extern "C" int pic_set_guess(pic_ctx* c, const double* phi0) {
    if (!c || !phi0) {
        set_error("pic_set_guess: null argument");
        return PIC_ERR_ARG;
    }
    PIC_HIP(hipMemcpy(c->guess, phi0, c->G * sizeof(double), hipMemcpyHostToDevice));
    c->guess_set = 1;
    return PIC_OK;
}

extern "C" int pic_update_fields_from_phi(pic_ctx* c) {
    if (!c) {
        set_error("pic_update_fields_from_phi: null context");
        return PIC_ERR_ARG;
    }
    PIC_TRY(launch_overwrite_phi(c));
    PIC_TRY(launch_efield(c));
    PIC_HIP(hipStreamSynchronize(c->stream));
    c->have_prev_phi = 1;
    return PIC_OK;
}

extern "C" int pic_push(pic_ctx* c, int s, int rebin) {
    PIC_TRY(check_species(c, s, "pic_push"));
    if (rebin)
        PIC_TRY(launch_move(c, s));
    else
        PIC_TRY(launch_push(c, s));
    PIC_HIP(hipStreamSynchronize(c->stream));
    return PIC_OK;
}

extern "C" int pic_step(pic_ctx* c, int iteration, pic_solve_stats* st) {
    if (!c) {
        set_error("pic_step: null context");
        return PIC_ERR_ARG;
    }
    // both species' tiles in one launch; their node pass is the solve's first kernel
    PIC_TRY(launch_deposit_all(c, true));
    // Poisson and E; a direct solve's check is still in flight: the pushes
    // are enqueued behind it (they do nothing unless the check passed), so
    // the device has work while the host waits for the check
    const int ns = (int)c->sp.size();
    PIC_TRY(poisson_solve(c, iteration, st, 1));
    const bool early = c->tail.active && ns > 0;
    if (early) PIC_TRY(move_push_all(c, c->flags));
    int redo = 0;
    PIC_TRY(poisson_finish(c, st, &redo));
    // the PCG corrected phi and E (push again), or there was no check
    if (!early || redo) PIC_TRY(move_push_all(c));
    if (!c->inj.on) {
        if (ns == 2) {
            // both scans in one set of launches, both scatters in one launch
            PIC_TRY(move_scan_pair(c));
            return move_scatter_pair(c);
        }
        for (int s = 0; s < ns; ++s) PIC_TRY(move_finish(c, s));
        return PIC_OK;
    }
    // with injection (main.cpp:232-245): after the pushes of every species,
    // stage the new particles (their count depends on both species' exits),
    // count them into the cells, then scan and scatter pushed and new
    // particles together
    hipEvent_t ev;
    PIC_TRY(timer_begin(c, PIC_STAGE_BIN, &ev));
    PIC_TRY(inject_generate(c));
    PIC_TRY(inject_count(c, 0));
    PIC_TRY(inject_count(c, 1));
    PIC_TRY(timer_end(c, PIC_STAGE_BIN, ev));
    if (ns == 2) {
        PIC_TRY(move_scan_pair(c));
        PIC_TRY(move_scatter_pair(c));
    } else {
        for (int s = 0; s < ns; ++s) PIC_TRY(move_finish(c, s));
    }
    PIC_TRY(timer_begin(c, PIC_STAGE_SCATTER, &ev));
    for (int s = 0; s < std::min(ns, 2); ++s) PIC_TRY(inject_scatter(c, s, c->sp[s].cur));
    PIC_TRY(timer_end(c, PIC_STAGE_SCATTER, ev));
    return inject_publish(c);
}

static double* grid_ptr(pic_ctx* c, int which) {
    switch (which) {
    case PIC_GRID_RHO: return c->rho;
    case PIC_GRID_PHI: return c->phi;
    case PIC_GRID_EX: return c->ex;
    case PIC_GRID_EY: return c->ey;
    case PIC_GRID_BZ: return c->bz;
    default: break;
    }
    if (which >= PIC_GRID_FINAL0 && which < PIC_GRID_FINAL0 + (int)c->sp.size())
        return c->sp[which - PIC_GRID_FINAL0].final_mesh;
    if (which >= PIC_GRID_ENERGY0 && which < PIC_GRID_ENERGY0 + (int)c->sp.size())
        return c->sp[which - PIC_GRID_ENERGY0].energy_mesh;
    return nullptr;
}

extern "C" int pic_get_grid(pic_ctx* c, int which, double* out) {
    if (!c || !out) {
        set_error("pic_get_grid: null argument");
        return PIC_ERR_ARG;
    }
    double* p = grid_ptr(c, which);
    if (!p) {
        set_error("pic_get_grid: unknown grid %d", which);
        return PIC_ERR_ARG;
    }
    if (which >= PIC_GRID_FINAL0) PIC_TRY(materialize_meshes(c));   // final_mesh / energy_mesh
    if (which == PIC_GRID_EX || which == PIC_GRID_EY) PIC_TRY(materialize_e(c));
    PIC_HIP(hipStreamSynchronize(c->stream));
    PIC_HIP(hipMemcpy(out, p, c->G * sizeof(double), hipMemcpyDeviceToHost));
    return PIC_OK;
}

__global__ void k_pack_fld(int G, const double* ex, const double* ey, const double* bz, double4* fld) {
    int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n < G) fld[n] = make_double4(ex[n], ey[n], bz[n], 0.0);
}

__global__ void k_unpack_e(int G, const double4* fld, double* ex, double* ey) {
    int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n < G) {
        const double4 v = fld[n];
        ex[n] = v.x;
        ey[n] = v.y;
    }
}

namespace pic {
int materialize_e(pic_ctx* c) {
    if (!c->e_stale) return PIC_OK;
    hipLaunchKernelGGL(k_unpack_e, dim3((c->G + 255) / 256), dim3(256), 0, c->stream, c->G, (const double4*)c->fld,
                       c->ex, c->ey);
    PIC_HIP(hipGetLastError());
    c->e_stale = 0;
    return PIC_OK;
}
} // namespace pic

extern "C" int pic_set_grid(pic_ctx* c, int which, const double* in) {
    if (!c || !in) {
        set_error("pic_set_grid: null argument");
        return PIC_ERR_ARG;
    }
    double* p = grid_ptr(c, which);
    if (!p) {
        set_error("pic_set_grid: unknown grid %d", which);
        return PIC_ERR_ARG;
    }
    // the lagging grids first, so that the caller's values are not overwritten
    // later and a repack of fld sees current ex / ey
    if (which >= PIC_GRID_FINAL0) PIC_TRY(materialize_meshes(c));
    if (which == PIC_GRID_EX || which == PIC_GRID_EY || which == PIC_GRID_BZ) PIC_TRY(materialize_e(c));
    PIC_HIP(hipStreamSynchronize(c->stream));
    PIC_HIP(hipMemcpy(p, in, c->G * sizeof(double), hipMemcpyHostToDevice));
    if (which == PIC_GRID_BZ) {
        // the push's constant-rotation mode needs one value on every node
        c->bz0 = in[0];
        c->bz_uniform = 1;
        for (int n = 1; n < c->G; ++n)
            if (in[n] != in[0]) {
                c->bz_uniform = 0;
                break;
            }
    }
    if (which == PIC_GRID_EX || which == PIC_GRID_EY || which == PIC_GRID_BZ) {
        hipLaunchKernelGGL(k_pack_fld, dim3((c->G + 255) / 256), dim3(256), 0, c->stream, c->G, c->ex, c->ey, c->bz,
                           c->fld);
        PIC_HIP(hipGetLastError());
        PIC_HIP(hipStreamSynchronize(c->stream));
    }
    if (which == PIC_GRID_PHI) c->have_prev_phi = 1;
    return PIC_OK;
}

extern "C" int pic_get_absorbed(pic_ctx* c, int s, int64_t* left, int64_t* total) {
    PIC_TRY(check_species(c, s, "pic_get_absorbed"));
    unsigned long long v[2] = {0, 0};
    PIC_HIP(hipStreamSynchronize(c->stream));
    PIC_HIP(hipMemcpy(v, c->sp[s].absorbed, sizeof(v), hipMemcpyDeviceToHost));
    if (left) *left = (int64_t)v[0];
    if (total) *total = (int64_t)v[1];
    return PIC_OK;
}

extern "C" int pic_kinetic_energy(pic_ctx* c, int s, double* energy) {
    PIC_TRY(check_species(c, s, "pic_kinetic_energy"));
    if (!energy) {
        set_error("pic_kinetic_energy: null output");
        return PIC_ERR_ARG;
    }
    return launch_kinetic_energy(c, s, energy);
}

extern "C" void pic_injection_default(pic_injection* q) {
    if (!q) return;
    q->reinject = 1;
    q->pairs = 1;
    q->x_inject = 0.024;          // utils.cpp:400
    q->te_reinject_eV = 10;       // electronTemp (types.cpp:89) as the int argument of utils.cpp:381
    q->te_eV = 10;
    q->ti_eV = 0.026;             // ionTemp (types.cpp:90)
    q->x11 = 0.0025;              // types.cpp:103-105
    q->x22 = 0.01;
    q->xm = 0.00625;
    q->ionimax = 5.23e23;         // utils.cpp:619
    q->max_reinject = 1 << 20;
    q->seed = 1;
}

extern "C" int pic_set_injection(pic_ctx* c, const pic_injection* cfg) {
    if (!c) {
        set_error("pic_set_injection: null context");
        return PIC_ERR_ARG;
    }
    return inject_setup(c, cfg);
}

extern "C" int pic_inject(pic_ctx* c) {
    if (!c) {
        set_error("pic_inject: null context");
        return PIC_ERR_ARG;
    }
    PIC_TRY(inject_append_rebin(c));
    PIC_HIP(hipStreamSynchronize(c->stream));
    return PIC_OK;
}

extern "C" int pic_injection_target(pic_ctx* c, int64_t* target, int64_t* attempts) {
    if (!c) {
        set_error("pic_injection_target: null context");
        return PIC_ERR_ARG;
    }
    if (target) *target = c->inj.on ? c->inj.target : 0;
    if (attempts) *attempts = c->inj.on ? c->inj.attempts : 0;
    return PIC_OK;
}

extern "C" int pic_injection_stats(pic_ctx* c, int64_t* injections, int64_t* reinjected, int64_t* pairs,
                                   int64_t* dropped, int64_t* last_electrons, int64_t* last_ions) {
    if (!c) {
        set_error("pic_injection_stats: null context");
        return PIC_ERR_ARG;
    }
    int64_t v[6];
    PIC_TRY(inject_stats(c, v));
    int64_t* outs[6] = {injections, reinjected, pairs, dropped, last_electrons, last_ions};
    for (int i = 0; i < 6; ++i)
        if (outs[i]) *outs[i] = v[i];
    return PIC_OK;
}

extern "C" int pic_sync(pic_ctx* c) {
    if (!c) {
        set_error("pic_sync: null context");
        return PIC_ERR_ARG;
    }
    PIC_HIP(hipStreamSynchronize(c->stream));
    return PIC_OK;
}

extern "C" int pic_set_timing(pic_ctx* c, int on) {
    if (!c) {
        set_error("pic_set_timing: null context");
        return PIC_ERR_ARG;
    }
    if (on < 0 || on > 3) {
        set_error("pic_set_timing: mode %d (0 off, 1 stages, 2 particle kernels, 3 scatter kernel)", on);
        return PIC_ERR_ARG;
    }
    c->timing = on;
    return PIC_OK;
}

extern "C" int pic_stage_times(pic_ctx* c, double* ms, int64_t* launches) {
    if (!c) {
        set_error("pic_stage_times: null context");
        return PIC_ERR_ARG;
    }
    PIC_HIP(hipStreamSynchronize(c->stream));
    for (int k = 0; k < PIC_NSTAGES; ++k) {
        auto& t = c->timers[k];
        for (auto& e : t.pending) {
            float m = 0;
            PIC_HIP(hipEventElapsedTime(&m, e.first, e.second));
            t.ms += m;
            c->event_pool.push_back(e.first);
            c->event_pool.push_back(e.second);
        }
        t.pending.clear();
        if (ms) ms[k] = t.ms;
        if (launches) launches[k] = t.launches;
        t.ms = 0;
        t.launches = 0;
    }
    return PIC_OK;
}

extern "C" int pic_grid_shape(pic_ctx* c, int* gx, int* gy) {
    if (!c) {
        set_error("pic_grid_shape: null context");
        return PIC_ERR_ARG;
    }
    if (gx) *gx = c->g.gx;
    if (gy) *gy = c->g.gy;
    return PIC_OK;
}

extern "C" int pic_learned_enable(pic_ctx* c, int train_steps, int min_trained, float lr, uint64_t seed) {
    if (!c || train_steps < 0 || min_trained < 0 || !(lr > 0.f)) {
        set_error("pic_learned_enable: invalid argument");
        return PIC_ERR_ARG;
    }
    if (c->p.periodic) {
        set_error("pic_learned_enable: the learned guess is trained on the non-periodic operator");
        return PIC_ERR_UNSUPPORTED;
    }
    const int n = (std::min(c->p.NX, c->p.NY) / 128) * 128;
    if (n < 128) {
        set_error("pic_learned_enable: the grid needs at least 128 x 128 cells (got %d x %d)", c->p.NX, c->p.NY);
        return PIC_ERR_UNSUPPORTED;
    }
    if (c->unet) {
        pic_unet_destroy(c->unet);
        c->unet = nullptr;
    }
    PIC_TRY(pic_unet_create(n, seed, &c->unet));
    c->learn_train = train_steps;
    c->learn_min = min_trained;
    c->learn_lr = lr;
    c->learn_trained = c->learn_tried = c->learn_used = 0;
    c->learn_loss = 0.f;
    return PIC_OK;
}

extern "C" int pic_learned_stats(pic_ctx* c, int* trained, int* tried, int* used, float* last_loss) {
    if (!c) {
        set_error("pic_learned_stats: null context");
        return PIC_ERR_ARG;
    }
    if (trained) *trained = c->learn_trained;
    if (tried) *tried = c->learn_tried;
    if (used) *used = c->learn_used;
    if (last_loss) *last_loss = c->learn_loss;
    return PIC_OK;
}
