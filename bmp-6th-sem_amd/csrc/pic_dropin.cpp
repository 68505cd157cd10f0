// pic_dropin.cpp — the reference's plain C wrapper surface
// (SYCL-PIC-poissonsolver/src/sycl_pic_ctypes.cpp:46-272 and the extra calls
// of sycl_pic_module.cpp) over one global pic_ctx.
//
// rho and phi live on the device.  The pointers returned by
// sycl_pic_get_rho_ptr / sycl_pic_get_phi_ptr are pinned host mirrors: once a
// caller has asked for one, every call that changes that grid refreshes the
// mirror before returning, and calls that read it (poisson_solve for rho,
// update_fields_from_phi for phi) first copy the caller's writes back to the
// device — the observable behaviour of the reference's shared allocations.
#include <hip/hip_runtime.h>

#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <string>
#include <sys/stat.h>
#include <vector>

#include "pic_config.h"

namespace {

struct DropIn {
    pic_ctx* ctx = nullptr;
    pic::Picdat dat;
    int gx = 0, gy = 0, G = 0;
    double* rho_host = nullptr;
    double* phi_host = nullptr;
    bool rho_view = false, phi_view = false;
    std::vector<double> scratch;
};

DropIn g;

double now() {
    return std::chrono::duration<double>(std::chrono::high_resolution_clock::now().time_since_epoch()).count();
}

void report(int rc, const char* where) {
    if (rc != PIC_OK) std::fprintf(stderr, "%s failed (%d): %s\n", where, rc, pic_last_error());
}

void pull(int which, double* host) {
    if (host) report(pic_get_grid(g.ctx, which, host), "copy grid to host");
}

void push_grid(int which, double* host) {
    if (host) report(pic_set_grid(g.ctx, which, host), "copy grid to device");
}

bool dir_exists(const std::string& p) {
    struct stat st;
    return stat(p.c_str(), &st) == 0 && S_ISDIR(st.st_mode);
}

// Text grid writer with the reference's "%lf" row format (utils.cpp:829-1025).
bool write_grid(const std::string& path, const double* v, double scale) {
    FILE* fd = std::fopen(path.c_str(), "w");
    if (!fd) return false;
    for (int i = 0; i < g.gy; ++i) {
        int j;
        for (j = 0; j < g.gx - 1; ++j) std::fprintf(fd, "%lf ", v[i * g.gx + j] * scale);
        std::fprintf(fd, "%lf", v[i * g.gx + j] * scale);
        std::fprintf(fd, "\n");
    }
    std::fclose(fd);
    return true;
}

// printNumParticles (utils.cpp:993-1025): per-key counts, "%d " rows.
void write_counts(const char* dir, int iteration, int species) {
    std::string d = std::string(dir) + (species == 0 ? "/electronNumber/" : "/ionNumber/");
    if (!dir_exists(d)) return;
    std::vector<int32_t> cnt(g.G);
    if (pic_get_cell_counts(g.ctx, species, cnt.data()) != PIC_OK) return;
    FILE* fd = std::fopen((d + std::to_string(iteration)).c_str(), "w");
    if (!fd) return;
    for (int i = 0; i < g.gy; ++i) {
        int j;
        for (j = 0; j < g.gx - 1; ++j) std::fprintf(fd, "%d ", cnt[i * g.gx + j]);
        std::fprintf(fd, "%d", cnt[i * g.gx + j]);
        std::fprintf(fd, "\n");
    }
    std::fclose(fd);
}

} // namespace

extern "C" {

int sycl_pic_init(const char* dev_type, int wg, int ps, const char* input1, const char* input2) {
    (void)wg;
    (void)ps;
    (void)input1;
    (void)input2;
    // The reference picks a SYCL CPU or GPU queue and returns -1 for anything
    // else (sycl_pic_ctypes.cpp:53-60).  Both values run on the GPU here: the
    // library drives no other device, and the results do not depend on it.
    std::string dt(dev_type ? dev_type : "");
    if (dt != "gpu" && dt != "cpu") {
        std::fprintf(stderr, "sycl_pic_init: device must be 'cpu' or 'gpu'\n");
        return -1;
    }
    if (dt == "cpu") std::fprintf(stderr, "sycl_pic_init: device 'cpu' requested; running on the GPU\n");
    if (g.ctx) {
        std::fprintf(stderr, "sycl_pic_init: already initialised; call sycl_pic_finalize first\n");
        return PIC_ERR_STATE;
    }
    std::string err;
    int rc = pic::read_picdat("M_PICDAT.DAT", g.dat, err);
    if (rc != PIC_OK) {
        std::fprintf(stderr, "%s\n", err.c_str());
        return rc;
    }
    const pic::Picdat& d = g.dat;
    pic_params p;
    pic_params_default(&p);
    p.NX = d.NX;
    p.NY = d.NY;
    p.width = d.width;
    p.height = d.height;
    p.dt = d.dt;
    p.epsilon = d.epsilon;
    p.b0 = d.B0;
    p.surf_area = d.surf_area;
    p.periodic = d.periodic;
    p.layout = PIC_LAYOUT_F64;
    rc = pic_create(&p, &g.ctx);
    if (rc != PIC_OK) {
        report(rc, "sycl_pic_init: pic_create");
        return rc;
    }
    rc = pic_read_geometry(g.ctx, "./geometry.conf");
    if (rc != PIC_OK) {
        report(rc, "sycl_pic_init: readGeometry");
        pic_destroy(g.ctx);
        g.ctx = nullptr;
        return rc;
    }
    pic_grid_shape(g.ctx, &g.gx, &g.gy);
    g.G = g.gx * g.gy;
    const char* seed_env = std::getenv("BMP_PIC_SEED");
    uint64_t seed = seed_env ? std::strtoull(seed_env, nullptr, 10) : 1;
    const double dx = d.width / d.NX, dy = d.height / d.NY;
    (void)dx;
    (void)dy;
    for (int s = 0; s < d.num_type; ++s) {
        const pic::PicdatSpecies& row = d.species[s];
        int64_t n = (int64_t)row.sim;
        int id = 0;
        // storage for five times the initial particles (space, particles.cpp:183)
        rc = pic_add_species(g.ctx, row.charge, row.mass, row.sim, row.actual, 5 * n, &id);
        if (rc != PIC_OK) {
            report(rc, "sycl_pic_init: species");
            sycl_pic_finalize();
            return rc;
        }
        std::vector<double> x(n), y(n), vx(n), vy(n), vz(n);
        double temp = s == 0 ? d.Te : (s == 1 ? d.Ti : 0.0);   // initParticlesRandomly (:143-148)
        double x0 = 0, x1 = d.width, y0 = 0, y1 = d.height;
        if (!d.uniform) {
            x0 = d.fl * d.width;
            x1 = d.fr * d.width;
            y0 = d.ft * d.height;
            y1 = d.fb * d.height;
        }
        pic::generate_box(seed, s, n, x0, x1, y0, y1, d.width, d.height, nullptr, std::fabs(row.charge / row.mass),
                          temp, x.data(), y.data(), vx.data(), vy.data(), vz.data());
        rc = pic_upload_particles(g.ctx, id, n, x.data(), y.data(), vx.data(), vy.data(), vz.data());
        if (rc != PIC_OK) {
            report(rc, "sycl_pic_init: upload");
            sycl_pic_finalize();
            return rc;
        }
    }
    if (d.num_type >= 2) {
        // generate_new_electrons gets electronTemp as an int (utils.cpp:381);
        // generate_pairs uses the double globals (utils.cpp:646,664)
        pic_injection inj;
        pic_injection_default(&inj);
        inj.te_reinject_eV = (double)(int)d.Te;
        inj.te_eV = d.Te;
        inj.ti_eV = d.Ti;
        inj.seed = seed;
        rc = pic_set_injection(g.ctx, &inj);
        if (rc != PIC_OK) {
            report(rc, "sycl_pic_init: injection");
            sycl_pic_finalize();
            return rc;
        }
    }
    if (hipHostMalloc((void**)&g.rho_host, g.G * sizeof(double), hipHostMallocDefault) != hipSuccess ||
        hipHostMalloc((void**)&g.phi_host, g.G * sizeof(double), hipHostMallocDefault) != hipSuccess) {
        std::fprintf(stderr, "sycl_pic_init: pinned host allocation failed\n");
        sycl_pic_finalize();
        return PIC_ERR_ALLOC;
    }
    std::memset(g.rho_host, 0, g.G * sizeof(double));
    pull(PIC_GRID_PHI, g.phi_host);
    return 0;
}

int sycl_pic_get_grid_size(void) { return g.G; }
int sycl_pic_get_grid_x(void) { return g.gx; }
int sycl_pic_get_grid_y(void) { return g.gy; }
int sycl_pic_get_timesteps(void) { return g.dat.timesteps; }
int sycl_pic_get_print_interval(void) { return g.dat.print_interval; }
int sycl_pic_get_num_species(void) { return g.dat.num_type; }

int sycl_pic_get_particle_count(int species) {
    if (!g.ctx) return 0;
    return (int)pic_particle_count(g.ctx, species);
}

double* sycl_pic_get_rho_ptr(void) {
    if (!g.ctx) return nullptr;
    if (!g.rho_view) pull(PIC_GRID_RHO, g.rho_host);
    g.rho_view = true;
    return g.rho_host;
}

double* sycl_pic_get_phi_ptr(void) {
    if (!g.ctx) return nullptr;
    if (!g.phi_view) pull(PIC_GRID_PHI, g.phi_host);
    g.phi_view = true;
    return g.phi_host;
}

void sycl_pic_zero_rho(void) {
    if (!g.ctx) return;
    report(pic_zero_rho(g.ctx), "sycl_pic_zero_rho");
    if (g.rho_host) std::memset(g.rho_host, 0, g.G * sizeof(double));
}

double sycl_pic_charge_deposition(int species) {
    if (!g.ctx) return 0.0;
    double t0 = now();
    if (g.rho_view) push_grid(PIC_GRID_RHO, g.rho_host);
    report(pic_deposit(g.ctx, species), "sycl_pic_charge_deposition");
    if (g.rho_view) pull(PIC_GRID_RHO, g.rho_host);
    return now() - t0;
}

double sycl_pic_poisson_solve(int iteration) {
    if (!g.ctx) return 0.0;
    double t0 = now();
    if (g.rho_view) push_grid(PIC_GRID_RHO, g.rho_host);
    if (g.phi_view) push_grid(PIC_GRID_PHI, g.phi_host);
    report(pic_poisson_solve(g.ctx, iteration, nullptr), "sycl_pic_poisson_solve");
    if (g.rho_view) pull(PIC_GRID_RHO, g.rho_host);
    if (g.phi_view) pull(PIC_GRID_PHI, g.phi_host);
    return now() - t0;
}

void sycl_pic_update_fields_from_phi(void) {
    if (!g.ctx) return;
    if (g.phi_view) push_grid(PIC_GRID_PHI, g.phi_host);
    report(pic_update_fields_from_phi(g.ctx), "sycl_pic_update_fields_from_phi");
    if (g.phi_view) pull(PIC_GRID_PHI, g.phi_host);
}

double sycl_pic_move_particles(int species, int iteration) {
    if (!g.ctx) return 0.0;
    double t0 = now();
    report(pic_push(g.ctx, species, 1), "sycl_pic_move_particles");
    double t1 = now();
    // printNumParticles every printInterval steps (mover.cpp:421-425), when ./out exists
    if (g.dat.print_interval > 0 && iteration % g.dat.print_interval == 0) write_counts("./out", iteration, species);
    return t1 - t0;
}

double sycl_pic_compact_particles(int species, double* out_alloc_t, double* out_scan_t) {
    // The counting sort in move_particles already stores the live particles
    // contiguously; there is nothing left to compact.
    (void)species;
    if (out_alloc_t) *out_alloc_t = 0.0;
    if (out_scan_t) *out_scan_t = 0.0;
    return 0.0;
}

double sycl_pic_cleanup_map(int species) {
    // No hashmap: the sort leaves no deleted slots to reclaim.
    (void)species;
    return 0.0;
}

double sycl_pic_generate_new_particles(void) {
    // sycl_pic_ctypes.cpp:212-236: re-inject electrons for the x < 0 exits
    // since the last call, add this step's ionization pairs
    if (!g.ctx || g.dat.num_type < 2) return 0.0;
    auto t0 = std::chrono::high_resolution_clock::now();
    int rc = pic_inject(g.ctx);
    if (rc != PIC_OK) report(rc, "sycl_pic_generate_new_particles");
    auto t1 = std::chrono::high_resolution_clock::now();
    return std::chrono::duration<double>(t1 - t0).count();
}

void sycl_pic_print_diagnostics(const char* dir, int iteration) {
    if (!g.ctx || !dir) return;
    const pic::Picdat& d = g.dat;
    std::vector<double> f0(g.G), f1(g.G), e0(g.G), e1(g.G), ex(g.G), ey(g.G), ph(g.G);
    int ns = d.num_type;
    pull(PIC_GRID_FINAL0, f0.data());
    pull(PIC_GRID_ENERGY0, e0.data());
    if (ns > 1) {
        pull(PIC_GRID_FINAL0 + 1, f1.data());
        pull(PIC_GRID_ENERGY0 + 1, e1.data());
    }
    pull(PIC_GRID_EX, ex.data());
    pull(PIC_GRID_EY, ey.data());
    pull(PIC_GRID_PHI, ph.data());
    std::string it = std::to_string(iteration);
    std::string base(dir);
    auto must = [&](bool ok, const char* what) {
        if (!ok) std::fprintf(stderr, "File not Created: %s/%s/%s\n", dir, what, it.c_str());
    };
    // printNumberDensity (utils.cpp:942-991): final_mesh * scaleFactor
    must(write_grid(base + "/electronDensity/" + it, f0.data(), d.scale_factor), "electronDensity");
    if (ns > 1) must(write_grid(base + "/ionDensity/" + it, f1.data(), d.scale_factor), "ionDensity");
    // printEnergy (utils.cpp:897-940): 0.66667 * energy / (density + 1); the
    // reference's last electron column divides by energy + 1 instead, kept.
    auto energy_file = [&](const std::string& path, const std::vector<double>& en, const std::vector<double>& fm,
                           bool last_col_quirk) {
        FILE* fd = std::fopen(path.c_str(), "w");
        if (!fd) return false;
        for (int i = 0; i < g.gy; ++i) {
            int j;
            for (j = 0; j < g.gx - 1; ++j)
                std::fprintf(fd, "%lf ", 0.66667 * en[i * g.gx + j] / (fm[i * g.gx + j] + 1.0));
            double den = last_col_quirk ? en[i * g.gx + j] : fm[i * g.gx + j];
            std::fprintf(fd, "%lf\n", 0.66667 * en[i * g.gx + j] / (den + 1.0));
        }
        std::fclose(fd);
        return true;
    };
    must(energy_file(base + "/electronenergy/" + it, e0, f0, true), "electronenergy");
    if (ns > 1) must(energy_file(base + "/ionenergy/" + it, e1, f1, false), "ionenergy");
    // printElectricField (utils.cpp:852-895), printPhi (:829-850)
    must(write_grid(base + "/Ex/" + it, ex.data(), 1.0), "Ex");
    must(write_grid(base + "/Ey/" + it, ey.data(), 1.0), "Ey");
    must(write_grid(base + "/potential/" + it, ph.data(), 1.0), "potential");
}

void sycl_pic_save_meshes(void) {
    if (!g.ctx) return;
    for (int s = 0; s < g.dat.num_type; ++s) {
        std::vector<double> f(g.G);
        pull(PIC_GRID_FINAL0 + s, f.data());
        std::string file = "Mesh_" + std::to_string(s) + ".out";
        std::ofstream out(file);
        if (!out.is_open()) {
            std::fprintf(stderr, "Error opening output file: %s\n", file.c_str());
            continue;
        }
        for (int y = 0; y < g.gy; ++y) {
            for (int x = 0; x < g.gx; ++x) out << f[(size_t)y * g.gx + x] << " ";
            out << "\n";
        }
    }
}

void sycl_pic_finalize(void) {
    if (g.ctx) pic_destroy(g.ctx);
    if (g.rho_host) (void)hipHostFree(g.rho_host);
    if (g.phi_host) (void)hipHostFree(g.phi_host);
    g = DropIn{};
}

} // extern "C"
