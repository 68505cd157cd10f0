// pic_config.cpp — M_PICDAT.DAT parser and the project's seeded particle
// generator (host code).
#include "pic_config.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <thread>
#include <vector>

namespace pic {

// ------------------------------------------------------------- generator
// splitmix64 finaliser; the per-particle stream is a hash of
// (seed, stream, index), so output does not depend on n or thread count.
static inline uint64_t mix64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

struct Rng {
    uint64_t s;
    explicit Rng(uint64_t seed, uint64_t stream, uint64_t idx)
        : s(mix64(seed ^ mix64(0x5EEDull + stream * 0x100000001B3ull)) ^ mix64(idx * 0xD1B54A32D192ED03ull)) {}
    uint64_t next() { return mix64(s += 0x9E3779B97F4A7C15ull); }
    double u01() { return (double)(next() >> 11) * 0x1.0p-53; }              // [0, 1)
    double u01_open0() { return (double)((next() >> 11) + 1) * 0x1.0p-53; }  // (0, 1]
};

void generate_box(uint64_t seed, int stream, int64_t n, double x0, double x1, double y0, double y1,
                  double xmax, double ymax, const double* excl_m, double qm_abs, double temp_eV,
                  double* x, double* y, double* vx, double* vy, double* vz) {
    const double v = std::sqrt(2 * qm_abs * temp_eV);
    const double two_pi = 6.283185307179586;
    auto work = [&](int64_t a, int64_t b) {
        for (int64_t i = a; i < b; ++i) {
            Rng r(seed, (uint64_t)stream, (uint64_t)i);
            double px = 0, py = 0;
            for (int tries = 0; tries < 256; ++tries) {
                px = x0 + (x1 - x0) * r.u01();
                py = y0 + (y1 - y0) * r.u01();
                if (px >= xmax || py >= ymax) continue;
                if (excl_m && px >= excl_m[0] && px <= excl_m[2] && py >= excl_m[1] && py <= excl_m[3]) continue;
                break;
            }
            double r1 = std::sqrt(-std::log(r.u01_open0()));
            double a1 = two_pi * r.u01();
            double r2 = std::sqrt(-std::log(r.u01_open0()));
            double a2 = two_pi * r.u01();
            x[i] = px;
            y[i] = py;
            vx[i] = v * r1 * std::cos(a1);
            vy[i] = v * r1 * std::sin(a1);
            vz[i] = v * r2 * std::cos(a2);
        }
    };
    int nt = (int)std::min<int64_t>(std::max(1u, std::thread::hardware_concurrency()), 32);
    if (n < 100000) nt = 1;
    if (nt == 1) {
        work(0, n);
        return;
    }
    std::vector<std::thread> th;
    int64_t chunk = (n + nt - 1) / nt;
    for (int t = 0; t < nt; ++t) {
        int64_t a = t * chunk, b = std::min<int64_t>(n, a + chunk);
        if (a < b) th.emplace_back(work, a, b);
    }
    for (auto& t : th) t.join();
}

// --------------------------------------------------------------- PICDAT
// Same field order and scanf formats as readPICDAT (particles.cpp:308-633).
int read_picdat(const char* path, Picdat& d, std::string& err) {
    FILE* fd = fopen(path, "r");
    if (!fd) {
        err = std::string("File not found: ") + path;
        return PIC_ERR_FILE;
    }
    static char u1[8000], u2[8000];
    char t;
    auto fail = [&](const char* what) {
        fclose(fd);
        err = std::string("Inappropriate file Format: ") + path + " (" + what + ")";
        return PIC_ERR_FORMAT;
    };
    auto skip2 = [&](bool trailing) {
        return fscanf(fd, trailing ? "%7999[^\n] %7999[^\n] " : "%7999[^\n] %7999[^\n]", u1, u2) >= 2;
    };
    if (!skip2(false)) return fail("header");
    if (fscanf(fd, "%d %c", &d.timesteps, &t) < 2) return fail("TIMESTEPS");
    if (!skip2(false)) return fail("header 2");
    if (fscanf(fd, "%lf %lf %d %d %c", &d.width, &d.height, &d.NX, &d.NY, &t) < 4) return fail("GRID_WIDTH, GRID_HEIGHT, NX, NY");
    if (!skip2(false)) return fail("header 3");
    if (fscanf(fd, "%lf %c", &d.dt, &t) < 2) return fail("dt");
    if (!skip2(true)) return fail("header 4");
    if (fscanf(fd, "%lf %lf %lf %c", &d.B0, &d.sigma, &d.mean, &t) < 4) return fail("B0, sigma, mean");
    d.B0 = d.B0 / 1000;             // mT -> T
    d.mean = d.mean * d.width;
    d.sigma = d.sigma / 4;
    if (!skip2(true)) return fail("header 5");
    if (fscanf(fd, "%lf %lf %lf %c", &d.Te, &d.Ti, &d.Tg, &t) < 4) return fail("electronTemp, ionTemp, gasTemp");
    if (!skip2(true)) return fail("header 6");
    if (fscanf(fd, "%lf %c", &d.pressure, &t) < 2) return fail("pressure");
    if (!skip2(true)) return fail("header 7");
    if (fscanf(fd, "%d %c", &d.uniform, &t) < 2) return fail("Uniform Initialization Flag");
    if (!skip2(true)) return fail("header 8");
    if (fscanf(fd, "%lf %lf %lf %lf %c", &d.fl, &d.fr, &d.ft, &d.fb, &t) < 5) return fail("xx1 xx2 yy1 yy2");
    d.surf_area = d.uniform ? d.width * d.height : (d.fr - d.fl) * d.width * (d.fb - d.ft) * d.height;
    if (!skip2(true)) return fail("header 9");
    if (fscanf(fd, "%d %c", &d.reinjection, &t) < 2) return fail("reinjection flag");
    if (!skip2(true)) return fail("header 10");
    if (fscanf(fd, "%d %c", &d.scale_factor, &t) < 2) return fail("scale factor");
    if (!skip2(true)) return fail("header 11");
    if (!skip2(true)) return fail("header 12");
    if (fscanf(fd, "%d %c", &d.init_flag, &t) < 2) return fail("initParticleFlag");
    if (!skip2(true)) return fail("header 13");
    if (fscanf(fd, "%lf %c", &d.epsilon, &t) < 2) return fail("epsilon");
    if (!skip2(true)) return fail("header 14");
    if (fscanf(fd, "%d %c", &d.periodic, &t) < 2) return fail("periodic boundary flag");
    if (!skip2(true)) return fail("header 15");
    if (fscanf(fd, "%29[^\n] %c", d.gas, &t) < 1) return fail("gasname");
    for (int i = (int)strlen(d.gas) - 1; i >= 0 && (d.gas[i] == ' ' || d.gas[i] == '\t' || d.gas[i] == '\r'); --i)
        d.gas[i] = 0;
    if (!skip2(true)) return fail("header 16");
    if (fscanf(fd, "%lf %c", &d.attachment, &t) < 2) return fail("attachmentFactor");
    if (!skip2(true)) return fail("header 17");
    if (fscanf(fd, "%d %c", &d.collision, &t) < 2) return fail("collision flag");
    if (!skip2(true)) return fail("header 18");
    if (fscanf(fd, "%d %lf %lf %c", &d.power_mode, &d.heat_freq, &d.power, &t) < 4) return fail("heating variables");
    if (!skip2(true)) return fail("header 99");
    if (fscanf(fd, "%d %c", &d.print_interval, &t) <= 1) return fail("printInterval");
    if (!skip2(true)) return fail("header 99b");
    if (fscanf(fd, "%d %d %c", &d.start_avg, &d.end_avg, &t) < 2) return fail("start and end print spaceavg");
    if (!skip2(true)) return fail("header 19");
    if (fscanf(fd, "%d %c", &d.num_type, &t) <= 1) return fail("no of species");
    if (fscanf(fd, "%7999[^\n] %7999[^\n] ", u1, u2) <= 1) return fail("header 20");
    if (fscanf(fd, "%7999[^\n] ", u1) <= 0) return fail("header 21");
    if (d.num_type < 0 || d.num_type > 8) return fail("species count");
    d.species.assign(d.num_type, PicdatSpecies{});
    for (int i = 0; i < d.num_type; ++i) {
        PicdatSpecies& s = d.species[i];
        if (fscanf(fd, "%d %lf %lf %lf %lf", &s.index, &s.sim, &s.actual, &s.charge, &s.mass) <= 4)
            return fail("tabular data");
    }
    fclose(fd);
    return PIC_OK;
}

} // namespace pic

extern "C" int pic_generate_particles(uint64_t seed, int stream, int64_t n, double width, double height, double dx,
                                      double dy, const int* excl, double qm_abs, double temp_eV, double* x,
                                      double* y, double* vx, double* vy, double* vz) {
    if (n < 0 || !(width > 0) || !(height > 0) || (n > 0 && (!x || !y || !vx || !vy || !vz))) {
        pic::set_error("pic_generate_particles: invalid argument");
        return PIC_ERR_ARG;
    }
    double em[4];
    const double* ep = nullptr;
    if (excl) {
        em[0] = excl[0] * dx;
        em[1] = excl[1] * dy;
        em[2] = excl[2] * dx;
        em[3] = excl[3] * dy;
        ep = em;
    }
    pic::generate_box(seed, stream, n, 0.0, width, 0.0, height, width, height, ep, qm_abs, temp_eV, x, y, vx, vy, vz);
    return PIC_OK;
}
