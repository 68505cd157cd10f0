// pic_config.h — configuration-file parsing and particle generation (host).
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/bmp_pic.h"

namespace pic {

void set_error(const char* fmt, ...);

struct PicdatSpecies {
    int index = 0;
    double sim = 0, actual = 0, charge = 0, mass = 0;
};

// Fields of M_PICDAT.DAT in file order (particles.cpp:308-633).
struct Picdat {
    int timesteps = 0;
    double width = 0, height = 0;
    int NX = 0, NY = 0;
    double dt = 0;
    double B0 = 0, sigma = 0, mean = 0;   // B0 in T after the /1000
    double Te = 0, Ti = 0, Tg = 0;
    double pressure = 0;
    int uniform = 1;
    double fl = 0, fr = 1, ft = 0, fb = 1;
    double surf_area = 0;
    int reinjection = 0, scale_factor = 1, init_flag = 1;
    double epsilon = 8.854e-12;
    int periodic = 0;
    char gas[30] = {0};
    double attachment = 0;
    int collision = 0, power_mode = 0;
    double heat_freq = 0, power = 0;
    int print_interval = 1, start_avg = 0, end_avg = 0;
    int num_type = 0;
    std::vector<PicdatSpecies> species;
};

int read_picdat(const char* path, Picdat& d, std::string& err);

// Particles uniform in [x0, x1) x [y0, y1) (clipped below xmax, ymax),
// rejecting the box excl_m = (x1, y1, x2, y2) in metres when non-null.
void generate_box(uint64_t seed, int stream, int64_t n, double x0, double x1, double y0, double y1,
                  double xmax, double ymax, const double* excl_m, double qm_abs, double temp_eV,
                  double* x, double* y, double* vx, double* vy, double* vz);

} // namespace pic
