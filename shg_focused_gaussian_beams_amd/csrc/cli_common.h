// cli_common.h — pieces shared by the SHG and SHG_sweep programs: the seeded SH noise,
// the reference's .dat writers (with an optional binary sidecar per file) and the
// --key=value overrides.
#pragma once
#include <sys/time.h>

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <string>
#include <vector>

#include "shg.h"

namespace shgcli {

// --sidecar=1: next to every <name>_r.dat / _i.dat pair also write <name>.c64 (the plane
// as raw interleaved complex64, row-major, x fastest) and next to every Tfinal_<z>.dat
// also Tfinal_<z>.f32 (raw float32): the exact floats, for programs and tests that
// should not parse 20-digit text.
inline bool g_sidecar = false;

bool save_raw(const std::string &path, const float *a, size_t count) {
    std::ofstream f(path, std::ios::binary);
    if (!f) return false;
    f.write(reinterpret_cast<const char *>(a), (std::streamsize)(count * sizeof(float)));
    return (bool)f;
}

double seconds() {
    timeval tp;
    gettimeofday(&tp, nullptr);
    return (double)tp.tv_sec + (double)tp.tv_usec * 1.e-6;
}

// Reports a failed library call of program `prog`; returns the exit status 1.
int die(const char *what, int rc, const char *prog = "SHG") {
    std::cerr << prog << ": " << what << " failed (" << rc << "): " << shg_last_error() << "\n";
    return 1;
}

// Seeded complex Gaussian noise (splitmix64 + Box-Muller in double, then float): the
// reference draws SH noise from a clock-seeded std::normal_distribution
// (src/headers/Efields.h:129-144), whose output is implementation-defined.
void gaussian_noise(std::vector<float> &out, double sigma, uint64_t seed) {
    uint64_t s = seed;
    auto next = [&s]() {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    };
    for (size_t i = 0; i < out.size(); i += 2) {
        const double u1 = ((next() >> 11) + 1.0) * (1.0 / 9007199254740993.0);
        const double u2 = (next() >> 11) * (1.0 / 9007199254740992.0);
        const double r = std::sqrt(-2.0 * std::log(u1)), th = 2.0 * M_PI * u2;
        out[i] = (float)(sigma * r * std::cos(th));
        if (i + 1 < out.size()) out[i + 1] = (float)(sigma * r * std::sin(th));
    }
}

// saveMatrixComplex (src/headers/Files.h:36-57): NY rows of NX tab-terminated values.
bool save_complex(const std::string &dir, const std::string &name, const std::vector<float> &a, int nx, int ny) {
    if (g_sidecar && !save_raw(dir + name + ".c64", a.data(), 2 * (size_t)nx * ny)) return false;
    for (int part = 0; part < 2; part++) {
        std::ofstream f(dir + name + (part == 0 ? "_r.dat" : "_i.dat"));
        if (!f) return false;
        for (int iy = 0; iy < ny; iy++) {
            for (int ix = 0; ix < nx; ix++) f << std::setprecision(20) << a[2 * ((size_t)iy * nx + ix) + part] << "\t";
            f << "\n";
        }
    }
    return true;
}

// saveTensorReal (src/headers/Files.h:60-79): one file per z slice, default precision.
bool save_tensor(const std::string &dir, const std::string &name, const std::vector<float> &t, int nx, int ny, int nz) {
    for (int iz = 0; iz < nz; iz++) {
        if (g_sidecar && !save_raw(dir + name + "_" + std::to_string(iz) + ".f32", t.data() + (size_t)iz * ny * nx,
                                   (size_t)ny * nx))
            return false;
        std::ofstream f(dir + name + "_" + std::to_string(iz) + ".dat");
        if (!f) return false;
        for (int iy = 0; iy < ny; iy++) {
            for (int ix = 0; ix < nx; ix++) f << t[((size_t)iz * ny + iy) * nx + ix] << "\t";
            f << "\n";
        }
    }
    return true;
}

bool parse_override(shg_params &p, const std::string &key, const std::string &val, uint64_t &seed, double &noise,
                    std::string &outdir) {
    const double d = std::atof(val.c_str());
    const int i = std::atoi(val.c_str());
    if (key == "nx") { p.nx = i; p.ny = i; }
    else if (key == "nz") p.nz = i;
    else if (key == "lx") p.lx = d;
    else if (key == "ly") p.ly = d;
    else if (key == "lz") p.lz = d;
    else if (key == "lambda") p.lambda_p = d;
    else if (key == "deff") p.deff = d;
    else if (key == "period") p.qpm_period = d;
    else if (key == "alpha_p") p.alpha_p = d;
    else if (key == "alpha_s") p.alpha_s = d;
    else if (key == "beta_s") p.beta_s = d;
    else if (key == "rho_p") p.rho_p = d;
    else if (key == "rho_s") p.rho_s = d;
    else if (key == "k_th") p.k_th = d;
    else if (key == "h_conv") p.h_conv = d;
    else if (key == "focus") p.focus = d;
    else if (key == "t_inf") p.t_inf = d;
    else if (key == "crystal") p.crystal = val == "ppln" ? SHG_CRYSTAL_PPLN : SHG_CRYSTAL_SPPLT;
    else if (key == "scheme") p.scheme = val == "strang" ? SHG_SCHEME_STRANG : SHG_SCHEME_REFERENCE;
    else if (key == "thermal") p.thermal = i;
    else if (key == "thermal_bc")
        p.thermal_bc = val == "top_open" ? SHG_BC_TOP_OPEN : val == "oven_surrounded" ? SHG_BC_OVEN_SURROUNDED : -1;
    else if (key == "heat_nz") p.heat_nz = i;
    else if (key == "jacobi_check") p.jacobi_check = i;
    else if (key == "max_outer") p.max_outer = i;
    else if (key == "jacobi_max") p.jacobi_max = i;
    else if (key == "jacobi_tol") p.jacobi_tol = d;
    else if (key == "seed") seed = (uint64_t)std::strtoull(val.c_str(), nullptr, 10);
    else if (key == "noise") noise = d;
    else if (key == "sidecar") g_sidecar = i != 0;
    else if (key == "outdir") { outdir = val; if (!outdir.empty() && outdir.back() != '/') outdir += "/"; }
    else return false;
    return true;
}

}  // namespace shgcli
