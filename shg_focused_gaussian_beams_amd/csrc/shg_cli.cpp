// shg_cli.cpp — the `SHG` program: a drop-in for the reference's command line
// (src/SHG.cpp:24-102) on top of the C ABI in include/shg.h.
//
//   SHG <Power W> <waist um> <T1 C> <T2 C> <save_only_last 0|1> <save_temperature 0|1> [--key=value ...]
//
// argv[3] is both the crystal temperature and oven 1 (src/SHG.cpp:42-44); the focal
// point is 0.5*LZ and the ambient 25 C unless overridden (src/SHG.cpp:39,43).  Output
// files have the reference's names and format (src/headers/Files.h:36-79):
//   save_only_last = 1: Pump_out_{r,i}.dat, Signal_out_{r,i}.dat (after nz-1 steps)
//   save_only_last = 0: Pump_<s>_{r,i}.dat, Signal_<s>_{r,i}.dat before every step s
// and the elapsed seconds are appended to CPUTime_NZ_<NZ>_N_<NX>_T.txt (src/SHG.cpp:85-97).
// Overrides (defaults are the reference's constants): --nx --nz --lx --ly --lz --lambda
// --deff --period --alpha_p --alpha_s --beta_s --rho_p --rho_s --k_th --h_conv --focus
// --t_inf --crystal=spplt|ppln --scheme=reference|strang --thermal=0|1 --heat_nz
// --jacobi_check --max_outer --jacobi_max --jacobi_tol --seed --noise --outdir
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

namespace {

double seconds() {
    timeval tp;
    gettimeofday(&tp, nullptr);
    return (double)tp.tv_sec + (double)tp.tv_usec * 1.e-6;
}

void usage() {
    std::cerr << "usage: SHG <Power W> <waist um> <T1 C> <T2 C> <save_only_last 0|1> <save_temperature 0|1>"
                 " [--key=value ...]\n";
}

int die(const char *what, int rc) {
    std::cerr << "SHG: " << what << " failed (" << rc << "): " << shg_last_error() << "\n";
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
    else if (key == "heat_nz") p.heat_nz = i;
    else if (key == "jacobi_check") p.jacobi_check = i;
    else if (key == "max_outer") p.max_outer = i;
    else if (key == "jacobi_max") p.jacobi_max = i;
    else if (key == "jacobi_tol") p.jacobi_tol = d;
    else if (key == "seed") seed = (uint64_t)std::strtoull(val.c_str(), nullptr, 10);
    else if (key == "noise") noise = d;
    else if (key == "outdir") { outdir = val; if (!outdir.empty() && outdir.back() != '/') outdir += "/"; }
    else return false;
    return true;
}

}  // namespace

int main(int argc, char *argv[]) {
    std::cout << "#######---SHG efficiency calculator---#######\n\n" << std::endl;
    const double t_start = seconds();
    if (argc < 7) { usage(); return 1; }
    shg_params p;
    shg_default_params(&p);
    p.power = std::atof(argv[1]);
    p.waist = std::atof(argv[2]);
    p.temp = std::atof(argv[3]);
    p.t_peltier1 = std::atof(argv[3]);
    p.t_peltier2 = std::atof(argv[4]);
    const bool save_only_last = std::atoi(argv[5]) != 0;
    const bool save_temperature = std::atoi(argv[6]) != 0;
    uint64_t seed = 1;
    double noise = 1.0e-20;                          // reference SH noise sigma
    std::string outdir;
    bool focus_set = false;
    for (int a = 7; a < argc; a++) {
        std::string s = argv[a];
        const size_t eq = s.find('=');
        if (s.rfind("--", 0) != 0 || eq == std::string::npos) { usage(); return 1; }
        const std::string key = s.substr(2, eq - 2), val = s.substr(eq + 1);
        if (key == "focus") focus_set = true;
        if (!parse_override(p, key, val, seed, noise, outdir)) {
            std::cerr << "SHG: unknown option --" << key << "\n";
            return 1;
        }
    }
    if (!focus_set) p.focus = 0.5 * (double)(float)p.lz;   // focalpoint = 0.5*LZ (src/SHG.cpp:39)

    std::cout << "\n\nPump power =  " << (float)p.power << " W" << std::endl;
    std::cout << "Beam waist = " << (float)p.waist << " μm" << std::endl;
    std::cout << "Focal distance = " << (float)p.focus << " μm" << std::endl;
    std::cout << "Phase-matching temperature =  " << (float)p.temp << " ºC" << std::endl;
    std::cout << "Environment temperature =  " << (float)p.t_inf << " ºC" << std::endl;
    std::cout << "Peltier 1 temperature =  " << (float)p.t_peltier1 << " ºC" << std::endl;
    std::cout << "Peltier 2 temperature =  " << (float)p.t_peltier2 << " ºC\n" << std::endl;
    std::cout << "Grid = " << p.nx << " x " << p.ny << " x " << p.nz << ", crystal = "
              << (p.crystal == SHG_CRYSTAL_PPLN ? "MgO:PPLN" : "MgO:sPPLT") << ", scheme = "
              << (p.scheme == SHG_SCHEME_STRANG ? "strang" : "reference") << std::endl;

    const size_t n2 = (size_t)p.nx * p.ny;
    std::vector<float> fh(2 * n2), sh(2 * n2);
    int rc = shg_make_pump(&p, p.temp, fh.data());
    if (rc) return die("shg_make_pump", rc);
    gaussian_noise(sh, noise, seed);

    shg_ctx *ctx = nullptr;
    if (p.thermal && !save_only_last) {
        std::cerr << "SHG: THERMAL runs write the output plane only (save_only_last=1)\n";
        return 1;
    }
    if ((rc = shg_create(&p, &ctx))) return die("shg_create", rc);
    if ((rc = shg_set_fields(ctx, fh.data(), sh.data()))) return die("shg_set_fields", rc);
    std::cout << "\n\nEvolution in the crystal...\n" << std::endl;
    if (p.thermal) {
        // THERMAL branch of Solver::run (src/headers/Solver.h:289-340): Jacobi, dk(T), the
        // march and the heat-source feedback, up to max_outer global iterations.
        shg_thermal_stats st{};
        if ((rc = shg_thermal_run(ctx, 1, &st))) return die("shg_thermal_run", rc);
        std::cout << "Global iterations: " << st.outer_iterations << ", Jacobi sweeps: " << st.sweeps_total
                  << ", last mean squared update: " << st.residual_last << std::endl;
        if ((rc = shg_get_fields(ctx, fh.data(), sh.data()))) return die("shg_get_fields", rc);
        if (!save_complex(outdir, "Pump_out", fh, p.nx, p.ny) || !save_complex(outdir, "Signal_out", sh, p.nx, p.ny)) {
            std::cerr << "SHG: cannot write output files\n";
            return 1;
        }
        if (save_temperature) {
            const int hz = p.heat_nz > 0 ? p.heat_nz : p.nz;
            std::vector<float> T(n2 * (size_t)hz);
            if ((rc = shg_get_temperature(ctx, T.data()))) return die("shg_get_temperature", rc);
            if (!save_tensor(outdir, "Tfinal", T, p.nx, p.ny, hz)) {
                std::cerr << "SHG: cannot write temperature files\n";
                return 1;
            }
        }
    } else if (save_only_last) {
        if ((rc = shg_forward(ctx))) return die("shg_forward", rc);
        if ((rc = shg_get_fields(ctx, fh.data(), sh.data()))) return die("shg_get_fields", rc);
        if (!save_complex(outdir, "Pump_out", fh, p.nx, p.ny) || !save_complex(outdir, "Signal_out", sh, p.nx, p.ny)) {
            std::cerr << "SHG: cannot write output files\n";
            return 1;
        }
    } else {
        for (int s = 0; s < p.nz; s++) {
            if ((rc = shg_get_fields(ctx, fh.data(), sh.data()))) return die("shg_get_fields", rc);
            if (!save_complex(outdir, "Pump_" + std::to_string(s), fh, p.nx, p.ny) ||
                !save_complex(outdir, "Signal_" + std::to_string(s), sh, p.nx, p.ny)) {
                std::cerr << "SHG: cannot write output files\n";
                return 1;
            }
            if (s < p.nz - 1 && (rc = shg_march(ctx, s, 1))) return die("shg_march", rc);
        }
    }
    double pin = 0, pout = 0, eta = 0;
    if ((rc = shg_powers(ctx, &pin, &pout, &eta))) return die("shg_powers", rc);
    std::cout << std::setprecision(10) << "P_FH(z=0) = " << pin << " W, P_SH(out) = " << pout
              << " W, efficiency = " << eta << std::endl;
    shg_destroy(ctx);
    std::cout << "Deallocating memory" << std::endl;

    const double elapsed = seconds() - t_start;
    if (elapsed < 60.) std::cout << "\n\n...time elapsed " << elapsed << " seconds\n\n " << std::endl;
    else if (elapsed < 3600.) std::cout << "\n\n...time elapsed " << elapsed / 60 << " minutes\n\n " << std::endl;
    else std::cout << "\n\n...time elapsed " << elapsed / 3600 << " hours\n\n " << std::endl;
    std::ofstream tf(outdir + "CPUTime_NZ_" + std::to_string(p.nz) + "_N_" + std::to_string(p.nx) + "_T.txt",
                     std::ios::app);
    if (tf.is_open()) {
        tf << elapsed << std::endl;
        std::cout << "Data successfully written to the end of the file." << std::endl;
    } else {
        std::cerr << "Unable to open the file for writing." << std::endl;
    }
    return 0;
}
