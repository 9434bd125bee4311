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
// --thermal_bc=top_open|oven_surrounded --jacobi_check --max_outer --jacobi_max
// --jacobi_tol --seed --noise --outdir --sidecar=1 (raw .c64/.f32 copies of the outputs)
#include <algorithm>
#include <atomic>
#include <thread>

#include "cli_common.h"

using namespace shgcli;

namespace {

// save_only_last = 0 (src/headers/Solver.h:257-262): Pump_<s>, Signal_<s> before every
// step s.  The march runs once on the GPU keeping every plane; the text files (20
// significant digits, the bulk of the time) are formatted by a pool of writer threads.
int save_every_step(shg_ctx *ctx, const shg_params &p, const std::string &outdir) {
    const size_t n2 = (size_t)p.nx * p.ny, per = 2 * n2;
    std::vector<float> fh(per * p.nz), sh(per * p.nz);
    if (int rc = shg_forward_snapshots(ctx, fh.data(), sh.data())) return die("shg_forward_snapshots", rc);
    const int workers = std::max(1, std::min<int>(16, (int)std::thread::hardware_concurrency()));
    std::atomic<int> next{0};
    std::atomic<bool> ok{true};
    auto work = [&] {
        std::vector<float> a(per), b(per);
        for (int s = next++; s < p.nz && ok; s = next++) {
            std::copy(fh.begin() + per * s, fh.begin() + per * (s + 1), a.begin());
            std::copy(sh.begin() + per * s, sh.begin() + per * (s + 1), b.begin());
            if (!save_complex(outdir, "Pump_" + std::to_string(s), a, p.nx, p.ny) ||
                !save_complex(outdir, "Signal_" + std::to_string(s), b, p.nx, p.ny))
                ok = false;
        }
    };
    std::vector<std::thread> pool;
    for (int w = 0; w < workers; w++) pool.emplace_back(work);
    for (auto &t : pool) t.join();
    if (!ok) {
        std::cerr << "SHG: cannot write output files\n";
        return 1;
    }
    return 0;
}

void usage() {
    std::cerr << "usage: SHG <Power W> <waist um> <T1 C> <T2 C> <save_only_last 0|1> <save_temperature 0|1>"
                 " [--key=value ...]\n";
}

}  // namespace

// Prints the reference's THERMAL progress lines (src/headers/Solver.h:299-333): the outer
// iteration banner and tolerance (the reference switches stdout to 20 digits after its
// first tolerance line), the mean squared update at every check, the sweep count, and
// the start of the CWES march.  The residual is printed in float, as the reference sums it.
static void print_thermal_event(void *, int event, int32_t outer, int64_t sweeps, double value) {
    switch (event) {
    case SHG_THERMAL_EVENT_OUTER:
        std::cout << "\n\nGlobal iteration #" << outer << "\n\n" << std::endl;
        std::cout << "Solving 3D Heat Equation in the steady-state...\n" << std::endl;
        std::cout << "\nTemperature tolerance = " << (float)value << "\n" << std::endl;
        std::cout.precision(20);
        break;
    case SHG_THERMAL_EVENT_CHECK:
        std::cout << "\u03A3|Tf-Ti|\u00B2/N\u00B3 at #" << sweeps << " iteration is: " << (float)value << std::endl;
        break;
    case SHG_THERMAL_EVENT_STEADY:
        std::cout << sweeps << " iterations -> steady-state." << std::endl;
        break;
    case SHG_THERMAL_EVENT_MARCH:
        std::cout << "\n\nSolving Coupled-Wave Equations (CWEs)...\n" << std::endl;
        break;
    default:
        break;
    }
}

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
    if ((rc = shg_create(&p, &ctx))) return die("shg_create", rc);
    if ((rc = shg_set_fields(ctx, fh.data(), sh.data()))) return die("shg_set_fields", rc);
    if (!p.thermal) std::cout << "\n\nEvolution in the crystal...\n" << std::endl;
    if (p.thermal) {
        // THERMAL branch of Solver::run (src/headers/Solver.h:289-340): Jacobi, dk(T), the
        // march and the heat-source feedback, up to max_outer global iterations, with the
        // reference's progress lines.
        if ((rc = shg_set_thermal_log(ctx, print_thermal_event, nullptr))) return die("shg_set_thermal_log", rc);
        shg_thermal_stats st{};
        if ((rc = shg_thermal_run(ctx, 1, &st))) return die("shg_thermal_run", rc);
        std::cout.precision(6);
        std::cout << "Global iterations: " << st.outer_iterations << ", Jacobi sweeps: " << st.sweeps_total
                  << ", last mean squared update: " << st.residual_last << std::endl;
        if (save_only_last) {
            if ((rc = shg_get_fields(ctx, fh.data(), sh.data()))) return die("shg_get_fields", rc);
            if (!save_complex(outdir, "Pump_out", fh, p.nx, p.ny) ||
                !save_complex(outdir, "Signal_out", sh, p.nx, p.ny)) {
                std::cerr << "SHG: cannot write output files\n";
                return 1;
            }
        } else {
            // the files the last outer iteration's CWES writes: its march again at the
            // final temperature (the march only reads T, and DKint restarts from zero)
            if ((rc = save_every_step(ctx, p, outdir))) return rc;
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
        if ((rc = save_every_step(ctx, p, outdir))) return rc;
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
