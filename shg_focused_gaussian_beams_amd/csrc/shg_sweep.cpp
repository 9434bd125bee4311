// shg_sweep.cpp — the `SHG_sweep` program: the parameter sweeps that src/SHG.sh:39-101
// runs as one `SHG` process per point, as one batched launch (plain) or one context per
// point (THERMAL), with the same per-point folder and file names.
//
//   SHG_sweep --power=LIST --waist=LIST --t1=LIST [--t2=LIST] [--save_temperature=0|1]
//             [--folder=Simulations] [--thermal_batch=0|1] [--workers=N] [--key=value ...]
//
// LIST is `v1,v2,...` or `start:stop:step` (inclusive, as the awk loops of SHG.sh).  The
// points are the product power x waist x t1 (x t2; by default t2 = t1, as SHG.sh sets
// TOVEN2 = TOVEN1).  Point (P, w, T1) writes <folder>/<crystal>_P_<P>_W_T1_<T1>_WAIST_<w>/
// with Pump_out_{r,i}.dat, Signal_out_{r,i}.dat (the state after nz-1 steps), the
// THERMAL temperature Tfinal_<z>.dat when asked, and <same name>.txt with the point's
// powers and efficiency (src/SHG.sh:65-80).  Every SHG override (--nx, --nz, --thermal,
// --seed, --noise, ...) applies to all points; the SH noise of point k uses seed + k.
// THERMAL points run as one lockstep batch (shg_thermal_batch: one Jacobi launch and one
// march launch cover every point still running), or with --thermal_batch=0 as one
// context each, N at a time from N host threads (--workers, default 8), each context on
// its own stream.  Both give every point the result of its own SHG run.
#include <atomic>
#include <filesystem>
#include <sstream>
#include <thread>
#include <utility>

#include "cli_common.h"

using namespace shgcli;

namespace {

bool parse_list(const std::string &v, std::vector<double> &out) {
    out.clear();
    const size_t c1 = v.find(':');
    if (c1 != std::string::npos) {
        const size_t c2 = v.find(':', c1 + 1);
        if (c2 == std::string::npos) return false;
        const double a = std::atof(v.substr(0, c1).c_str()), b = std::atof(v.substr(c1 + 1, c2 - c1 - 1).c_str());
        const double h = std::atof(v.substr(c2 + 1).c_str());
        if (!(h > 0) || b < a) return false;
        for (int i = 0; a + i * h <= b + 1e-9 * std::fabs(h); i++) out.push_back(a + i * h);
        return !out.empty();
    }
    std::stringstream ss(v);
    std::string item;
    while (std::getline(ss, item, ','))
        if (!item.empty()) out.push_back(std::atof(item.c_str()));
    return !out.empty();
}

std::string num(double v) {
    char b[64];
    std::snprintf(b, sizeof(b), "%g", v);
    return b;
}

void usage() {
    std::cerr << "usage: SHG_sweep --power=LIST --waist=LIST --t1=LIST [--t2=LIST] [--save_temperature=0|1]"
                 " [--folder=DIR] [--thermal_batch=0|1] [--workers=N] [--key=value ...]"
                 "   (LIST: v1,v2,... or start:stop:step)\n";
}

struct Point {
    double power, waist, t1, t2;
    std::string name;
};

}  // namespace

int main(int argc, char *argv[]) {
    const double t_start = seconds();
    std::vector<double> powers{30.0}, waists{29.0}, t1s{55.66}, t2s;   // the SHG.sh defaults
    bool save_temperature = false;
    std::string folder = "Simulations/";
    shg_params base;
    shg_default_params(&base);
    uint64_t seed = 1;
    double noise = 1.0e-20;
    std::string unused_outdir;
    bool focus_set = false;
    int workers = 8;
    bool thermal_batch = true;
    for (int a = 1; a < argc; a++) {
        const std::string s = argv[a];
        const size_t eq = s.find('=');
        if (s.rfind("--", 0) != 0 || eq == std::string::npos) { usage(); return 1; }
        const std::string key = s.substr(2, eq - 2), val = s.substr(eq + 1);
        bool ok = true;
        if (key == "power") ok = parse_list(val, powers);
        else if (key == "waist") ok = parse_list(val, waists);
        else if (key == "t1") ok = parse_list(val, t1s);
        else if (key == "t2") ok = parse_list(val, t2s);
        else if (key == "save_temperature") save_temperature = std::atoi(val.c_str()) != 0;
        else if (key == "save_only_last") ok = std::atoi(val.c_str()) != 0;   // sweeps save the last slice
        else if (key == "folder") { folder = val; if (!folder.empty() && folder.back() != '/') folder += "/"; }
        else if (key == "workers") ok = (workers = std::atoi(val.c_str())) >= 1 && workers <= 64;
        else if (key == "thermal_batch") thermal_batch = std::atoi(val.c_str()) != 0;
        else {
            if (key == "focus") focus_set = true;
            ok = parse_override(base, key, val, seed, noise, unused_outdir);
        }
        if (!ok) {
            std::cerr << "SHG_sweep: bad or unsupported option --" << key << "=" << val << "\n";
            return 1;
        }
    }
    if (!focus_set) base.focus = 0.5 * (double)(float)base.lz;   // focalpoint = 0.5*LZ (src/SHG.cpp:39)
    const std::string crystal = base.crystal == SHG_CRYSTAL_PPLN ? "PPLN" : "sPPLT";
    std::vector<Point> pts;
    for (double P : powers)
        for (double w : waists)
            for (double T1 : t1s)
                for (double T2 : (t2s.empty() ? std::vector<double>{T1} : t2s))
                    pts.push_back({P, w, T1, T2,
                                   crystal + "_P_" + num(P) + "_W_T1_" + num(T1) + "_WAIST_" + num(w) +
                                       (t2s.empty() ? "" : "_T2_" + num(T2))});
    const int count = (int)pts.size();
    std::vector<shg_params> ps(count, base);
    for (int k = 0; k < count; k++) {
        ps[k].power = pts[k].power;
        ps[k].waist = pts[k].waist;
        ps[k].temp = ps[k].t_peltier1 = pts[k].t1;
        ps[k].t_peltier2 = pts[k].t2;
    }
    const size_t n2 = (size_t)base.nx * base.ny;
    std::cout << "#######---SHG sweep: " << count << " points, " << base.nx << " x " << base.ny << " x " << base.nz
              << (base.thermal ? ", THERMAL" : "") << "---#######" << std::endl;

    std::vector<double> eta(count), pin(count), pout(count);
    std::vector<std::vector<float>> outs_fh(count), outs_sh(count), temps(count);
    int rc;
    if (!base.thermal) {
        shg_batch *b = nullptr;
        if ((rc = shg_batch_create(ps.data(), count, &b))) return die("shg_batch_create", rc, "SHG_sweep");
        for (int k = 0; k < count; k++) {
            std::vector<float> fh(2 * n2), sh(2 * n2);
            if ((rc = shg_make_pump(&ps[k], ps[k].temp, fh.data()))) return die("shg_make_pump", rc, "SHG_sweep");
            gaussian_noise(sh, noise, seed + (uint64_t)k);
            if ((rc = shg_batch_set_fields(b, k, fh.data(), sh.data()))) return die("shg_batch_set_fields", rc, "SHG_sweep");
        }
        if ((rc = shg_batch_run(b))) return die("shg_batch_run", rc, "SHG_sweep");
        if ((rc = shg_batch_powers(b, pin.data(), pout.data(), eta.data()))) return die("shg_batch_powers", rc, "SHG_sweep");
        for (int k = 0; k < count; k++) {
            outs_fh[k].resize(2 * n2);
            outs_sh[k].resize(2 * n2);
            if ((rc = shg_batch_get_fields(b, k, outs_fh[k].data(), outs_sh[k].data())))
                return die("shg_batch_get_fields", rc, "SHG_sweep");
        }
        shg_batch_destroy(b);
    } else if (thermal_batch) {
        shg_thermal_batch *b = nullptr;
        if ((rc = shg_thermal_batch_create(ps.data(), count, &b))) return die("shg_thermal_batch_create", rc, "SHG_sweep");
        for (int k = 0; k < count && !rc; k++) {
            std::vector<float> sh(2 * n2);
            gaussian_noise(sh, noise, seed + (uint64_t)k);
            rc = shg_thermal_batch_set_signal(b, k, sh.data());
        }
        std::vector<shg_thermal_stats> st(count);
        if (!rc) rc = shg_thermal_batch_run(b, st.data());
        if (!rc) rc = shg_thermal_batch_powers(b, pin.data(), pout.data(), eta.data());
        for (int k = 0; k < count && !rc; k++) {
            outs_fh[k].resize(2 * n2);
            outs_sh[k].resize(2 * n2);
            rc = shg_thermal_batch_get_fields(b, k, outs_fh[k].data(), outs_sh[k].data());
            if (!rc && save_temperature) {
                temps[k].resize(n2 * (size_t)(base.heat_nz > 0 ? base.heat_nz : base.nz));
                rc = shg_thermal_batch_get_temperature(b, k, temps[k].data());
            }
        }
        if (rc) {
            std::cerr << "SHG_sweep: THERMAL batch: " << shg_last_error() << "\n";
            shg_thermal_batch_destroy(b);
            return 1;
        }
        shg_thermal_batch_destroy(b);
    } else {
        std::atomic<int> next{0};
        std::atomic<bool> failed{false};
        std::vector<std::string> errors(count);
        // context of point k with its z = 0 fields (SH noise of seed + k)
        auto prepare = [&](int k, shg_ctx **out) -> bool {
            std::vector<float> fh(2 * n2), sh(2 * n2);
            shg_ctx *ctx = nullptr;
            if (shg_make_pump(&ps[k], ps[k].temp, fh.data())) {
                errors[k] = std::string("shg_make_pump: ") + shg_last_error();
                return false;
            }
            gaussian_noise(sh, noise, seed + (uint64_t)k);
            if (shg_create(&ps[k], &ctx) || shg_set_fields(ctx, fh.data(), sh.data())) {
                errors[k] = std::string("shg_create: ") + shg_last_error();
                shg_destroy(ctx);
                return false;
            }
            *out = ctx;
            return true;
        };
        // One tuning for the whole sweep, on point 0 before the workers start: every point
        // has the same grid and so the same fastest pass layouts.
        shg_ctx *first = nullptr;
        int rv = -1, cv = -1;
        const char *nt = std::getenv("SHG_NO_TUNE");
        if (!(nt && nt[0] == '1')) {
            if (!prepare(0, &first)) return die("shg_create", SHG_E_DEVICE, "SHG_sweep");
            if ((rc = shg_tune(first, 20, &rv, &cv))) {
                shg_destroy(first);
                return die("shg_tune", rc, "SHG_sweep");
            }
        }
        auto point = [&](int k) -> bool {
            shg_ctx *ctx = k == 0 ? std::exchange(first, nullptr) : nullptr;
            if (!ctx && !prepare(k, &ctx)) return false;
            std::vector<float> fh(2 * n2), sh(2 * n2);
            shg_thermal_stats st{};
            bool ok = !((rv >= 0 && shg_set_variants(ctx, rv, cv)) || shg_thermal_run(ctx, 1, &st) ||
                        shg_get_fields(ctx, fh.data(), sh.data()) || shg_powers(ctx, &pin[k], &pout[k], &eta[k]));
            if (ok && save_temperature) {
                temps[k].resize(n2 * (size_t)(ps[k].heat_nz > 0 ? ps[k].heat_nz : ps[k].nz));
                ok = shg_get_temperature(ctx, temps[k].data()) == SHG_OK;
            }
            if (!ok) errors[k] = std::string("THERMAL point: ") + shg_last_error();
            outs_fh[k] = fh;
            outs_sh[k] = sh;
            shg_destroy(ctx);
            return ok;
        };
        auto work = [&] {
            for (int k = next++; k < count && !failed; k = next++)
                if (!point(k)) failed = true;
        };
        std::vector<std::thread> pool;
        for (int w = 0; w < std::min(workers, count); w++) pool.emplace_back(work);
        for (auto &t : pool) t.join();
        shg_destroy(first);                            // set only if point 0 never ran
        if (failed) {
            for (int k = 0; k < count; k++)
                if (!errors[k].empty()) std::cerr << "SHG_sweep: " << pts[k].name << ": " << errors[k] << "\n";
            return 1;
        }
    }
    const double elapsed = seconds() - t_start;
    for (int k = 0; k < count; k++) {
        const std::string dir = folder + pts[k].name + "/";
        std::error_code ec;
        std::filesystem::create_directories(dir, ec);
        if (ec || !save_complex(dir, "Pump_out", outs_fh[k], base.nx, base.ny) ||
            !save_complex(dir, "Signal_out", outs_sh[k], base.nx, base.ny) ||
            (!temps[k].empty() && !save_tensor(dir, "Tfinal", temps[k], base.nx, base.ny,
                                               (int)(temps[k].size() / n2)))) {
            std::cerr << "SHG_sweep: cannot write " << dir << "\n";
            return 1;
        }
        std::ofstream log(dir + pts[k].name + ".txt");
        log << std::setprecision(10) << "Pump power = " << pts[k].power << " W\nBeam waist = " << pts[k].waist
            << " um\nPeltier 1 temperature = " << pts[k].t1 << " C\nPeltier 2 temperature = " << pts[k].t2
            << " C\nP_FH(z=0) = " << pin[k] << " W, P_SH(out) = " << pout[k] << " W, efficiency = " << eta[k] << "\n";
        std::cout << pts[k].name << "\tefficiency = " << std::setprecision(8) << eta[k] << std::endl;
    }
    std::cout << "\n...time elapsed " << elapsed << " seconds (" << count << " points)" << std::endl;
    return 0;
}
