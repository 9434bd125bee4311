// shg_thermal_batch.cpp — batched THERMAL sweeps (C ABI in include/shg.h).
//
// The reference's sweep script builds the THERMAL solver and runs one process per sweep
// point (src/SHG.sh:23, 39-81), each the THERMAL branch of Solver::run
// (src/headers/Solver.h:289-340).  A sweep varies the beam and the ovens; the crystal, the
// grid and the heat-equation constants are the build's.  Here the instances of such a
// sweep share one device state laid out instance after instance, and run that outer loop
// in lockstep: the Jacobi sweeps of every instance still converging in one launch, the
// residual checks at the same sweep counts (each instance stops by its own residual and
// its own stagnation and early-stop rules), and the z-marches of every instance in the
// same launches (a per-instance flag lets the instances that have stopped sit a march
// out).  Each instance's sweeps, checks and march are the ones shg_thermal_run performs
// for it alone, so its result is that run's, bit for bit with the same pass layouts.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <functional>
#include <string>
#include <utility>
#include <vector>

#include "../../include/shg.h"
#include "guard.h"
#include "context.h"
#include "physics.h"
#include "ssfm_kernels.h"
#include "thermal_kernels.h"

namespace {

#define TB_TRY(expr)                                                                                              \
    do {                                                                                                          \
        hipError_t e_ = (expr);                                                                                   \
        if (e_ != hipSuccess) return shg_internal_fail(SHG_E_DEVICE, std::string(#expr ": ") + hipGetErrorString(e_)); \
    } while (0)

#define TB_RC(expr)            \
    do {                       \
        int rc_ = (expr);      \
        if (rc_) return rc_;   \
    } while (0)

// Whether two instances share everything but the beam and the ovens (power, waist,
// focus, temp, t_peltier1, t_peltier2), field by field (the struct's padding bytes are
// not compared: callers need not clear them).
bool same_build(const shg_params &a, const shg_params &b) {
    return a.nx == b.nx && a.ny == b.ny && a.nz == b.nz && a.crystal == b.crystal && a.thermal == b.thermal &&
           a.scheme == b.scheme && a.lx == b.lx && a.ly == b.ly && a.lz == b.lz && a.lambda_p == b.lambda_p &&
           a.deff == b.deff && a.qpm_period == b.qpm_period && a.alpha_p == b.alpha_p && a.alpha_s == b.alpha_s &&
           a.beta_s == b.beta_s && a.rho_p == b.rho_p && a.rho_s == b.rho_s && a.k_th == b.k_th &&
           a.h_conv == b.h_conv && a.t_inf == b.t_inf && a.heat_nz == b.heat_nz && a.jacobi_check == b.jacobi_check &&
           a.max_outer == b.max_outer && a.jacobi_max == b.jacobi_max && a.jacobi_tol == b.jacobi_tol &&
           a.thermal_bc == b.thermal_bc;
}

}  // namespace

struct shg_thermal_batch {
    int count = 0;
    std::vector<shg_params> p;
    int n = 0, nz = 0, hz = 0;
    size_t n2 = 0, n3 = 0;
    hipStream_t stream = nullptr;
    float4 *U = nullptr, *Uout = nullptr;
    float2 *tw = nullptr, *G = nullptr, *NT = nullptr;
    float *T[2] = {nullptr, nullptr};
    float *Q = nullptr, *DKint = nullptr;
    shg::InstDesc *desc = nullptr;
    int *active = nullptr;                 // device flags: the instance marches
    int *ids = nullptr;                    // device list of the instances the Jacobi launches sweep
    int nids = 0;
    std::vector<int> host_ids;
    shg::Jacobi2Plan jp{};                 // the two-sweep plan of layout jp_v for nids instances
    int jp_v = -2;
    double *jpart = nullptr, *jres = nullptr;
    double2 *ppart = nullptr, *pred = nullptr;
    int njb = 0;                           // residual partials per instance
    int tcur = 0;
    shg::NlConst nl{};
    shg::JacobiConst jc{};
    shg::ThermConst tc{};
    std::vector<float> sh0;                // per instance, the SH input (interleaved complex)
    std::vector<float> stage;              // float4 staging of one instance's plane
    std::vector<double> p_fh_in;
    std::vector<float> t_centre;
    std::vector<shg_thermal_stats> stats;
    int row_v = -1, col_v = -1;
    int jac_v = -1;                        // Jacobi layout: a two-sweep plan, or single sweeps
    shg::Jacobi2Plan jplan[shg::kJacobi2VariantCount] = {};
    hipGraphExec_t gmarch = nullptr;       // the batched march (reads the flags from `active`)
    hipEvent_t e0 = nullptr, e1 = nullptr;
    bool ran = false;
};

namespace {

void tb_release(shg_thermal_batch *b) {
    if (!b) return;
    if (b->stream) (void)hipStreamSynchronize(b->stream);   // work an error return left in flight
    if (b->gmarch) (void)hipGraphExecDestroy(b->gmarch);
    for (void *ptr : {(void *)b->U, (void *)b->Uout, (void *)b->tw, (void *)b->G, (void *)b->NT, (void *)b->T[0],
                      (void *)b->T[1], (void *)b->Q, (void *)b->DKint, (void *)b->desc, (void *)b->active, (void *)b->ids,
                      (void *)b->jpart, (void *)b->jres, (void *)b->ppart, (void *)b->pred})
        if (ptr) (void)hipFree(ptr);
    if (b->e0) (void)hipEventDestroy(b->e0);
    if (b->e1) (void)hipEventDestroy(b->e1);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
}

int set_active(shg_thermal_batch *b, const std::vector<int> &on) {
    TB_TRY(shg_internal_copy(b->active, on.data(), on.size() * sizeof(int), hipMemcpyHostToDevice, b->stream));
    return SHG_OK;
}

// The instances the next Jacobi launches sweep (those with on[i] set), and the two-sweep
// plan sized for that many: instances that have stopped take no workgroups.
int set_sweeping(shg_thermal_batch *b, const std::vector<int> &on) {
    std::vector<int> list;
    for (int i = 0; i < b->count; i++)
        if (on[i]) list.push_back(i);
    if (list == b->host_ids && b->jp_v == b->jac_v && b->nids == (int)list.size()) return SHG_OK;
    b->host_ids = list;
    b->nids = (int)list.size();
    b->jp_v = b->jac_v;
    if (!b->nids) return SHG_OK;
    TB_TRY(shg_internal_copy(b->ids, list.data(), list.size() * sizeof(int), hipMemcpyHostToDevice, b->stream));
    if (b->jac_v >= 0 && b->jac_v < shg::kJacobi2VariantCount &&
        shg::jacobi2_plan(b->n, b->n, b->hz, b->jac_v, &b->jp, b->nids) != 0)
        return shg_internal_fail(SHG_E_DEVICE, "shg_thermal_batch_run: Jacobi layout query failed");
    return SHG_OK;
}

// initTemperature + the bottom (or surrounding) ovens of every instance (Tfield.h:31-110), Q = 0.
int init_temperatures(shg_thermal_batch *b) {
    std::vector<float> t(b->n3);
    const int n = b->n, hz = b->hz;
    const float nzf = (float)(unsigned)hz;
    for (int i = 0; i < b->count; i++) {
        const shg_params &p = b->p[i];
        std::fill(t.begin(), t.end(), (float)p.temp);
        for (int z = 0; z < hz; z++) {
            const float v = (z < nzf / 2.0) ? (float)p.t_peltier1 : (float)p.t_peltier2;
            float *tz = t.data() + (size_t)z * b->n2;
            if (p.thermal_bc == SHG_BC_OVEN_SURROUNDED) {
                for (int j = 0; j < n; j++) {
                    tz[j] = tz[(size_t)(n - 1) * n + j] = v;
                    tz[(size_t)j * n] = tz[(size_t)j * n + n - 1] = v;
                }
            } else {
                for (int x = 0; x < n; x++) tz[x] = v;
            }
        }
        for (int k = 0; k < 2; k++)
            TB_TRY(shg_internal_copy(b->T[k] + (size_t)i * b->n3, t.data(), b->n3 * sizeof(float), hipMemcpyHostToDevice,
                                     b->stream));
    }
    TB_TRY(hipMemsetAsync(b->Q, 0, (size_t)b->count * b->n3 * sizeof(float), b->stream));
    TB_TRY(hipStreamSynchronize(b->stream));
    b->tcur = 0;
    return SHG_OK;
}

// The batched THERMAL march of Solver::CWES (src/headers/Solver.h:245-276) for the flagged
// instances: the schedule of one context's march, with every launch covering them all.
int enqueue_march(shg_thermal_batch *b) {
    const shg_params &p = b->p[0];
    shg::ThermRowArgs a{};
    a.k = b->tc;
    a.inst_cells = b->n2;
    a.inst_voxels = b->n3;
    a.active = b->active;
    a.cur = shg_internal_therm_step_of(p, b->hz, 0);
    a.flags = shg::TROW_FIRST | shg::TROW_NCUR | shg::TROW_BWD;
    const int rv = b->row_v < 0 ? 0 : b->row_v, cv = b->col_v < 0 ? 0 : b->col_v;
    TB_TRY(shg::launch_thermal_row_pass(b->n, b->U, b->Uout, b->tw, b->NT, b->Q, b->DKint, a, b->stream, rv, b->count));
    int w = 0;                                      // DKint holds the prefix of step w
    for (int s = 0; s < b->nz; s++) {
        TB_TRY(shg::launch_col_pass(b->n, b->desc, b->count, b->tw, b->stream, cv));
        a.prev = shg_internal_therm_step_of(p, b->hz, s);
        a.dk_pending = s - w;
        if (s < b->nz - 1) {
            a.cur = shg_internal_therm_step_of(p, b->hz, s + 1);
            a.flags = shg::TROW_FWD | shg::TROW_NPREV | shg::TROW_NCUR | shg::TROW_BWD;
            if (a.prev.slice != a.cur.slice) a.flags |= shg::TROW_QPREV;
            if (a.prev.slice == a.cur.slice && s + 1 < b->nz - 1 && s + 1 - w < shg::kMaxDkPending)
                a.flags |= shg::TROW_DKKEEP;
            else w = s + 1;
            if (s + 1 == b->nz - 1) a.flags |= shg::TROW_SNAP;
        } else {
            a.flags = shg::TROW_FWD | shg::TROW_NPREV | shg::TROW_QPREV;
        }
        TB_TRY(shg::launch_thermal_row_pass(b->n, b->U, b->Uout, b->tw, b->NT, b->Q, b->DKint, a, b->stream, rv,
                                            b->count));
    }
    return SHG_OK;
}

// Times every (row, column) layout pair of the batched march as the march runs them and
// keeps the fastest (all instances flagged; the state is rebuilt afterwards by run()).
int tune(shg_thermal_batch *b) {
    const shg_params &p = b->p[0];
    TB_RC(set_active(b, std::vector<int>(b->count, 1)));
    TB_TRY(shg::launch_index_table(b->T[b->tcur], b->NT, (int64_t)b->count * b->n3, b->tc, b->stream));
    shg::ThermRowArgs a{};
    a.k = b->tc;
    a.inst_cells = b->n2;
    a.inst_voxels = b->n3;
    a.active = b->active;
    a.prev = shg_internal_therm_step_of(p, b->hz, 0);
    a.cur = shg_internal_therm_step_of(p, b->hz, 1);
    a.flags = shg::TROW_FWD | shg::TROW_NPREV | shg::TROW_QPREV | shg::TROW_NCUR | shg::TROW_BWD;
    float best = 1e30f;
    int br = 0, bc = 0;
    for (int rv = 0; rv < shg::kThermRowVariantCount; rv++)
        for (int cv = 0; cv < shg::kColVariantCount; cv++) {
            float ms = 0.f;
            TB_RC(shg_internal_time_graph(b->stream, b->e0, b->e1, 10, [&] {
                hipError_t e = shg::launch_thermal_row_pass(b->n, b->U, b->Uout, b->tw, b->NT, b->Q, b->DKint, a,
                                                            b->stream, rv, b->count);
                if (e == hipSuccess) e = shg::launch_col_pass(b->n, b->desc, b->count, b->tw, b->stream, cv);
                return e == hipSuccess ? (int)SHG_OK
                                       : shg_internal_fail(SHG_E_DEVICE, std::string("tuning pass: ") + hipGetErrorString(e));
            }, &ms));
            if (ms < best) { best = ms; br = rv; bc = cv; }
        }
    b->row_v = br;
    b->col_v = bc;
    return SHG_OK;
}

// Times every two-sweep layout (all instances flagged) and two single sweeps, each
// replayed from a graph, and keeps the fastest.  The timed launches read the current
// temperatures and write the other buffer only.
int tune_jacobi(shg_thermal_batch *b) {
    const char *nt = std::getenv("SHG_NO_TUNE");
    if (nt && nt[0] == '1') {
        b->jac_v = 0;
        return SHG_OK;
    }
    TB_RC(set_sweeping(b, std::vector<int>(b->count, 1)));
    float *Ti = b->T[b->tcur], *Tf = b->T[b->tcur ^ 1];
    float best = 1e30f;
    int bv = 0;
    for (int v = 0; v <= shg::kJacobi2VariantCount; v++) {
        float ms = 0.f;
        TB_RC(shg_internal_time_graph(b->stream, b->e0, b->e1, 4, [&] {
            hipError_t e;
            if (v < shg::kJacobi2VariantCount) {
                e = shg::launch_jacobi2_batch(Tf, Ti, b->Q, b->n, b->n, b->hz, b->nids, b->ids, b->jc, b->stream,
                                              b->jplan[v]);
            } else {
                e = shg::launch_jacobi_batch(Tf, Ti, b->Q, b->n, b->n, b->hz, b->nids, b->ids, b->jc, nullptr,
                                             b->stream);
                if (e == hipSuccess)
                    e = shg::launch_jacobi_batch(Tf, Ti, b->Q, b->n, b->n, b->hz, b->nids, b->ids, b->jc, nullptr,
                                                 b->stream);
            }
            return e == hipSuccess ? (int)SHG_OK
                                   : shg_internal_fail(SHG_E_DEVICE, std::string("tuning Jacobi: ") + hipGetErrorString(e));
        }, &ms));
        if (ms < best) { best = ms; bv = v; }
    }
    b->jac_v = bv;
    return SHG_OK;
}

// `nsweeps` sweeps of the instances set_sweeping chose, without the residual: two per
// pass with a two-sweep layout, then a single one if the count is odd.
int sweeps(shg_thermal_batch *b, int64_t nsweeps) {
    if (!b->nids) return SHG_OK;
    int64_t i = 0;
    if (b->jac_v < shg::kJacobi2VariantCount)
        for (; nsweeps - i >= 2; i += 2) {
            TB_TRY(shg::launch_jacobi2_batch(b->T[b->tcur ^ 1], b->T[b->tcur], b->Q, b->n, b->n, b->hz, b->nids,
                                             b->ids, b->jc, b->stream, b->jp));
            b->tcur ^= 1;
        }
    for (; i < nsweeps; i++) {
        TB_TRY(shg::launch_jacobi_batch(b->T[b->tcur ^ 1], b->T[b->tcur], b->Q, b->n, b->n, b->hz, b->nids, b->ids,
                                        b->jc, nullptr, b->stream));
        b->tcur ^= 1;
    }
    return SHG_OK;
}

// One single sweep of the instances set_sweeping chose, with each one's mean squared
// update (FP64) in resid[instance] (zero for the others).
int sweep_checked(shg_thermal_batch *b, std::vector<double> *resid) {
    resid->assign(b->count, 0.0);
    if (!b->nids) return SHG_OK;
    float *Ti = b->T[b->tcur], *Tf = b->T[b->tcur ^ 1];
    TB_TRY(shg::launch_jacobi_batch(Tf, Ti, b->Q, b->n, b->n, b->hz, b->nids, b->ids, b->jc, b->jpart, b->stream));
    b->tcur ^= 1;
    TB_TRY(shg::launch_sum_partials_batch(b->jpart, b->njb, b->nids, b->jres, b->stream));
    std::vector<double> r(b->nids);
    TB_TRY(shg_internal_copy(r.data(), b->jres, r.size() * sizeof(double), hipMemcpyDeviceToHost, b->stream));
    for (int j = 0; j < b->nids; j++) (*resid)[b->host_ids[j]] = r[j] / (double)b->n3;
    return SHG_OK;
}

// FP64 sums of |FH|^2 and |SH|^2 of one instance's plane.
int plane_sums(shg_thermal_batch *b, const float4 *U, double *sf, double *ss) {
    TB_TRY(shg::launch_power(U, (int64_t)b->n2, b->ppart, 256, b->pred, b->stream));
    double2 h;
    TB_TRY(shg_internal_copy(&h, b->pred, sizeof(double2), hipMemcpyDeviceToHost, b->stream));
    *sf = h.x;
    *ss = h.y;
    return SHG_OK;
}

double power_of(const shg_thermal_batch *b, int i, int which, double sum, double temp) {
    const shg::phys::Consts k(b->p[i]);
    const double lam = which == 0 ? (double)k.lp : (double)k.ls;
    const double nn = shg::phys::index_d(b->p[i].crystal, lam, temp);
    return 0.5 * 8.8541878128e-6 * 299.792458 * nn * sum * (double)k.dx * (double)k.dy;
}

}  // namespace

extern "C" {

int shg_thermal_batch_create(const shg_params *params, int count, shg_thermal_batch **out) {
    return shg_internal_guard(__func__, [&]() -> int {
        if (!params || !out || count < 1 || count > 65535)
            return shg_internal_fail(SHG_E_INVALID, "shg_thermal_batch_create: bad argument (count must be 1..65535)");
        *out = nullptr;
        const shg_params &p0 = params[0];
        if (!p0.thermal) return shg_internal_fail(SHG_E_INVALID, "shg_thermal_batch_create: instances must be THERMAL");
        for (int i = 1; i < count; i++)
            if (!same_build(p0, params[i]))
                return shg_internal_fail(SHG_E_INVALID, "shg_thermal_batch_create: instance " + std::to_string(i) +
                                                            " differs from instance 0 in more than power, waist, focus, "
                                                            "temp, t_peltier1, t_peltier2");
        // the single-instance constraints, checked by creating (and dropping) a context
        shg_ctx *probe = nullptr;
        TB_RC(shg_create(&p0, &probe));
        shg_destroy(probe);
        shg_thermal_batch *b = new shg_thermal_batch();
        auto bad = [&](int rc) { tb_release(b); return rc; };
        b->count = count;
        b->p.assign(params, params + count);
        b->n = p0.nx;
        b->nz = p0.nz;
        b->hz = p0.heat_nz > 0 ? p0.heat_nz : p0.nz;
        b->n2 = (size_t)b->n * b->n;
        b->n3 = b->n2 * b->hz;
        b->njb = shg::jacobi_blocks(b->n, b->n, b->hz);
        for (int v = 0; v < shg::kJacobi2VariantCount; v++)
            if (shg::jacobi2_plan(b->n, b->n, b->hz, v, &b->jplan[v], count) != 0)
                return bad(shg_internal_fail(SHG_E_DEVICE, "shg_thermal_batch_create: Jacobi layout query failed"));
        const size_t C = (size_t)count;
        hipError_t e;
        if ((e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking)) != hipSuccess ||
            (e = hipEventCreate(&b->e0)) != hipSuccess || (e = hipEventCreate(&b->e1)) != hipSuccess)
            return bad(shg_internal_fail(SHG_E_DEVICE, std::string("shg_thermal_batch_create: ") + hipGetErrorString(e)));
        if (hipMalloc(&b->U, C * b->n2 * sizeof(float4)) != hipSuccess ||
            hipMalloc(&b->Uout, C * b->n2 * sizeof(float4)) != hipSuccess ||
            hipMalloc(&b->tw, (size_t)b->n * sizeof(float2)) != hipSuccess ||
            hipMalloc(&b->G, C * 4 * b->n * sizeof(float2)) != hipSuccess ||
            hipMalloc(&b->NT, C * b->n3 * sizeof(float2)) != hipSuccess ||
            hipMalloc(&b->T[0], C * b->n3 * sizeof(float)) != hipSuccess ||
            hipMalloc(&b->T[1], C * b->n3 * sizeof(float)) != hipSuccess ||
            hipMalloc(&b->Q, C * b->n3 * sizeof(float)) != hipSuccess ||
            hipMalloc(&b->DKint, C * b->n2 * sizeof(float)) != hipSuccess ||
            hipMalloc(&b->desc, C * sizeof(shg::InstDesc)) != hipSuccess || hipMalloc(&b->active, C * sizeof(int)) != hipSuccess ||
            hipMalloc(&b->ids, C * sizeof(int)) != hipSuccess ||
            hipMalloc(&b->jpart, C * b->njb * sizeof(double)) != hipSuccess || hipMalloc(&b->jres, C * sizeof(double)) != hipSuccess ||
            hipMalloc(&b->ppart, 256 * sizeof(double2)) != hipSuccess || hipMalloc(&b->pred, sizeof(double2)) != hipSuccess)
            return bad(shg_internal_fail(SHG_E_NOMEM, "shg_thermal_batch_create: device allocation failed"));
        shg_internal_thermal_consts(p0, b->hz, &b->nl, &b->jc, &b->tc);
        const auto tw = shg::phys::twiddle_table(b->n);
        if (shg_internal_copy(b->tw, tw.data(), tw.size() * sizeof(float), hipMemcpyHostToDevice, b->stream) != hipSuccess)
            return bad(shg_internal_fail(SHG_E_DEVICE, "shg_thermal_batch_create: table upload failed"));
        std::vector<shg::InstDesc> d(C);
        for (size_t i = 0; i < C; i++) {
            shg::JacobiConst jc;
            shg::ThermConst tc;
            d[i] = shg::InstDesc{};
            d[i].U = b->U + i * b->n2;
            d[i].G = b->G + i * 4 * b->n;
            shg_internal_thermal_consts(b->p[i], b->hz, &d[i].nl, &jc, &tc);   // n_s(temp) of its own
            d[i].steps = b->nz;
        }
        if (shg_internal_copy(b->desc, d.data(), C * sizeof(shg::InstDesc), hipMemcpyHostToDevice, b->stream) != hipSuccess)
            return bad(shg_internal_fail(SHG_E_DEVICE, "shg_thermal_batch_create: table upload failed"));
        // defined contents for the layout tuner's passes (run before any instance's fields exist)
        for (std::pair<void *, size_t> z : {std::make_pair((void *)b->U, C * b->n2 * sizeof(float4)),
                                            std::make_pair((void *)b->Uout, C * b->n2 * sizeof(float4)),
                                            std::make_pair((void *)b->G, C * 4 * b->n * sizeof(float2)),
                                            std::make_pair((void *)b->DKint, C * b->n2 * sizeof(float))})
            if (hipMemsetAsync(z.first, 0, z.second, b->stream) != hipSuccess)
                return bad(shg_internal_fail(SHG_E_DEVICE, "shg_thermal_batch_create: hipMemsetAsync failed"));
        b->sh0.assign(C * 2 * b->n2, 0.f);
        b->stage.resize(4 * b->n2);
        b->p_fh_in.assign(C, 0.0);
        b->t_centre.assign(C, 0.f);
        b->stats.assign(C, shg_thermal_stats{});
        if (int rc = init_temperatures(b)) return bad(rc);
        *out = b;
        return SHG_OK;
    });
}

void shg_thermal_batch_destroy(shg_thermal_batch *b) { tb_release(b); }

int shg_thermal_batch_set_signal(shg_thermal_batch *b, int index, const float *sh) {
    return shg_internal_guard(__func__, [&]() -> int {
        if (!b || !sh || index < 0 || index >= b->count)
            return shg_internal_fail(SHG_E_INVALID, "shg_thermal_batch_set_signal: bad argument");
        std::memcpy(b->sh0.data() + (size_t)index * 2 * b->n2, sh, 2 * b->n2 * sizeof(float));
        return SHG_OK;
    });
}

int shg_thermal_batch_set_variants(shg_thermal_batch *b, int row_variant, int col_variant) {
    return shg_internal_guard(__func__, [&]() -> int {
        if (!b || row_variant < 0 || row_variant >= shg::kThermRowVariantCount || col_variant < 0 ||
            col_variant >= shg::kColVariantCount)
            return shg_internal_fail(SHG_E_INVALID, "shg_thermal_batch_set_variants: variant out of range");
        if (b->gmarch && (row_variant != b->row_v || col_variant != b->col_v)) {
            (void)hipGraphExecDestroy(b->gmarch);
            b->gmarch = nullptr;
        }
        b->row_v = row_variant;
        b->col_v = col_variant;
        return SHG_OK;
    });
}

int shg_thermal_batch_get_variants(const shg_thermal_batch *b, int *row_variant, int *col_variant) {
    return shg_internal_guard(__func__, [&]() -> int {
        if (!b || !row_variant || !col_variant) return shg_internal_fail(SHG_E_INVALID, "shg_thermal_batch_get_variants: bad argument");
        *row_variant = b->row_v;
        *col_variant = b->col_v;
        return SHG_OK;
    });
}

int shg_thermal_batch_set_jacobi_layout(shg_thermal_batch *b, int layout) {
    return shg_internal_guard(__func__, [&]() -> int {
        if (!b || layout < 0 || layout > shg::kJacobi2VariantCount)
            return shg_internal_fail(SHG_E_INVALID, "shg_thermal_batch_set_jacobi_layout: layout out of range");
        b->jac_v = layout;
        return SHG_OK;
    });
}

int shg_thermal_batch_get_jacobi_layout(const shg_thermal_batch *b, int *layout) {
    return shg_internal_guard(__func__, [&]() -> int {
        if (!b || !layout) return shg_internal_fail(SHG_E_INVALID, "shg_thermal_batch_get_jacobi_layout: bad argument");
        *layout = b->jac_v;
        return SHG_OK;
    });
}

int shg_thermal_batch_run(shg_thermal_batch *b, shg_thermal_stats *stats) {
    return shg_internal_guard(__func__, [&]() -> int {
        if (!b) return shg_internal_fail(SHG_E_INVALID, "shg_thermal_batch_run: null batch");
        const int C = b->count;
        const shg_params &p0 = b->p[0];
        if (b->row_v < 0) {
            const char *nt = std::getenv("SHG_NO_TUNE");
            if (nt && nt[0] == '1') {
                b->row_v = 0;
                b->col_v = 0;
            } else {
                TB_RC(tune(b));
            }
        }
        TB_RC(init_temperatures(b));
        if (b->jac_v < 0) TB_RC(tune_jacobi(b));
        const int check = p0.jacobi_check > 0 ? p0.jacobi_check : 5000;
        const int num_of_iter = p0.max_outer > 0 ? p0.max_outer : 5;
        const double tol = p0.jacobi_tol > 0 ? (double)(float)p0.jacobi_tol : (double)(float)(5e-4 / (double)b->n3);
        std::vector<int> global(C, 0), done(C, 0);
        // which of the two buffers holds each instance's latest temperature: a sweep flips
        // the pair for every instance, but only the instances it sweeps move
        std::vector<int> where(C, b->tcur);
        auto bring_to_current = [&](int i) -> int {
            if (where[i] != b->tcur) {
                TB_TRY(hipMemcpyAsync(b->T[b->tcur] + (size_t)i * b->n3, b->T[where[i]] + (size_t)i * b->n3,
                                      b->n3 * sizeof(float), hipMemcpyDeviceToDevice, b->stream));
                where[i] = b->tcur;
            }
            return SHG_OK;
        };
        std::vector<shg_thermal_stats> st(C, shg_thermal_stats{});
        double jac_ms = 0.0, march_ms = 0.0;
        auto all_done = [&] { return std::all_of(done.begin(), done.end(), [](int d) { return d != 0; }); };
        const shg::phys::Consts k(p0);
        while (!all_done()) {
            // Jacobi of every instance in this outer iteration, in lockstep: the same sweep
            // counts and checks as each one's own loop (src/headers/Solver.h:302-325)
            std::vector<double> reduced(C, 1.0), aux(C, 0.0);
            std::vector<int> going(C);
            std::vector<int64_t> counter_at_stop(C, 0);
            for (int i = 0; i < C; i++) going[i] = !done[i];
            int64_t counter = 0;
            float ms = 0.f;
            TB_TRY(hipEventRecord(b->e0, b->stream));
            while (std::any_of(going.begin(), going.end(), [](int g) { return g != 0; })) {
                TB_RC(set_sweeping(b, going));
                if (counter % check == 0) {
                    std::vector<double> r;
                    TB_RC(sweep_checked(b, &r));
                    for (int i = 0; i < C; i++) {
                        if (!going[i]) continue;
                        if (counter > 0) aux[i] = reduced[i];
                        reduced[i] = r[i];
                        if (counter > 0 && reduced[i] == aux[i]) reduced[i] = 0.5 * tol;
                        st[i].residual_last = r[i];
                    }
                    counter++;
                } else {
                    const int64_t next_check = (counter / check + 1) * (int64_t)check;
                    int64_t burst = next_check - counter;
                    if (p0.jacobi_max > 0) burst = std::min<int64_t>(burst, p0.jacobi_max - counter);
                    TB_RC(sweeps(b, burst));
                    counter += burst;
                }
                for (int i = 0; i < C; i++)
                    if (going[i]) where[i] = b->tcur;
                for (int i = 0; i < C; i++) {
                    if (!going[i]) continue;
                    const bool capped = p0.jacobi_max > 0 && counter >= p0.jacobi_max;
                    if (!(reduced[i] >= tol) || capped) {
                        going[i] = 0;
                        counter_at_stop[i] = counter;
                    }
                }
            }
            TB_TRY(hipEventRecord(b->e1, b->stream));
            TB_TRY(hipEventSynchronize(b->e1));
            TB_TRY(hipEventElapsedTime(&ms, b->e0, b->e1));
            jac_ms += ms;
            // each instance of this iteration stopped after its own number of sweeps, so its
            // temperature sits in one buffer or the other: bring the marching ones to T[tcur]
            const int tnow = b->tcur;
            for (int i = 0; i < C; i++)
                if (!done[i]) TB_RC(bring_to_current(i));
            // fields and propagators of the marching instances at their grid-centre
            // temperature of slice 0 (Solver.h:329-331), then one batched march
            std::vector<int> marching(C);
            for (int i = 0; i < C; i++) {
                marching[i] = !done[i];
                if (!marching[i]) continue;
                float tcen = 0.f;
                TB_TRY(shg_internal_copy(&tcen, b->T[tnow] + (size_t)i * b->n3 + (size_t)(b->n / 2) * b->n + b->n / 2,
                                         sizeof(float), hipMemcpyDeviceToHost, b->stream));
                const shg_params &p = b->p[i];
                std::vector<float> fh(2 * b->n2);
                shg::phys::make_pump(k, (float)p.power, (float)p.focus, (float)p.waist, tcen, fh.data());
                const float *sh = b->sh0.data() + (size_t)i * 2 * b->n2;
                float *h = b->stage.data();
                for (size_t c = 0; c < b->n2; c++) {
                    h[4 * c] = fh[2 * c];
                    h[4 * c + 1] = fh[2 * c + 1];
                    h[4 * c + 2] = sh[2 * c];
                    h[4 * c + 3] = sh[2 * c + 1];
                }
                float4 *Ui = b->U + (size_t)i * b->n2;
                TB_TRY(shg_internal_copy(Ui, h, b->n2 * sizeof(float4), hipMemcpyHostToDevice, b->stream));
                const auto G = shg::phys::propagator_factors(k, tcen, (double)k.dz);
                TB_TRY(shg_internal_copy(b->G + (size_t)i * 4 * b->n, G.data(), G.size() * sizeof(float),
                                         hipMemcpyHostToDevice, b->stream));
                double sf, ss;
                TB_RC(plane_sums(b, Ui, &sf, &ss));
                b->t_centre[i] = tcen;
                b->p_fh_in[i] = power_of(b, i, 0, sf, tcen);
            }
            TB_RC(set_active(b, marching));
            if (!b->gmarch) TB_RC(shg_internal_capture(b->stream, [&] { return enqueue_march(b); }, &b->gmarch));
            TB_TRY(hipEventRecord(b->e0, b->stream));
            TB_TRY(shg::launch_index_table(b->T[tnow], b->NT, (int64_t)C * b->n3, b->tc, b->stream));
            TB_TRY(hipGraphLaunch(b->gmarch, b->stream));
            TB_TRY(hipEventRecord(b->e1, b->stream));
            TB_TRY(hipEventSynchronize(b->e1));
            TB_TRY(hipEventElapsedTime(&ms, b->e0, b->e1));
            march_ms += ms;
            for (int i = 0; i < C; i++) {
                if (!marching[i]) continue;
                st[i].outer_iterations++;
                st[i].sweeps_last = counter_at_stop[i];
                st[i].sweeps_total += counter_at_stop[i];
                if (counter_at_stop[i] == 1 && global[i] != 0) global[i] = num_of_iter;   // (Solver.h:336)
                global[i]++;
                if (global[i] >= num_of_iter) done[i] = 1;
            }
        }
        for (int i = 0; i < C; i++) {            // instances that stopped early sat out the later sweeps
            TB_RC(bring_to_current(i));
            st[i].jacobi_ms = jac_ms;
            st[i].march_ms = march_ms;
        }
        TB_TRY(hipStreamSynchronize(b->stream));
        b->stats = st;
        b->ran = true;
        if (stats) std::copy(st.begin(), st.end(), stats);
        return SHG_OK;
    });
}

int shg_thermal_batch_get_fields(shg_thermal_batch *b, int index, float *fh, float *sh) {
    return shg_internal_guard(__func__, [&]() -> int {
        if (!b || !fh || !sh || index < 0 || index >= b->count)
            return shg_internal_fail(SHG_E_INVALID, "shg_thermal_batch_get_fields: bad argument");
        if (!b->ran) return shg_internal_fail(SHG_E_STATE, "shg_thermal_batch_get_fields: call shg_thermal_batch_run first");
        float *h = b->stage.data();
        TB_TRY(shg_internal_copy(h, b->Uout + (size_t)index * b->n2, b->n2 * sizeof(float4), hipMemcpyDeviceToHost,
                                 b->stream));
        for (size_t c = 0; c < b->n2; c++) {
            fh[2 * c] = h[4 * c];
            fh[2 * c + 1] = h[4 * c + 1];
            sh[2 * c] = h[4 * c + 2];
            sh[2 * c + 1] = h[4 * c + 3];
        }
        return SHG_OK;
    });
}

int shg_thermal_batch_get_temperature(shg_thermal_batch *b, int index, float *T) {
    return shg_internal_guard(__func__, [&]() -> int {
        if (!b || !T || index < 0 || index >= b->count)
            return shg_internal_fail(SHG_E_INVALID, "shg_thermal_batch_get_temperature: bad argument");
        TB_TRY(shg_internal_copy(T, b->T[b->tcur] + (size_t)index * b->n3, b->n3 * sizeof(float), hipMemcpyDeviceToHost,
                                 b->stream));
        return SHG_OK;
    });
}

int shg_thermal_batch_powers(shg_thermal_batch *b, double *p_fh_in, double *p_sh_out, double *eta) {
    return shg_internal_guard(__func__, [&]() -> int {
        if (!b) return shg_internal_fail(SHG_E_INVALID, "shg_thermal_batch_powers: null batch");
        if (!b->ran) return shg_internal_fail(SHG_E_STATE, "shg_thermal_batch_powers: call shg_thermal_batch_run first");
        for (int i = 0; i < b->count; i++) {
            double sf, ss;
            TB_RC(plane_sums(b, b->Uout + (size_t)i * b->n2, &sf, &ss));
            if (!std::isfinite(sf) || !std::isfinite(ss))
                return shg_internal_fail(SHG_E_NONFINITE, "shg_thermal_batch_powers: non-finite field in instance " +
                                                              std::to_string(i));
            const double psh = power_of(b, i, 1, ss, b->t_centre[i]);
            if (p_fh_in) p_fh_in[i] = b->p_fh_in[i];
            if (p_sh_out) p_sh_out[i] = psh;
            if (eta) eta[i] = b->p_fh_in[i] > 0 ? psh / b->p_fh_in[i] : 0.0;
        }
        return SHG_OK;
    });
}

}  // extern "C"
