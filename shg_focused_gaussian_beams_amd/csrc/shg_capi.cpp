// shg_capi.cpp — the C ABI (include/shg.h): contexts, device state and the z-march
// schedules of the plain and THERMAL builds.  Host code only; device code lives in
// ssfm_kernels.hip and thermal_kernels.hip.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <string>
#include <utility>
#include <vector>

#include "../../include/shg.h"
#include "guard.h"
#include "physics.h"
#include "ssfm_kernels.h"
#include "thermal_kernels.h"
#include "context.h"

namespace {

thread_local std::string g_err;

int fail(int code, const std::string &msg) {
    g_err = msg;
    return code;
}

#define HIP_TRY(expr)                                                                                      \
    do {                                                                                                   \
        hipError_t e_ = (expr);                                                                            \
        if (e_ != hipSuccess) return fail(SHG_E_DEVICE, std::string(#expr ": ") + hipGetErrorString(e_)); \
    } while (0)

#define RC_TRY(expr)           \
    do {                       \
        int rc_ = (expr);      \
        if (rc_) return rc_;   \
    } while (0)

bool size_supported(int n) { return n == 64 || n == 128 || n == 256 || n == 512; }

}  // namespace

int shg_internal_fail(int code, const std::string &msg) { return fail(code, msg); }

void shg_internal_thermal_consts(const shg_params &p, int hz, shg::NlConst *nl, shg::JacobiConst *jc,
                                 shg::ThermConst *tc) {
    const shg::phys::Consts k(p);
    const float T = (float)p.temp;
    const float h = p.scheme == SHG_SCHEME_STRANG ? 0.5f * k.dz : k.dz;
    nl->lossP = (float)(0.5 * (double)k.alpha_p);
    nl->lossS = (float)(0.5 * (double)k.alpha_s);
    nl->tpa = (float)(0.5 * ((double)k.beta_s * 0.5 * (double)k.c * (double)k.eps0 * (double)k.n_s(T)));
    nl->h6 = h / 6.0f;
    if (!p.thermal) return;
    const float hdz = shg::phys::heat_dz(k, hz);
    *jc = shg::phys::jacobi_const<shg::JacobiConst>(k, hdz, (float)p.k_th, (float)p.h_conv, (float)p.t_inf);
    jc->bc = p.thermal_bc;
    tc->pi = k.pi; tc->dq = k.dq; tc->lp = k.lp; tc->ls = k.ls; tc->inv_period = 1.0f / k.period;
    tc->rlp = 1.0f / k.lp; tc->rls = 1.0f / k.ls;
    tc->lossP = nl->lossP; tc->lossS = nl->lossS; tc->h6 = nl->h6;
    tc->tpa_base = (double)k.beta_s * 0.5 * (double)k.eps0 * (double)k.c;
    tc->q_p = (double)k.alpha_p * 0.5 * (double)k.c * (double)k.eps0;
    tc->q_s = (double)k.alpha_s * 0.5 * (double)k.c * (double)k.eps0;
    tc->q_i = 0.5 * (double)k.c * (double)k.eps0;
    tc->beta_s = k.beta_s;
    tc->crystal = p.crystal;
}

shg::ThermStep shg_internal_therm_step_of(const shg_params &p, int hz, int s) {
    const shg::phys::Consts k(p);
    shg::ThermStep st{};
    const float incs[3] = {0.0f, 0.25f, 0.5f};     // 0.5 * {0, .5, 1} (src/headers/Solver.h:201-212)
    for (int i = 0; i < 3; i++) st.zfac[i] = (double)(k.dz * incs[i]) + (s + 1.0) * (double)k.dz;
    st.sp1 = s + 1.0;
    for (int i = 0; i < 3; i++) st.zr[i] = st.zfac[i] / st.sp1;
    st.slice = shg::phys::heat_slice(s, p.nz, hz);
    st.step = s;
    return st;
}

hipError_t shg_internal_copy(void *dst, const void *src, size_t bytes, hipMemcpyKind kind, hipStream_t st) {
    const hipError_t e = hipMemcpyAsync(dst, src, bytes, kind, st);
    return e == hipSuccess ? hipStreamSynchronize(st) : e;
}

int shg_internal_capture(hipStream_t st, const std::function<int()> &enqueue, hipGraphExec_t *out) {
    *out = nullptr;
    hipGraph_t graph = nullptr;
    HIP_TRY(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    const int rc = enqueue();
    const hipError_t ee = hipStreamEndCapture(st, &graph);
    if (rc) {
        if (graph) (void)hipGraphDestroy(graph);
        return rc;
    }
    if (ee != hipSuccess) return fail(SHG_E_DEVICE, std::string("hipStreamEndCapture: ") + hipGetErrorString(ee));
    const hipError_t ei = hipGraphInstantiate(out, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (ei != hipSuccess) {
        *out = nullptr;
        return fail(SHG_E_DEVICE, std::string("hipGraphInstantiate: ") + hipGetErrorString(ei));
    }
    return SHG_OK;
}

int shg_internal_time_graph(hipStream_t st, hipEvent_t e0, hipEvent_t e1, int reps, const std::function<int()> &one,
                            float *ms) {
    hipGraphExec_t g = nullptr;
    const int rc = shg_internal_capture(st, [&] {
        for (int r = 0; r < reps; r++)
            if (int e = one()) return e;
        return (int)SHG_OK;
    }, &g);
    if (rc) return rc;
    hipError_t e = hipGraphLaunch(g, st);                      // warm-up replay
    float best = 1e30f;
    for (int round = 0; round < 3 && e == hipSuccess; round++) {   // the fastest of three replays
        float t = 0.f;
        e = hipEventRecord(e0, st);
        if (e == hipSuccess) e = hipGraphLaunch(g, st);
        if (e == hipSuccess) e = hipEventRecord(e1, st);
        if (e == hipSuccess) e = hipEventSynchronize(e1);
        if (e == hipSuccess) e = hipEventElapsedTime(&t, e0, e1);
        best = std::min(best, t);
    }
    (void)hipGraphExecDestroy(g);
    if (e != hipSuccess) return fail(SHG_E_DEVICE, std::string("timing replay: ") + hipGetErrorString(e));
    *ms = best / reps;
    return SHG_OK;
}


namespace {

int upload_propagators(shg_ctx *c, float temp) {
    const shg::phys::Consts k(c->p);
    const auto G = shg::phys::propagator_factors(k, temp, (double)k.dz);
    HIP_TRY(hipMemcpyAsync(c->G, G.data(), G.size() * sizeof(float), hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    c->t_centre = temp;
    return SHG_OK;
}

// Plain build: enqueues steps [first, first + count) on the context stream.  The field
// is in real space before and after.
// With `snaps` (nz planes), the row pass after step s writes the state before step s+1
// to snaps[s+1].
int enqueue_march_plain(shg_ctx *c, int first, int count, float4 *snaps = nullptr) {
    if (count <= 0) return SHG_OK;
    const int rv = c->row_v < 0 ? 0 : c->row_v, cv = c->col_v < 0 ? 0 : c->col_v;
    const size_t n2 = (size_t)c->n * c->n;
    HIP_TRY(shg::launch_row_pass(c->n, c->d_desc, 1, c->tw, first, shg::ROW_START, c->stream, rv));
    const int last = first + count - 1;
    for (int s = first; s <= last; s++) {
        HIP_TRY(shg::launch_col_pass(c->n, c->d_desc, 1, c->tw, c->stream, cv));
        HIP_TRY(shg::launch_row_pass(c->n, c->d_desc, 1, c->tw, s, s < last ? shg::ROW_MID : shg::ROW_END, c->stream,
                                     rv, snaps ? snaps + (size_t)(s + 1) * n2 : nullptr));
    }
    return SHG_OK;
}

// Times `reps` back-to-back launches of every plain-pass layout variant on the context's
// data and keeps the fastest row and column variants.  Clobbers U (the next march
// restores it from U0).
shg::ThermStep therm_step(const shg_ctx *c, int s);

// Row-pass arguments of a THERMAL middle step (every stage of the pass switched on).
static shg::ThermRowArgs therm_mid_args(const shg_ctx *c) {
    shg::ThermRowArgs a{};
    a.k = c->tc;
    a.prev = therm_step(c, 0);
    a.cur = therm_step(c, 1);
    a.flags = shg::TROW_FWD | shg::TROW_NPREV | shg::TROW_QPREV | shg::TROW_NCUR | shg::TROW_BWD;
    return a;
}

// Times every (row layout, column layout) pair on this context's size as the march runs
// them, a row pass then a column pass replayed from a graph, and keeps the fastest pair:
// timed alone, a pass finds its input in the L2 of the XCD that wrote it, but in the
// march every row pass reads what the column pass scattered over all XCDs and vice
// versa; launched one by one from the host, the launch cost hides the passes' own.  The
// passes run on the context's buffers: the caller restores the fields afterwards (a
// THERMAL march rewrites DKint and every heat-source slice it touches from scratch).
int tune_variants(shg_ctx *c, int reps) {
    const int mode = c->steps > 1 ? shg::ROW_MID : shg::ROW_END;
    const int nrow = c->p.thermal ? shg::kThermRowVariantCount : shg::kRowVariantCount;
    shg::ThermRowArgs a{};
    if (c->p.thermal) {
        HIP_TRY(shg::launch_index_table(c->T[c->tcur], c->ntab, (int64_t)c->n * c->n * c->hz, c->tc, c->stream));
        a = therm_mid_args(c);
    }
    float best = 1e30f;
    int br = 0, bc = 0;
    for (int rv = 0; rv < nrow; rv++) {
        for (int cv = 0; cv < shg::kColVariantCount; cv++) {
            float ms = 0.f;
            RC_TRY(shg_internal_time_graph(c->stream, c->e0, c->e1, reps, [&] {
                hipError_t e = c->p.thermal
                    ? shg::launch_thermal_row_pass(c->n, c->U, c->Uout, c->tw, c->ntab, c->Q, c->DKint, a,
                                                   c->stream, rv)
                    : shg::launch_row_pass(c->n, c->d_desc, 1, c->tw, 0, mode, c->stream, rv);
                if (e == hipSuccess) e = shg::launch_col_pass(c->n, c->d_desc, 1, c->tw, c->stream, cv);
                return e == hipSuccess ? (int)SHG_OK : fail(SHG_E_DEVICE, std::string("tuning pass: ") +
                                                                              hipGetErrorString(e));
            }, &ms));
            if (ms < best) { best = ms; br = rv; bc = cv; }
        }
    }
    if (c->gexec && (br != c->row_v || bc != c->col_v)) {
        (void)hipGraphExecDestroy(c->gexec);
        c->gexec = nullptr;
    }
    c->row_v = br;
    c->col_v = bc;
    return SHG_OK;
}

shg::ThermStep therm_step(const shg_ctx *c, int s) { return shg_internal_therm_step_of(c->p, c->hz, s); }

// THERMAL build: all nz steps of Solver::CWES (src/headers/Solver.h:245-276) with the
// heat-source update after every step; Uout receives the fields before the last step.
// Reads the index table (launch_index_table of the current temperature first).
int enqueue_march_thermal(shg_ctx *c, float4 *snaps = nullptr) {
    const int nz = c->p.nz;
    const size_t n2 = (size_t)c->n * c->n;
    shg::ThermRowArgs a{};
    a.k = c->tc;
    a.cur = therm_step(c, 0);
    a.flags = shg::TROW_FIRST | shg::TROW_NCUR | shg::TROW_BWD;
    const float2 *NT = c->ntab;
    const int tv = c->row_v < 0 ? 0 : c->row_v;
    HIP_TRY(shg::launch_thermal_row_pass(c->n, c->U, c->Uout, c->tw, NT, c->Q, c->DKint, a, c->stream, tv));
    int w = 0;                                      // DKint holds the prefix of step w
    for (int s = 0; s < nz; s++) {
        HIP_TRY(shg::launch_col_pass(c->n, c->d_desc, 1, c->tw, c->stream, c->col_v < 0 ? 0 : c->col_v));
        a.prev = therm_step(c, s);
        a.dk_pending = s - w;
        float4 *snap = c->Uout;
        if (s < nz - 1) {
            a.cur = therm_step(c, s + 1);
            // the heat source of slice(s) only from the last step that maps to it: the
            // reference's setQ after every step leaves the same final Q, since the later
            // steps of a slice overwrite the earlier ones (heat grid coarser than the march)
            a.flags = shg::TROW_FWD | shg::TROW_NPREV | shg::TROW_NCUR | shg::TROW_BWD;
            if (a.prev.slice != a.cur.slice) a.flags |= shg::TROW_QPREV;
            // DKint is stored where the heat slice changes (and at the last step); within
            // a slice the next pass re-adds the slice's dk to the stored prefix
            if (a.prev.slice == a.cur.slice && s + 1 < nz - 1 && s + 1 - w < shg::kMaxDkPending)
                a.flags |= shg::TROW_DKKEEP;
            else w = s + 1;
            if (s + 1 == nz - 1 || snaps) a.flags |= shg::TROW_SNAP;
            if (snaps) snap = snaps + (size_t)(s + 1) * n2;     // the state before step s+1
        } else {
            a.flags = shg::TROW_FWD | shg::TROW_NPREV | shg::TROW_QPREV;
        }
        HIP_TRY(shg::launch_thermal_row_pass(c->n, c->U, snap, c->tw, NT, c->Q, c->DKint, a, c->stream, tv));
    }
    return SHG_OK;
}

int reduce_powers(shg_ctx *c, const float4 *U, double *sum_fh, double *sum_sh) {
    HIP_TRY(shg::launch_power(U, (int64_t)c->n * c->n, c->part, c->nparts, c->red, c->stream));
    double2 h;
    HIP_TRY(hipMemcpyAsync(&h, c->red, sizeof(double2), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    *sum_fh = h.x;
    *sum_sh = h.y;
    return SHG_OK;
}

double power_of(const shg_ctx *c, int which, double sum, double temp) {
    const shg::phys::Consts k(c->p);
    const double eps0 = 8.8541878128e-6, cl = 299.792458;
    const double lam = which == 0 ? (double)k.lp : (double)k.ls;
    const double n = shg::phys::index_d(c->p.crystal, lam, temp);
    return 0.5 * eps0 * cl * n * sum * (double)k.dx * (double)k.dy;
}

void release(shg_ctx *c) {
    if (!c) return;
    // work an error return left in flight ends before its graphs and buffers go
    if (c->adj.ts) (void)hipStreamSynchronize(c->adj.ts);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    if (c->gexec) (void)hipGraphExecDestroy(c->gexec);
    for (void *ptr : {(void *)c->U, (void *)c->U0, (void *)c->Uout, (void *)c->tw, (void *)c->G, (void *)c->part,
                      (void *)c->d_coef, (void *)c->d_desc,
                      (void *)c->red, (void *)c->T[0], (void *)c->T[1], (void *)c->Q, (void *)c->DKint,
                      (void *)c->ntab, (void *)c->jpart, (void *)c->jres, (void *)c->adj.lam, (void *)c->adj.ckpt,
                      (void *)c->adj.a2, (void *)c->adj.spec, (void *)c->adj.g1, (void *)c->adj.g2, (void *)c->adj.gH,
                      (void *)c->adj.gs, (void *)c->adj.qxy, (void *)c->adj.dlam, (void *)c->adj.Ap, (void *)c->adj.As,
                      (void *)c->adj.Gc, (void *)c->adj.Dc, (void *)c->adj.dpart, (void *)c->adj.dsum,
                      (void *)c->adj.dkt, (void *)c->adj.dkc, (void *)c->adj.pump_df, (void *)c->adj.pump_dn,
                      (void *)c->adj.grad, (void *)c->adj.gpart, (void *)c->adj.gsum})
        if (ptr) (void)hipFree(ptr);
    if (c->adj.host) (void)hipHostFree(c->adj.host);
    for (hipGraphExec_t g : c->adj.g_adj)
        if (g) (void)hipGraphExecDestroy(g);
    for (hipEvent_t e : {c->adj.ev_u, c->adj.ev_tape[0], c->adj.ev_tape[1], c->adj.ev_free[0], c->adj.ev_free[1],
                         c->adj.ev_end})
        if (e) (void)hipEventDestroy(e);
    if (c->adj.ts) (void)hipStreamDestroy(c->adj.ts);
    if (c->e0) (void)hipEventDestroy(c->e0);
    if (c->e1) (void)hipEventDestroy(c->e1);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

int set_u_from_host(shg_ctx *c, float4 *dst, const float *fh, const float *sh) {
    const size_t ncell = (size_t)c->n * c->n;
    float *h = c->host_stage.data();
    for (size_t i = 0; i < ncell; i++) {
        h[4 * i] = fh[2 * i];
        h[4 * i + 1] = fh[2 * i + 1];
        h[4 * i + 2] = sh[2 * i];
        h[4 * i + 3] = sh[2 * i + 1];
    }
    HIP_TRY(hipMemcpyAsync(dst, h, ncell * sizeof(float4), hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return SHG_OK;
}

// Two Jacobi sweeps of the current temperature as `variant` runs them: one two-sweep pass
// into the other buffer (< kJacobi2VariantCount), or two single sweeps there and back
// (== kJacobi2VariantCount).  With `part`, the second sweep's partial residuals; *nparts
// receives how many.
int jacobi_pair(shg_ctx *c, int variant, double *part, int *nparts) {
    const int n = c->n, hz = c->hz;
    float *Ti = c->T[c->tcur], *Tf = c->T[c->tcur ^ 1];
    if (variant < shg::kJacobi2VariantCount) {
        HIP_TRY(shg::launch_jacobi2(Tf, Ti, c->Q, n, n, hz, c->jc, part, c->njpart, c->stream, c->jplan[variant]));
        c->tcur ^= 1;
        *nparts = c->jplan[variant].blocks;
    } else {
        HIP_TRY(shg::launch_jacobi(Tf, Ti, c->Q, n, n, hz, c->jc, nullptr, c->njpart, c->stream));
        HIP_TRY(shg::launch_jacobi(Ti, Tf, c->Q, n, n, hz, c->jc, part, c->njpart, c->stream));
        *nparts = shg::jacobi_blocks(n, n, hz);
    }
    return SHG_OK;
}

// Times every two-sweep layout (and two single sweeps) on this context's heat grid, each
// replayed from a graph, and keeps the fastest.  Every timed launch reads the current
// temperature and writes the other buffer only, so the state is unchanged.
int tune_jacobi(shg_ctx *c) {
    const char *nt = std::getenv("SHG_NO_TUNE");
    if (nt && nt[0] == '1') {
        c->jac_v = 0;
        return SHG_OK;
    }
    const int n = c->n, hz = c->hz;
    float *Ti = c->T[c->tcur], *Tf = c->T[c->tcur ^ 1];
    float best = 1e30f;
    int bv = 0;
    for (int v = 0; v <= shg::kJacobi2VariantCount; v++) {
        float ms = 0.f;
        RC_TRY(shg_internal_time_graph(c->stream, c->e0, c->e1, 4, [&] {
            hipError_t e;
            if (v < shg::kJacobi2VariantCount) {
                e = shg::launch_jacobi2(Tf, Ti, c->Q, n, n, hz, c->jc, nullptr, c->njpart, c->stream, c->jplan[v]);
            } else {
                e = shg::launch_jacobi(Tf, Ti, c->Q, n, n, hz, c->jc, nullptr, c->njpart, c->stream);
                if (e == hipSuccess) e = shg::launch_jacobi(Tf, Ti, c->Q, n, n, hz, c->jc, nullptr, c->njpart, c->stream);
            }
            return e == hipSuccess ? (int)SHG_OK : fail(SHG_E_DEVICE, std::string("tuning Jacobi: ") + hipGetErrorString(e));
        }, &ms));
        if (ms < best) { best = ms; bv = v; }
    }
    c->jac_v = bv;
    return SHG_OK;
}

// Jacobi sweeps on the current temperature, two per pass where it can (the layout tuned
// on first use); with `want_resid` the mean squared update (FP64) of the last sweep is
// returned in *resid.
int jacobi_run(shg_ctx *c, int64_t nsweeps, bool want_resid, double *resid) {
    const size_t n3 = (size_t)c->n * c->n * c->hz;
    if (nsweeps >= 2 && c->jac_v < 0) RC_TRY(tune_jacobi(c));
    int64_t i = 0;
    int nparts = 0;
    bool have_part = false;
    while (nsweeps - i >= 2) {
        const bool r = want_resid && nsweeps - i == 2;
        RC_TRY(jacobi_pair(c, c->jac_v, r ? c->jpart : nullptr, &nparts));
        have_part = r;
        i += 2;
    }
    if (i < nsweeps) {
        float *Ti = c->T[c->tcur], *Tf = c->T[c->tcur ^ 1];
        HIP_TRY(shg::launch_jacobi(Tf, Ti, c->Q, c->n, c->n, c->hz, c->jc, want_resid ? c->jpart : nullptr, c->njpart,
                                   c->stream));
        c->tcur ^= 1;
        nparts = shg::jacobi_blocks(c->n, c->n, c->hz);
        have_part = want_resid;
    }
    if (have_part) {
        HIP_TRY(shg::launch_sum_partials(c->jpart, nparts, c->jres, c->stream));
        double h;
        HIP_TRY(hipMemcpyAsync(&h, c->jres, sizeof(double), hipMemcpyDeviceToHost, c->stream));
        HIP_TRY(hipStreamSynchronize(c->stream));
        *resid = h / (double)n3;
    }
    return SHG_OK;
}

int thermal_init_temperature(shg_ctx *c) {
    const int n = c->n, hz = c->hz;
    const size_t n2 = (size_t)n * n;
    std::vector<float> t(n2 * hz, (float)c->p.temp);          // initTemperature (Tfield.h:31-45)
    const float nzf = (float)(unsigned)hz;
    for (int z = 0; z < hz; z++) {
        const float v = (z < nzf / 2.0) ? (float)c->p.t_peltier1 : (float)c->p.t_peltier2;
        float *tz = t.data() + (size_t)z * n2;
        if (c->p.thermal_bc == SHG_BC_OVEN_SURROUNDED) {
            for (int i = 0; i < n; i++) {                       // initOvenSurrounded (:71-110)
                tz[i] = tz[(size_t)(n - 1) * n + i] = v;        // y = 0, y = ny-1
                tz[(size_t)i * n] = tz[(size_t)i * n + n - 1] = v;   // x = 0, x = nx-1
            }
        } else {
            for (int x = 0; x < n; x++) tz[x] = v;              // initBottomOvensTemperature (:48-68)
        }
    }
    for (int b = 0; b < 2; b++)
        HIP_TRY(hipMemcpyAsync(c->T[b], t.data(), t.size() * sizeof(float), hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipMemsetAsync(c->Q, 0, t.size() * sizeof(float), c->stream));   // initQ0 (:261-274)
    HIP_TRY(hipStreamSynchronize(c->stream));
    c->tcur = 0;
    return SHG_OK;
}

// Picks the pass layouts on first use (SHG_NO_TUNE=1: the first layouts).  Tuning runs
// the passes on the context's buffers, so the fields are restored from U0 (the
// shg_set_fields state) afterwards; callers that stage other fields in U tune first.
int ensure_tuned(shg_ctx *c) {
    if (c->row_v >= 0) return SHG_OK;
    const char *nt = std::getenv("SHG_NO_TUNE");
    if (nt && nt[0] == '1') {
        c->row_v = 0;
        c->col_v = 0;
        return SHG_OK;
    }
    RC_TRY(tune_variants(c, 20));
    HIP_TRY(hipMemcpyAsync(c->U, c->U0, (size_t)c->n * c->n * sizeof(float4), hipMemcpyDeviceToDevice, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return SHG_OK;
}

int timed_march(shg_ctx *c) {
    RC_TRY(ensure_tuned(c));
    auto enqueue = [c] { return c->p.thermal ? enqueue_march_thermal(c) : enqueue_march_plain(c, 0, c->steps); };
    if (c->use_graph && !c->gexec) RC_TRY(shg_internal_capture(c->stream, enqueue, &c->gexec));
    HIP_TRY(hipEventRecord(c->e0, c->stream));
    if (c->p.thermal)
        HIP_TRY(shg::launch_index_table(c->T[c->tcur], c->ntab, (int64_t)c->n * c->n * c->hz, c->tc, c->stream));
    if (c->use_graph) {
        HIP_TRY(hipGraphLaunch(c->gexec, c->stream));
    } else {
        RC_TRY(enqueue());
    }
    HIP_TRY(hipEventRecord(c->e1, c->stream));
    HIP_TRY(hipEventSynchronize(c->e1));
    float ms = 0.f;
    HIP_TRY(hipEventElapsedTime(&ms, c->e0, c->e1));
    c->last_ms = ms;
    return SHG_OK;
}

}  // namespace

shg::ThermStep shg_internal_therm_step(const shg_ctx *c, int s) { return therm_step(c, s); }
int shg_internal_reduce_powers(shg_ctx *c, const float4 *U, double *sum_fh, double *sum_sh) {
    return reduce_powers(c, U, sum_fh, sum_sh);
}
double shg_internal_power_of(const shg_ctx *c, int which, double sum, double temp) {
    return power_of(c, which, sum, temp);
}

extern "C" {

const char *shg_last_error(void) { return g_err.c_str(); }

void shg_default_params(shg_params *p) {
    if (!p) return;
    std::memset(p, 0, sizeof(*p));
    p->nx = 64; p->ny = 64; p->nz = 200;                       // src/SHG.cpp:11-13
    p->crystal = SHG_CRYSTAL_SPPLT;
    p->thermal = 0;
    p->scheme = SHG_SCHEME_REFERENCE;
    p->lx = 2000.0; p->ly = 1000.0; p->lz = 30000.0;           // src/headers/Crystal.h:31-33
    p->lambda_p = 1.064;                                       // :25
    p->deff = 11.00e-6; p->qpm_period = 7.97;                  // :41-43
    p->alpha_p = 0.17e-6; p->alpha_s = 1.57e-6; p->beta_s = 5e-5;
    p->rho_p = 0.0; p->rho_s = 0.0;
    p->k_th = 8.4e-6; p->h_conv = 10e-12;                      // :52-53
    p->power = 30.0; p->waist = 29.0;                          // src/SHG.sh:31-32
    p->focus = 0.5 * 30000.0;                                  // src/SHG.cpp:39
    p->temp = 55.66; p->t_inf = 25.0;                          // src/SHG.sh:33, src/SHG.cpp:43
    p->t_peltier1 = 55.66; p->t_peltier2 = 55.66;
    p->heat_nz = 0; p->jacobi_check = 5000; p->max_outer = 5; p->jacobi_max = 0; p->jacobi_tol = 0.0;
    p->thermal_bc = SHG_BC_TOP_OPEN;                           // src/headers/Solver.h:290-291
}

int shg_create(const shg_params *p, shg_ctx **out) {
    return shg_internal_guard(__func__, [&]() -> int {
        if (!p || !out) return fail(SHG_E_INVALID, "shg_create: null argument");
        *out = nullptr;
        if (p->nx != p->ny) return fail(SHG_E_INVALID, "shg_create: nx must equal ny");
        if (!size_supported(p->nx)) return fail(SHG_E_INVALID, "shg_create: nx must be one of 64, 128, 256, 512");
        if (p->nz < 2) return fail(SHG_E_INVALID, "shg_create: nz must be >= 2");
        if (!(p->lx > 0 && p->ly > 0 && p->lz > 0)) return fail(SHG_E_INVALID, "shg_create: box lengths must be > 0");
        if (!(p->lambda_p > 0)) return fail(SHG_E_INVALID, "shg_create: lambda_p must be > 0");
        if (p->crystal != SHG_CRYSTAL_SPPLT && p->crystal != SHG_CRYSTAL_PPLN)
            return fail(SHG_E_INVALID, "shg_create: unknown crystal");
        if (p->scheme != SHG_SCHEME_REFERENCE && p->scheme != SHG_SCHEME_STRANG)
            return fail(SHG_E_INVALID, "shg_create: unknown scheme");
        if (p->thermal) {
            const int hz = p->heat_nz > 0 ? p->heat_nz : p->nz;
            if (hz < 2 || hz > p->nz) return fail(SHG_E_INVALID, "shg_create: heat_nz must be in [2, nz]");
            if (!(p->k_th > 0)) return fail(SHG_E_INVALID, "shg_create: k_th must be > 0");
            if (p->thermal_bc != SHG_BC_TOP_OPEN && p->thermal_bc != SHG_BC_OVEN_SURROUNDED)
                return fail(SHG_E_INVALID, "shg_create: unknown thermal_bc");
        }

        shg_ctx *c = new shg_ctx();
        c->p = *p;
        c->n = p->nx;
        c->steps = p->thermal ? p->nz : p->nz - 1;
        c->hz = p->thermal ? (p->heat_nz > 0 ? p->heat_nz : p->nz) : 0;
        const char *g = std::getenv("SHG_NO_GRAPH");
        c->use_graph = (g && g[0] == '1') ? 0 : 1;
        const size_t ncell = (size_t)c->n * c->n;
        auto bad = [&](int code) { release(c); return code; };
        hipError_t e;
        if ((e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking)) != hipSuccess)
            return bad(fail(SHG_E_DEVICE, std::string("hipStreamCreate: ") + hipGetErrorString(e)));
        c->nparts = 256;
        if (hipMalloc(&c->U, ncell * sizeof(float4)) != hipSuccess || hipMalloc(&c->U0, ncell * sizeof(float4)) != hipSuccess ||
            hipMalloc(&c->tw, (size_t)c->n * sizeof(float2)) != hipSuccess ||
            hipMalloc(&c->G, 4 * (size_t)c->n * sizeof(float2)) != hipSuccess ||
            hipMalloc(&c->part, c->nparts * sizeof(double2)) != hipSuccess || hipMalloc(&c->red, sizeof(double2)) != hipSuccess)
            return bad(fail(SHG_E_NOMEM, "shg_create: device allocation failed"));
        if (hipEventCreate(&c->e0) != hipSuccess || hipEventCreate(&c->e1) != hipSuccess)
            return bad(fail(SHG_E_DEVICE, "shg_create: hipEventCreate failed"));

        const shg::phys::Consts k(*p);
        const float T = (float)p->temp;
        const auto tw = shg::phys::twiddle_table(c->n);
        if (shg_internal_copy(c->tw, tw.data(), tw.size() * sizeof(float), hipMemcpyHostToDevice, c->stream) != hipSuccess)
            return bad(fail(SHG_E_DEVICE, "shg_create: table upload failed"));
        shg_internal_thermal_consts(*p, c->hz, &c->nl, &c->jc, &c->tc);
        c->host_stage.resize(4 * ncell);
        if (hipMalloc(&c->d_desc, sizeof(shg::InstDesc)) != hipSuccess)
            return bad(fail(SHG_E_NOMEM, "shg_create: device allocation failed"));
        if (!p->thermal) {
            c->coef = shg::phys::step_coefficients(k, T, c->steps);
            if (hipMalloc(&c->d_coef, c->coef.size() * sizeof(float)) != hipSuccess)
                return bad(fail(SHG_E_NOMEM, "shg_create: device allocation failed"));
            if (shg_internal_copy(c->d_coef, c->coef.data(), c->coef.size() * sizeof(float), hipMemcpyHostToDevice,
                                  c->stream) != hipSuccess)
                return bad(fail(SHG_E_DEVICE, "shg_create: table upload failed"));
        } else {
            const size_t n3 = ncell * c->hz;
            c->njpart = shg::jacobi_blocks(c->n, c->n, c->hz);
            for (int v = 0; v < shg::kJacobi2VariantCount; v++) {
                if (shg::jacobi2_plan(c->n, c->n, c->hz, v, &c->jplan[v]) != 0)
                    return bad(fail(SHG_E_DEVICE, "shg_create: Jacobi layout query failed"));
                c->njpart = std::max(c->njpart, c->jplan[v].blocks);
            }
            if (hipMalloc(&c->Uout, ncell * sizeof(float4)) != hipSuccess ||
                hipMalloc(&c->T[0], n3 * sizeof(float)) != hipSuccess || hipMalloc(&c->T[1], n3 * sizeof(float)) != hipSuccess ||
                hipMalloc(&c->Q, n3 * sizeof(float)) != hipSuccess || hipMalloc(&c->DKint, ncell * sizeof(float)) != hipSuccess ||
                hipMalloc(&c->ntab, n3 * sizeof(float2)) != hipSuccess ||
                hipMalloc(&c->jpart, (size_t)c->njpart * sizeof(double)) != hipSuccess ||
                hipMalloc(&c->jres, sizeof(double)) != hipSuccess)
                return bad(fail(SHG_E_NOMEM, "shg_create: device allocation failed"));
            if (int rc = thermal_init_temperature(c)) return bad(rc);
        }
        shg::InstDesc hd{};
        hd.U = c->U;
        hd.G = c->G;
        hd.coef = c->d_coef;
        hd.nl = c->nl;
        hd.steps = c->steps;
        if (shg_internal_copy(c->d_desc, &hd, sizeof(hd), hipMemcpyHostToDevice, c->stream) != hipSuccess)
            return bad(fail(SHG_E_DEVICE, "shg_create: table upload failed"));
        if (int rc = upload_propagators(c, T)) return bad(rc);
        *out = c;
        return SHG_OK;
    });
}

void shg_destroy(shg_ctx *ctx) { release(ctx); }

int shg_make_pump(const shg_params *p, double t_centre, float *fh) {
    return shg_internal_guard(__func__, [&]() -> int {
        if (!p || !fh) return fail(SHG_E_INVALID, "shg_make_pump: null argument");
        if (p->nx < 2 || p->ny < 2) return fail(SHG_E_INVALID, "shg_make_pump: bad grid");
        const shg::phys::Consts k(*p);
        shg::phys::make_pump(k, (float)p->power, (float)p->focus, (float)p->waist, (float)t_centre, fh);
        return SHG_OK;
    });
}

int shg_set_fields(shg_ctx *c, const float *fh, const float *sh) {
    return shg_internal_guard(__func__, [&]() -> int {
        if (!c || !fh || !sh) return fail(SHG_E_INVALID, "shg_set_fields: null argument");
        const size_t ncell = (size_t)c->n * c->n;
        RC_TRY(set_u_from_host(c, c->U0, fh, sh));
        HIP_TRY(hipMemcpyAsync(c->U, c->U0, ncell * sizeof(float4), hipMemcpyDeviceToDevice, c->stream));
        double sf, ss;
        RC_TRY(reduce_powers(c, c->U0, &sf, &ss));
        if (!std::isfinite(sf) || !std::isfinite(ss)) return fail(SHG_E_NONFINITE, "shg_set_fields: non-finite input field");
        c->p_fh_in = power_of(c, 0, sf, c->t_centre);
        c->have_fields = true;
        return SHG_OK;
    });
}

int shg_forward(shg_ctx *c) {
    return shg_internal_guard(__func__, [&]() -> int {
        if (!c) return fail(SHG_E_INVALID, "shg_forward: null context");
        if (!c->have_fields) return fail(SHG_E_STATE, "shg_forward: call shg_set_fields first");
        HIP_TRY(hipMemcpyAsync(c->U, c->U0, (size_t)c->n * c->n * sizeof(float4), hipMemcpyDeviceToDevice, c->stream));
        return timed_march(c);
    });
}

int shg_reset_fields(shg_ctx *c) {
    return shg_internal_guard(__func__, [&]() -> int {
        if (!c) return fail(SHG_E_INVALID, "shg_reset_fields: null context");
        if (!c->have_fields) return fail(SHG_E_STATE, "shg_reset_fields: call shg_set_fields first");
        HIP_TRY(hipMemcpyAsync(c->U, c->U0, (size_t)c->n * c->n * sizeof(float4), hipMemcpyDeviceToDevice, c->stream));
        HIP_TRY(hipStreamSynchronize(c->stream));
        return SHG_OK;
    });
}

int shg_march(shg_ctx *c, int first_step, int nsteps) {
    return shg_internal_guard(__func__, [&]() -> int {
        if (!c) return fail(SHG_E_INVALID, "shg_march: null context");
        if (c->p.thermal) return fail(SHG_E_INVALID, "shg_march: THERMAL contexts march whole solves (shg_forward)");
        if (!c->have_fields) return fail(SHG_E_STATE, "shg_march: call shg_set_fields first");
        if (first_step < 0 || nsteps < 0 || first_step + nsteps > c->steps)
            return fail(SHG_E_INVALID, "shg_march: step range outside [0, nz-1)");
        RC_TRY(enqueue_march_plain(c, first_step, nsteps));
        HIP_TRY(hipStreamSynchronize(c->stream));
        return SHG_OK;
    });
}

int shg_forward_snapshots(shg_ctx *c, float *fh_steps, float *sh_steps) {
    return shg_internal_guard(__func__, [&]() -> int {
        if (!c || !fh_steps || !sh_steps) return fail(SHG_E_INVALID, "shg_forward_snapshots: null argument");
        if (!c->have_fields) return fail(SHG_E_STATE, "shg_forward_snapshots: call shg_set_fields first");
        const int nz = c->p.nz;
        const size_t n2 = (size_t)c->n * c->n;
        if (c->row_v < 0) {                                     // the layouts shg_forward would pick
            RC_TRY(timed_march(c));
        }
        float4 *snaps = nullptr;
        if (hipMalloc(&snaps, (size_t)nz * n2 * sizeof(float4)) != hipSuccess)
            return fail(SHG_E_NOMEM, "shg_forward_snapshots: device allocation failed");
        auto run = [&]() -> int {
            HIP_TRY(hipMemcpyAsync(c->U, c->U0, n2 * sizeof(float4), hipMemcpyDeviceToDevice, c->stream));
            HIP_TRY(hipMemcpyAsync(snaps, c->U0, n2 * sizeof(float4), hipMemcpyDeviceToDevice, c->stream));
            if (c->p.thermal) {
                HIP_TRY(shg::launch_index_table(c->T[c->tcur], c->ntab, (int64_t)n2 * c->hz, c->tc, c->stream));
                RC_TRY(enqueue_march_thermal(c, snaps));
                HIP_TRY(hipMemcpyAsync(c->Uout, snaps + (size_t)(nz - 1) * n2, n2 * sizeof(float4),
                                       hipMemcpyDeviceToDevice, c->stream));
            } else {
                RC_TRY(enqueue_march_plain(c, 0, c->steps, snaps));
            }
            std::vector<float> h(4 * n2);
            for (int s = 0; s < nz; s++) {
                HIP_TRY(hipMemcpyAsync(h.data(), snaps + (size_t)s * n2, n2 * sizeof(float4), hipMemcpyDeviceToHost,
                                       c->stream));
                HIP_TRY(hipStreamSynchronize(c->stream));
                float *fh = fh_steps + 2 * (size_t)s * n2, *sh = sh_steps + 2 * (size_t)s * n2;
                for (size_t i = 0; i < n2; i++) {
                    fh[2 * i] = h[4 * i];
                    fh[2 * i + 1] = h[4 * i + 1];
                    sh[2 * i] = h[4 * i + 2];
                    sh[2 * i + 1] = h[4 * i + 3];
                }
            }
            return SHG_OK;
        };
        const int rc = run();
        (void)hipFree(snaps);
        return rc;
    });
}

int shg_get_fields(shg_ctx *c, float *fh, float *sh) {
    return shg_internal_guard(__func__, [&]() -> int {
        if (!c || !fh || !sh) return fail(SHG_E_INVALID, "shg_get_fields: null argument");
        const size_t ncell = (size_t)c->n * c->n;
        float *h = c->host_stage.data();
        const float4 *src = c->p.thermal ? c->Uout : c->U;
        HIP_TRY(hipMemcpyAsync(h, src, ncell * sizeof(float4), hipMemcpyDeviceToHost, c->stream));
        HIP_TRY(hipStreamSynchronize(c->stream));
        for (size_t i = 0; i < ncell; i++) {
            fh[2 * i] = h[4 * i];
            fh[2 * i + 1] = h[4 * i + 1];
            sh[2 * i] = h[4 * i + 2];
            sh[2 * i + 1] = h[4 * i + 3];
        }
        return SHG_OK;
    });
}

int shg_powers(shg_ctx *c, double *p_fh_in, double *p_sh_out, double *eta) {
    return shg_internal_guard(__func__, [&]() -> int {
        if (!c) return fail(SHG_E_INVALID, "shg_powers: null context");
        if (!c->have_fields) return fail(SHG_E_STATE, "shg_powers: no fields");
        double sf, ss;
        RC_TRY(reduce_powers(c, c->p.thermal ? c->Uout : c->U, &sf, &ss));
        if (!std::isfinite(sf) || !std::isfinite(ss)) return fail(SHG_E_NONFINITE, "shg_powers: non-finite field");
        const double psh = power_of(c, 1, ss, c->t_centre);
        if (p_fh_in) *p_fh_in = c->p_fh_in;
        if (p_sh_out) *p_sh_out = psh;
        if (eta) *eta = c->p_fh_in > 0 ? psh / c->p_fh_in : 0.0;
        return SHG_OK;
    });
}

int shg_last_forward_ms(shg_ctx *c, double *ms) {
    return shg_internal_guard(__func__, [&]() -> int {
        if (!c || !ms) return fail(SHG_E_INVALID, "shg_last_forward_ms: null argument");
        *ms = c->last_ms;
        return SHG_OK;
    });
}

int64_t shg_cell_steps(const shg_ctx *c) { return c ? (int64_t)c->n * c->n * c->steps : 0; }

int shg_tune(shg_ctx *c, int reps, int *row_variant, int *col_variant) {
    return shg_internal_guard(__func__, [&]() -> int {
        if (!c || reps <= 0) return fail(SHG_E_INVALID, "shg_tune: bad argument");
        RC_TRY(tune_variants(c, reps));
        if (c->have_fields)
            HIP_TRY(hipMemcpyAsync(c->U, c->U0, (size_t)c->n * c->n * sizeof(float4), hipMemcpyDeviceToDevice, c->stream));
        HIP_TRY(hipStreamSynchronize(c->stream));
        if (row_variant) *row_variant = c->row_v;
        if (col_variant) *col_variant = c->col_v;
        return SHG_OK;
    });
}

int shg_set_variants(shg_ctx *c, int row_variant, int col_variant) {
    return shg_internal_guard(__func__, [&]() -> int {
        const int nrow = c && c->p.thermal ? shg::kThermRowVariantCount : shg::kRowVariantCount;
        if (!c || row_variant < 0 || row_variant >= nrow || col_variant < 0 || col_variant >= shg::kColVariantCount)
            return fail(SHG_E_INVALID, "shg_set_variants: variant out of range");
        if (c->gexec && (row_variant != c->row_v || col_variant != c->col_v)) {
            (void)hipGraphExecDestroy(c->gexec);
            c->gexec = nullptr;
        }
        c->row_v = row_variant;
        c->col_v = col_variant;
        return SHG_OK;
    });
}

int shg_get_variants(const shg_ctx *c, int *row_variant, int *col_variant) {
    return shg_internal_guard(__func__, [&]() -> int {
        if (!c || !row_variant || !col_variant) return fail(SHG_E_INVALID, "shg_get_variants: bad argument");
        *row_variant = c->row_v;
        *col_variant = c->col_v;
        return SHG_OK;
    });
}

int shg_set_jacobi_layout(shg_ctx *c, int layout) {
    return shg_internal_guard(__func__, [&]() -> int {
        static_assert(SHG_JACOBI_LAYOUTS == shg::kJacobi2VariantCount + 1, "Jacobi layout count");
        if (!c || !c->p.thermal) return fail(SHG_E_INVALID, "shg_set_jacobi_layout: needs a THERMAL context");
        if (layout < 0 || layout >= SHG_JACOBI_LAYOUTS) return fail(SHG_E_INVALID, "shg_set_jacobi_layout: layout out of range");
        c->jac_v = layout;
        return SHG_OK;
    });
}

int shg_get_jacobi_layout(const shg_ctx *c, int *layout) {
    return shg_internal_guard(__func__, [&]() -> int {
        if (!c || !layout) return fail(SHG_E_INVALID, "shg_get_jacobi_layout: bad argument");
        *layout = c->jac_v;
        return SHG_OK;
    });
}

int shg_profile_kernels(shg_ctx *c, int reps, double *row_ms, double *col_ms) {
    return shg_internal_guard(__func__, [&]() -> int {
        if (!c || reps <= 0 || !row_ms || !col_ms) return fail(SHG_E_INVALID, "shg_profile_kernels: bad argument");
        if (c->p.thermal)
            HIP_TRY(shg::launch_index_table(c->T[c->tcur], c->ntab, (int64_t)c->n * c->n * c->hz, c->tc, c->stream));
        const int rv = c->row_v < 0 ? 0 : c->row_v, cv = c->col_v < 0 ? 0 : c->col_v;
        const int mode = c->steps > 1 ? shg::ROW_MID : shg::ROW_END;
        const shg::ThermRowArgs a = c->p.thermal ? therm_mid_args(c) : shg::ThermRowArgs{};
        auto check = [](hipError_t e) {
            return e == hipSuccess ? (int)SHG_OK : fail(SHG_E_DEVICE, std::string("profiled pass: ") + hipGetErrorString(e));
        };
        float ms = 0.f;
        RC_TRY(shg_internal_time_graph(c->stream, c->e0, c->e1, reps, [&] {
            return check(c->p.thermal ? shg::launch_thermal_row_pass(c->n, c->U, c->Uout, c->tw, c->ntab, c->Q, c->DKint,
                                                                     a, c->stream, rv)
                                      : shg::launch_row_pass(c->n, c->d_desc, 1, c->tw, 0, mode, c->stream, rv));
        }, &ms));
        *row_ms = ms;
        RC_TRY(shg_internal_time_graph(c->stream, c->e0, c->e1, reps, [&] {
            return check(shg::launch_col_pass(c->n, c->d_desc, 1, c->tw, c->stream, cv));
        }, &ms));
        *col_ms = ms;
        return SHG_OK;
    });
}

// ----- THERMAL --------------------------------------------------------------

int shg_thermal_run(shg_ctx *c, int pump_from_temperature, shg_thermal_stats *st) {
    return shg_internal_guard(__func__, [&]() -> int {
        if (!c) return fail(SHG_E_INVALID, "shg_thermal_run: null context");
        if (!c->p.thermal) return fail(SHG_E_INVALID, "shg_thermal_run: context was created with thermal = 0");
        if (!c->have_fields) return fail(SHG_E_STATE, "shg_thermal_run: call shg_set_fields first (SH input)");
        const shg::phys::Consts k(c->p);
        const size_t ncell = (size_t)c->n * c->n, n3 = ncell * c->hz;
        const int check = c->p.jacobi_check > 0 ? c->p.jacobi_check : 5000;
        const int num_of_iter = c->p.max_outer > 0 ? c->p.max_outer : 5;
        const double tol = c->p.jacobi_tol > 0 ? (double)(float)c->p.jacobi_tol : (double)(float)(5e-4 / (double)n3);
        // Tune before the loop: each iteration stages its own pump in U, which tuning
        // inside the first march would overwrite with the shg_set_fields state.  Tuning
        // also writes heat-source slices, which the initialisation below clears.
        RC_TRY(ensure_tuned(c));
        RC_TRY(thermal_init_temperature(c));
        // the caller's z = 0 fields (SH input; FH used when pump_from_temperature == 0)
        std::vector<float> fh0(2 * ncell), sh0(2 * ncell), stage(4 * ncell);
        HIP_TRY(hipMemcpyAsync(stage.data(), c->U0, ncell * sizeof(float4), hipMemcpyDeviceToHost, c->stream));
        HIP_TRY(hipStreamSynchronize(c->stream));
        for (size_t i = 0; i < ncell; i++) {
            fh0[2 * i] = stage[4 * i]; fh0[2 * i + 1] = stage[4 * i + 1];
            sh0[2 * i] = stage[4 * i + 2]; sh0[2 * i + 1] = stage[4 * i + 3];
        }
        shg_thermal_stats s{};
        double jac_ms = 0.0, march_ms = 0.0;
        int global_count = 0;
        auto log = [c](int event, int outer, int64_t sweeps, double value) {
            if (c->log_fn) c->log_fn(c->log_user, event, outer, sweeps, value);
        };
        while (global_count < num_of_iter) {
            log(SHG_THERMAL_EVENT_OUTER, global_count, 0, tol);
            // Jacobi to steady state, residual checked every `check` sweeps (src/headers/Solver.h:302-325)
            double reduced = 1.0, aux = 0.0;
            int64_t counter = 0;
            float ms = 0.f;
            HIP_TRY(hipEventRecord(c->e0, c->stream));
            while (reduced >= tol) {
                if (counter % check == 0) {
                    double r = 0.0;
                    RC_TRY(jacobi_run(c, 1, true, &r));
                    log(SHG_THERMAL_EVENT_CHECK, global_count, counter, r);
                    if (counter > 0) aux = reduced;
                    reduced = r;
                    if (counter > 0 && reduced == aux) reduced = 0.5 * tol;
                    s.residual_last = r;
                    counter++;
                } else {
                    const int64_t next_check = (counter / check + 1) * (int64_t)check;
                    int64_t burst = next_check - counter;
                    if (c->p.jacobi_max > 0) burst = std::min<int64_t>(burst, c->p.jacobi_max - counter);
                    RC_TRY(jacobi_run(c, burst, false, nullptr));
                    counter += burst;
                }
                if (c->p.jacobi_max > 0 && counter >= c->p.jacobi_max) break;
            }
            HIP_TRY(hipEventRecord(c->e1, c->stream));
            HIP_TRY(hipEventSynchronize(c->e1));
            HIP_TRY(hipEventElapsedTime(&ms, c->e0, c->e1));
            jac_ms += ms;
            s.sweeps_last = counter;
            s.sweeps_total += counter;
            log(SHG_THERMAL_EVENT_STEADY, global_count, counter, s.residual_last);
            // fields and propagators at the temperature of the grid centre, slice 0 (Solver.h:329-331)
            float tcen = 0.f;
            HIP_TRY(shg_internal_copy(&tcen, c->T[c->tcur] + (size_t)(c->n / 2) * c->n + c->n / 2, sizeof(float),
                                      hipMemcpyDeviceToHost, c->stream));
            if (pump_from_temperature)
                shg::phys::make_pump(k, (float)c->p.power, (float)c->p.focus, (float)c->p.waist, tcen, fh0.data());
            RC_TRY(set_u_from_host(c, c->U, fh0.data(), sh0.data()));
            RC_TRY(upload_propagators(c, tcen));
            log(SHG_THERMAL_EVENT_MARCH, global_count, counter, (double)tcen);
            RC_TRY(timed_march(c));
            march_ms += c->last_ms;
            s.outer_iterations++;
            if (counter == 1 && global_count != 0) global_count = num_of_iter;   // (Solver.h:336)
            global_count++;
        }
        // input power of the last outer iteration's pump
        RC_TRY(set_u_from_host(c, c->U0, fh0.data(), sh0.data()));
        double sf, ss;
        RC_TRY(reduce_powers(c, c->U0, &sf, &ss));
        c->p_fh_in = power_of(c, 0, sf, c->t_centre);
        s.jacobi_ms = jac_ms;
        s.march_ms = march_ms;
        if (st) *st = s;
        return SHG_OK;
    });
}

int shg_set_thermal_log(shg_ctx *c, shg_thermal_log_fn fn, void *user) {
    return shg_internal_guard(__func__, [&]() -> int {
        if (!c) return fail(SHG_E_INVALID, "shg_set_thermal_log: null context");
        c->log_fn = fn;
        c->log_user = user;
        return SHG_OK;
    });
}

int shg_jacobi(shg_ctx *c, int64_t nsweeps, double *mean_sq_update) {
    return shg_internal_guard(__func__, [&]() -> int {
        if (!c || nsweeps < 0) return fail(SHG_E_INVALID, "shg_jacobi: bad argument");
        if (!c->p.thermal) return fail(SHG_E_INVALID, "shg_jacobi: context was created with thermal = 0");
        double r = 0.0;
        RC_TRY(jacobi_run(c, nsweeps, mean_sq_update != nullptr, &r));
        HIP_TRY(hipStreamSynchronize(c->stream));
        if (mean_sq_update) *mean_sq_update = r;
        return SHG_OK;
    });
}

int shg_get_temperature(shg_ctx *c, float *T) {
    return shg_internal_guard(__func__, [&]() -> int {
        if (!c || !T || !c->p.thermal) return fail(SHG_E_INVALID, "shg_get_temperature: bad argument");
        HIP_TRY(shg_internal_copy(T, c->T[c->tcur], (size_t)c->n * c->n * c->hz * sizeof(float), hipMemcpyDeviceToHost,
                                  c->stream));
        return SHG_OK;
    });
}

int shg_set_temperature(shg_ctx *c, const float *T) {
    return shg_internal_guard(__func__, [&]() -> int {
        if (!c || !T || !c->p.thermal) return fail(SHG_E_INVALID, "shg_set_temperature: bad argument");
        const size_t bytes = (size_t)c->n * c->n * c->hz * sizeof(float);
        HIP_TRY(shg_internal_copy(c->T[0], T, bytes, hipMemcpyHostToDevice, c->stream));
        HIP_TRY(shg_internal_copy(c->T[1], T, bytes, hipMemcpyHostToDevice, c->stream));
        c->tcur = 0;
        return upload_propagators(c, T[(size_t)(c->n / 2) * c->n + c->n / 2]);
    });
}

int shg_get_heat_source(shg_ctx *c, float *Q) {
    return shg_internal_guard(__func__, [&]() -> int {
        if (!c || !Q || !c->p.thermal) return fail(SHG_E_INVALID, "shg_get_heat_source: bad argument");
        HIP_TRY(shg_internal_copy(Q, c->Q, (size_t)c->n * c->n * c->hz * sizeof(float), hipMemcpyDeviceToHost, c->stream));
        return SHG_OK;
    });
}

int shg_set_heat_source(shg_ctx *c, const float *Q) {
    return shg_internal_guard(__func__, [&]() -> int {
        if (!c || !Q || !c->p.thermal) return fail(SHG_E_INVALID, "shg_set_heat_source: bad argument");
        HIP_TRY(shg_internal_copy(c->Q, Q, (size_t)c->n * c->n * c->hz * sizeof(float), hipMemcpyHostToDevice, c->stream));
        return SHG_OK;
    });
}

}  // extern "C"
