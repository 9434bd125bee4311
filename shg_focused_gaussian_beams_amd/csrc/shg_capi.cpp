// shg_capi.cpp — the C ABI (include/shg.h): contexts, device state and the z-march
// schedule.  Host code only; device code lives in ssfm_kernels.hip.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/shg.h"
#include "physics.h"
#include "ssfm_kernels.h"

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

bool pow2_supported(int n) { return n == 64 || n == 128 || n == 256 || n == 512; }

}  // namespace

struct shg_ctx {
    shg_params p{};
    int n = 0;
    int steps = 0;                   // z-steps executed by shg_forward (plain: nz - 1)
    hipStream_t stream = nullptr;
    float4 *U = nullptr, *U0 = nullptr;
    float2 *tw = nullptr, *G = nullptr;
    double2 *part = nullptr, *red = nullptr;
    int nparts = 0;
    std::vector<float> coef;         // (c1, c2) per step, 4 floats
    shg::NlConst nl{};
    bool have_fields = false;
    double p_fh_in = 0.0;
    double sum_fh_in = 0.0;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    double last_ms = 0.0;
    hipGraphExec_t gexec = nullptr;
    int use_graph = 1;
    std::vector<float> host_stage;   // 4 floats per cell
};

namespace {

float4 coef_at(const shg_ctx *c, int s) {
    return make_float4(c->coef[4 * s], c->coef[4 * s + 1], c->coef[4 * s + 2], c->coef[4 * s + 3]);
}

// Enqueues steps [first, first + count) of the plain z-march on the context stream.
// The field is in real space before and after.
int enqueue_march(shg_ctx *c, int first, int count) {
    if (count <= 0) return SHG_OK;
    shg::RowArgs a{};
    a.nl = c->nl;
    a.ccur = coef_at(c, first);
    a.flags = shg::ROW_NCUR | shg::ROW_BWD;
    HIP_TRY(shg::launch_row_pass(c->n, c->U, c->tw, a, c->stream));
    const int last = first + count - 1;
    for (int s = first; s <= last; s++) {
        HIP_TRY(shg::launch_col_pass(c->n, c->U, c->tw, c->G, c->stream));
        a.cprev = coef_at(c, s);
        if (s < last) {
            a.ccur = coef_at(c, s + 1);
            a.flags = shg::ROW_FWD | shg::ROW_NPREV | shg::ROW_NCUR | shg::ROW_BWD;
        } else {
            a.flags = shg::ROW_FWD | shg::ROW_NPREV;
        }
        HIP_TRY(shg::launch_row_pass(c->n, c->U, c->tw, a, c->stream));
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

double power_of(const shg_ctx *c, int which, double sum) {
    const shg::phys::Consts k(c->p);
    const double eps0 = 8.8541878128e-6, cl = 299.792458;
    const double lam = which == 0 ? (double)k.lp : (double)k.ls;
    const double n = shg::phys::index_d(c->p.crystal, lam, c->p.temp);
    return 0.5 * eps0 * cl * n * sum * (double)k.dx * (double)k.dy;
}

void release(shg_ctx *c) {
    if (!c) return;
    if (c->gexec) (void)hipGraphExecDestroy(c->gexec);
    for (void *ptr : {(void *)c->U, (void *)c->U0, (void *)c->tw, (void *)c->G, (void *)c->part, (void *)c->red})
        if (ptr) (void)hipFree(ptr);
    if (c->e0) (void)hipEventDestroy(c->e0);
    if (c->e1) (void)hipEventDestroy(c->e1);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

}  // namespace

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
}

int shg_create(const shg_params *p, shg_ctx **out) {
    if (!p || !out) return fail(SHG_E_INVALID, "shg_create: null argument");
    *out = nullptr;
    if (p->nx != p->ny) return fail(SHG_E_INVALID, "shg_create: nx must equal ny");
    if (!pow2_supported(p->nx)) return fail(SHG_E_INVALID, "shg_create: nx must be one of 64, 128, 256, 512");
    if (p->nz < 2) return fail(SHG_E_INVALID, "shg_create: nz must be >= 2");
    if (!(p->lx > 0 && p->ly > 0 && p->lz > 0)) return fail(SHG_E_INVALID, "shg_create: box lengths must be > 0");
    if (!(p->lambda_p > 0)) return fail(SHG_E_INVALID, "shg_create: lambda_p must be > 0");
    if (p->crystal != SHG_CRYSTAL_SPPLT && p->crystal != SHG_CRYSTAL_PPLN)
        return fail(SHG_E_INVALID, "shg_create: unknown crystal");
    if (p->scheme != SHG_SCHEME_REFERENCE && p->scheme != SHG_SCHEME_STRANG)
        return fail(SHG_E_INVALID, "shg_create: unknown scheme");
    if (p->thermal) return fail(SHG_E_INVALID, "shg_create: THERMAL contexts use shg_thermal_create");

    shg_ctx *c = new shg_ctx();
    c->p = *p;
    c->n = p->nx;
    c->steps = p->nz - 1;
    const char *g = std::getenv("SHG_NO_GRAPH");
    c->use_graph = (g && g[0] == '1') ? 0 : 1;
    const size_t ncell = (size_t)c->n * c->n;
    auto bad = [&](int code) { release(c); return code; };
    hipError_t e;
    if ((e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking)) != hipSuccess)
        return bad(fail(SHG_E_DEVICE, std::string("hipStreamCreate: ") + hipGetErrorString(e)));
    if (hipMalloc(&c->U, ncell * sizeof(float4)) != hipSuccess || hipMalloc(&c->U0, ncell * sizeof(float4)) != hipSuccess ||
        hipMalloc(&c->tw, (size_t)c->n * sizeof(float2)) != hipSuccess ||
        hipMalloc(&c->G, 4 * (size_t)c->n * sizeof(float2)) != hipSuccess)
        return bad(fail(SHG_E_NOMEM, "shg_create: device allocation failed"));
    c->nparts = 256;
    if (hipMalloc(&c->part, c->nparts * sizeof(double2)) != hipSuccess || hipMalloc(&c->red, sizeof(double2)) != hipSuccess)
        return bad(fail(SHG_E_NOMEM, "shg_create: device allocation failed"));
    if (hipEventCreate(&c->e0) != hipSuccess || hipEventCreate(&c->e1) != hipSuccess)
        return bad(fail(SHG_E_DEVICE, "shg_create: hipEventCreate failed"));

    const shg::phys::Consts k(*p);
    const float T = (float)p->temp;
    const auto tw = shg::phys::twiddle_table(c->n);
    const auto G = shg::phys::propagator_factors(k, T, (double)k.dz);
    c->coef = shg::phys::step_coefficients(k, T, c->steps);
    const float h = p->scheme == SHG_SCHEME_STRANG ? 0.5f * k.dz : k.dz;
    const float ns = k.n_s(T);
    c->nl.lossP = (float)(0.5 * (double)k.alpha_p);
    c->nl.lossS = (float)(0.5 * (double)k.alpha_s);
    c->nl.tpa = (float)(0.5 * ((double)k.beta_s * 0.5 * (double)k.c * (double)k.eps0 * (double)ns));
    c->nl.h6 = h / 6.0f;
    if (hipMemcpy(c->tw, tw.data(), tw.size() * sizeof(float), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(c->G, G.data(), G.size() * sizeof(float), hipMemcpyHostToDevice) != hipSuccess)
        return bad(fail(SHG_E_DEVICE, "shg_create: table upload failed"));
    c->host_stage.resize(4 * ncell);
    *out = c;
    return SHG_OK;
}

void shg_destroy(shg_ctx *ctx) { release(ctx); }

int shg_make_pump(const shg_params *p, double t_centre, float *fh) {
    if (!p || !fh) return fail(SHG_E_INVALID, "shg_make_pump: null argument");
    if (p->nx < 2 || p->ny < 2) return fail(SHG_E_INVALID, "shg_make_pump: bad grid");
    const shg::phys::Consts k(*p);
    shg::phys::make_pump(k, (float)p->power, (float)p->focus, (float)p->waist, (float)t_centre, fh);
    return SHG_OK;
}

int shg_set_fields(shg_ctx *c, const float *fh, const float *sh) {
    if (!c || !fh || !sh) return fail(SHG_E_INVALID, "shg_set_fields: null argument");
    const size_t ncell = (size_t)c->n * c->n;
    float *h = c->host_stage.data();
    for (size_t i = 0; i < ncell; i++) {
        h[4 * i] = fh[2 * i];
        h[4 * i + 1] = fh[2 * i + 1];
        h[4 * i + 2] = sh[2 * i];
        h[4 * i + 3] = sh[2 * i + 1];
    }
    HIP_TRY(hipMemcpyAsync(c->U0, h, ncell * sizeof(float4), hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipMemcpyAsync(c->U, c->U0, ncell * sizeof(float4), hipMemcpyDeviceToDevice, c->stream));
    double sf, ss;
    int rc = reduce_powers(c, c->U0, &sf, &ss);
    if (rc) return rc;
    if (!std::isfinite(sf) || !std::isfinite(ss)) return fail(SHG_E_NONFINITE, "shg_set_fields: non-finite input field");
    c->sum_fh_in = sf;
    c->p_fh_in = power_of(c, 0, sf);
    c->have_fields = true;
    return SHG_OK;
}

int shg_forward(shg_ctx *c) {
    if (!c) return fail(SHG_E_INVALID, "shg_forward: null context");
    if (!c->have_fields) return fail(SHG_E_STATE, "shg_forward: call shg_set_fields first");
    const size_t ncell = (size_t)c->n * c->n;
    HIP_TRY(hipMemcpyAsync(c->U, c->U0, ncell * sizeof(float4), hipMemcpyDeviceToDevice, c->stream));
    if (c->use_graph && !c->gexec) {
        hipGraph_t graph;
        HIP_TRY(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
        int rc = enqueue_march(c, 0, c->steps);
        hipError_t ee = hipStreamEndCapture(c->stream, &graph);
        if (rc) return rc;
        if (ee != hipSuccess) return fail(SHG_E_DEVICE, std::string("hipStreamEndCapture: ") + hipGetErrorString(ee));
        ee = hipGraphInstantiate(&c->gexec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (ee != hipSuccess) {
            c->gexec = nullptr;
            return fail(SHG_E_DEVICE, std::string("hipGraphInstantiate: ") + hipGetErrorString(ee));
        }
    }
    HIP_TRY(hipEventRecord(c->e0, c->stream));
    if (c->use_graph) {
        HIP_TRY(hipGraphLaunch(c->gexec, c->stream));
    } else {
        int rc = enqueue_march(c, 0, c->steps);
        if (rc) return rc;
    }
    HIP_TRY(hipEventRecord(c->e1, c->stream));
    HIP_TRY(hipEventSynchronize(c->e1));
    float ms = 0.f;
    HIP_TRY(hipEventElapsedTime(&ms, c->e0, c->e1));
    c->last_ms = ms;
    return SHG_OK;
}

int shg_get_fields(shg_ctx *c, float *fh, float *sh) {
    if (!c || !fh || !sh) return fail(SHG_E_INVALID, "shg_get_fields: null argument");
    const size_t ncell = (size_t)c->n * c->n;
    float *h = c->host_stage.data();
    HIP_TRY(hipMemcpyAsync(h, c->U, ncell * sizeof(float4), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    for (size_t i = 0; i < ncell; i++) {
        fh[2 * i] = h[4 * i];
        fh[2 * i + 1] = h[4 * i + 1];
        sh[2 * i] = h[4 * i + 2];
        sh[2 * i + 1] = h[4 * i + 3];
    }
    return SHG_OK;
}

int shg_powers(shg_ctx *c, double *p_fh_in, double *p_sh_out, double *eta) {
    if (!c) return fail(SHG_E_INVALID, "shg_powers: null context");
    if (!c->have_fields) return fail(SHG_E_STATE, "shg_powers: no fields");
    double sf, ss;
    int rc = reduce_powers(c, c->U, &sf, &ss);
    if (rc) return rc;
    if (!std::isfinite(sf) || !std::isfinite(ss)) return fail(SHG_E_NONFINITE, "shg_powers: non-finite field");
    const double psh = power_of(c, 1, ss);
    if (p_fh_in) *p_fh_in = c->p_fh_in;
    if (p_sh_out) *p_sh_out = psh;
    if (eta) *eta = c->p_fh_in > 0 ? psh / c->p_fh_in : 0.0;
    return SHG_OK;
}

int shg_last_forward_ms(shg_ctx *c, double *ms) {
    if (!c || !ms) return fail(SHG_E_INVALID, "shg_last_forward_ms: null argument");
    *ms = c->last_ms;
    return SHG_OK;
}

int64_t shg_cell_steps(const shg_ctx *c) { return c ? (int64_t)c->n * c->n * c->steps : 0; }

// Average device time of one row pass (the mid-march form: R_fwd, two RK4, R_bwd) and
// one column pass, each launched `reps` times back to back between two events on the
// context stream.  The field values are overwritten; the next shg_forward restores them.
int shg_profile_kernels(shg_ctx *c, int reps, double *row_ms, double *col_ms) {
    if (!c || reps <= 0 || !row_ms || !col_ms) return fail(SHG_E_INVALID, "shg_profile_kernels: bad argument");
    shg::RowArgs a{};
    a.nl = c->nl;
    a.cprev = coef_at(c, 0);
    a.ccur = coef_at(c, c->steps > 1 ? 1 : 0);
    a.flags = shg::ROW_FWD | shg::ROW_NPREV | shg::ROW_NCUR | shg::ROW_BWD;
    float ms = 0.f;
    HIP_TRY(hipEventRecord(c->e0, c->stream));
    for (int r = 0; r < reps; r++) HIP_TRY(shg::launch_row_pass(c->n, c->U, c->tw, a, c->stream));
    HIP_TRY(hipEventRecord(c->e1, c->stream));
    HIP_TRY(hipEventSynchronize(c->e1));
    HIP_TRY(hipEventElapsedTime(&ms, c->e0, c->e1));
    *row_ms = ms / reps;
    HIP_TRY(hipEventRecord(c->e0, c->stream));
    for (int r = 0; r < reps; r++) HIP_TRY(shg::launch_col_pass(c->n, c->U, c->tw, c->G, c->stream));
    HIP_TRY(hipEventRecord(c->e1, c->stream));
    HIP_TRY(hipEventSynchronize(c->e1));
    HIP_TRY(hipEventElapsedTime(&ms, c->e0, c->e1));
    *col_ms = ms / reps;
    return SHG_OK;
}

}  // extern "C"
