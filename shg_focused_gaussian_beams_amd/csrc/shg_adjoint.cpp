// shg_adjoint.cpp — gradients of the SH efficiency (C ABI in include/shg.h), of a plain
// context and of a THERMAL one with its temperature field held fixed.  New: the
// reference has no derivatives.
//
// J = eta = P_SH(out) / P_FH(z = 0).  The z-march is differentiated discretely
// (adjoint_kernels.hip): a forward that keeps the state every K steps (checkpoints) and
// tapes the last segment, then per segment, last first, the adjoint march through it.
// The taped recomputation of segment j-1 runs on a second stream while the adjoint of
// segment j runs (two tape buffers): both are latency-bound passes over one field, so
// overlapping them fills the GPU that either leaves mostly idle.  Parameter sensitivities:
//   - the z = 0 pump field: lambda_0 plus the direct P_in term;
//   - the focal position: through the reference pump setPump (src/headers/Efields.h:74-104)
//     differentiated analytically;
//   - the crystal temperature: the step coefficients c1, c2 (kappa and the phase z*dk),
//     the TPA coefficient, the propagator (k_j = 2 pi n_j / lambda_j), the pump (n_p in
//     Ap0 and zR) and the n_j of the two powers.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/shg.h"
#include "guard.h"
#include "adjoint_kernels.h"
#include "thermal_adjoint_kernels.h"
#include "context.h"
#include "physics.h"

namespace {

#define AHIP_TRY(expr)                                                                                               \
    do {                                                                                                             \
        hipError_t e_ = (expr);                                                                                      \
        if (e_ != hipSuccess) return shg_internal_fail(SHG_E_DEVICE, std::string(#expr ": ") + hipGetErrorString(e_)); \
    } while (0)

#define ARC_TRY(expr)          \
    do {                       \
        int rc_ = (expr);      \
        if (rc_) return rc_;   \
    } while (0)

using cd = std::complex<double>;
constexpr int kMapParts = 1024;                // block sums of the THERMAL dJ/dT map
constexpr int kGradParts = 256;                // block sums of the pump projections

// Layout of the pinned result staging: the gradient field (n2 float2), the output power
// sums, the pump projections, the dJ/dT map total, the per-step sums (steps x 6).
struct HostOffsets {
    size_t grad, sums, gsum, dsum, gs, end;
};
HostOffsets host_offsets(size_t n2, int steps) {
    HostOffsets o;
    o.grad = 0;
    o.sums = n2 * sizeof(float2);
    o.gsum = o.sums + sizeof(double2);
    o.dsum = o.gsum + sizeof(double2);
    o.gs = o.dsum + sizeof(double2);
    o.end = o.gs + (size_t)steps * 6 * sizeof(double);
    return o;
}
template <class T> T *host_at(AdjointBuffers &b, size_t off) {
    return reinterpret_cast<T *>(static_cast<char *>(b.host) + off);
}

// Device allocation of the one-time adjoint buffers.  SHG_ADJ_FAIL_ALLOC=k makes the
// k-th such allocation of a setup fail (a test hook for the recovery path below).
hipError_t adj_malloc(void **ptr, size_t bytes, int *count) {
    const char *f = std::getenv("SHG_ADJ_FAIL_ALLOC");
    if (f && std::atoi(f) == ++*count) return hipErrorOutOfMemory;
    return hipMalloc(ptr, bytes);
}

// Frees the one-time buffers of ensure_buffers and resets them, so that a setup that
// failed part way leaves nothing behind and the next call starts again from scratch.
void free_base_buffers(AdjointBuffers &b) {
    for (void **ptr : {(void **)&b.lam, (void **)&b.g1, (void **)&b.g2, (void **)&b.gH, (void **)&b.gs,
                       (void **)&b.qxy, (void **)&b.dlam, (void **)&b.grad, (void **)&b.gpart, (void **)&b.gsum})
        if (*ptr) { (void)hipFree(*ptr); *ptr = nullptr; }
    if (b.host) { (void)hipHostFree(b.host); b.host = nullptr; }
    b.host_bytes = 0;
    for (hipEvent_t *e : {&b.ev_end, &b.ev_u, &b.ev_tape[0], &b.ev_tape[1], &b.ev_free[0], &b.ev_free[1]})
        if (*e) { (void)hipEventDestroy(*e); *e = nullptr; }
    if (b.ts) { (void)hipStreamDestroy(b.ts); b.ts = nullptr; }
    b.base_ready = false;
}

int setup_base_buffers(shg_ctx *c) {
    AdjointBuffers &b = c->adj;
    const size_t n2 = (size_t)c->n * c->n;
    const int M = c->steps;
    int count = 0;
    if (adj_malloc((void **)&b.lam, n2 * sizeof(float4), &count) != hipSuccess ||
        adj_malloc((void **)&b.g1, (size_t)M * shg::adj_row_parts(c->n) * 5 * sizeof(double), &count) != hipSuccess ||
        adj_malloc((void **)&b.g2, (size_t)M * shg::adj_row_parts(c->n) * 5 * sizeof(double), &count) != hipSuccess ||
        adj_malloc((void **)&b.gH, (size_t)M * shg::adj_col_parts(c->n) * sizeof(double), &count) != hipSuccess ||
        adj_malloc((void **)&b.gs, (size_t)M * 6 * sizeof(double), &count) != hipSuccess ||
        adj_malloc((void **)&b.qxy, (size_t)c->n * sizeof(float2), &count) != hipSuccess ||
        adj_malloc((void **)&b.dlam, sizeof(shg::InstDesc), &count) != hipSuccess ||
        adj_malloc((void **)&b.grad, n2 * sizeof(float2), &count) != hipSuccess ||
        adj_malloc((void **)&b.gpart, kGradParts * sizeof(double2), &count) != hipSuccess ||
        adj_malloc((void **)&b.gsum, sizeof(double2), &count) != hipSuccess)
        return shg_internal_fail(SHG_E_NOMEM, "shg_adjoint: device allocation failed");
    b.host_bytes = host_offsets(n2, M).end;
    if (hipHostMalloc(&b.host, b.host_bytes, hipHostMallocDefault) != hipSuccess) {
        b.host = nullptr;
        return shg_internal_fail(SHG_E_NOMEM, "shg_adjoint: pinned host allocation failed");
    }
    AHIP_TRY(hipEventCreate(&b.ev_end));
    AHIP_TRY(hipStreamCreateWithFlags(&b.ts, hipStreamNonBlocking));
    for (hipEvent_t *e : {&b.ev_u, &b.ev_tape[0], &b.ev_tape[1], &b.ev_free[0], &b.ev_free[1]})
        AHIP_TRY(hipEventCreateWithFlags(e, hipEventDisableTiming));
    // 2 pi^2 f^2 on the fftshifted index grid (DC at 0), x and y spacings
    const shg::phys::Consts k(c->p);
    std::vector<float> q(2 * (size_t)c->n);
    const double PI = 3.14159265358979323846;
    for (int i = 0; i < c->n; i++) {
        const int orig = (i + c->n / 2) % c->n;
        const double fx = (orig - 0.5 * c->n) / ((double)k.dx * c->n), fy = (orig - 0.5 * c->n) / ((double)k.dy * c->n);
        q[2 * i] = (float)(2 * PI * PI * fx * fx);
        q[2 * i + 1] = (float)(2 * PI * PI * fy * fy);
    }
    AHIP_TRY(shg_internal_copy(b.qxy, q.data(), q.size() * sizeof(float), hipMemcpyHostToDevice, c->stream));
    shg::InstDesc d{};
    AHIP_TRY(shg_internal_copy(&d, c->d_desc, sizeof(d), hipMemcpyDeviceToHost, c->stream));
    d.U = b.lam;
    AHIP_TRY(shg_internal_copy(b.dlam, &d, sizeof(d), hipMemcpyHostToDevice, c->stream));
    b.base_ready = true;
    return SHG_OK;
}

int ensure_buffers(shg_ctx *c, int K, int nseg) {
    AdjointBuffers &b = c->adj;
    const size_t n2 = (size_t)c->n * c->n;
    if (!b.base_ready) {
        const int rc = setup_base_buffers(c);
        if (rc) {
            free_base_buffers(b);
            return rc;
        }
    }
    if (b.k != K || b.nseg != nseg) {
        for (hipGraphExec_t &g : b.g_adj)
            if (g) { (void)hipGraphExecDestroy(g); g = nullptr; }
        for (float4 **ptr : {&b.ckpt, &b.a2, &b.spec})
            if (*ptr) { (void)hipFree(*ptr); *ptr = nullptr; }
        const int slots = nseg > 1 ? 2 : 1;
        if (hipMalloc(&b.ckpt, (size_t)nseg * n2 * sizeof(float4)) != hipSuccess ||
            hipMalloc(&b.a2, (size_t)slots * K * n2 * sizeof(float4)) != hipSuccess ||
            hipMalloc(&b.spec, (size_t)slots * K * n2 * sizeof(float4)) != hipSuccess) {
            b.k = 0;
            return shg_internal_fail(SHG_E_NOMEM, "shg_adjoint: tape allocation failed");
        }
        b.k = K;
        b.nseg = nseg;
    }
    return SHG_OK;
}

// Taped forward of the segment [s0, s1) from the state in the context's field buffer.
int enqueue_tape(shg_ctx *c, const shg::TapeArgs &tp, int s1, hipStream_t st) {
    const int n = c->n;
    AHIP_TRY(shg::launch_row_tape(n, c->d_desc, c->tw, tp.s0, shg::ROW_START, tp, st));
    for (int s = tp.s0; s < s1; s++) {
        AHIP_TRY(shg::launch_col_tape(n, c->d_desc, c->tw, s, tp, st));
        AHIP_TRY(shg::launch_row_tape(n, c->d_desc, c->tw, s, s < s1 - 1 ? shg::ROW_MID : shg::ROW_END, tp, st));
    }
    return SHG_OK;
}

// dn/dT of the double-precision index (central difference; the Sellmeier form is smooth).
double dn_dT(int crystal, double lam, double T) {
    const double h = 1e-3;
    return (shg::phys::index_d(crystal, lam, T + h) - shg::phys::index_d(crystal, lam, T - h)) / (2 * h);
}

// The derivative fields of the reference pump (setPump, src/headers/Efields.h:74-104,
// differentiated analytically at temperature T) with respect to the focus and to n_p:
// they depend on the context's parameters and T only, so they are built once per T and
// kept on the device.
int ensure_pump_fields(shg_ctx *c, double np, double T) {
    AdjointBuffers &b = c->adj;
    const int n = c->n;
    const size_t n2 = (size_t)n * n;
    if (b.pump_df && b.pump_dn && b.pump_T == T) return SHG_OK;
    b.pump_T = -1e300;                         // valid again only once both fields hold T's values
    for (double2 **ptr : {&b.pump_df, &b.pump_dn})
        if (!*ptr && hipMalloc(ptr, n2 * sizeof(double2)) != hipSuccess) {
            *ptr = nullptr;
            return shg_internal_fail(SHG_E_NOMEM, "shg_adjoint: device allocation failed");
        }
    const shg::phys::Consts k(c->p);
    const double eps0 = 8.8541878128e-6, cl = 299.792458;
    const double w0 = c->p.waist, f = c->p.focus, lam0 = (double)k.lp;
    const double zr = M_PI * np * w0 * w0 / lam0;
    const double ap0 = std::sqrt(4.0 * c->p.power / (eps0 * cl * M_PI * np * w0 * w0));
    const cd Mq(1.0, -f / zr);
    std::vector<cd> pdf(n2), pdn(n2);
    for (int iy = 0; iy < n; iy++)
        for (int ix = 0; ix < n; ix++) {
            const double xx = (ix - 0.5 * n) * (double)k.dx, yy = (iy - 0.5 * n) * (double)k.dy;
            const double r2 = xx * xx + yy * yy;
            const cd A = ap0 / Mq * std::exp(-r2 / (w0 * w0 * Mq));
            const cd dA_dM = A * (-1.0 / Mq + r2 / (w0 * w0 * Mq * Mq));
            pdf[(size_t)iy * n + ix] = dA_dM * cd(0.0, -1.0 / zr);
            pdn[(size_t)iy * n + ix] = -A / (2.0 * np) + dA_dM * cd(0.0, f / (zr * np));
        }
    AHIP_TRY(shg_internal_copy(b.pump_df, pdf.data(), n2 * sizeof(double2), hipMemcpyHostToDevice, c->stream));
    AHIP_TRY(shg_internal_copy(b.pump_dn, pdn.data(), n2 * sizeof(double2), hipMemcpyHostToDevice, c->stream));
    b.pump_T = T;
    return SHG_OK;
}

// After the adjoint march (lambda_0 in b.lam): d eta / d A0 = lambda_0 (FH part) + the
// d(1/P_in)/dA0 term, and its projections on the pump derivative fields, on the device;
// then the results are copied to the pinned staging (the field only when wanted).  The
// caller synchronises the stream before reading them.
int enqueue_gradient_results(shg_ctx *c, double cs, double cp, bool want_field, hipStream_t st) {
    AdjointBuffers &b = c->adj;
    const size_t n2 = (size_t)c->n * c->n;
    const HostOffsets o = host_offsets(n2, c->steps);
    AHIP_TRY(shg::launch_field_grad(b.lam, c->U0, c->red, cs, cp, c->p_fh_in, b.pump_df, b.pump_dn, b.grad, b.gpart,
                                    kGradParts, (int64_t)n2, st));
    AHIP_TRY(shg::launch_fold2(b.gpart, kGradParts, b.gsum, st));
    AHIP_TRY(hipEventRecord(b.ev_end, st));
    AHIP_TRY(hipMemcpyAsync(host_at<double2>(b, o.sums), c->red, sizeof(double2), hipMemcpyDeviceToHost, st));
    AHIP_TRY(hipMemcpyAsync(host_at<double2>(b, o.gsum), b.gsum, sizeof(double2), hipMemcpyDeviceToHost, st));
    AHIP_TRY(hipMemcpyAsync(host_at<double>(b, o.gs), b.gs, (size_t)c->steps * 6 * sizeof(double),
                            hipMemcpyDeviceToHost, st));
    if (want_field)
        AHIP_TRY(hipMemcpyAsync(host_at<float2>(b, o.grad), b.grad, n2 * sizeof(float2), hipMemcpyDeviceToHost, st));
    return SHG_OK;
}

void copy_field_out(shg_ctx *c, float *d_eta_d_fh) {
    if (!d_eta_d_fh) return;
    const size_t n2 = (size_t)c->n * c->n;
    std::memcpy(d_eta_d_fh, host_at<float2>(c->adj, host_offsets(n2, c->steps).grad), n2 * sizeof(float2));
}

// ----- THERMAL, temperature held fixed -------------------------------------------
int ensure_thermal_buffers(shg_ctx *c, int K, int nseg) {
    AdjointBuffers &b = c->adj;
    const size_t n2 = (size_t)c->n * c->n, n3 = n2 * c->hz;
    if (!b.th_ready) {
        int count = 0;
        if (adj_malloc((void **)&b.Ap, n3 * sizeof(double), &count) != hipSuccess ||
            adj_malloc((void **)&b.As, n3 * sizeof(double), &count) != hipSuccess ||
            adj_malloc((void **)&b.Gc, n2 * sizeof(double), &count) != hipSuccess ||
            adj_malloc((void **)&b.Dc, n2 * sizeof(double), &count) != hipSuccess ||
            adj_malloc((void **)&b.dpart, kMapParts * sizeof(double), &count) != hipSuccess ||
            adj_malloc((void **)&b.dsum, sizeof(double), &count) != hipSuccess) {
            for (double **ptr : {&b.Ap, &b.As, &b.Gc, &b.Dc, &b.dpart, &b.dsum})
                if (*ptr) { (void)hipFree(*ptr); *ptr = nullptr; }
            return shg_internal_fail(SHG_E_NOMEM, "shg_adjoint: device allocation failed");
        }
        b.th_ready = true;
    }
    if (b.th_k != K || b.th_nseg != nseg) {
        for (float **ptr : {&b.dkt, &b.dkc})
            if (*ptr) { (void)hipFree(*ptr); *ptr = nullptr; }
        const int slots = nseg > 1 ? 2 : 1;
        if (hipMalloc(&b.dkt, (size_t)slots * K * n2 * sizeof(float)) != hipSuccess ||
            hipMalloc(&b.dkc, (size_t)nseg * n2 * sizeof(float)) != hipSuccess) {
            b.th_k = 0;
            return shg_internal_fail(SHG_E_NOMEM, "shg_adjoint: tape allocation failed");
        }
        b.th_k = K;
        b.th_nseg = nseg;
    }
    return SHG_OK;
}

// THERMAL forward of steps [s0, s1) from the real-space state A_{s0} in U (and, for s0 > 0,
// DKint_{s0-1} in DKint), without the heat-source update; with a tape it records the
// fields before every second RK4, DKint of every step and the column-pass spectra.
int enqueue_thermal_segment(shg_ctx *c, int s0, int s1, const shg::TapeArgs *tp, float *dk_tape, hipStream_t st) {
    const int n = c->n, rv = c->row_v < 0 ? 0 : c->row_v, cv = c->col_v < 0 ? 0 : c->col_v;
    shg::ThermRowArgs a{};
    a.k = c->tc;
    const int tflag = tp ? shg::TROW_TAPE : 0;
    if (tp) {
        a.tape_s0 = s0;
        a.tape_a2 = tp->a2;
        a.tape_dk = dk_tape;
    }
    a.cur = shg_internal_therm_step(c, s0);
    a.flags = shg::TROW_NCUR | shg::TROW_BWD | (s0 == 0 ? shg::TROW_FIRST : 0) | tflag;
    AHIP_TRY(shg::launch_thermal_row_pass(n, c->U, c->Uout, c->tw, c->ntab, c->Q, c->DKint, a, st, rv));
    for (int s = s0; s < s1; s++) {
        if (tp) {
            AHIP_TRY(shg::launch_col_tape(n, c->d_desc, c->tw, s, *tp, st));
        } else {
            AHIP_TRY(shg::launch_col_pass(n, c->d_desc, 1, c->tw, st, cv));
        }
        a.prev = shg_internal_therm_step(c, s);
        if (s < s1 - 1) {
            a.cur = shg_internal_therm_step(c, s + 1);
            a.flags = shg::TROW_FWD | shg::TROW_NPREV | shg::TROW_NCUR | shg::TROW_BWD | tflag;
        } else {
            a.flags = shg::TROW_FWD | shg::TROW_NPREV | tflag;
        }
        AHIP_TRY(shg::launch_thermal_row_pass(n, c->U, c->Uout, c->tw, c->ntab, c->Q, c->DKint, a, st, rv));
    }
    return SHG_OK;
}

// Gradients of eta = P_SH(Signal_out) / P_FH(z = 0) of a THERMAL context's march at its
// current temperature field, which is held fixed: field, focus, and temperature (a
// uniform shift of the field; per voxel in AdjointBuffers::Ap).  The heat source is not
// touched.
// sens = false: the field gradient and eta only (no focus, temperature or dJ/dT map).
// Checkpoint interval of the adjoint for `checkpoint_interval` <= 0: the whole march as one
// taped segment (no recomputation) when its tape (the fields before every second RK4 and
// the column spectra, plus DKint per step in THERMAL) takes at most a quarter of the free
// device memory, else ceil(sqrt(M)) with per-segment recomputation.  A context that
// already holds a whole-march tape keeps it.
int auto_interval(const shg_ctx *c, int M, size_t n2, bool thermal) {
    if (c->adj.k == M && c->adj.nseg == 1) return M;
    const double tape = (double)M * (double)n2 * (2.0 * sizeof(float4) + (thermal ? sizeof(float) : 0.0));
    size_t free_b = 0, total_b = 0;
    if (hipMemGetInfo(&free_b, &total_b) == hipSuccess && tape <= 0.25 * (double)free_b) return M;
    return (int)std::ceil(std::sqrt((double)M));
}

// The tape buffers for interval K (and the THERMAL ones); an automatic whole-march tape that
// does not fit after all (other contexts allocated meanwhile) falls back to sqrt(M).
int adjoint_buffers(shg_ctx *c, int M, bool auto_k, bool thermal, int *K, int *nseg) {
    *K = std::max(1, std::min(*K, M));
    *nseg = (M + *K - 1) / *K;
    int rc = ensure_buffers(c, *K, *nseg);
    if (!rc && thermal) rc = ensure_thermal_buffers(c, *K, *nseg);
    if (rc == SHG_E_NOMEM && auto_k && *K == M && M > 1) {
        (void)hipGetLastError();  // the failed hipMalloc's status must not reach the next launch check
        *K = std::max(1, std::min((int)std::ceil(std::sqrt((double)M)), M));
        *nseg = (M + *K - 1) / *K;
        rc = ensure_buffers(c, *K, *nseg);
        if (!rc && thermal) rc = ensure_thermal_buffers(c, *K, *nseg);
    }
    return rc;
}

int thermal_adjoint(shg_ctx *c, int checkpoint_interval, float *d_eta_d_fh, shg_adjoint_result *res, bool sens) {
    const int M = c->p.nz - 1, n = c->n;
    const size_t n2 = (size_t)n * n, n3 = n2 * c->hz;
    int K = checkpoint_interval > 0 ? checkpoint_interval : auto_interval(c, M, n2, true), nseg = 0;
    ARC_TRY(adjoint_buffers(c, M, checkpoint_interval <= 0, true, &K, &nseg));
    AdjointBuffers &b = c->adj;
    b.map_valid = false;
    hipStream_t st = c->stream;
    const int cb = shg::adj_col_parts(n);
    const shg::phys::Consts k(c->p);
    const double T = c->t_centre;                              // propagators, powers, pump
    const double eps0 = 8.8541878128e-6, cl = 299.792458;
    const double np = shg::phys::index_d(c->p.crystal, (double)k.lp, T), ns = shg::phys::index_d(c->p.crystal, (double)k.ls, T);
    const double cp = 0.5 * eps0 * cl * np * (double)k.dx * (double)k.dy;
    const double cs = 0.5 * eps0 * cl * ns * (double)k.dx * (double)k.dy;
    const double pin = c->p_fh_in;
    const double dnp = dn_dT(c->p.crystal, (double)k.lp, T), dns = dn_dT(c->p.crystal, (double)k.ls, T);
    const double kp = 2 * M_PI * np / (double)k.lp, ks = 2 * M_PI * ns / (double)k.ls;
    const double dkp = 2 * M_PI * dnp / (double)k.lp, dks = 2 * M_PI * dns / (double)k.ls;
    const float a_p = (float)((double)k.dz * dkp / (kp * kp)), a_s = (float)((double)k.dz * dks / (ks * ks));
    ARC_TRY(ensure_pump_fields(c, np, T));
    auto slot = [&](int j) { return (size_t)(nseg > 1 ? (j & 1) : 0); };
    auto tape_of = [&](int j) {
        const size_t off = slot(j) * K * n2;
        return shg::TapeArgs{b.ckpt + (size_t)j * n2, b.a2 + off, b.spec + off, nullptr, j * K, 0};
    };
    auto dk_tape = [&](int j) { return b.dkt + slot(j) * K * n2; };
    auto copy_floats = [&](float *dst, const float *src, hipStream_t s) {
        return shg::launch_copy_cells(reinterpret_cast<float4 *>(dst), reinterpret_cast<const float4 *>(src),
                                      (int64_t)(n2 / 4), s);
    };

    // 1. forward at the fixed temperature: checkpoints (fields, DKint) of every segment,
    //    the last segment taped, the output powers
    AHIP_TRY(hipEventRecord(c->e0, st));
    AHIP_TRY(shg::launch_index_table(c->T[c->tcur], c->ntab, (int64_t)n3, c->tc, st));
    AHIP_TRY(shg::launch_copy_cells(c->U, c->U0, (int64_t)n2, st));
    for (int j = 0; j < nseg; j++) {
        const int s0 = j * K, s1 = std::min(M, s0 + K);
        AHIP_TRY(shg::launch_copy_cells(b.ckpt + (size_t)j * n2, c->U, (int64_t)n2, st));
        if (j > 0) AHIP_TRY(copy_floats(b.dkc + (size_t)j * n2, c->DKint, st));
        if (j < nseg - 1) {
            ARC_TRY(enqueue_thermal_segment(c, s0, s1, nullptr, nullptr, st));
        } else {
            const shg::TapeArgs tp = tape_of(j);
            ARC_TRY(enqueue_thermal_segment(c, s0, s1, &tp, dk_tape(j), st));
        }
    }
    AHIP_TRY(shg::launch_power(c->U, (int64_t)n2, c->part, c->nparts, c->red, st));
    AHIP_TRY(hipEventRecord(c->e1, st));

    // 2. seed, then per segment, last first: the adjoint on `st`, the previous segment's
    //    taped recomputation on `ts` (no host round trip after the forward: the seed
    //    scale needs P_in only, and eta is read back with the gradients)
    AHIP_TRY(shg::launch_adj_seed(c->U, b.lam, (int64_t)n2, (float)(2.0 * cs / pin), st));
    if (sens) {
        for (double *ptr : {b.Ap, b.As})
            AHIP_TRY(shg::launch_fill_zero(ptr, (int64_t)n3, st));
        for (double *ptr : {b.Gc, b.Dc})
            AHIP_TRY(shg::launch_fill_zero(ptr, (int64_t)n2, st));
    }
    if (nseg > 1) {
        AHIP_TRY(hipEventRecord(b.ev_u, st));
        AHIP_TRY(hipStreamWaitEvent(b.ts, b.ev_u, 0));
    }
    shg::ThermAdjArgs aa{};
    aa.k = c->tc;
    aa.NT = c->ntab;
    aa.Ap = b.Ap;
    aa.As = b.As;
    aa.G = b.Gc;
    aa.Dp = b.Dc;
    for (int j = nseg - 1; j >= 0; j--) {
        if (j >= 1) {
            if (j + 1 <= nseg - 1) AHIP_TRY(hipStreamWaitEvent(b.ts, b.ev_free[(j - 1) & 1], 0));
            AHIP_TRY(shg::launch_copy_cells(c->U, b.ckpt + (size_t)(j - 1) * n2, (int64_t)n2, b.ts));
            if (j - 1 > 0) AHIP_TRY(copy_floats(c->DKint, b.dkc + (size_t)(j - 1) * n2, b.ts));
            const shg::TapeArgs tpp = tape_of(j - 1);
            ARC_TRY(enqueue_thermal_segment(c, (j - 1) * K, j * K, &tpp, dk_tape(j - 1), b.ts));
            AHIP_TRY(hipEventRecord(b.ev_tape[(j - 1) & 1], b.ts));
        }
        if (j < nseg - 1) AHIP_TRY(hipStreamWaitEvent(st, b.ev_tape[j & 1], 0));
        const int s0 = j * K, s1 = std::min(M, s0 + K);
        const shg::TapeArgs tp = tape_of(j);
        aa.a0 = tp.a;
        aa.a2 = tp.a2;
        aa.dk = dk_tape(j);
        aa.s0 = s0;
        aa.cur = shg_internal_therm_step(c, s1 - 1);
        AHIP_TRY(shg::launch_therm_adj_row(n, b.dlam, c->tw, shg::ADJ_START, aa, st, sens));
        for (int s = s1 - 1; s >= s0; s--) {
            AHIP_TRY(shg::launch_adj_col(n, b.dlam, c->tw, s, tp, b.qxy, a_p, a_s, b.gH, st, sens));
            aa.cur = shg_internal_therm_step(c, s);
            if (s > s0) aa.prev = shg_internal_therm_step(c, s - 1);
            AHIP_TRY(shg::launch_therm_adj_row(n, b.dlam, c->tw, s > s0 ? shg::ADJ_MID : shg::ADJ_END, aa, st, sens));
        }
        if (j >= 2) AHIP_TRY(hipEventRecord(b.ev_free[j & 1], st));
    }
    if (sens) {
        AHIP_TRY(shg::launch_adj_reduce(nullptr, nullptr, 0, b.gH, cb, M, b.gs, st));
        AHIP_TRY(shg::launch_therm_dtmap(b.Ap, b.As, c->T[c->tcur], (int64_t)n3, c->p.crystal, (double)k.lp,
                                         (double)k.ls, b.dpart, kMapParts, st));
        AHIP_TRY(shg::launch_sum_partials(b.dpart, kMapParts, b.dsum, st));
    }
    ARC_TRY(enqueue_gradient_results(c, cs, cp, d_eta_d_fh != nullptr, st));
    const HostOffsets o = host_offsets(n2, M);
    if (sens) AHIP_TRY(hipMemcpyAsync(host_at<double>(b, o.dsum), b.dsum, sizeof(double), hipMemcpyDeviceToHost, st));
    AHIP_TRY(hipStreamSynchronize(st));
    float fwd_ms = 0.f, adj_ms = 0.f;
    AHIP_TRY(hipEventElapsedTime(&fwd_ms, c->e0, c->e1));
    AHIP_TRY(hipEventElapsedTime(&adj_ms, c->e1, b.ev_end));

    // 3. host assembly (FP64)
    const double J = cs * host_at<double2>(b, o.sums)->y / pin;
    copy_field_out(c, d_eta_d_fh);
    res->eta = J;
    res->forward_ms = fwd_ms;
    res->adjoint_ms = adj_ms;
    res->checkpoint_interval = K;
    res->segments = nseg;
    if (!sens) return SHG_OK;
    const double *gs = host_at<double>(b, o.gs);
    const double dT_map = *host_at<double>(b, o.dsum);
    const double d_focus = host_at<double2>(b, o.gsum)->x, dT_pump = host_at<double2>(b, o.gsum)->y * dnp;
    double dT_H = 0.0;
    for (int s = 0; s < M; s++) dT_H += gs[(size_t)s * 6 + 5];
    const double dT_powers = J * (dns / ns - dnp / np);
    b.map_centre_extra = dT_H + dT_powers;
    b.map_valid = true;

    res->d_eta_d_focus = d_focus;
    res->d_eta_d_temp_crystal = dT_map + dT_H + dT_powers;
    res->d_eta_d_temp = res->d_eta_d_temp_crystal + dT_pump;
    return SHG_OK;
}

// The plain adjoint; sens = false: the field gradient and eta only.
int plain_adjoint(shg_ctx *c, int checkpoint_interval, float *d_eta_d_fh, shg_adjoint_result *res, bool sens) {
    const int M = c->steps, n = c->n;
    const size_t n2 = (size_t)n * n;
    int K = checkpoint_interval > 0 ? checkpoint_interval : auto_interval(c, M, n2, false), nseg = 0;
    ARC_TRY(adjoint_buffers(c, M, checkpoint_interval <= 0, false, &K, &nseg));
    AdjointBuffers &b = c->adj;
    hipStream_t st = c->stream;
    const int rb = shg::adj_row_parts(n), cb = shg::adj_col_parts(n);

    // tape of segment j: slot j % 2, its start state is checkpoint j
    auto tape_of = [&](int j, bool write_ckpt) {
        const size_t off = (size_t)(nseg > 1 ? (j & 1) : 0) * K * n2;
        return shg::TapeArgs{b.ckpt + (size_t)j * n2, b.a2 + off, b.spec + off,
                             write_ckpt ? b.ckpt + (size_t)j * n2 : nullptr, j * K, 0};
    };

    // 1. forward: one uninterrupted march whose row passes at the segment boundaries
    //    leave the checkpoints (the state after step jK - 1), the last segment taped, the
    //    powers
    auto enqueue_forward = [&]() -> int {
        AHIP_TRY(shg::launch_copy_cells(c->U, c->U0, (int64_t)n2, st));
        if (nseg == 1) {
            ARC_TRY(enqueue_tape(c, tape_of(0, true), M, st));
        } else {
            const int rv = c->row_v < 0 ? 0 : c->row_v, cv = c->col_v < 0 ? 0 : c->col_v;
            const int s_tape = (nseg - 1) * K;             // first step of the taped segment
            AHIP_TRY(shg::launch_copy_cells(b.ckpt, c->U0, (int64_t)n2, st));
            AHIP_TRY(shg::launch_row_pass(n, c->d_desc, 1, c->tw, 0, shg::ROW_START, st, rv));
            for (int s = 0; s < s_tape; s++) {
                AHIP_TRY(shg::launch_col_pass(n, c->d_desc, 1, c->tw, st, cv));
                float4 *snap = (s + 1) % K == 0 ? b.ckpt + (size_t)((s + 1) / K) * n2 : nullptr;
                AHIP_TRY(shg::launch_row_pass(n, c->d_desc, 1, c->tw, s, shg::ROW_MID, st, rv, snap));
            }
            const shg::TapeArgs tp = tape_of(nseg - 1, false);
            for (int s = s_tape; s < M; s++) {
                AHIP_TRY(shg::launch_col_tape(n, c->d_desc, c->tw, s, tp, st));
                AHIP_TRY(shg::launch_row_tape(n, c->d_desc, c->tw, s, s < M - 1 ? shg::ROW_MID : shg::ROW_END, tp,
                                              st));
            }
        }
        AHIP_TRY(shg::launch_power(c->U, (int64_t)n2, c->part, c->nparts, c->red, st));
        return SHG_OK;
    };

    // 2. per segment, last first: the adjoint on `st`, the next segment's tape on `ts`
    const shg::phys::Consts k(c->p);
    const double T = c->p.temp;
    const double eps0 = 8.8541878128e-6, cl = 299.792458;
    const double np = shg::phys::index_d(c->p.crystal, (double)k.lp, T), ns = shg::phys::index_d(c->p.crystal, (double)k.ls, T);
    const double cp = 0.5 * eps0 * cl * np * (double)k.dx * (double)k.dy;
    const double cs = 0.5 * eps0 * cl * ns * (double)k.dx * (double)k.dy;
    const double dnp = dn_dT(c->p.crystal, (double)k.lp, T), dns = dn_dT(c->p.crystal, (double)k.ls, T);
    const double kp = 2 * M_PI * np / (double)k.lp, ks = 2 * M_PI * ns / (double)k.ls;
    const double dkp = 2 * M_PI * dnp / (double)k.lp, dks = 2 * M_PI * dns / (double)k.ls;
    const float a_p = (float)((double)k.dz * dkp / (kp * kp)), a_s = (float)((double)k.dz * dks / (ks * ks));
    auto enqueue_adjoint = [&]() -> int {
        if (nseg > 1) {                                     // forks `ts` off `st`
            AHIP_TRY(hipEventRecord(b.ev_u, st));
            AHIP_TRY(hipStreamWaitEvent(b.ts, b.ev_u, 0));
        }
        for (int j = nseg - 1; j >= 0; j--) {
            if (j >= 1) {
                // slot (j-1) % 2 was last read by the adjoint of segment j+1
                if (j + 1 <= nseg - 1) AHIP_TRY(hipStreamWaitEvent(b.ts, b.ev_free[(j - 1) & 1], 0));
                AHIP_TRY(shg::launch_copy_cells(c->U, b.ckpt + (size_t)(j - 1) * n2, (int64_t)n2, b.ts));
                ARC_TRY(enqueue_tape(c, tape_of(j - 1, false), j * K, b.ts));
                AHIP_TRY(hipEventRecord(b.ev_tape[(j - 1) & 1], b.ts));
            }
            if (j < nseg - 1) AHIP_TRY(hipStreamWaitEvent(st, b.ev_tape[j & 1], 0));   // joins `ts` at j = 0
            const int s0 = j * K, s1 = std::min(M, s0 + K);
            const shg::TapeArgs tp = tape_of(j, false);
            AHIP_TRY(shg::launch_adj_row(n, b.dlam, c->tw, s1 - 1, shg::ADJ_START, tp, b.g1, b.g2, st, sens));
            for (int s = s1 - 1; s >= s0; s--) {
                AHIP_TRY(shg::launch_adj_col(n, b.dlam, c->tw, s, tp, b.qxy, a_p, a_s, b.gH, st, sens));
                AHIP_TRY(shg::launch_adj_row(n, b.dlam, c->tw, s, s > s0 ? shg::ADJ_MID : shg::ADJ_END, tp, b.g1,
                                             b.g2, st, sens));
            }
            if (j >= 2) AHIP_TRY(hipEventRecord(b.ev_free[j & 1], st));
        }
        if (sens) AHIP_TRY(shg::launch_adj_reduce(b.g1, b.g2, rb, b.gH, cb, M, b.gs, st));
        return SHG_OK;
    };

    // The adjoint march is replayed from a graph; the forward is launched directly: its
    // ~800 launches replayed from a captured graph ran 0.45 ms slower than launched one by
    // one (measured at 256 x 256, nz 400), unlike the plain march's graph.
    hipGraphExec_t &g_adj = b.g_adj[sens ? 0 : 1];
    if (c->use_graph && !g_adj) ARC_TRY(shg_internal_capture(st, enqueue_adjoint, &g_adj));

    // forward, seed and adjoint back to back: the seed scale needs P_in only, and eta is
    // read back with the gradients, so the host waits once
    const double pin = c->p_fh_in;
    ARC_TRY(ensure_pump_fields(c, np, T));
    AHIP_TRY(hipEventRecord(c->e0, st));
    ARC_TRY(enqueue_forward());
    AHIP_TRY(hipEventRecord(c->e1, st));
    AHIP_TRY(shg::launch_adj_seed(c->U, b.lam, (int64_t)n2, (float)(2.0 * cs / pin), st));
    if (c->use_graph) {
        AHIP_TRY(hipGraphLaunch(g_adj, st));
    } else {
        ARC_TRY(enqueue_adjoint());
    }
    ARC_TRY(enqueue_gradient_results(c, cs, cp, d_eta_d_fh != nullptr, st));
    AHIP_TRY(hipStreamSynchronize(st));
    float fwd_ms = 0.f, adj_ms = 0.f;
    AHIP_TRY(hipEventElapsedTime(&fwd_ms, c->e0, c->e1));
    AHIP_TRY(hipEventElapsedTime(&adj_ms, c->e1, b.ev_end));

    // 3. gradients on the host (FP64)
    const HostOffsets o = host_offsets(n2, M);
    const double J = cs * host_at<double2>(b, o.sums)->y / pin;
    copy_field_out(c, d_eta_d_fh);
    res->eta = J;
    res->forward_ms = fwd_ms;
    res->adjoint_ms = adj_ms;
    res->checkpoint_interval = K;
    res->segments = nseg;
    if (!sens) return SHG_OK;
    const double *gs = host_at<double>(b, o.gs);
    const double d_focus = host_at<double2>(b, o.gsum)->x, dT_pump = host_at<double2>(b, o.gsum)->y * dnp;

    // temperature through the step coefficients and the TPA coefficient
    const double kap = (double)k.kappa_p((float)T), kas = (double)k.kappa_s((float)T);
    const double dkap = -kap * dnp / np, dkas = -kas * dns / ns;
    const double ddk = 2 * M_PI * (2 * dnp / (double)k.lp - dns / (double)k.ls);
    const double dtpa = 0.5 * (double)k.beta_s * 0.5 * (double)k.c * (double)k.eps0 * dns;
    const double dkf = (double)k.dk((float)T);
    double dT_coef = 0.0, dT_H = 0.0;
    float z = 0.0f;
    for (int s = 0; s < M; s++) {
        const double *x = &gs[(size_t)s * 6];
        const cd c1b(x[0], x[1]), c2b(x[2], x[3]);
        const double tb = x[4];
        const double zs = (double)z, th = (double)(float)(z * (float)dkf);
        const cd I(0, 1);
        const cd dc1 = I * std::exp(-I * th) * (dkap - I * kap * zs * ddk);
        const cd dc2 = I * std::exp(I * th) * (dkas + I * kas * zs * ddk);
        dT_coef += (std::conj(c1b) * dc1).real() + (std::conj(c2b) * dc2).real() + tb * dtpa;
        dT_H += x[5];
        z += k.dz;
    }
    const double dT_powers = J * (dns / ns - dnp / np);

    res->d_eta_d_focus = d_focus;
    res->d_eta_d_temp_crystal = dT_coef + dT_H + dT_powers;
    res->d_eta_d_temp = res->d_eta_d_temp_crystal + dT_pump;
    return SHG_OK;
}

}  // namespace

extern "C" {

int shg_adjoint(shg_ctx *c, int checkpoint_interval, float *d_eta_d_fh, shg_adjoint_result *res) {
    return shg_internal_guard(__func__, [&]() -> int {
        if (!c || !res) return shg_internal_fail(SHG_E_INVALID, "shg_adjoint: null argument");
        if (!c->have_fields) return shg_internal_fail(SHG_E_STATE, "shg_adjoint: call shg_set_fields first");
        return c->p.thermal ? thermal_adjoint(c, checkpoint_interval, d_eta_d_fh, res, true)
                            : plain_adjoint(c, checkpoint_interval, d_eta_d_fh, res, true);
    });
}

int shg_adjoint_field(shg_ctx *c, int checkpoint_interval, float *d_eta_d_fh, double *eta) {
    return shg_internal_guard(__func__, [&]() -> int {
        if (!c || !d_eta_d_fh) return shg_internal_fail(SHG_E_INVALID, "shg_adjoint_field: null argument");
        if (!c->have_fields) return shg_internal_fail(SHG_E_STATE, "shg_adjoint_field: call shg_set_fields first");
        shg_adjoint_result r{};
        const int rc = c->p.thermal ? thermal_adjoint(c, checkpoint_interval, d_eta_d_fh, &r, false)
                                    : plain_adjoint(c, checkpoint_interval, d_eta_d_fh, &r, false);
        if (rc == SHG_OK && eta) *eta = r.eta;
        return rc;
    });
}

int shg_adjoint_temperature_map(shg_ctx *c, double *map) {
    return shg_internal_guard(__func__, [&]() -> int {
        if (!c || !map) return shg_internal_fail(SHG_E_INVALID, "shg_adjoint_temperature_map: null argument");
        if (!c->p.thermal || !c->adj.map_valid)
            return shg_internal_fail(SHG_E_STATE, "shg_adjoint_temperature_map: run shg_adjoint on a THERMAL context first");
        const size_t n2 = (size_t)c->n * c->n;
        AHIP_TRY(shg_internal_copy(map, c->adj.Ap, n2 * c->hz * sizeof(double), hipMemcpyDeviceToHost, c->stream));
        map[(size_t)(c->n / 2) * c->n + c->n / 2] += c->adj.map_centre_extra;   // slice 0, grid centre
        return SHG_OK;
    });
}

}  // extern "C"
