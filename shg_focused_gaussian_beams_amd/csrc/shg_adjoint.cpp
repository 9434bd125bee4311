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
#include <string>
#include <vector>

#include "../../include/shg.h"
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

int ensure_buffers(shg_ctx *c, int K, int nseg) {
    AdjointBuffers &b = c->adj;
    const size_t n2 = (size_t)c->n * c->n;
    const int M = c->steps;
    if (!b.lam) {
        if (hipMalloc(&b.lam, n2 * sizeof(float4)) != hipSuccess ||
            hipMalloc(&b.g1, (size_t)M * shg::adj_row_parts(c->n) * 5 * sizeof(double)) != hipSuccess ||
            hipMalloc(&b.g2, (size_t)M * shg::adj_row_parts(c->n) * 5 * sizeof(double)) != hipSuccess ||
            hipMalloc(&b.gH, (size_t)M * shg::adj_col_parts(c->n) * sizeof(double)) != hipSuccess ||
            hipMalloc(&b.gs, (size_t)M * 6 * sizeof(double)) != hipSuccess ||
            hipMalloc(&b.qxy, (size_t)c->n * sizeof(float2)) != hipSuccess ||
            hipMalloc(&b.dlam, sizeof(shg::InstDesc)) != hipSuccess)
            return shg_internal_fail(SHG_E_NOMEM, "shg_adjoint: device allocation failed");
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
        AHIP_TRY(hipMemcpy(b.qxy, q.data(), q.size() * sizeof(float), hipMemcpyHostToDevice));
        shg::InstDesc d{};
        AHIP_TRY(hipMemcpy(&d, c->d_desc, sizeof(d), hipMemcpyDeviceToHost));
        d.U = b.lam;
        AHIP_TRY(hipMemcpy(b.dlam, &d, sizeof(d), hipMemcpyHostToDevice));
    }
    if (b.k != K || b.nseg != nseg) {
        for (hipGraphExec_t *g : {&b.g_fwd, &b.g_adj})
            if (*g) { (void)hipGraphExecDestroy(*g); *g = nullptr; }
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

// d eta / d A0 = lambda_0 (FH part) + the d(1/P_in)/dA0 term, written to d_eta_d_fh (if
// set), and its projections on the derivative fields of the reference pump (setPump,
// src/headers/Efields.h:74-104, differentiated analytically at temperature T): the
// focus sensitivity and the temperature sensitivity through n_p(T) in Ap0 and zR.
int field_and_pump_gradients(shg_ctx *c, double J, double pin, double cp, double np, double dnp, double T,
                             float *d_eta_d_fh, double *d_focus, double *dT_pump) {
    AdjointBuffers &b = c->adj;
    const int n = c->n;
    const size_t n2 = (size_t)n * n;
    std::vector<float> lam(4 * n2), u0(4 * n2);
    AHIP_TRY(hipMemcpy(lam.data(), b.lam, n2 * sizeof(float4), hipMemcpyDeviceToHost));
    AHIP_TRY(hipMemcpy(u0.data(), c->U0, n2 * sizeof(float4), hipMemcpyDeviceToHost));
    std::vector<cd> grad(n2);
    const double direct = -J / pin * 2.0 * cp;
    for (size_t i = 0; i < n2; i++) grad[i] = cd(lam[4 * i], lam[4 * i + 1]) + direct * cd(u0[4 * i], u0[4 * i + 1]);
    if (d_eta_d_fh)
        for (size_t i = 0; i < n2; i++) {
            d_eta_d_fh[2 * i] = (float)grad[i].real();
            d_eta_d_fh[2 * i + 1] = (float)grad[i].imag();
        }
    const shg::phys::Consts k(c->p);
    const double eps0 = 8.8541878128e-6, cl = 299.792458;
    const double w0 = c->p.waist, f = c->p.focus, lam0 = (double)k.lp;
    const double zr = M_PI * np * w0 * w0 / lam0;
    const double ap0 = std::sqrt(4.0 * c->p.power / (eps0 * cl * M_PI * np * w0 * w0));
    const cd Mq(1.0, -f / zr);
    // the derivative fields depend on the context's parameters and T only: built once
    std::vector<cd> &pdf = b.pump_df, &pdn = b.pump_dn;
    if (pdf.size() != n2 || b.pump_T != T) {
        pdf.resize(n2);
        pdn.resize(n2);
        for (int iy = 0; iy < n; iy++)
            for (int ix = 0; ix < n; ix++) {
                const double xx = (ix - 0.5 * n) * (double)k.dx, yy = (iy - 0.5 * n) * (double)k.dy;
                const double r2 = xx * xx + yy * yy;
                const cd A = ap0 / Mq * std::exp(-r2 / (w0 * w0 * Mq));
                const cd dA_dM = A * (-1.0 / Mq + r2 / (w0 * w0 * Mq * Mq));
                pdf[(size_t)iy * n + ix] = dA_dM * cd(0.0, -1.0 / zr);
                pdn[(size_t)iy * n + ix] = -A / (2.0 * np) + dA_dM * cd(0.0, f / (zr * np));
            }
        b.pump_T = T;
    }
    double df = 0.0, dt = 0.0;
    for (size_t i = 0; i < n2; i++) {
        df += grad[i].real() * pdf[i].real() + grad[i].imag() * pdf[i].imag();
        dt += (grad[i].real() * pdn[i].real() + grad[i].imag() * pdn[i].imag()) * dnp;
    }
    *d_focus = df;
    *dT_pump = dt;
    return SHG_OK;
}

// ----- THERMAL, temperature held fixed -------------------------------------------
int ensure_thermal_buffers(shg_ctx *c, int K, int nseg) {
    AdjointBuffers &b = c->adj;
    const size_t n2 = (size_t)c->n * c->n, n3 = n2 * c->hz;
    if (!b.Ap) {
        if (hipMalloc(&b.Ap, n3 * sizeof(double)) != hipSuccess || hipMalloc(&b.As, n3 * sizeof(double)) != hipSuccess ||
            hipMalloc(&b.Gc, n2 * sizeof(double)) != hipSuccess || hipMalloc(&b.Dc, n2 * sizeof(double)) != hipSuccess ||
            hipMalloc(&b.dpart, kMapParts * sizeof(double)) != hipSuccess || hipMalloc(&b.dsum, sizeof(double)) != hipSuccess)
            return shg_internal_fail(SHG_E_NOMEM, "shg_adjoint: device allocation failed");
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
    a.ct = c->ctab;
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
int thermal_adjoint(shg_ctx *c, int checkpoint_interval, float *d_eta_d_fh, shg_adjoint_result *res) {
    const int M = c->p.nz - 1, n = c->n;
    const size_t n2 = (size_t)n * n, n3 = n2 * c->hz;
    int K = checkpoint_interval > 0 ? checkpoint_interval : (int)std::ceil(std::sqrt((double)M));
    K = std::max(1, std::min(K, M));
    const int nseg = (M + K - 1) / K;
    ARC_TRY(ensure_buffers(c, K, nseg));
    ARC_TRY(ensure_thermal_buffers(c, K, nseg));
    AdjointBuffers &b = c->adj;
    b.map_valid = false;
    hipStream_t st = c->stream;
    const int cb = shg::adj_col_parts(n);
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
    AHIP_TRY(shg::launch_index_table(c->T[c->tcur], c->ntab, c->ctab, (int64_t)n3, c->tc, st));
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
    double2 sums;
    AHIP_TRY(hipMemcpyAsync(&sums, c->red, sizeof(double2), hipMemcpyDeviceToHost, st));
    AHIP_TRY(hipStreamSynchronize(st));
    float fwd_ms = 0.f;
    AHIP_TRY(hipEventElapsedTime(&fwd_ms, c->e0, c->e1));

    const shg::phys::Consts k(c->p);
    const double T = c->t_centre;                              // propagators, powers, pump
    const double eps0 = 8.8541878128e-6, cl = 299.792458;
    const double np = shg::phys::index_d(c->p.crystal, (double)k.lp, T), ns = shg::phys::index_d(c->p.crystal, (double)k.ls, T);
    const double cp = 0.5 * eps0 * cl * np * (double)k.dx * (double)k.dy;
    const double cs = 0.5 * eps0 * cl * ns * (double)k.dx * (double)k.dy;
    const double pin = c->p_fh_in, J = cs * sums.y / pin;
    const double dnp = dn_dT(c->p.crystal, (double)k.lp, T), dns = dn_dT(c->p.crystal, (double)k.ls, T);
    const double kp = 2 * M_PI * np / (double)k.lp, ks = 2 * M_PI * ns / (double)k.ls;
    const double dkp = 2 * M_PI * dnp / (double)k.lp, dks = 2 * M_PI * dns / (double)k.ls;
    const float a_p = (float)((double)k.dz * dkp / (kp * kp)), a_s = (float)((double)k.dz * dks / (ks * ks));

    // 2. seed, then per segment, last first: the adjoint on `st`, the previous segment's
    //    taped recomputation on `ts`
    AHIP_TRY(hipEventRecord(c->e0, st));
    AHIP_TRY(shg::launch_adj_seed(c->U, b.lam, (int64_t)n2, (float)(2.0 * cs / pin), st));
    for (double *ptr : {b.Ap, b.As})
        AHIP_TRY(shg::launch_fill_zero(ptr, (int64_t)n3, st));
    for (double *ptr : {b.Gc, b.Dc})
        AHIP_TRY(shg::launch_fill_zero(ptr, (int64_t)n2, st));
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
        AHIP_TRY(shg::launch_therm_adj_row(n, b.dlam, c->tw, shg::ADJ_START, aa, st));
        for (int s = s1 - 1; s >= s0; s--) {
            AHIP_TRY(shg::launch_adj_col(n, b.dlam, c->tw, s, tp, b.qxy, a_p, a_s, b.gH, st));
            aa.cur = shg_internal_therm_step(c, s);
            if (s > s0) aa.prev = shg_internal_therm_step(c, s - 1);
            AHIP_TRY(shg::launch_therm_adj_row(n, b.dlam, c->tw, s > s0 ? shg::ADJ_MID : shg::ADJ_END, aa, st));
        }
        if (j >= 2) AHIP_TRY(hipEventRecord(b.ev_free[j & 1], st));
    }
    AHIP_TRY(shg::launch_adj_reduce(nullptr, nullptr, 0, b.gH, cb, M, b.gs, st));
    AHIP_TRY(shg::launch_therm_dtmap(b.Ap, b.As, c->T[c->tcur], (int64_t)n3, c->p.crystal, (double)k.lp,
                                     (double)k.ls, b.dpart, kMapParts, st));
    AHIP_TRY(shg::launch_sum_partials(b.dpart, kMapParts, b.dsum, st));
    AHIP_TRY(hipEventRecord(c->e1, st));
    AHIP_TRY(hipEventSynchronize(c->e1));
    float adj_ms = 0.f;
    AHIP_TRY(hipEventElapsedTime(&adj_ms, c->e0, c->e1));

    // 3. host assembly (FP64)
    std::vector<double> gs((size_t)M * 6);
    double dT_map = 0.0;
    AHIP_TRY(hipMemcpy(gs.data(), b.gs, gs.size() * sizeof(double), hipMemcpyDeviceToHost));
    AHIP_TRY(hipMemcpy(&dT_map, b.dsum, sizeof(double), hipMemcpyDeviceToHost));
    double d_focus = 0.0, dT_pump = 0.0;
    ARC_TRY(field_and_pump_gradients(c, J, pin, cp, np, dnp, T, d_eta_d_fh, &d_focus, &dT_pump));
    double dT_H = 0.0;
    for (int s = 0; s < M; s++) dT_H += gs[(size_t)s * 6 + 5];
    const double dT_powers = J * (dns / ns - dnp / np);
    b.map_centre_extra = dT_H + dT_powers;
    b.map_valid = true;

    res->eta = J;
    res->d_eta_d_focus = d_focus;
    res->d_eta_d_temp_crystal = dT_map + dT_H + dT_powers;
    res->d_eta_d_temp = res->d_eta_d_temp_crystal + dT_pump;
    res->forward_ms = fwd_ms;
    res->adjoint_ms = adj_ms;
    res->checkpoint_interval = K;
    res->segments = nseg;
    return SHG_OK;
}

}  // namespace

extern "C" {

int shg_adjoint(shg_ctx *c, int checkpoint_interval, float *d_eta_d_fh, shg_adjoint_result *res) {
    if (!c || !res) return shg_internal_fail(SHG_E_INVALID, "shg_adjoint: null argument");
    if (!c->have_fields) return shg_internal_fail(SHG_E_STATE, "shg_adjoint: call shg_set_fields first");
    if (c->p.thermal) return thermal_adjoint(c, checkpoint_interval, d_eta_d_fh, res);
    const int M = c->steps, n = c->n;
    const size_t n2 = (size_t)n * n;
    int K = checkpoint_interval > 0 ? checkpoint_interval : (int)std::ceil(std::sqrt((double)M));
    K = std::max(1, std::min(K, M));
    const int nseg = (M + K - 1) / K;
    ARC_TRY(ensure_buffers(c, K, nseg));
    AdjointBuffers &b = c->adj;
    hipStream_t st = c->stream;
    const int rb = shg::adj_row_parts(n), cb = shg::adj_col_parts(n);

    // tape of segment j: slot j % 2, its start state is checkpoint j
    auto tape_of = [&](int j, bool write_ckpt) {
        const size_t off = (size_t)(nseg > 1 ? (j & 1) : 0) * K * n2;
        return shg::TapeArgs{b.ckpt + (size_t)j * n2, b.a2 + off, b.spec + off,
                             write_ckpt ? b.ckpt + (size_t)j * n2 : nullptr, j * K, 0};
    };

    // 1. forward: checkpoints of segments 0 .. nseg-2, the last segment taped, the powers
    auto enqueue_forward = [&]() -> int {
        AHIP_TRY(shg::launch_copy_cells(c->U, c->U0, (int64_t)n2, st));
        for (int j = 0; j < nseg - 1; j++) {
            AHIP_TRY(shg::launch_copy_cells(b.ckpt + (size_t)j * n2, c->U, (int64_t)n2, st));
            ARC_TRY(shg_internal_enqueue_march_plain(c, j * K, K));
        }
        ARC_TRY(enqueue_tape(c, tape_of(nseg - 1, true), M, st));
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
            AHIP_TRY(shg::launch_adj_row(n, b.dlam, c->tw, s1 - 1, shg::ADJ_START, tp, b.g1, b.g2, st));
            for (int s = s1 - 1; s >= s0; s--) {
                AHIP_TRY(shg::launch_adj_col(n, b.dlam, c->tw, s, tp, b.qxy, a_p, a_s, b.gH, st));
                AHIP_TRY(shg::launch_adj_row(n, b.dlam, c->tw, s, s > s0 ? shg::ADJ_MID : shg::ADJ_END, tp, b.g1,
                                             b.g2, st));
            }
            if (j >= 2) AHIP_TRY(hipEventRecord(b.ev_free[j & 1], st));
        }
        AHIP_TRY(shg::launch_adj_reduce(b.g1, b.g2, rb, b.gH, cb, M, b.gs, st));
        return SHG_OK;
    };

    const int rv = c->row_v < 0 ? 0 : c->row_v, cv = c->col_v < 0 ? 0 : c->col_v;
    if (c->use_graph && (b.g_row_v != rv || b.g_col_v != cv))
        for (hipGraphExec_t *g : {&b.g_fwd, &b.g_adj})
            if (*g) { (void)hipGraphExecDestroy(*g); *g = nullptr; }
    if (c->use_graph && !b.g_fwd) {
        ARC_TRY(shg_internal_capture(st, enqueue_forward, &b.g_fwd));
        ARC_TRY(shg_internal_capture(st, enqueue_adjoint, &b.g_adj));
        b.g_row_v = rv;
        b.g_col_v = cv;
    }

    AHIP_TRY(hipEventRecord(c->e0, st));
    if (c->use_graph) {
        AHIP_TRY(hipGraphLaunch(b.g_fwd, st));
    } else {
        ARC_TRY(enqueue_forward());
    }
    AHIP_TRY(hipEventRecord(c->e1, st));
    double2 sums;
    AHIP_TRY(hipMemcpyAsync(&sums, c->red, sizeof(double2), hipMemcpyDeviceToHost, st));
    AHIP_TRY(hipStreamSynchronize(st));
    float fwd_ms = 0.f;
    AHIP_TRY(hipEventElapsedTime(&fwd_ms, c->e0, c->e1));
    const double pin = c->p_fh_in, J = cs * sums.y / pin;

    AHIP_TRY(hipEventRecord(c->e0, st));
    AHIP_TRY(shg::launch_adj_seed(c->U, b.lam, (int64_t)n2, (float)(2.0 * cs / pin), st));
    if (c->use_graph) {
        AHIP_TRY(hipGraphLaunch(b.g_adj, st));
    } else {
        ARC_TRY(enqueue_adjoint());
    }
    AHIP_TRY(hipEventRecord(c->e1, st));
    AHIP_TRY(hipEventSynchronize(c->e1));
    float adj_ms = 0.f;
    AHIP_TRY(hipEventElapsedTime(&adj_ms, c->e0, c->e1));

    // 3. gradients on the host (FP64)
    std::vector<double> gs((size_t)M * 6);
    AHIP_TRY(hipMemcpy(gs.data(), b.gs, gs.size() * sizeof(double), hipMemcpyDeviceToHost));
    double d_focus = 0.0, dT_pump = 0.0;
    ARC_TRY(field_and_pump_gradients(c, J, pin, cp, np, dnp, T, d_eta_d_fh, &d_focus, &dT_pump));

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

    res->eta = J;
    res->d_eta_d_focus = d_focus;
    res->d_eta_d_temp_crystal = dT_coef + dT_H + dT_powers;
    res->d_eta_d_temp = res->d_eta_d_temp_crystal + dT_pump;
    res->forward_ms = fwd_ms;
    res->adjoint_ms = adj_ms;
    res->checkpoint_interval = K;
    res->segments = nseg;
    return SHG_OK;
}

int shg_adjoint_temperature_map(shg_ctx *c, double *map) {
    if (!c || !map) return shg_internal_fail(SHG_E_INVALID, "shg_adjoint_temperature_map: null argument");
    if (!c->p.thermal || !c->adj.map_valid)
        return shg_internal_fail(SHG_E_STATE, "shg_adjoint_temperature_map: run shg_adjoint on a THERMAL context first");
    const size_t n2 = (size_t)c->n * c->n;
    AHIP_TRY(hipMemcpy(map, c->adj.Ap, n2 * c->hz * sizeof(double), hipMemcpyDeviceToHost));
    map[(size_t)(c->n / 2) * c->n + c->n / 2] += c->adj.map_centre_extra;   // slice 0, grid centre
    return SHG_OK;
}

}  // extern "C"
