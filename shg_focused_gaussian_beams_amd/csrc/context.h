// context.h — the solver context behind the opaque `shg_ctx` of include/shg.h (private
// to the library's translation units).
#pragma once
#include <hip/hip_runtime.h>

#include <complex>
#include <functional>
#include <string>
#include <vector>

#include "../../include/shg.h"
#include "ssfm_kernels.h"
#include "thermal_kernels.h"

// Device buffers of the adjoint (allocated on first use, kept with the context).
struct AdjointBuffers {
    bool base_ready = false;                   // the one-time buffers below are all set up
    bool th_ready = false;                     // the same for the THERMAL buffers (Ap .. dsum)
    int k = 0, nseg = 0;                       // checkpoint interval and segment count they were sized for
    float4 *lam = nullptr;                     // adjoint state
    float4 *ckpt = nullptr;                    // nseg checkpoints (segment-start states)
    float4 *a2 = nullptr, *spec = nullptr;     // two segment tapes of K steps each (double buffer)
    double *g1 = nullptr, *g2 = nullptr, *gH = nullptr;    // FP64 parameter-sensitivity partials
    double *gs = nullptr;                      // their per-step sums (6 per step)
    float2 *qxy = nullptr;                     // 2 pi^2 fx^2, 2 pi^2 fy^2 per index
    shg::InstDesc *dlam = nullptr;             // instance table whose U is `lam`
    hipStream_t ts = nullptr;                  // stream of the tape recomputation
    hipEvent_t ev_u = nullptr;                 // the forward state buffer is free for `ts`
    hipEvent_t ev_tape[2] = {nullptr, nullptr};  // tape slot k written
    hipEvent_t ev_free[2] = {nullptr, nullptr};  // tape slot k read by the adjoint
    hipGraphExec_t g_adj[2] = {nullptr, nullptr};   // captured adjoint march for interval k:
                                                    // [0] with the sensitivities, [1] lambda only
    double2 *pump_df = nullptr, *pump_dn = nullptr;  // d(pump)/d(focus), d(pump)/d(n_p) per cell (device)
    double pump_T = -1e300;                    // temperature the pump derivative fields were built at
    float2 *grad = nullptr;                    // d eta / d FH(z = 0) per cell
    double2 *gpart = nullptr, *gsum = nullptr; // block sums of its pump projections and their total
    hipEvent_t ev_end = nullptr;               // end of the adjoint (timing)
    // pinned host staging of the results: the gradient field, then the FP64 scalars
    void *host = nullptr;
    size_t host_bytes = 0;
    // THERMAL (temperature held fixed)
    double *Ap = nullptr, *As = nullptr;       // per heat voxel dJ/dn_p, dJ/dn_s; Ap then holds dJ/dT
    double *Gc = nullptr, *Dc = nullptr;       // per cell DKint suffix sum and pending part
    double *dpart = nullptr, *dsum = nullptr;  // block sums of the dJ/dT map and their total
    float *dkt = nullptr;                      // two tapes of DKint per step (K steps each)
    float *dkc = nullptr;                      // DKint at every checkpoint (step s0 - 1)
    int th_k = 0, th_nseg = 0;                 // interval and segment count the THERMAL buffers hold
    bool map_valid = false;
    double map_centre_extra = 0.0;             // propagator and power terms, at the grid-centre voxel
};

struct shg_ctx {
    shg_params p{};
    int n = 0;
    int steps = 0;                   // z-steps executed by shg_forward (plain nz-1, THERMAL nz)
    int hz = 0;                      // heat-grid slices (THERMAL)
    hipStream_t stream = nullptr;
    float4 *U = nullptr, *U0 = nullptr, *Uout = nullptr;
    float2 *tw = nullptr, *G = nullptr;
    double2 *part = nullptr, *red = nullptr;
    int nparts = 0;
    std::vector<float> coef;         // plain: (c1, c2) per step, 4 floats
    float4 *d_coef = nullptr;        // the same on the device
    shg::InstDesc *d_desc = nullptr; // one-entry instance table for the passes
    shg::NlConst nl{};
    // THERMAL state
    float *T[2] = {nullptr, nullptr};
    int tcur = 0;                    // index of the buffer holding the latest temperature
    float *Q = nullptr, *DKint = nullptr;
    float2 *ntab = nullptr;          // {n_p(T), n_s(T)} per heat voxel, refreshed per march
    double *jpart = nullptr, *jres = nullptr;
    int njpart = 0;
    shg::JacobiConst jc{};
    shg::ThermConst tc{};
    float t_centre = 0.f;            // temperature used for the propagators and powers
    bool have_fields = false;
    double p_fh_in = 0.0;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    double last_ms = 0.0;
    hipGraphExec_t gexec = nullptr;
    int use_graph = 1;
    int row_v = -1, col_v = -1;      // plain-pass layout variants (-1: not tuned yet)
    int jac_v = -1;                  // Jacobi pair layout (launch_jacobi2 variant, or two single sweeps)
    shg::Jacobi2Plan jplan[shg::kJacobi2VariantCount] = {};   // tiles and z-chunks of each two-sweep layout
    std::vector<float> host_stage;   // 4 floats per cell
    shg_thermal_log_fn log_fn = nullptr;   // progress events of shg_thermal_run
    void *log_user = nullptr;
    AdjointBuffers adj;
};

// Shared helpers of the library's translation units (defined in shg_capi.cpp).
int shg_internal_fail(int code, const std::string &msg);
shg::ThermStep shg_internal_therm_step(const shg_ctx *c, int s);
// The step-independent constants of a context's passes (THERMAL ones only if p.thermal)
// and the descriptor of THERMAL field step s on a heat grid of hz slices.
void shg_internal_thermal_consts(const shg_params &p, int hz, shg::NlConst *nl, shg::JacobiConst *jc,
                                 shg::ThermConst *tc);
shg::ThermStep shg_internal_therm_step_of(const shg_params &p, int hz, int s);
int shg_internal_reduce_powers(shg_ctx *c, const float4 *U, double *sum_fh, double *sum_sh);
double shg_internal_power_of(const shg_ctx *c, int which, double sum, double temp);
// A copy ordered on `st` that has finished on return.  The library copies through its own
// (non-blocking) streams only, never the legacy null stream: a legacy-stream copy issued by
// one host thread invalidates a graph capture running in another, so contexts driven from
// several threads at once (one per point of a sweep) must not use it.
hipError_t shg_internal_copy(void *dst, const void *src, size_t bytes, hipMemcpyKind kind, hipStream_t st);
// Captures the work `enqueue` issues on `st` (and on streams forked from it through
// events) into an executable graph.
int shg_internal_capture(hipStream_t st, const std::function<int()> &enqueue, hipGraphExec_t *out);
// Average time in ms of `one` (a sequence of launches on `st`), replayed from a graph of
// `reps` repetitions: the fastest of three replays after a warm-up replay.
int shg_internal_time_graph(hipStream_t st, hipEvent_t e0, hipEvent_t e1, int reps, const std::function<int()> &one,
                            float *ms);
