// shg_batch.cpp — batched parameter sweeps (C ABI in include/shg.h).
//
// The reference runs a sweep as one serial process per parameter point (src/SHG.sh:39-81).
// Here every instance of a sweep marches in one device schedule: per z-step, one column-
// pass and one row-pass launch per transverse size over a table of instance descriptors.
// Each size's table is sorted by step count, longest first, so the instances still
// marching at step s are a prefix and a launch covers exactly them.  The whole schedule
// is captured once in a hipGraph and replayed by shg_batch_run.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <numeric>
#include <string>
#include <vector>

#include "../../include/shg.h"
#include "guard.h"
#include "batch_kernels.h"
#include "physics.h"
#include "ssfm_kernels.h"

// Error reporting shared with shg_capi.cpp.
int shg_internal_fail(int code, const std::string &msg);
hipError_t shg_internal_copy(void *dst, const void *src, size_t bytes, hipMemcpyKind kind, hipStream_t st);
int shg_internal_capture(hipStream_t st, const std::function<int()> &enqueue, hipGraphExec_t *out);
int shg_internal_time_graph(hipStream_t st, hipEvent_t e0, hipEvent_t e1, int reps, const std::function<int()> &one,
                            float *ms);

namespace {

#define BHIP_TRY(expr)                                                                                          \
    do {                                                                                                        \
        hipError_t e_ = (expr);                                                                                 \
        if (e_ != hipSuccess) return shg_internal_fail(SHG_E_DEVICE, std::string(#expr ": ") + hipGetErrorString(e_)); \
    } while (0)

const int kSizes[4] = {64, 128, 256, 512};
constexpr int kMaxStreams = 4, kDefaultGroupMiB = 128, kDefaultStreams = 3;

int size_class(int n) {
    for (int c = 0; c < 4; c++)
        if (kSizes[c] == n) return c;
    return -1;
}

}  // namespace

struct shg_batch {
    int count = 0;
    std::vector<shg_params> p;             // caller order
    std::vector<int> order;                // slot -> caller index (slots sorted by size, steps desc)
    std::vector<int> slot_of;              // caller index -> slot
    std::vector<int> n, steps;             // per slot
    std::vector<size_t> cell_off;          // per slot, in cells
    int cls_off[4] = {0, 0, 0, 0}, cls_cnt[4] = {0, 0, 0, 0};
    std::vector<int> cls_steps[4];         // steps of the slots of each class (descending)
    size_t total_cells = 0;
    int max_steps = 0;
    hipStream_t stream = nullptr;
    float4 *U = nullptr, *U0 = nullptr;
    float2 *G = nullptr, *tw[4] = {nullptr, nullptr, nullptr, nullptr};
    float4 *coef = nullptr;
    shg::InstDesc *desc = nullptr;
    int *ncell = nullptr;
    double2 *pow = nullptr;
    std::vector<double> p_in;              // per slot, FP64 input pump power
    std::vector<float> t_slot;             // per slot, crystal temperature
    bool have_fields = false;
    hipGraphExec_t gexec = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    double last_ms = 0.0;
    int row_v[4] = {-1, -1, -1, -1}, col_v[4] = {-1, -1, -1, -1};   // layout variant per size class
    // schedule: each size class is cut into groups of at most group_mb MiB of fields; a
    // group marches all its steps before the next group of its stream starts, so its
    // fields stay in the last-level cache between the passes; groups are spread over
    // nstreams streams (longest first onto the least-loaded one), which overlaps the
    // latency-bound small grids with the large ones.
    struct Group {
        int c, first, count, steps;
        double work;
        int stream;
    };
    int group_mb = kDefaultGroupMiB, nstreams = kDefaultStreams;
    std::vector<Group> groups;
    hipStream_t xs[kMaxStreams - 1] = {};
    hipEvent_t ev_fork = nullptr, ev_join[kMaxStreams - 1] = {};
};

namespace {

void batch_release(shg_batch *b) {
    if (!b) return;
    // work an error return left in flight ends before its graph and buffers go
    for (hipStream_t x : b->xs)
        if (x) (void)hipStreamSynchronize(x);
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    if (b->gexec) (void)hipGraphExecDestroy(b->gexec);
    for (hipStream_t x : b->xs)
        if (x) (void)hipStreamDestroy(x);
    for (hipEvent_t e : b->ev_join)
        if (e) (void)hipEventDestroy(e);
    if (b->ev_fork) (void)hipEventDestroy(b->ev_fork);
    for (void *ptr : {(void *)b->U, (void *)b->U0, (void *)b->G, (void *)b->coef, (void *)b->desc, (void *)b->ncell,
                      (void *)b->pow, (void *)b->tw[0], (void *)b->tw[1], (void *)b->tw[2], (void *)b->tw[3]})
        if (ptr) (void)hipFree(ptr);
    if (b->e0) (void)hipEventDestroy(b->e0);
    if (b->e1) (void)hipEventDestroy(b->e1);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
}

double slot_power(const shg_batch *b, int slot, int which, double sum) {
    const shg_params &p = b->p[b->order[slot]];
    const shg::phys::Consts k(p);
    const double lam = which == 0 ? (double)k.lp : (double)k.ls;
    const double n = shg::phys::index_d(p.crystal, lam, p.temp);
    return 0.5 * 8.8541878128e-6 * 299.792458 * n * sum * (double)k.dx * (double)k.dy;
}

int batch_sums(shg_batch *b, std::vector<double2> &sums) {
    BHIP_TRY(shg::launch_power_batch(b->desc, b->ncell, b->count, b->pow, b->stream));
    sums.resize(b->count);
    BHIP_TRY(hipMemcpyAsync(sums.data(), b->pow, b->count * sizeof(double2), hipMemcpyDeviceToHost, b->stream));
    BHIP_TRY(hipStreamSynchronize(b->stream));
    return SHG_OK;
}

int after_fields_set(shg_batch *b) {
    BHIP_TRY(hipMemcpyAsync(b->U, b->U0, b->total_cells * sizeof(float4), hipMemcpyDeviceToDevice, b->stream));
    std::vector<double2> sums;
    if (int rc = batch_sums(b, sums)) return rc;
    b->p_in.resize(b->count);
    for (int s = 0; s < b->count; s++) {
        if (!std::isfinite(sums[s].x) || !std::isfinite(sums[s].y))
            return shg_internal_fail(SHG_E_NONFINITE, "shg_batch: non-finite input field");
        b->p_in[s] = slot_power(b, s, 0, sums[s].x);
    }
    b->have_fields = true;
    return SHG_OK;
}

// Per size class, times every (row layout, column layout) pair as the schedule runs them
// (a row pass then a column pass over a group of the class, replayed from a graph) and
// keeps the fastest pair (clobbers U; shg_batch_run restores it from U0).
int tune_batch(shg_batch *b, int reps) {
    for (int c = 0; c < 4; c++) {
        if (!b->cls_cnt[c]) continue;
        int cnt = 1;
        for (const auto &gr : b->groups)
            if (gr.c == c) cnt = std::max(cnt, gr.count);
        const int n = kSizes[c];
        const shg::InstDesc *d = b->desc + b->cls_off[c];
        float best = 1e30f;
        for (int rv = 0; rv < shg::kRowVariantCount; rv++) {
            for (int cv = 0; cv < shg::kColVariantCount; cv++) {
                float ms = 0.f;
                const int rc = shg_internal_time_graph(b->stream, b->e0, b->e1, reps, [&] {
                    hipError_t e = shg::launch_row_pass(n, d, cnt, b->tw[c], 0, shg::ROW_AUTO, b->stream, rv);
                    if (e == hipSuccess) e = shg::launch_col_pass(n, d, cnt, b->tw[c], b->stream, cv);
                    return e == hipSuccess ? (int)SHG_OK
                                           : shg_internal_fail(SHG_E_DEVICE, std::string("tuning pass: ") +
                                                                                 hipGetErrorString(e));
                }, &ms);
                if (rc) return rc;
                if (ms < best) { best = ms; b->row_v[c] = rv; b->col_v[c] = cv; }
            }
        }
    }
    return SHG_OK;
}

// Cuts every size class into groups (slots of a class are sorted by steps, descending)
// and assigns them to streams.
void build_groups(shg_batch *b) {
    b->groups.clear();
    const double cap = (double)b->group_mb * (1 << 20);
    for (int c = 0; c < 4; c++) {
        if (!b->cls_cnt[c]) continue;
        const double inst = (double)kSizes[c] * kSizes[c] * sizeof(float4);
        const int g = std::max(1, (int)(cap / inst));
        for (int first = 0; first < b->cls_cnt[c]; first += g) {
            shg_batch::Group gr{c, first, std::min(g, b->cls_cnt[c] - first), b->cls_steps[c][first], 0.0, 0};
            for (int i = first; i < first + gr.count; i++) gr.work += (double)kSizes[c] * kSizes[c] * b->cls_steps[c][i];
            b->groups.push_back(gr);
        }
    }
    std::stable_sort(b->groups.begin(), b->groups.end(),
                     [](const shg_batch::Group &x, const shg_batch::Group &y) { return x.work > y.work; });
    std::vector<double> load(b->nstreams, 0.0);
    for (auto &gr : b->groups) {
        gr.stream = (int)(std::min_element(load.begin(), load.end()) - load.begin());
        load[gr.stream] += gr.work;
    }
}

// One group's march: START row pass, then per step a column pass and an AUTO row pass
// (MID while more steps follow, END on an instance's last step) over the group's
// instances still marching.
int enqueue_group(shg_batch *b, const shg_batch::Group &gr, hipStream_t st) {
    const int n = kSizes[gr.c], rv = b->row_v[gr.c] < 0 ? 0 : b->row_v[gr.c], cv = b->col_v[gr.c] < 0 ? 0 : b->col_v[gr.c];
    const shg::InstDesc *d = b->desc + b->cls_off[gr.c] + gr.first;
    const int *steps = b->cls_steps[gr.c].data() + gr.first;   // descending
    BHIP_TRY(shg::launch_row_pass(n, d, gr.count, b->tw[gr.c], 0, shg::ROW_START, st, rv));
    int active = gr.count;
    for (int s = 0; s < gr.steps; s++) {
        while (active > 0 && steps[active - 1] <= s) active--;
        BHIP_TRY(shg::launch_col_pass(n, d, active, b->tw[gr.c], st, cv));
        BHIP_TRY(shg::launch_row_pass(n, d, active, b->tw[gr.c], s, shg::ROW_AUTO, st, rv));
    }
    return SHG_OK;
}

// The whole schedule on b->stream, forking the other streams off it and joining them.
int enqueue_batch(shg_batch *b) {
    const int S = b->nstreams;
    if (S > 1) {
        BHIP_TRY(hipEventRecord(b->ev_fork, b->stream));
        for (int k = 1; k < S; k++) BHIP_TRY(hipStreamWaitEvent(b->xs[k - 1], b->ev_fork, 0));
    }
    for (const auto &gr : b->groups)
        if (int rc = enqueue_group(b, gr, gr.stream == 0 ? b->stream : b->xs[gr.stream - 1])) return rc;
    for (int k = 1; k < S; k++) {
        BHIP_TRY(hipEventRecord(b->ev_join[k - 1], b->xs[k - 1]));
        BHIP_TRY(hipStreamWaitEvent(b->stream, b->ev_join[k - 1], 0));
    }
    return SHG_OK;
}

}  // namespace

extern "C" {

int shg_batch_create(const shg_params *params, int count, shg_batch **out) {
    return shg_internal_guard(__func__, [&]() -> int {
        if (!params || !out || count <= 0) return shg_internal_fail(SHG_E_INVALID, "shg_batch_create: bad argument");
        *out = nullptr;
        for (int i = 0; i < count; i++) {
            const shg_params &p = params[i];
            if (p.nx != p.ny || size_class(p.nx) < 0)
                return shg_internal_fail(SHG_E_INVALID, "shg_batch_create: instance " + std::to_string(i) +
                                                            ": nx == ny in {64, 128, 256, 512} required");
            if (p.nz < 2 || !(p.lx > 0 && p.ly > 0 && p.lz > 0) || p.thermal ||
                (p.crystal != SHG_CRYSTAL_SPPLT && p.crystal != SHG_CRYSTAL_PPLN) ||
                (p.scheme != SHG_SCHEME_REFERENCE && p.scheme != SHG_SCHEME_STRANG))
                return shg_internal_fail(SHG_E_INVALID, "shg_batch_create: instance " + std::to_string(i) +
                                                            ": unsupported parameters (batches are plain SSFM)");
        }
        if (count > 65535) return shg_internal_fail(SHG_E_INVALID, "shg_batch_create: at most 65535 instances");
        shg_batch *b = new shg_batch();
        auto bad = [&](int rc) { batch_release(b); return rc; };
        b->count = count;
        b->p.assign(params, params + count);
        b->order.resize(count);
        std::iota(b->order.begin(), b->order.end(), 0);
        std::stable_sort(b->order.begin(), b->order.end(), [&](int x, int y) {
            const int cx = size_class(params[x].nx), cy = size_class(params[y].nx);
            if (cx != cy) return cx < cy;
            return params[x].nz > params[y].nz;
        });
        b->slot_of.resize(count);
        b->n.resize(count);
        b->steps.resize(count);
        b->cell_off.resize(count);
        b->t_slot.resize(count);
        size_t coef_total = 0, g_total = 0;
        std::vector<size_t> coef_off(count), g_off(count);
        for (int s = 0; s < count; s++) {
            const shg_params &p = params[b->order[s]];
            b->slot_of[b->order[s]] = s;
            b->n[s] = p.nx;
            b->steps[s] = p.nz - 1;
            b->cell_off[s] = b->total_cells;
            b->total_cells += (size_t)p.nx * p.nx;
            coef_off[s] = coef_total;
            coef_total += b->steps[s];
            g_off[s] = g_total;
            g_total += 4 * (size_t)p.nx;
            b->max_steps = std::max(b->max_steps, b->steps[s]);
            const int c = size_class(p.nx);
            if (b->cls_cnt[c] == 0) b->cls_off[c] = s;
            b->cls_cnt[c]++;
            b->cls_steps[c].push_back(b->steps[s]);
            b->t_slot[s] = (float)p.temp;
        }
        bool ok = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) == hipSuccess &&
                  hipEventCreate(&b->e0) == hipSuccess && hipEventCreate(&b->e1) == hipSuccess &&
                  hipEventCreateWithFlags(&b->ev_fork, hipEventDisableTiming) == hipSuccess;
        for (int k = 0; ok && k < kMaxStreams - 1; k++)
            ok = hipStreamCreateWithFlags(&b->xs[k], hipStreamNonBlocking) == hipSuccess &&
                 hipEventCreateWithFlags(&b->ev_join[k], hipEventDisableTiming) == hipSuccess;
        if (!ok) return bad(shg_internal_fail(SHG_E_DEVICE, "shg_batch_create: stream/event creation failed"));
        build_groups(b);
        if (hipMalloc(&b->U, b->total_cells * sizeof(float4)) != hipSuccess ||
            hipMalloc(&b->U0, b->total_cells * sizeof(float4)) != hipSuccess ||
            hipMalloc(&b->G, g_total * sizeof(float2)) != hipSuccess ||
            hipMalloc(&b->coef, coef_total * sizeof(float4)) != hipSuccess ||
            hipMalloc(&b->desc, count * sizeof(shg::InstDesc)) != hipSuccess ||
            hipMalloc(&b->ncell, count * sizeof(int)) != hipSuccess || hipMalloc(&b->pow, count * sizeof(double2)) != hipSuccess)
            return bad(shg_internal_fail(SHG_E_NOMEM, "shg_batch_create: device allocation failed"));
        for (int c = 0; c < 4; c++) {
            if (!b->cls_cnt[c]) continue;
            const auto tw = shg::phys::twiddle_table(kSizes[c]);
            if (hipMalloc(&b->tw[c], tw.size() * sizeof(float)) != hipSuccess ||
                shg_internal_copy(b->tw[c], tw.data(), tw.size() * sizeof(float), hipMemcpyHostToDevice, b->stream) != hipSuccess)
                return bad(shg_internal_fail(SHG_E_DEVICE, "shg_batch_create: twiddle upload failed"));
        }
        std::vector<float> gh(2 * g_total), ch(4 * coef_total);
        std::vector<shg::InstDesc> dh(count);
        std::vector<int> nc(count);
        for (int s = 0; s < count; s++) {
            const shg_params &p = params[b->order[s]];
            const shg::phys::Consts k(p);
            const float T = (float)p.temp;
            const auto g = shg::phys::propagator_factors(k, T, (double)k.dz);
            std::copy(g.begin(), g.end(), gh.begin() + 2 * g_off[s]);
            const auto cf = shg::phys::step_coefficients(k, T, b->steps[s]);
            std::copy(cf.begin(), cf.begin() + 4 * (size_t)b->steps[s], ch.begin() + 4 * coef_off[s]);
            shg::InstDesc &d = dh[s];
            d.U = b->U + b->cell_off[s];
            d.G = b->G + g_off[s];
            d.coef = b->coef + coef_off[s];
            const float h = p.scheme == SHG_SCHEME_STRANG ? 0.5f * k.dz : k.dz;
            d.nl.lossP = (float)(0.5 * (double)k.alpha_p);
            d.nl.lossS = (float)(0.5 * (double)k.alpha_s);
            d.nl.tpa = (float)(0.5 * ((double)k.beta_s * 0.5 * (double)k.c * (double)k.eps0 * (double)k.n_s(T)));
            d.nl.h6 = h / 6.0f;
            d.steps = b->steps[s];
            nc[s] = p.nx * p.nx;
        }
        hipStream_t st = b->stream;
        if (shg_internal_copy(b->G, gh.data(), gh.size() * sizeof(float), hipMemcpyHostToDevice, st) != hipSuccess ||
            shg_internal_copy(b->coef, ch.data(), ch.size() * sizeof(float), hipMemcpyHostToDevice, st) != hipSuccess ||
            shg_internal_copy(b->desc, dh.data(), dh.size() * sizeof(shg::InstDesc), hipMemcpyHostToDevice, st) !=
                hipSuccess ||
            shg_internal_copy(b->ncell, nc.data(), nc.size() * sizeof(int), hipMemcpyHostToDevice, st) != hipSuccess)
            return bad(shg_internal_fail(SHG_E_DEVICE, "shg_batch_create: table upload failed"));
        *out = b;
        return SHG_OK;
    });
}

void shg_batch_destroy(shg_batch *b) { batch_release(b); }

int shg_batch_init_fields(shg_batch *b, double noise_rel, uint64_t seed) {
    return shg_internal_guard(__func__, [&]() -> int {
        if (!b || !(noise_rel >= 0)) return shg_internal_fail(SHG_E_INVALID, "shg_batch_init_fields: bad argument");
        std::vector<shg::PumpArgs> a(b->count);
        int max_cells = 0;
        for (int s = 0; s < b->count; s++) {
            const int i = b->order[s];
            const shg_params &p = b->p[i];
            const shg::phys::Consts k(p);
            // the prologue of setPump (src/headers/Efields.h:77-92) in float
            const float T = (float)p.temp, w2 = (float)p.waist * (float)p.waist;
            const float np = k.n_p(T);
            const float zr = k.pi * np * w2 / k.lp;
            const float ap0 = sqrtf(4.0f * (float)p.power / (k.eps0 * k.c * k.pi * np * w2));
            const float eta = (float)p.focus / zr;
            const float mr = 1.0f, mi = -eta;
            const float mden = mr * mr + mi * mi;
            shg::PumpArgs &q = a[s];
            q.U = b->U0 + b->cell_off[s];
            q.ar = (ap0 * mr) / mden;
            q.ai = (-ap0 * mi) / mden;
            q.qr = w2 * mr;
            q.qi = w2 * mi;
            q.qden = q.qr * q.qr + q.qi * q.qi;
            q.dx = k.dx; q.dy = k.dy;
            q.ux = (float)(0.5 * p.nx); q.uy = (float)(0.5 * p.ny);
            q.sigma = (float)(noise_rel * std::sqrt((double)q.ar * q.ar + (double)q.ai * q.ai));
            q.n = p.nx;
            q.seed = seed * 0x100000001B3ull + (uint64_t)i;
            max_cells = std::max(max_cells, p.nx * p.nx);
        }
        shg::PumpArgs *d = nullptr;
        BHIP_TRY(hipMalloc(&d, a.size() * sizeof(shg::PumpArgs)));
        hipError_t e = shg_internal_copy(d, a.data(), a.size() * sizeof(shg::PumpArgs), hipMemcpyHostToDevice, b->stream);
        if (e == hipSuccess) e = shg::launch_pump_init(d, b->count, max_cells, b->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(b->stream);
        (void)hipFree(d);
        if (e != hipSuccess) return shg_internal_fail(SHG_E_DEVICE, std::string("shg_batch_init_fields: ") + hipGetErrorString(e));
        return after_fields_set(b);
    });
}

int shg_batch_set_fields(shg_batch *b, int index, const float *fh, const float *sh) {
    return shg_internal_guard(__func__, [&]() -> int {
        if (!b || !fh || !sh || index < 0 || index >= b->count)
            return shg_internal_fail(SHG_E_INVALID, "shg_batch_set_fields: bad argument");
        const int s = b->slot_of[index];
        const size_t nc = (size_t)b->n[s] * b->n[s];
        std::vector<float> h(4 * nc);
        for (size_t i = 0; i < nc; i++) {
            h[4 * i] = fh[2 * i]; h[4 * i + 1] = fh[2 * i + 1];
            h[4 * i + 2] = sh[2 * i]; h[4 * i + 3] = sh[2 * i + 1];
        }
        BHIP_TRY(shg_internal_copy(b->U0 + b->cell_off[s], h.data(), nc * sizeof(float4), hipMemcpyHostToDevice, b->stream));
        return after_fields_set(b);
    });
}

int shg_batch_set_schedule(shg_batch *b, int group_mib, int streams) {
    return shg_internal_guard(__func__, [&]() -> int {
        if (!b || group_mib < 1 || streams < 1 || streams > kMaxStreams)
            return shg_internal_fail(SHG_E_INVALID, "shg_batch_set_schedule: group_mib >= 1, 1 <= streams <= " +
                                                        std::to_string(kMaxStreams));
        b->group_mb = group_mib;
        b->nstreams = streams;
        build_groups(b);
        if (b->gexec) {
            (void)hipGraphExecDestroy(b->gexec);
            b->gexec = nullptr;
        }
        for (int c = 0; c < 4; c++) b->row_v[c] = b->col_v[c] = -1;   // group sizes changed: tune again
        return SHG_OK;
    });
}

int shg_batch_run(shg_batch *b) {
    return shg_internal_guard(__func__, [&]() -> int {
        if (!b) return shg_internal_fail(SHG_E_INVALID, "shg_batch_run: null batch");
        if (!b->have_fields) return shg_internal_fail(SHG_E_STATE, "shg_batch_run: set or init the fields first");
        if (b->row_v[0] < 0 && b->row_v[1] < 0 && b->row_v[2] < 0 && b->row_v[3] < 0) {
            const char *nt = std::getenv("SHG_NO_TUNE");
            if (nt && nt[0] == '1') {
                for (int c = 0; c < 4; c++) b->row_v[c] = b->col_v[c] = 0;
            } else if (int rc = tune_batch(b, 5)) {
                return rc;
            }
        }
        BHIP_TRY(hipMemcpyAsync(b->U, b->U0, b->total_cells * sizeof(float4), hipMemcpyDeviceToDevice, b->stream));
        const char *ng = std::getenv("SHG_NO_GRAPH");
        const bool graph = !(ng && ng[0] == '1');
        if (graph && !b->gexec)
            if (int rc = shg_internal_capture(b->stream, [b] { return enqueue_batch(b); }, &b->gexec)) return rc;
        BHIP_TRY(hipEventRecord(b->e0, b->stream));
        if (graph) {
            BHIP_TRY(hipGraphLaunch(b->gexec, b->stream));
        } else if (int rc = enqueue_batch(b)) {
            return rc;
        }
        BHIP_TRY(hipEventRecord(b->e1, b->stream));
        BHIP_TRY(hipEventSynchronize(b->e1));
        float ms = 0.f;
        BHIP_TRY(hipEventElapsedTime(&ms, b->e0, b->e1));
        b->last_ms = ms;
        return SHG_OK;
    });
}

static int batch_copy_out(shg_batch *b, const float4 *src_base, int index, float *fh, float *sh) {
    const int s = b->slot_of[index];
    const size_t nc = (size_t)b->n[s] * b->n[s];
    std::vector<float> h(4 * nc);
    BHIP_TRY(shg_internal_copy(h.data(), src_base + b->cell_off[s], nc * sizeof(float4), hipMemcpyDeviceToHost,
                               b->stream));
    for (size_t i = 0; i < nc; i++) {
        fh[2 * i] = h[4 * i]; fh[2 * i + 1] = h[4 * i + 1];
        sh[2 * i] = h[4 * i + 2]; sh[2 * i + 1] = h[4 * i + 3];
    }
    return SHG_OK;
}

int shg_batch_get_fields(shg_batch *b, int index, float *fh, float *sh) {
    return shg_internal_guard(__func__, [&]() -> int {
        if (!b || !fh || !sh || index < 0 || index >= b->count)
            return shg_internal_fail(SHG_E_INVALID, "shg_batch_get_fields: bad argument");
        return batch_copy_out(b, b->U, index, fh, sh);
    });
}

int shg_batch_get_input(shg_batch *b, int index, float *fh, float *sh) {
    return shg_internal_guard(__func__, [&]() -> int {
        if (!b || !fh || !sh || index < 0 || index >= b->count)
            return shg_internal_fail(SHG_E_INVALID, "shg_batch_get_input: bad argument");
        return batch_copy_out(b, b->U0, index, fh, sh);
    });
}

int shg_batch_powers(shg_batch *b, double *p_fh_in, double *p_sh_out, double *eta) {
    return shg_internal_guard(__func__, [&]() -> int {
        if (!b) return shg_internal_fail(SHG_E_INVALID, "shg_batch_powers: null batch");
        if (!b->have_fields) return shg_internal_fail(SHG_E_STATE, "shg_batch_powers: no fields");
        std::vector<double2> sums;
        if (int rc = batch_sums(b, sums)) return rc;
        for (int i = 0; i < b->count; i++) {
            const int s = b->slot_of[i];
            if (!std::isfinite(sums[s].y)) return shg_internal_fail(SHG_E_NONFINITE, "shg_batch_powers: non-finite field");
            const double po = slot_power(b, s, 1, sums[s].y);
            if (p_fh_in) p_fh_in[i] = b->p_in[s];
            if (p_sh_out) p_sh_out[i] = po;
            if (eta) eta[i] = b->p_in[s] > 0 ? po / b->p_in[s] : 0.0;
        }
        return SHG_OK;
    });
}

int64_t shg_batch_cell_steps(const shg_batch *b) {
    if (!b) return 0;
    int64_t t = 0;
    for (int s = 0; s < b->count; s++) t += (int64_t)b->n[s] * b->n[s] * b->steps[s];
    return t;
}

int shg_batch_last_ms(shg_batch *b, double *ms) {
    return shg_internal_guard(__func__, [&]() -> int {
        if (!b || !ms) return shg_internal_fail(SHG_E_INVALID, "shg_batch_last_ms: bad argument");
        *ms = b->last_ms;
        return SHG_OK;
    });
}

}  // extern "C"
