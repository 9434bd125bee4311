// shg_noalloc.hip — shg_jacobi_noalloc (include/shg.h): Jacobi sweeps of the THERMAL heat
// grid on caller-owned device buffers, for callers that cannot allocate device memory
// (code under HIP graph capture, serving loops).  The residual reduction's partials live
// in `scratch`, a fixed device array of this library: it is resolved to its per-device
// address once per device and shared by every call on that device.
#include <hip/hip_runtime.h>

#include <mutex>

#include "../../include/shg.h"
#include "guard.h"
#include "physics.h"
#include "thermal_kernels.h"

namespace {

// Residual partials (one FP64 sum per workgroup of the sweep that reports the residual)
// and, in the last slot, their total: the single-sweep kernel writes one per workgroup
// of 64 x 4 columns of jacobi_planes planes (8 on a 512 x 512 grid), so 2^15 partials
// cover every heat grid up to 512 x 512 x 256.
constexpr int kScratchParts = 1 << 15;
constexpr int kMaxDevices = 64;

static __device__ __align__(256) double scratch[kScratchParts + 1];

std::mutex g_mu;
double *g_scratch[kMaxDevices] = {};          // the array's address on each device, once resolved

// The device address of `scratch` on the current device (resolved on this device's
// first call; the array's host-side name is not a device address).
int scratch_on_device(double **out) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return shg_internal_fail(SHG_E_DEVICE, "shg_jacobi_noalloc: hipGetDevice failed");
    if (dev < 0 || dev >= kMaxDevices) return shg_internal_fail(SHG_E_DEVICE, "shg_jacobi_noalloc: device index out of range");
    std::lock_guard<std::mutex> lock(g_mu);
    if (!g_scratch[dev]) {
        double *ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress((void **)&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr)
            return shg_internal_fail(SHG_E_DEVICE, std::string("shg_jacobi_noalloc: hipGetSymbolAddress: ") +
                                                       hipGetErrorString(e));
        g_scratch[dev] = ptr;
    }
    *out = g_scratch[dev];
    return SHG_OK;
}

// Voxels a sweep never writes (the ovens) copied from T to the second buffer, so that
// both buffers of the ping-pong hold them.
__global__ __launch_bounds__(256) void k_copy_fixed(float *__restrict__ dst, const float *__restrict__ src, int nx,
                                                    int ny, int64_t n3, int bc) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n3) return;
    const int x = (int)(i % nx), y = (int)((i / nx) % ny);
    const bool updated = bc == SHG_BC_OVEN_SURROUNDED ? (x >= 1 && x < nx - 1 && y >= 1 && y < ny - 1) : y >= 1;
    if (!updated) dst[i] = src[i];
}

__global__ __launch_bounds__(256) void k_copy(float *__restrict__ dst, const float *__restrict__ src, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) dst[i] = src[i];
}

__global__ void k_mean(const double *__restrict__ total, double n3, double *__restrict__ out) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *out = *total / n3;
}

}  // namespace

extern "C" int shg_jacobi_noalloc(const shg_params *p, float *T, float *T_work, const float *Q, int64_t nsweeps,
                                  int layout, double *mean_sq_update, void *stream) {
    return shg_internal_guard(__func__, [&]() -> int {
        if (!p || !T || !T_work || !Q || nsweeps < 0) return shg_internal_fail(SHG_E_INVALID, "shg_jacobi_noalloc: bad argument");
        if (p->nx != p->ny || (p->nx != 64 && p->nx != 128 && p->nx != 256 && p->nx != 512))
            return shg_internal_fail(SHG_E_INVALID, "shg_jacobi_noalloc: nx == ny must be one of 64, 128, 256, 512");
        const int n = p->nx, hz = p->heat_nz > 0 ? p->heat_nz : p->nz;
        if (hz < 2 || p->nz < 2 || hz > p->nz) return shg_internal_fail(SHG_E_INVALID, "shg_jacobi_noalloc: heat_nz must be in [2, nz]");
        if (!(p->k_th > 0) || !(p->lx > 0 && p->ly > 0 && p->lz > 0))
            return shg_internal_fail(SHG_E_INVALID, "shg_jacobi_noalloc: box lengths and k_th must be > 0");
        if (p->thermal_bc != SHG_BC_TOP_OPEN && p->thermal_bc != SHG_BC_OVEN_SURROUNDED)
            return shg_internal_fail(SHG_E_INVALID, "shg_jacobi_noalloc: unknown thermal_bc");
        if (layout < -1 || layout >= SHG_JACOBI_LAYOUTS)
            return shg_internal_fail(SHG_E_INVALID, "shg_jacobi_noalloc: layout out of range");
        const bool pairs = layout >= 0 && layout < shg::kJacobi2VariantCount;
        shg::Jacobi2Plan plan{};
        if (pairs && shg::jacobi2_plan(n, n, hz, layout, &plan) != 0)
            return shg_internal_fail(SHG_E_DEVICE, "shg_jacobi_noalloc: Jacobi layout query failed");
        // the residual's partials must fit the scratch array
        const int last_parts = (pairs && nsweeps % 2 == 0) ? plan.blocks : shg::jacobi_blocks(n, n, hz);
        if (mean_sq_update && nsweeps > 0 && last_parts > kScratchParts)
            return shg_internal_fail(SHG_E_INVALID, "shg_jacobi_noalloc: the heat grid is larger than the residual "
                                                    "scratch (at most 2^15 partials: 512 x 512 x 256)");
        double *scr = nullptr;
        if (int rc = scratch_on_device(&scr)) return rc;
        hipStream_t st = (hipStream_t)stream;
        const shg::phys::Consts k(*p);
        shg::JacobiConst jc = shg::phys::jacobi_const<shg::JacobiConst>(k, shg::phys::heat_dz(k, hz), (float)p->k_th,
                                                                        (float)p->h_conv, (float)p->t_inf);
        jc.bc = p->thermal_bc;
        const int64_t n3 = (int64_t)n * n * hz;
        const unsigned g3 = (unsigned)((n3 + 255) / 256);
        auto check = [](hipError_t e) {
            return e == hipSuccess ? (int)SHG_OK
                                   : shg_internal_fail(SHG_E_DEVICE, std::string("shg_jacobi_noalloc: ") + hipGetErrorString(e));
        };
        if (nsweeps == 0) return SHG_OK;
        hipLaunchKernelGGL(k_copy_fixed, dim3(g3), dim3(256), 0, st, T_work, (const float *)T, n, n, n3, p->thermal_bc);
        if (int rc = check(hipGetLastError())) return rc;
        float *buf[2] = {T, T_work};
        int cur = 0, nparts = 0;
        int64_t i = 0;
        if (pairs) {
            for (; nsweeps - i >= 2; i += 2) {
                const bool r = mean_sq_update && nsweeps - i == 2;
                if (int rc = check(shg::launch_jacobi2(buf[cur ^ 1], buf[cur], Q, n, n, hz, jc, r ? scr : nullptr,
                                                       kScratchParts, st, plan)))
                    return rc;
                cur ^= 1;
                nparts = plan.blocks;
            }
        }
        for (; i < nsweeps; i++) {
            const bool r = mean_sq_update && i == nsweeps - 1;
            if (int rc = check(shg::launch_jacobi(buf[cur ^ 1], buf[cur], Q, n, n, hz, jc, r ? scr : nullptr,
                                                  kScratchParts, st)))
                return rc;
            cur ^= 1;
            nparts = shg::jacobi_blocks(n, n, hz);
        }
        if (mean_sq_update) {
            if (int rc = check(shg::launch_sum_partials(scr, nparts, scr + kScratchParts, st))) return rc;
            hipLaunchKernelGGL(k_mean, dim3(1), dim3(64), 0, st, (const double *)(scr + kScratchParts), (double)n3,
                               mean_sq_update);
            if (int rc = check(hipGetLastError())) return rc;
        }
        if (cur != 0) {                                 // the result is in T_work: back to T
            hipLaunchKernelGGL(k_copy, dim3(g3), dim3(256), 0, st, T, (const float *)T_work, n3);
            if (int rc = check(hipGetLastError())) return rc;
        }
        return SHG_OK;
    });
}
