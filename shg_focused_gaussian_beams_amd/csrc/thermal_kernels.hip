// thermal_kernels.hip — THERMAL mode on gfx950: the steady-state heat-equation Jacobi
// sweep, and the row pass whose nonlinear step reads the temperature, accumulates the
// phase mismatch and writes the heat source.
//
// Arithmetic that the reference evaluates in float (the Jacobi weights and sums, the
// refractive index and the phase mismatch) is evaluated here with FP contraction off
// and the same float/double typing, so a sweep and a dk(T) match the reference's
// float results bit for bit; the mismatch dk is a difference of O(4) terms whose
// rounding is amplified about a thousandfold.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <type_traits>

#include "fft_lds.h"
#include "nl_step.h"
#include "thermal_cell.h"
#include "pass_cfg.h"
#include "thermal_kernels.h"
#include "../../include/shg.h"

namespace shg {

// ---------------------------------------------------------------------------
// Boundary class of a coordinate along x or z: 0 inside, +1 on the low face (its one
// neighbour at +1), -1 on the high face (neighbour at -1).
__device__ __forceinline__ int jacobi_cls(int i, int n) { return i == 0 ? 1 : (i == n - 1 ? -1 : 0); }

// One Jacobi sweep of heatEquationTopOpen (src/headers/Tfield.h:113-221) at a voxel with
// y >= 1 of x class xc and z class zc (jacobi_cls) below the top (yi) or on it: the new
// value from the old values of its neighbours, ld(dx, dy, dz) = Ti at (x+dx, y+dy, z+dz),
// and q() = Q at the voxel.  Every formula keeps the reference's float/double typing and
// operation order, with contraction off, so that whichever kernel calls it, a sweep is
// the reference's float sweep bit for bit.  The voxels kernelHeatEquationOvenSurrounded
// updates (:224-258: inner points, z = 0 and z = L faces inside the lateral ovens) use the
// same inner and z-face formulas.
template <class LD, class LQ>
__device__ __forceinline__ float jacobi_voxel(const JacobiConst &c, int xc, bool yi, int zc, LD ld, LQ q) {
#pragma clang fp contract(off)
    const bool xi = xc == 0, zi = zc == 0;
    const int xn = xc, zn = zc;                        // the one x (z) neighbour on a face
    float out;
    if (yi) {
        if (xi && zi) {
            out = c.L2 * (div_const(ld(1, 0, 0) + ld(-1, 0, 0), c.dxx, c.rdxx) +
                          div_const(ld(0, 1, 0) + ld(0, -1, 0), c.dyy, c.rdyy) +
                          div_const(ld(0, 0, 1) + ld(0, 0, -1), c.dzz, c.rdzz) + div_const(q(), c.kth, c.rkth));
        } else if (!xi && zi) {
            out = div_const(ld(xn, 0, 0) + c.ax * (ld(0, 1, 0) + ld(0, -1, 0)) + c.bx * (ld(0, 0, 1) + ld(0, 0, -1)) +
                                c.bixT, c.fdenx, c.rfdenx);
        } else if (xi && !zi) {
            out = div_const(ld(0, 0, zn) + c.az * (ld(1, 0, 0) + ld(-1, 0, 0)) + c.bz * (ld(0, 1, 0) + ld(0, -1, 0)) +
                                c.bizT, c.fdenz, c.rfdenz);
        } else {
            const float f = c.av * ld(xn, 0, 0) + c.bv * ld(0, 0, zn) + c.cv * (ld(0, 1, 0) + ld(0, -1, 0));
            out = (float)(((double)f + c.sv) * c.V);
        }
    } else {
        if (xi && zi) {
            out = div_const(ld(0, -1, 0) + c.ay * (ld(1, 0, 0) + ld(-1, 0, 0)) + c.by * (ld(0, 0, 1) + ld(0, 0, -1)) +
                                c.biyT, c.fdeny, c.rfdeny);
        } else if (xi && !zi) {
            const float f = c.ah * ld(0, -1, 0) + c.bh * ld(0, 0, zn) + c.ch * (ld(-1, 0, 0) + ld(1, 0, 0));
            out = (float)(((double)f + c.sh) * c.H);
        } else if (!xi && zi) {
            const float f = c.ad * ld(0, -1, 0) + c.bd * ld(xn, 0, 0) + c.cd * (ld(0, 0, 1) + ld(0, 0, -1));
            out = (float)(((double)f + c.sd) * c.D);
        } else {
            out = (c.AV * ld(xn, 0, 0) + c.BV * ld(0, -1, 0) + c.CV * ld(0, 0, zn) + c.tinfBiD) * c.GV;
        }
    }
    return out;
}

// Whether a sweep writes voxel (x, y): everything above the y = 0 ovens (top open), or
// the voxels inside the four lateral oven faces (surrounded).
__device__ __forceinline__ bool jacobi_updates(const JacobiConst &c, int x, int y, int nx, int ny) {
    if (c.bc == SHG_BC_OVEN_SURROUNDED) return x >= 1 && x < nx - 1 && y >= 1 && y < ny - 1;
    return y >= 1;
}

// One Jacobi sweep of the voxels (x, y, z0 .. z0+ZB-1) of one thread (a column of ZB
// planes: the planes' z neighbours are shared, so a block reads each of its planes once
// and its two halo planes again, within its own lifetime); returns the FP64 sum of
// (Tf - Ti)^2 over them with RESID.  Voxels the sweep does not update are never written.
template <int ZB, bool RESID>
__device__ __forceinline__ double jacobi_column(float *__restrict__ Tf, const float *__restrict__ Ti,
                                                const float *__restrict__ Q, int nx, int ny, int hz, int x, int y,
                                                int z0, const JacobiConst &c) {
    double r = 0.0;
    if (x < nx && y < ny && jacobi_updates(c, x, y, nx, ny)) {
        const ptrdiff_t sy = nx, sz = (ptrdiff_t)nx * ny;
        const int xc = jacobi_cls(x, nx);
        const bool yi = y < ny - 1;
#pragma unroll
        for (int j = 0; j < ZB; j++) {
            const int z = z0 + j;
            if (z < hz) {
                const ptrdiff_t i = z * sz + y * sy + x;
                const float out = jacobi_voxel(
                    c, xc, yi, jacobi_cls(z, hz), [&](int dx, int dy, int dz) { return Ti[i + dx + dy * sy + dz * sz]; },
                    [&] { return Q[i]; });
                Tf[i] = out;
                if (RESID) {
                    const double d = (double)(out - Ti[i]);
                    r += d * d;
                }
            }
        }
    }
    return r;
}

// One Jacobi sweep Tf <- J(Ti, Q): a workgroup of 64 x BY threads, each a column of ZB
// planes (jacobi_column; ZB = jacobi_planes of the grid).  With RESID the FP64 sum of
// (Tf - Ti)^2 over the block goes to part[block] (BY = 4 then, so that every caller sums
// the same partials).  Used for single sweeps (the residual checks of the reference's
// loop, odd sweep counts) and as a layout of the sweep bursts.
template <int ZB, bool RESID, int BY = 4>
__global__ __launch_bounds__(64 * BY) void k_jacobi(float *__restrict__ Tf, const float *__restrict__ Ti,
                                                    const float *__restrict__ Q, int nx, int ny, int hz, JacobiConst c,
                                                    double *__restrict__ part) {
    static_assert(!RESID || BY == 4, "residual partials come from 64 x 4 blocks");
    const int x = blockIdx.x * 64 + (threadIdx.x & 63);
    const int y = blockIdx.y * BY + (threadIdx.x >> 6);
    double r = jacobi_column<ZB, RESID>(Tf, Ti, Q, nx, ny, hz, x, y, blockIdx.z * ZB, c);
    if (RESID) {
        for (int off = 32; off > 0; off >>= 1) r += __shfl_down(r, off, 64);
        __shared__ double w[4];
        if ((threadIdx.x & 63) == 0) w[threadIdx.x >> 6] = r;
        __syncthreads();
        if (threadIdx.x == 0)
            part[((size_t)blockIdx.z * gridDim.y + blockIdx.y) * gridDim.x + blockIdx.x] = w[0] + w[1] + w[2] + w[3];
    }
}

// Cell i of a W x H region whose inner cells form the rectangle [ax, bx) x [ay, by): the
// rectangle's cells first, row by row, then the rows above it, the rows below it, and the
// columns left and right of it.  A bijection of [0, W H) onto the region.
__device__ __forceinline__ void region_cell(int i, int W, int H, int ax, int bx, int ay, int by, int &row, int &col) {
    const int wi = bx - ax, hi = by - ay, ni = wi * hi;
    if (i < ni) {
        row = ay + i / wi;
        col = ax + i - (i / wi) * wi;
        return;
    }
    i -= ni;
    if (i < ay * W) {
        row = i / W;
        col = i - row * W;
        return;
    }
    i -= ay * W;
    if (i < (H - by) * W) {
        row = by + i / W;
        col = i - (row - by) * W;
        return;
    }
    i -= (H - by) * W;
    if (i < ax * hi) {
        row = ay + i / ax;
        col = i - (row - ay) * ax;
        return;
    }
    i -= ax * hi;
    const int wr = W - bx;
    row = ay + i / wr;
    col = bx + i - (row - ay) * wr;
}

// Cell i of the (TX+2) x (TY+2) region around a TX x TY tile, tile first: i < TX TY is
// tile cell i (row by row, at region row 1 + i / TX, column 1 + i % TX), then the
// region's top row, its bottom row, its left column and its right column (without their
// corners).  A bijection of [0, (TX+2)(TY+2)) onto the region whose first TX TY entries
// are the tile in its own row-major order.
__device__ __forceinline__ void ring_cell_tile_first(int i, int TX, int TY, int &row, int &col) {
    const int W = TX + 2;
    if (i < TX * TY) {
        row = 1 + i / TX;
        col = 1 + i - (i / TX) * TX;
        return;
    }
    i -= TX * TY;
    if (i < 2 * W) {
        row = i < W ? 0 : TY + 1;
        col = i < W ? i : i - W;
        return;
    }
    i -= 2 * W;
    row = 1 + (i < TY ? i : i - TY);
    col = i < TY ? 0 : W - 1;
}

// A raw buffer descriptor over `bytes` bytes at `base` (wave-uniform).  Its loads and
// stores take a per-lane byte offset, with no per-lane address arithmetic, and the
// hardware range check makes a load at an offset past `bytes` return 0 and a store there
// do nothing (the descriptor word 3 of CDNA: 32-bit data format, raw addressing).
constexpr int kPastBuffer = 0x7ffffff0;               // a byte offset beyond any plane
__device__ __forceinline__ __amdgpu_buffer_rsrc_t plane_rsrc(const float *base, int bytes) {
    return __builtin_amdgcn_make_buffer_rsrc((void *)base, 0, bytes, 0x00020000);
}
__device__ __forceinline__ float buf_load(__amdgpu_buffer_rsrc_t r, int off) {
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, off, 0, 0));
}
__device__ __forceinline__ void buf_store(__amdgpu_buffer_rsrc_t r, int off, float v) {
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), r, off, 0, 0);
}

// Two Jacobi sweeps per pass, Tf <- J(J(Ti, Q), Q), marching in z (temporal blocking).
// Workgroup = one TX x TY tile of the x-y plane and one chunk of zc planes.  It keeps in
// LDS a ring of four planes of Ti over the tile plus a halo of 2 (region R2) and a ring
// of four planes of the first sweep's values T1 over the tile plus a halo of 1 (R1).
// Step k of the march stages Ti plane k, computes T1 plane k-2 on R1 (from Ti planes
// staged by earlier steps), and the second sweep's plane k-4 on the tile (from T1 planes
// of earlier steps), which it writes: no ring slot is written and read in the same step,
// so one barrier per step suffices.  Ti plane k+1 and Q plane k-1 (over R1 and over the
// tile, each thread its own cells) are loaded into registers one step ahead (interior
// tiles: two steps ahead, jacobi2_march), so that their latency overlaps the step's
// arithmetic; the tile's values wait two more steps in registers for the second sweep.
// Loads and stores are buffer operations on a per-plane descriptor.  Q never goes
// through LDS, which keeps the footprint to the two rings.  T1 never leaves the chip: a pass reads Ti once over R2 (plus the z halo of its
// chunk) and Q over R1, and writes Tf once, where two single sweeps read and write
// everything twice.  The halo values of T1 are recomputed by every tile that needs them,
// with the same formulas, so the result is two single sweeps bit for bit.  With RESID the
// FP64 sum of (T2 - T1)^2 over the block's voxels (the second sweep's mean squared
// update) goes to part[block].
// Block order: workgroups go to the 8 XCDs round-robin; the logical block (instance,
// tile, chunk), chunk fastest, is dealt out so that each XCD runs a contiguous range of
// them, whose z-halo and x-y halo planes its L2 then shares.
// Batched instances: slot j < gridDim.y runs instance ids[j] (at offset ids[j]*inst_voxels
// of Tf, Ti, Q), its partials at j*gridDim.x; without `ids` the one slot is instance 0.
// The march of one workgroup over its tile and z-chunk (k_jacobi2).  INNER: every cell of
// the tile and of its one-cell halo is an updated inner cell in x and y (the tile's
// two-cell halo lies inside the grid and off the ovens), so the sweeps need no per-cell
// class: the inner formula, or on the grid's first and last planes the z-face formula,
// for every cell.  The first sweep then visits R1 tile first (ring_cell_tile_first), so
// that a thread's first R0 slots of R1 are its tile cells of the second sweep and their
// Q values serve both sweeps (one Q load per R1 cell and plane).  The general form
// dispatches per cell class.  Each cell gets the same formula either way, so a tile's
// result does not depend on which form ran.
template <int TX, int TY, bool RESID, bool INNER>
__device__ __forceinline__ void jacobi2_march(float *__restrict__ Tf, const float *__restrict__ Ti,
                                              const float *__restrict__ Q, int nx, int ny, int hz,
                                              const JacobiConst &c, double *__restrict__ part, int lb, int x0,
                                              int y0, int za, int zb, size_t sz, float (*Li)[(TX + 4) * (TY + 4)],
                                              float (*L1)[(TX + 2) * (TY + 2)]) {
    constexpr int W2 = TX + 4, H2 = TY + 4, N2 = W2 * H2;
    constexpr int W1 = TX + 2, H1 = TY + 2, N1 = W1 * H1;
    constexpr int N0 = TX * TY;
    constexpr int R2 = (N2 + 255) / 256, R1 = (N1 + 255) / 256, R0 = (N0 + 255) / 256;
    const int tid = threadIdx.x;
    // Per thread, everything about its cells that does not change along the march: the
    // plane offsets of its R2 cells (Ti staging, row by row), or -1 outside the domain;
    // for its R1 cells (first sweep) their place in the R2 and R1 ring planes, plane
    // offset (-1 outside) and what the sweep does there; for its tile cells (second
    // sweep) their place in the R1 ring planes, plane offset and what the sweep does
    // there.  Cell kinds: 0 outside the domain, 1 an oven (kept), 2 updated; bits 2-3 the
    // x class + 1, bit 4 below the top.  The sweeps visit a region's cells inner-first
    // (region_cell): the cells of the one formula then fill whole wavefronts, which run
    // it without class dispatch.
    int off2[R2], c2r[R1], q1[R1], o1[R1], k1[R1], c1r[R0], g0[R0], k0[R0];
#pragma unroll
    for (int r = 0; r < R2; r++) {
        const int t = tid + 256 * r;
        const int row = t / W2, col = t - row * W2;
        const int gx = x0 - 2 + col, gy = y0 - 2 + row;
        off2[r] = (t < N2 && gx >= 0 && gx < nx && gy >= 0 && gy < ny) ? 4 * (gy * nx + gx) : kPastBuffer;
    }
    auto kind = [&](int gx, int gy) {
        if (gx < 0 || gx >= nx || gy < 0 || gy >= ny) return 0;
        return (jacobi_updates(c, gx, gy, nx, ny) ? 2 : 1) | ((jacobi_cls(gx, nx) + 1) << 2) | ((gy < ny - 1) << 4);
    };
    constexpr int kInnerKind = 2 | (1 << 2) | (1 << 4);   // updated, x inside, below the top
    // inner cells (x in [1, nx-2], y in [1, ny-2]) of R1 and of the tile, in region coordinates
    const int a1x = max(0, 2 - x0), b1x = max(a1x, min(W1, nx - x0)), a1y = max(0, 2 - y0),
              b1y = max(a1y, min(H1, ny - y0));
    const int a0x = max(0, 1 - x0), b0x = max(a0x, min(TX, nx - 1 - x0)), a0y = max(0, 1 - y0),
              b0y = max(a0y, min(TY, ny - 1 - y0));
#pragma unroll
    for (int r = 0; r < R1; r++) {
        const int t = tid + 256 * r;
        int row = 0, col = 0;
        if (t < N1) {
            if (INNER)
                ring_cell_tile_first(t, TX, TY, row, col);
            else
                region_cell(t, W1, H1, a1x, b1x, a1y, b1y, row, col);
        }
        const int gx = x0 - 1 + col, gy = y0 - 1 + row;
        c2r[r] = (row + 1) * W2 + col + 1;                 // the cell in R2
        q1[r] = row * W1 + col;                            // and in R1
        o1[r] = (t < N1 && gx >= 0 && gx < nx && gy >= 0 && gy < ny) ? 4 * (gy * nx + gx) : kPastBuffer;
        k1[r] = INNER ? 0 : (t < N1 ? kind(gx, gy) : 0);
    }
#pragma unroll
    for (int r = 0; r < R0; r++) {
        const int t = tid + 256 * r;
        int row = 0, col = 0;
        if (t < N0) region_cell(t, TX, TY, a0x, b0x, a0y, b0y, row, col);
        const int gx = x0 + col, gy = y0 + row;
        c1r[r] = (row + 1) * W1 + col + 1;                 // the cell in R1
        g0[r] = 4 * (gy * nx + gx);
        k0[r] = INNER ? (t < N0 ? 1 : 0) : (t < N0 ? kind(gx, gy) : 0);
        if (!k0[r]) g0[r] = kPastBuffer;                 // outside the grid: never stored, Q reads 0
    }
    // per slot, whether every lane of the wavefront holds an inner cell (wave-uniform: the
    // sweeps then run the inner formula with no per-lane class branches)
    bool f1[R1], f0[R0];
#pragma unroll
    for (int r = 0; r < R1; r++)
        f1[r] = INNER || __builtin_amdgcn_ballot_w64(tid + 256 * r < N1 && k1[r] == kInnerKind) == ~0ull;
#pragma unroll
    for (int r = 0; r < R0; r++)
        f0[r] = INNER || __builtin_amdgcn_ballot_w64(tid + 256 * r < N0 && k0[r] == kInnerKind) == ~0ull;

    const int pbytes = (int)(4 * sz);
    double res = 0.0;
    if constexpr (INNER) {
        // The march, step k = za-4 .. zb+3 (the first two steps only load).  Step k stages Ti
        // plane k into ring slot k&3, runs the first sweep on plane k-2 (Ti planes k-3..k-1,
        // staged by earlier steps; first-sweep slot (k-2)&3), loads Ti plane k+2 and Q plane k
        // (register set k&1, whose previous contents, Ti plane k and Q plane k-2, this step has
        // just used), and runs the second sweep on plane k-4 (first-sweep planes k-5..k-3, of
        // earlier steps).  No ring slot is written and read in one step (rings of four planes),
        // so one barrier per step, at its end, orders everything, and a load has a step and a
        // half to arrive.  The loop is unrolled by four so that every ring slot and register
        // set is fixed at compile time (no per-cell slot arithmetic, no register copies: a copy
        // of a loaded register, or a select at the load, makes the wave wait for the load right
        // there).  Q over R1 serves the first sweep, and its first R0 slots (the tile cells,
        // ring_cell_tile_first) are held two steps in qh[slot] for the second.  Plane indices
        // outside [0, hz) are clamped to the grid (values unused).  The general form below
        // keeps one step of loads in flight: unrolled, its per-class code outgrew the
        // instruction cache and its registers.
        float pre[2][R2] = {}, qn1[2][R1] = {}, qh[4][R0];
        auto step = [&](auto slot, int k) {
            constexpr int S = decltype(slot)::value, P = S & 1;
            if (k < za - 4 || k > zb + 3) return;
            const bool live = k >= za - 2;                   // the first two steps only load
            // (a) stage Ti plane k
            if (live && k >= 0 && k < hz) {
#pragma unroll
                for (int r = 0; r < R2; r++) {
                    const int t = tid + 256 * r;
                    if (t < N2) Li[S][t] = pre[P][r];
                }
            }
            // (b) first sweep, plane p1 = k - 2, over R1
            const int p1 = k - 2;
            if (live && p1 >= max(za - 1, 0) && p1 <= min(zb, hz - 1)) {
                const float *lm = Li[(S + 1) & 3], *l0 = Li[(S + 2) & 3], *lp = Li[(S + 3) & 3];
                float *dst = L1[(S + 2) & 3];
                const int zc1 = jacobi_cls(p1, hz);
#pragma unroll
                for (int r = 0; r < R1; r++) {
                    const int t = tid + 256 * r;
                    const int c2 = c2r[r], p = q1[r];
                    auto ld = [&](int dx, int dy, int dz) {
                        const float *pl = dz < 0 ? lm : (dz > 0 ? lp : l0);
                        return pl[c2 + dy * W2 + dx];
                    };
                    auto lqv = [&] { return qn1[P][r]; };
                    if (t < N1) dst[p] = jacobi_voxel(c, 0, true, zc1, ld, lqv);
                }
            }
            // (c) hold the tile's Q of plane k-2 for the second sweep at step k+2, then load
#pragma unroll
            for (int r = 0; r < R0; r++) qh[S][r] = qn1[P][r];
            if (k + 2 <= zb + 3) {
                const auto rt = plane_rsrc(Ti + (size_t)min(max(k + 2, 0), hz - 1) * sz, pbytes);
#pragma unroll
                for (int r = 0; r < R2; r++) pre[P][r] = buf_load(rt, off2[r]);
                const auto rq = plane_rsrc(Q + (size_t)min(max(k, 0), hz - 1) * sz, pbytes);
#pragma unroll
                for (int r = 0; r < R1; r++) qn1[P][r] = buf_load(rq, o1[r]);
            }
            // (d) second sweep, plane p2 = k - 4, over the tile
            const int p2 = k - 4;
            if (live && p2 >= za && p2 < zb) {
                const float *lm = L1[(S + 3) & 3], *l0 = L1[S], *lp = L1[(S + 1) & 3];
                const int zc2 = jacobi_cls(p2, hz);
                const auto rtf = plane_rsrc(Tf + (size_t)p2 * sz, pbytes);
#pragma unroll
                for (int r = 0; r < R0; r++) {
                    const int c1 = c1r[r];
                    auto ld = [&](int dx, int dy, int dz) {
                        const float *pl = dz < 0 ? lm : (dz > 0 ? lp : l0);
                        return pl[c1 + dy * W1 + dx];
                    };
                    auto lqv = [&] { return qh[(S + 2) & 3][r]; };
                    if (tid + 256 * r < N0) {
                        const float out = jacobi_voxel(c, 0, true, zc2, ld, lqv);
                        buf_store(rtf, g0[r], out);
                        if (RESID) {
                            const double d = (double)(out - l0[c1]);
                            res += d * d;
                        }
                    }
                }
            }
            if (live) __syncthreads();
        };
        for (int kb = (za - 4) & ~3; kb <= zb + 3; kb += 4) {
            step(std::integral_constant<int, 0>{}, kb);
            step(std::integral_constant<int, 1>{}, kb + 1);
            step(std::integral_constant<int, 2>{}, kb + 2);
            step(std::integral_constant<int, 3>{}, kb + 3);
        }
    } else {
        // Loads of step k (issued during step k-1): Ti plane k over R2, and Q plane k-2 at the
        // R1 cells and at the tile cells: the first sweep uses the R1 values at step k, the
        // second the tile values at step k+2 (held in registers meanwhile; the two reads of a
        // Q plane come together, so the second hits the cache).  Every step loads (a plane index outside
        // [0, hz) is clamped to the grid, its values unused): a kept old value, or a select at
        // the load, would make the wave wait for the load right there, leaving no step of
        // latency hidden.  Out-of-domain cells load at an offset past the plane, which the
        // buffer range check turns into 0.
        float pre[R2], qn1[R1] = {}, qn0[R0] = {}, qm0[R0] = {}, qo0[R0] = {};
        auto fetch = [&](int k) {
            const auto rt = plane_rsrc(Ti + (size_t)min(max(k, 0), hz - 1) * sz, pbytes);
#pragma unroll
            for (int r = 0; r < R2; r++) pre[r] = buf_load(rt, off2[r]);
            const auto rq = plane_rsrc(Q + (size_t)min(max(k - 2, 0), hz - 1) * sz, pbytes);
#pragma unroll
            for (int r = 0; r < R1; r++) qn1[r] = buf_load(rq, o1[r]);
#pragma unroll
            for (int r = 0; r < R0; r++) qn0[r] = buf_load(rq, g0[r]);
        };
        // Step k stages Ti plane k, runs the first sweep on plane k-2 (its Ti planes k-3..k-1
        // were staged by earlier steps) and the second on plane k-4 (its first-sweep planes
        // k-5..k-3 come from earlier steps): within a step no ring slot is both written and
        // read (rings of four planes), so one barrier per step, at its end, orders everything.
        fetch(za - 2);
        for (int k = za - 2; k <= zb + 3; k++) {
            // (a) stage Ti plane k, take this step's Q values, then load the next step's
            if (k >= 0 && k < hz) {
                float *dst = Li[(k + 4) & 3];
#pragma unroll
                for (int r = 0; r < R2; r++) {
                    const int t = tid + 256 * r;
                    if (t < N2) dst[t] = pre[r];
                }
            }
            float q1c[R1], q0c[R0];
#pragma unroll
            for (int r = 0; r < R1; r++) q1c[r] = qn1[r];
#pragma unroll
            for (int r = 0; r < R0; r++) {                 // tile Q: plane k-4 now, k-3 and k-2 held
                q0c[r] = qo0[r];
                qo0[r] = qm0[r];
                qm0[r] = qn0[r];
            }
            if (k + 1 <= zb + 3) fetch(k + 1);
            // (b) second sweep, plane p2 = k - 4, over the tile: its stores go out before the
            // first sweep, which gives them that sweep's time to complete before the next step
            // waits for its loads
            const int p2 = k - 4;
            if (p2 >= za && p2 < zb) {
                const float *lm = L1[(p2 + 3) & 3], *l0 = L1[(p2 + 4) & 3], *lp = L1[(p2 + 5) & 3];
                const int zc2 = jacobi_cls(p2, hz);
                const auto rtf = plane_rsrc(Tf + (size_t)p2 * sz, pbytes);
#pragma unroll
                for (int r = 0; r < R0; r++) {
                    const int c1 = c1r[r], kd = k0[r];
                    auto ld = [&](int dx, int dy, int dz) {
                        const float *pl = dz < 0 ? lm : (dz > 0 ? lp : l0);
                        return pl[c1 + dy * W1 + dx];
                    };
                    auto lqv = [&] { return q0c[r]; };
                    const bool fast = zc2 == 0 && f0[r];                     // a wavefront of inner cells
                    if (fast || (kd & 3) == 2) {
                        const float out = fast ? jacobi_voxel(c, 0, true, 0, ld, lqv)
                                               : jacobi_voxel(c, ((kd >> 2) & 3) - 1, (kd >> 4) & 1, zc2, ld, lqv);
                        buf_store(rtf, g0[r], out);
                        if (RESID) {
                            const double d = (double)(out - l0[c1]);
                            res += d * d;
                        }
                    }
                }
            }
            // (c) first sweep, plane p1 = k - 2, over R1
            const int p1 = k - 2;
            if (p1 >= max(za - 1, 0) && p1 <= min(zb, hz - 1)) {
                const float *lm = Li[(p1 + 3) & 3], *l0 = Li[(p1 + 4) & 3], *lp = Li[(p1 + 5) & 3];
                float *dst = L1[(p1 + 4) & 3];
                const int zc1 = jacobi_cls(p1, hz);
#pragma unroll
                for (int r = 0; r < R1; r++) {
                    const int t = tid + 256 * r;
                    const int c2 = c2r[r], kd = k1[r], p = q1[r];
                    auto ld = [&](int dx, int dy, int dz) {
                        const float *pl = dz < 0 ? lm : (dz > 0 ? lp : l0);
                        return pl[c2 + dy * W2 + dx];
                    };
                    auto lqv = [&] { return q1c[r]; };
                    if (zc1 == 0 && f1[r]) {                          // a wavefront of inner cells
                        dst[p] = jacobi_voxel(c, 0, true, 0, ld, lqv);
                    } else if (t < N1) {
                        float v = 0.f;
                        if ((kd & 3) == 2) {
                            v = jacobi_voxel(c, ((kd >> 2) & 3) - 1, (kd >> 4) & 1, zc1, ld, lqv);
                        } else if ((kd & 3) == 1) {
                            v = l0[c2];                                      // ovens keep their value
                        }
                        dst[p] = v;
                    }
                }
            }
            __syncthreads();
        }
    }
    if (RESID) {
        for (int off = 32; off > 0; off >>= 1) res += __shfl_down(res, off, 64);
        __shared__ double w[4];
        if ((tid & 63) == 0) w[tid >> 6] = res;
        __syncthreads();
        if (tid == 0) part[lb] = w[0] + w[1] + w[2] + w[3];
    }
}

template <int TX, int TY, bool RESID>
__global__ __launch_bounds__(256) void k_jacobi2(float *__restrict__ Tf, const float *__restrict__ Ti,
                                                 const float *__restrict__ Q, int nx, int ny, int hz, int zc,
                                                 int tiles_x, int nchunks, JacobiConst c, double *__restrict__ part,
                                                 const int *__restrict__ ids, size_t inst_voxels) {
    constexpr int W2 = TX + 4, H2 = TY + 4, N2 = W2 * H2;
    constexpr int W1 = TX + 2, H1 = TY + 2, N1 = W1 * H1;
    constexpr int N0 = TX * TY;
    constexpr int R2 = (N2 + 255) / 256, R1 = (N1 + 255) / 256, R0 = (N0 + 255) / 256;
    __shared__ float Li[4][N2];
    __shared__ float L1[4][N1];
    const int nb = gridDim.x * gridDim.y, b = blockIdx.y * gridDim.x + blockIdx.x;
    const int per = nb >> 3, rem = nb & 7, xcd = b & 7;
    const int gb = xcd * per + min(xcd, rem) + (b >> 3);
    const int slot = gb / (int)gridDim.x, lb = gb - slot * (int)gridDim.x;
    const size_t inst = ids ? (size_t)ids[slot] : 0;
    Tf += inst * inst_voxels;
    Ti += inst * inst_voxels;
    Q += inst * inst_voxels;
    if (RESID) part += (size_t)slot * gridDim.x;
    const int tile = lb / nchunks, chunk = lb - tile * nchunks;
    const int ty = tile / tiles_x, tx = tile - ty * tiles_x;
    const int x0 = tx * TX, y0 = ty * TY;
    const int za = chunk * zc, zb = min(hz, za + zc);
    const size_t sz = (size_t)nx * ny;

    if (x0 >= 2 && x0 + TX <= nx - 2 && y0 >= 2 && y0 + TY <= ny - 2)
        jacobi2_march<TX, TY, RESID, true>(Tf, Ti, Q, nx, ny, hz, c, part, lb, x0, y0, za, zb, sz, Li, L1);
    else
        jacobi2_march<TX, TY, RESID, false>(Tf, Ti, Q, nx, ny, hz, c, part, lb, x0, y0, za, zb, sz, Li, L1);
}

__global__ __launch_bounds__(256) void k_sum_partials(const double *__restrict__ part, int n, double *out) {
    __shared__ double s[256];
    double r = 0.0;
    for (int i = threadIdx.x; i < n; i += 256) r += part[i];
    s[threadIdx.x] = r;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = s[0];
}

// Batched single sweep: slot blockIdx.y runs instance ids[blockIdx.y]; blockIdx.x is the
// block of k_jacobi's grid flattened (x fastest, then y, then z), so that a slot's
// partials (at blockIdx.y * gridDim.x) come in k_jacobi's order; the same columns.
template <int ZB, bool RESID>
__global__ __launch_bounds__(256) void k_jacobi_batch(float *__restrict__ Tf, const float *__restrict__ Ti,
                                                      const float *__restrict__ Q, int nx, int ny, int hz,
                                                      const int *__restrict__ ids, JacobiConst c,
                                                      double *__restrict__ part) {
    const size_t inst = (size_t)ids[blockIdx.y];
    const int gx = nx / 64, gy = (ny + 3) / 4;
    const int bz = (int)blockIdx.x / (gx * gy), rxy = (int)blockIdx.x - bz * gx * gy;
    const int by = rxy / gx, bx = rxy - by * gx;
    const size_t off = inst * (size_t)nx * ny * hz;
    double r = jacobi_column<ZB, RESID>(Tf + off, Ti + off, Q + off, nx, ny, hz, bx * 64 + (threadIdx.x & 63),
                                        by * 4 + (threadIdx.x >> 6), bz * ZB, c);
    if (RESID) {
        for (int o = 32; o > 0; o >>= 1) r += __shfl_down(r, o, 64);
        __shared__ double w[4];
        if ((threadIdx.x & 63) == 0) w[threadIdx.x >> 6] = r;
        __syncthreads();
        if (threadIdx.x == 0) part[(size_t)blockIdx.y * gridDim.x + blockIdx.x] = w[0] + w[1] + w[2] + w[3];
    }
}

// Per-instance sums of consecutive ranges of `per` partials (block = instance), in the
// order and tree of k_sum_partials.
__global__ __launch_bounds__(256) void k_sum_partials_batch(const double *__restrict__ part, int per, double *out) {
    __shared__ double s[256];
    const double *p = part + (size_t)blockIdx.x * per;
    double r = 0.0;
    for (int i = threadIdx.x; i < per; i += 256) r += p[i];
    s[threadIdx.x] = r;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[blockIdx.x] = s[0];
}

// Refractive indices of every heat voxel, {n_p(T), n_s(T)}: the temperature is frozen
// during a march, so the FP64 Sellmeier fits run once per voxel instead of twice per cell
// and step.
__global__ __launch_bounds__(256) void k_index_table(const float *__restrict__ T, float2 *__restrict__ tab, int64_t n,
                                                     ThermConst k) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const float t = T[i];
    tab[i] = make_float2(index_dev(k.crystal, k.lp, t), index_dev(k.crystal, k.ls, t));
}

// Batched instances (blockIdx.y): moves the pass's array pointers to this block's instance
// and tells whether the instance sits this pass out.
__device__ __forceinline__ bool therm_batch_skip(float4 *__restrict__ &U, float4 *__restrict__ &Uout,
                                                 const float2 *__restrict__ &NT, float *__restrict__ &Q,
                                                 float *__restrict__ &DKint, const ThermRowArgs &a) {
    const int inst = blockIdx.y;
    if (a.active && !a.active[inst]) return true;
    U += inst * a.inst_cells;
    Uout += inst * a.inst_cells;
    DKint += inst * a.inst_cells;
    NT += inst * a.inst_voxels;
    Q += inst * a.inst_voxels;
    return false;
}

// Row pass of the THERMAL z-march (flags in ThermRowFlags):
//   FWD   R_fwd (end of the previous step's diffraction)
//   NPREV second RK4 of step s-1 with T(slice(s-1)) and DKint_{s-1}
//   QPREV heat source of slice(s-1) from the fields after step s-1
//   SNAP  copy of the fields after step s-1 to Uout (the reference's Pump_out state)
//   NCUR  DKint_s = DKint_{s-1} + dk(T(slice(s))), then the first RK4 of step s
//         (stored unless DKKEEP; a lagging stored prefix is brought up by dk_pending adds)
//   BWD   R_bwd (start of step s's diffraction)
// DKint holds the running float prefix sum of dk over the executed steps: the reference
// recomputes the same float sum from z' = 0 twice per step (src/headers/PhaseMatching.h:
// 25-38, Solver.h:198-199); the sequential float additions are identical.
//
// One cell of the pass between the transforms (idx: cell index in the plane), from its
// already loaded phase accumulator and index-table entries.
__device__ __forceinline__ void therm_cell_pre(float2 &P, float2 &S, size_t idx, size_t n2, size_t sprev, float dkprev,
                                               float2 nnp, float2 nnc, bool same, float4 *__restrict__ Uout,
                                               float *__restrict__ Q, float *__restrict__ DKint,
                                               const ThermRowArgs &a) {
    const ThermConst &k = a.k;
    CellPhys cp{};
    if (a.flags & (TROW_NPREV | TROW_QPREV)) {
        cp = cell_phys(k, nnp);
#pragma unroll
        for (int i = 0; i < kMaxDkPending; i++)                         // the steps since the stored prefix
            if (i < a.dk_pending) dkprev = dkprev + cp.dk;
        if ((a.flags & (TROW_TAPE | TROW_NPREV)) == (TROW_TAPE | TROW_NPREV))
            a.tape_a2[(size_t)(a.prev.step - a.tape_s0) * n2 + idx] = make_float4(P.x, P.y, S.x, S.y);
        if (a.flags & TROW_NPREV) rk4_thermal(P, S, cp, dkprev, a.prev, k);
        if (a.flags & TROW_QPREV) Q[sprev + idx] = heat_source(k, cp, P, S);
    }
    if (a.flags & TROW_SNAP) Uout[idx] = make_float4(P.x, P.y, S.x, S.y);
    if (a.flags & TROW_NCUR) {
        const CellPhys cc = same ? cp : cell_phys(k, nnc);
        const float dk = dkprev + cc.dk;
        if (!(a.flags & TROW_DKKEEP)) DKint[idx] = dk;
        if (a.flags & TROW_TAPE) a.tape_dk[(size_t)(a.cur.step - a.tape_s0) * n2 + idx] = dk;
        rk4_thermal(P, S, cc, dk, a.cur, k);
    }
}

// The same, loading the cell's inputs itself.
__device__ __forceinline__ void therm_cell(float2 &P, float2 &S, size_t idx, size_t n2, size_t sprev, size_t scur,
                                           float4 *__restrict__ Uout, const float2 *__restrict__ NT,
                                           float *__restrict__ Q, float *__restrict__ DKint, const ThermRowArgs &a) {
    const float dkprev = (a.flags & TROW_FIRST) ? 0.0f : DKint[idx];
    const bool pv = a.flags & (TROW_NPREV | TROW_QPREV), cu = a.flags & TROW_NCUR;
    const bool same = pv && a.prev.slice == a.cur.slice;
    const float2 nnp = pv ? NT[sprev + idx] : make_float2(0.f, 0.f);
    const float2 nnc = (cu && !same) ? NT[scur + idx] : nnp;
    therm_cell_pre(P, S, idx, n2, sprev, dkprev, nnp, nnc, same, Uout, Q, DKint, a);
}

// Layout <N, E, W> as in pass_cfg.h.  Between the two transforms the thread's E cells
// are parked in the row's LDS slots and stepped by a rolled loop: the per-cell work (two
// RK4 with three sincos each, the heat source) is too large to unroll E times, and a
// rolled loop over register arrays would move them to scratch memory.
template <int N, int E, int W>
__global__ __launch_bounds__(64 * W) void k_therm_row_pass(float4 *__restrict__ U, float4 *__restrict__ Uout,
                                                           const float2 *__restrict__ twg,
                                                           const float2 *__restrict__ NT, float *__restrict__ Q,
                                                           float *__restrict__ DKint, ThermRowArgs a) {
    using C = RowCfgT<N, E, W>;
    constexpr int TT = C::T;
    if (therm_batch_skip(U, Uout, NT, Q, DKint, a)) return;
    __shared__ float2 lds[2][C::ROWS][C::L];
    __shared__ float2 tw[N];
    const int lane = threadIdx.x;
    const int t = lane % TT, rr = lane / TT;
    const int row = blockIdx.x * C::ROWS + rr;
    const size_t n2 = (size_t)N * N;
    const size_t base = (size_t)row * N;
    // the row's loads go out before the twiddles are staged (the latencies overlap
    // instead of adding up across the barrier)
    float2 fa[E], fb[E];
#pragma unroll
    for (int j = 0; j < E; j++) {
        const float4 v = U[base + t + j * TT];
        fa[j] = make_float2(v.x, v.y);
        fb[j] = make_float2(v.z, v.w);
    }
    for (int j = lane; j < N; j += C::THREADS) tw[j] = twg[j];
    __syncthreads();
    auto sync = [] { __syncthreads(); };
    float2 *la = lds[0][rr], *lb = lds[1][rr];
    if (a.flags & TROW_FWD) fft2fe<N, E, -1>(fa, fb, la, lb, t, 1, tw, sync);
    __syncthreads();                          // the transform's last LDS reads are done
#pragma unroll
    for (int j = 0; j < E; j++) {
        la[lds_pad(t + j * TT)] = fa[j];
        lb[lds_pad(t + j * TT)] = fb[j];
    }
    // each thread reads back only the slots it wrote: no barrier
    const size_t sprev = (size_t)a.prev.slice * n2, scur = (size_t)a.cur.slice * n2;
#pragma unroll 1
    for (int j = 0; j < E; j++) {
        const int o = lds_pad(t + j * TT);
        float2 P = la[o], S = lb[o];
        therm_cell(P, S, base + t + j * TT, n2, sprev, scur, Uout, NT, Q, DKint, a);
        la[o] = P;
        lb[o] = S;
    }
#pragma unroll
    for (int j = 0; j < E; j++) {
        fa[j] = la[lds_pad(t + j * TT)];
        fb[j] = lb[lds_pad(t + j * TT)];
    }
    if (a.flags & TROW_BWD) fft2fe<N, E, +1>(fa, fb, la, lb, t, 1, tw, sync);
#pragma unroll
    for (int j = 0; j < E; j++) U[base + t + j * TT] = make_float4(fa[j].x, fa[j].y, fb[j].x, fb[j].y);
}

// Wide layout <N, E, W>: one row per workgroup of W waves.  The first T = N/E lanes (one
// wavefront or less) run the two transforms; between them every thread of the group
// steps N/(64 W) cells of the row, so the per-cell work spreads over W times as many
// wavefronts as the transforms need.  Each thread loads its cells' phase accumulator and
// index-table entries before the first transform, which hides their latency behind it.
template <int N, int E, int W>
__global__ __launch_bounds__(64 * W) void k_therm_row_wide(float4 *__restrict__ U, float4 *__restrict__ Uout,
                                                           const float2 *__restrict__ twg,
                                                           const float2 *__restrict__ NT, float *__restrict__ Q,
                                                           float *__restrict__ DKint, ThermRowArgs a) {
    constexpr int TT = N / E, THREADS = 64 * W, L = LdsLen<N>::value, CPT = N / THREADS;
    if (therm_batch_skip(U, Uout, NT, Q, DKint, a)) return;
    static_assert(TT <= 64 && N % THREADS == 0, "the transforms run in one wavefront");
    __shared__ float2 lds[2][L];
    __shared__ float2 tw[N];
    const int tid = threadIdx.x;
    const size_t n2 = (size_t)N * N, base = (size_t)blockIdx.x * N;
    const size_t sprev = (size_t)a.prev.slice * n2, scur = (size_t)a.cur.slice * n2;
    float dkp[CPT];
    float2 nnp[CPT], nnc[CPT];
    const bool pv = a.flags & (TROW_NPREV | TROW_QPREV), cu = a.flags & TROW_NCUR;
    const bool same = pv && a.prev.slice == a.cur.slice;      // consecutive steps share a heat slice
    // The row (transform lanes) and the twiddles are loaded and the twiddles staged first,
    // the table entries loaded after them, so that the table loads are in flight behind
    // the forward transform (a wait issued after them would also wait for them: loads
    // complete in issue order).
    float4 uv[E];
    if (tid < TT) {
#pragma unroll
        for (int j = 0; j < E; j++) uv[j] = U[base + tid + j * TT];
    }
    for (int j = tid; j < N; j += THREADS) tw[j] = twg[j];
#pragma unroll
    for (int q = 0; q < CPT; q++) {
        const size_t idx = base + tid + q * THREADS;
        dkp[q] = (a.flags & TROW_FIRST) ? 0.0f : DKint[idx];
        nnp[q] = pv ? NT[sprev + idx] : make_float2(0.f, 0.f);
        nnc[q] = (cu && !same) ? NT[scur + idx] : make_float2(0.f, 0.f);   // unused when same
    }
    __syncthreads();
    auto sync = [] { wave_sync(); };
    if (tid < TT) {
        float2 fa[E], fb[E];
#pragma unroll
        for (int j = 0; j < E; j++) {
            fa[j] = make_float2(uv[j].x, uv[j].y);
            fb[j] = make_float2(uv[j].z, uv[j].w);
        }
        if (a.flags & TROW_FWD) fft2fe<N, E, -1>(fa, fb, lds[0], lds[1], tid, 1, tw, sync);
        wave_sync();
#pragma unroll
        for (int j = 0; j < E; j++) {
            lds[0][lds_pad(tid + j * TT)] = fa[j];
            lds[1][lds_pad(tid + j * TT)] = fb[j];
        }
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < CPT; q++) {
        const int cc = tid + q * THREADS, o = lds_pad(cc);
        float2 P = lds[0][o], S = lds[1][o];
        therm_cell_pre(P, S, base + cc, n2, sprev, dkp[q], nnp[q], nnc[q], same, Uout, Q, DKint, a);
        lds[0][o] = P;
        lds[1][o] = S;
    }
    __syncthreads();
    if (tid < TT) {
        float2 fa[E], fb[E];
#pragma unroll
        for (int j = 0; j < E; j++) {
            fa[j] = lds[0][lds_pad(tid + j * TT)];
            fb[j] = lds[1][lds_pad(tid + j * TT)];
        }
        if (a.flags & TROW_BWD) fft2fe<N, E, +1>(fa, fb, lds[0], lds[1], tid, 1, tw, sync);
#pragma unroll
        for (int j = 0; j < E; j++) U[base + tid + j * TT] = make_float4(fa[j].x, fa[j].y, fb[j].x, fb[j].y);
    }
}

// Split layout <N, E, W>, W >= 2: one row per workgroup; the row is loaded one cell per
// thread into LDS, waves 0 and 1 transform FH and SH side by side, every thread steps its
// cells, and the row is stored one cell per thread.
template <int N, int E, int W>
__global__ __launch_bounds__(64 * W) void k_therm_row_split(float4 *__restrict__ U, float4 *__restrict__ Uout,
                                                            const float2 *__restrict__ twg,
                                                            const float2 *__restrict__ NT, float *__restrict__ Q,
                                                            float *__restrict__ DKint, ThermRowArgs a) {
    constexpr int T = N / E, THREADS = 64 * W, L = LdsLen<N>::value, CPT = (N + THREADS - 1) / THREADS;
    if (therm_batch_skip(U, Uout, NT, Q, DKint, a)) return;
    static_assert(W >= 2 && T <= 64, "one wavefront per field transform");
    __shared__ float2 lds[2][L];
    __shared__ float2 tw[N];
    const int tid = threadIdx.x, f = tid >> 6, lane = tid & 63;
    const size_t n2 = (size_t)N * N, base = (size_t)blockIdx.x * N;
    const size_t sprev = (size_t)a.prev.slice * n2, scur = (size_t)a.cur.slice * n2;
    float dkp[CPT];
    float2 nnp[CPT], nnc[CPT];
    const bool pv = a.flags & (TROW_NPREV | TROW_QPREV), cu = a.flags & TROW_NCUR;
    const bool same = pv && a.prev.slice == a.cur.slice;      // consecutive steps share a heat slice
    // The row and the twiddles are staged first and the table entries loaded after them,
    // so that the table loads are in flight behind the forward transform (a wait for the
    // row issued after them would also wait for them: loads complete in issue order).
    for (int j = tid; j < N; j += THREADS) tw[j] = twg[j];
    for (int c = tid; c < N; c += THREADS) {
        const float4 v = U[base + c];
        lds[0][lds_pad(c)] = make_float2(v.x, v.y);
        lds[1][lds_pad(c)] = make_float2(v.z, v.w);
    }
#pragma unroll
    for (int q = 0; q < CPT; q++) {
        const int c = tid + q * THREADS;
        const size_t idx = base + (c < N ? c : 0);
        dkp[q] = (a.flags & TROW_FIRST) ? 0.0f : DKint[idx];
        nnp[q] = pv ? NT[sprev + idx] : make_float2(0.f, 0.f);
        nnc[q] = (cu && !same) ? NT[scur + idx] : make_float2(0.f, 0.f);   // unused when same
    }
    __syncthreads();
    if ((a.flags & TROW_FWD) && f < 2 && lane < T) fft_lds_field<N, E, -1>(lds[f], lane, tw);
    __syncthreads();
#pragma unroll
    for (int q = 0; q < CPT; q++) {
        const int c = tid + q * THREADS;
        if (c < N) {
            float2 P = lds[0][lds_pad(c)], S = lds[1][lds_pad(c)];
            therm_cell_pre(P, S, base + c, n2, sprev, dkp[q], nnp[q], nnc[q], same, Uout, Q, DKint, a);
            lds[0][lds_pad(c)] = P;
            lds[1][lds_pad(c)] = S;
        }
    }
    __syncthreads();
    if ((a.flags & TROW_BWD) && f < 2 && lane < T) fft_lds_field<N, E, +1>(lds[f], lane, tw);
    __syncthreads();
    for (int c = tid; c < N; c += THREADS) {
        const float2 P = lds[0][lds_pad(c)], S = lds[1][lds_pad(c)];
        U[base + c] = make_float4(P.x, P.y, S.x, S.y);
    }
}

// ---------------------------------------------------------------------------
// Planes per thread of the single sweep: the most of {8, 4, 2} that still gives the grid
// 4096 workgroups (16 per CU), else 1.  A rule of the grid alone (not timed), so every
// context and batch of one grid sums its residual partials alike.  Measured per sweep
// (planes 1 / 2 / 4 / 8): 64x64x200 5.3 / 5.4 / 6.0 / 8.0 us, 256x256x100 22.1 / 21.7 /
// 20.6 / 21.4 us, 512x512x200 214 / 201 / 189 / 184 us.
static int planes_for(long cols, int hz) {
    for (int zb = 8; zb > 1; zb >>= 1)
        if (cols * ((hz + zb - 1) / zb) >= 4096) return zb;
    return 1;
}

int jacobi_planes(int nx, int ny, int hz) { return planes_for((long)(nx / 64) * ((ny + 3) / 4), hz); }

int jacobi_blocks(int nx, int ny, int hz) {
    const int zb = jacobi_planes(nx, ny, hz);
    return (nx / 64) * ((ny + 3) / 4) * ((hz + zb - 1) / zb);
}

// Sweeps without the residual on grids whose one-plane sweep has fewer than 4096 64 x 4
// workgroups run 64 x 8 ones: half as many, each twice the work, which such a grid's
// sweep time (bound by its launch and ramp rather than its bytes) favours — 64x64x200:
// 4.78 -> 4.40 us per sweep; at 128x128x100 and above no gain.  Every voxel's formula is
// the same, so the result is too.
template <int ZB>
static void jacobi_launch(float *Tf, const float *Ti, const float *Q, int nx, int ny, int hz, const JacobiConst &c,
                          double *part, hipStream_t st) {
    const dim3 grid(nx / 64, (ny + 3) / 4, (hz + ZB - 1) / ZB);
    if (part) {
        hipLaunchKernelGGL((k_jacobi<ZB, true>), grid, dim3(256), 0, st, Tf, Ti, Q, nx, ny, hz, c, part);
    } else if (ZB == 1 && (long)grid.x * grid.y * grid.z < 4096) {
        const dim3 grid8(nx / 64, (ny + 7) / 8, hz);
        hipLaunchKernelGGL((k_jacobi<1, false, 8>), grid8, dim3(512), 0, st, Tf, Ti, Q, nx, ny, hz, c, part);
    } else {
        hipLaunchKernelGGL((k_jacobi<ZB, false>), grid, dim3(256), 0, st, Tf, Ti, Q, nx, ny, hz, c, part);
    }
}

hipError_t launch_jacobi(float *Tf, const float *Ti, const float *Q, int nx, int ny, int hz, const JacobiConst &c,
                         double *part, int nparts, hipStream_t st) {
    if (nx % 64 != 0 || hz < 2 || ny < 2) return hipErrorInvalidValue;
    if (part && nparts < jacobi_blocks(nx, ny, hz)) return hipErrorInvalidValue;
    switch (jacobi_planes(nx, ny, hz)) {
        case 8: jacobi_launch<8>(Tf, Ti, Q, nx, ny, hz, c, part, st); break;
        case 4: jacobi_launch<4>(Tf, Ti, Q, nx, ny, hz, c, part, st); break;
        case 2: jacobi_launch<2>(Tf, Ti, Q, nx, ny, hz, c, part, st); break;
        default: jacobi_launch<1>(Tf, Ti, Q, nx, ny, hz, c, part, st); break;
    }
    return hipGetLastError();
}

// Layouts of the two-sweep pass: a tile shape and how many waves of workgroups the device
// runs it in.  The z-chunk depth is chosen so that the workgroups (tiles x chunks) fill
// that many waves of the device's resident slots (CUs x workgroups per CU, queried for
// the kernel): with one chunk depth for every grid, a last partial wave of workgroups ran
// nearly alone and cost as much as a full one.
static constexpr int kJ2Shape[3][2] = {{64, 16}, {64, 8}, {128, 8}};
static constexpr int kJ2Waves[3] = {1, 2, 4};

template <int TX, int TY>
static int j2_slots() {
    int dev = 0, cus = 0, per_cu = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void *)k_jacobi2<TX, TY, false>, 256, 0) !=
        hipSuccess)
        return 0;
    return cus * per_cu;
}

int jacobi2_plan(int nx, int ny, int hz, int variant, Jacobi2Plan *plan, int count) {
    if (variant < 0 || variant >= kJacobi2VariantCount || !plan || nx < 2 || ny < 2 || hz < 2 || count < 1) return -1;
    if ((int64_t)nx * ny * 4 > INT32_MAX) return -1;   // a plane is one buffer descriptor's range
    const int shape = variant / 3, waves = kJ2Waves[variant % 3];
    const int tx = kJ2Shape[shape][0], ty = kJ2Shape[shape][1];
    const int slots = shape == 0 ? j2_slots<64, 16>() : shape == 1 ? j2_slots<64, 8>() : j2_slots<128, 8>();
    if (slots <= 0) return -1;
    Jacobi2Plan p{};
    p.shape = shape;
    p.tiles_x = (nx + tx - 1) / tx;
    const int tiles = p.tiles_x * ((ny + ty - 1) / ty);
    const int chunks = std::max(1, waves * slots / (tiles * count));
    p.zc = std::max(4, (hz + chunks - 1) / chunks);
    p.nchunks = (hz + p.zc - 1) / p.zc;
    p.blocks = tiles * p.nchunks;
    *plan = p;
    return 0;
}

template <int TX, int TY>
static hipError_t j2_launch(float *Tf, const float *Ti, const float *Q, int nx, int ny, int hz, const Jacobi2Plan &p,
                            const JacobiConst &c, double *part, hipStream_t st, int nids, const int *ids) {
    const size_t stride = (size_t)nx * ny * hz;
    if (part)
        hipLaunchKernelGGL((k_jacobi2<TX, TY, true>), dim3(p.blocks, nids), dim3(256), 0, st, Tf, Ti, Q, nx, ny, hz,
                           p.zc, p.tiles_x, p.nchunks, c, part, ids, stride);
    else
        hipLaunchKernelGGL((k_jacobi2<TX, TY, false>), dim3(p.blocks, nids), dim3(256), 0, st, Tf, Ti, Q, nx, ny, hz,
                           p.zc, p.tiles_x, p.nchunks, c, part, ids, stride);
    return hipGetLastError();
}

hipError_t launch_jacobi2(float *Tf, const float *Ti, const float *Q, int nx, int ny, int hz, const JacobiConst &c,
                          double *part, int nparts, hipStream_t st, const Jacobi2Plan &p) {
    if (p.blocks <= 0 || p.zc < 1 || hz < 2 || ny < 2 || nx < 2) return hipErrorInvalidValue;
    if (p.nchunks * p.zc < hz || p.tiles_x * kJ2Shape[p.shape][0] < nx) return hipErrorInvalidValue;
    if (part && nparts < p.blocks) return hipErrorInvalidValue;
    switch (p.shape) {
        case 0: return j2_launch<64, 16>(Tf, Ti, Q, nx, ny, hz, p, c, part, st, 1, nullptr);
        case 1: return j2_launch<64, 8>(Tf, Ti, Q, nx, ny, hz, p, c, part, st, 1, nullptr);
        case 2: return j2_launch<128, 8>(Tf, Ti, Q, nx, ny, hz, p, c, part, st, 1, nullptr);
        default: return hipErrorInvalidValue;
    }
}

hipError_t launch_jacobi2_batch(float *Tf, const float *Ti, const float *Q, int nx, int ny, int hz, int nids,
                                const int *ids, const JacobiConst &c, hipStream_t st, const Jacobi2Plan &p) {
    if (p.blocks <= 0 || p.zc < 1 || hz < 2 || ny < 2 || nx < 2 || nids < 1 || nids > 65535 || !ids)
        return hipErrorInvalidValue;
    if (p.nchunks * p.zc < hz || p.tiles_x * kJ2Shape[p.shape][0] < nx) return hipErrorInvalidValue;
    if ((int64_t)p.blocks * nids > INT32_MAX) return hipErrorInvalidValue;
    switch (p.shape) {
        case 0: return j2_launch<64, 16>(Tf, Ti, Q, nx, ny, hz, p, c, nullptr, st, nids, ids);
        case 1: return j2_launch<64, 8>(Tf, Ti, Q, nx, ny, hz, p, c, nullptr, st, nids, ids);
        case 2: return j2_launch<128, 8>(Tf, Ti, Q, nx, ny, hz, p, c, nullptr, st, nids, ids);
        default: return hipErrorInvalidValue;
    }
}

template <int ZB>
static void jacobi_batch_launch(float *Tf, const float *Ti, const float *Q, int nx, int ny, int hz, int nids,
                                const int *ids, const JacobiConst &c, double *part, hipStream_t st) {
    const dim3 grid((unsigned)((nx / 64) * ((ny + 3) / 4) * ((hz + ZB - 1) / ZB)), (unsigned)nids);
    if (part)
        hipLaunchKernelGGL((k_jacobi_batch<ZB, true>), grid, dim3(256), 0, st, Tf, Ti, Q, nx, ny, hz, ids, c, part);
    else
        hipLaunchKernelGGL((k_jacobi_batch<ZB, false>), grid, dim3(256), 0, st, Tf, Ti, Q, nx, ny, hz, ids, c, part);
}

hipError_t launch_jacobi_batch(float *Tf, const float *Ti, const float *Q, int nx, int ny, int hz, int nids,
                               const int *ids, const JacobiConst &c, double *part, hipStream_t st) {
    if (nx % 64 != 0 || hz < 2 || ny < 2 || nids < 1 || nids > 65535 || !ids) return hipErrorInvalidValue;
    // a sweep with residual partials takes a single instance's plane count, so that each
    // instance's partials come as its own context's; the others count every instance's
    // workgroups (the swept temperatures are the same whatever the count)
    const long cols = (long)(nx / 64) * ((ny + 3) / 4);
    switch (part ? jacobi_planes(nx, ny, hz) : planes_for(cols * nids, hz)) {
        case 8: jacobi_batch_launch<8>(Tf, Ti, Q, nx, ny, hz, nids, ids, c, part, st); break;
        case 4: jacobi_batch_launch<4>(Tf, Ti, Q, nx, ny, hz, nids, ids, c, part, st); break;
        case 2: jacobi_batch_launch<2>(Tf, Ti, Q, nx, ny, hz, nids, ids, c, part, st); break;
        default: jacobi_batch_launch<1>(Tf, Ti, Q, nx, ny, hz, nids, ids, c, part, st); break;
    }
    return hipGetLastError();
}

hipError_t launch_sum_partials_batch(const double *part, int per_instance, int count, double *out, hipStream_t st) {
    hipLaunchKernelGGL(k_sum_partials_batch, dim3(count), dim3(256), 0, st, part, per_instance, out);
    return hipGetLastError();
}

hipError_t launch_sum_partials(const double *part, int nparts, double *out, hipStream_t st) {
    hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(256), 0, st, part, nparts, out);
    return hipGetLastError();
}

hipError_t launch_index_table(const float *T, float2 *tab, int64_t n, const ThermConst &k, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_index_table, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, T, tab, n, k);
    return hipGetLastError();
}

template <int N, int E, int W>
static hipError_t therm_row_launch(float4 *U, float4 *Uout, const float2 *tw, const float2 *NT, float *Q,
                                   float *DKint, const ThermRowArgs &a, hipStream_t st, int count) {
    using C = RowCfgT<N, E, W>;
    hipLaunchKernelGGL((k_therm_row_pass<N, E, W>), dim3(N / C::ROWS, count), dim3(C::THREADS), 0, st, U, Uout, tw, NT, Q,
                       DKint, a);
    return hipGetLastError();
}

template <int N, int E, int W>
static hipError_t therm_wide_launch(float4 *U, float4 *Uout, const float2 *tw, const float2 *NT, float *Q,
                                    float *DKint, const ThermRowArgs &a, hipStream_t st, int count) {
    hipLaunchKernelGGL((k_therm_row_wide<N, E, W>), dim3(N, count), dim3(64 * W), 0, st, U, Uout, tw, NT, Q, DKint, a);
    return hipGetLastError();
}

template <int N, int E, int W>
static hipError_t therm_split_launch(float4 *U, float4 *Uout, const float2 *tw, const float2 *NT, float *Q,
                                     float *DKint, const ThermRowArgs &a, hipStream_t st, int count) {
    hipLaunchKernelGGL((k_therm_row_split<N, E, W>), dim3(N, count), dim3(64 * W), 0, st, U, Uout, tw, NT, Q, DKint, a);
    return hipGetLastError();
}

#define THRV(N, E, W) return therm_row_launch<N, E, W>(U, Uout, tw, NT, Q, DKint, a, st, count)
#define THRS(N, E, W) return therm_split_launch<N, E, W>(U, Uout, tw, NT, Q, DKint, a, st, count)
#define THRW(N, E, W) return therm_wide_launch<N, E, W>(U, Uout, tw, NT, Q, DKint, a, st, count)

hipError_t launch_thermal_row_pass(int n, float4 *U, float4 *Uout, const float2 *tw, const float2 *NT, float *Q,
                                   float *DKint, const ThermRowArgs &a, hipStream_t st, int variant, int count) {
    if (variant < 0 || variant >= kThermRowVariantCount || count < 1 || (count > 1 && !a.inst_cells))
        return hipErrorInvalidValue;
    switch (n * 16 + variant) {
        case 64 * 16 + 0: THRV(64, 8, 1);
        case 64 * 16 + 1: THRV(64, 4, 1);
        case 64 * 16 + 2: THRV(64, 16, 1);
        case 64 * 16 + 3: THRV(64, 8, 2);
        case 64 * 16 + 4: THRW(64, 4, 1);
        case 64 * 16 + 5: THRW(64, 8, 1);
        case 64 * 16 + 6: THRS(64, 4, 2);
        case 128 * 16 + 0: THRV(128, 16, 1);
        case 128 * 16 + 1: THRV(128, 8, 1);
        case 128 * 16 + 2: THRV(128, 4, 1);
        case 128 * 16 + 3: THRV(128, 4, 2);
        case 128 * 16 + 4: THRW(128, 4, 2);
        case 128 * 16 + 5: THRW(128, 4, 1);
        case 128 * 16 + 6: THRS(128, 4, 2);
        case 256 * 16 + 0: THRV(256, 16, 1);
        case 256 * 16 + 1: THRV(256, 8, 1);
        case 256 * 16 + 2: THRV(256, 4, 1);
        case 256 * 16 + 3: THRV(256, 4, 4);
        case 256 * 16 + 4: THRW(256, 4, 4);
        case 256 * 16 + 5: THRW(256, 4, 2);
        case 256 * 16 + 6: THRS(256, 4, 4);
        case 512 * 16 + 0: THRV(512, 8, 1);
        case 512 * 16 + 1: THRV(512, 4, 2);
        case 512 * 16 + 2: THRV(512, 16, 1);
        case 512 * 16 + 3: THRV(512, 8, 2);
        case 512 * 16 + 4: THRW(512, 8, 8);
        case 512 * 16 + 5: THRW(512, 8, 4);
        case 512 * 16 + 6: THRS(512, 8, 8);
        default: return hipErrorInvalidValue;
    }
}
#undef THRV
#undef THRW
#undef THRS

}  // namespace shg
