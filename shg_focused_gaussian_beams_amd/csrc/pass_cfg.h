// pass_cfg.h — thread layouts of the row and column passes.
//
// Row pass <N, E, W>: W waves per workgroup, T = N/E threads per row, ROWS rows per
// workgroup.  Column pass <N, E, CW>: CW adjacent columns per workgroup (a CW*16-byte
// segment of every row), T threads per column.  E is the FFT's elements per thread.
#pragma once
#include "fft_lds.h"

namespace shg {

// Global-memory view of a pointer.  Buffers reached through pointers stored in memory
// (instance tables, argument structs) are generic (flat) addresses to the compiler, and
// flat accesses count against the LDS wait counter too, so every LDS wait of the FFT
// exchanges would also wait for the loads in flight.  The cast states the address space.
// (The host compilation pass, which generates no kernel code, sees plain pointers.)
#if defined(__HIP_DEVICE_COMPILE__)
template <class T> using gptr = __attribute__((address_space(1))) T *;
#else
template <class T> using gptr = T *;
#endif
template <class T> __device__ __forceinline__ gptr<T> gp(T *p) { return (gptr<T>)p; }

// Workgroups are handed to the 8 XCDs round-robin (XCD = linear workgroup id mod 8), and
// every XCD has its own L2.  Workgroups whose lines share 128-byte cache lines (adjacent
// column groups) are given consecutive logical indices on the same XCD: logical group
// (b mod 8) * (nb / 8) + b / 8 for workgroup b of nb (identity unless 8 divides nb).
__device__ __forceinline__ int xcd_group(int b, int nb) {
    return (nb & 7) ? b : (b & 7) * (nb >> 3) + (b >> 3);
}

template <int N, int E_, int W> struct RowCfgT {
    static constexpr int E = E_;
    static constexpr int T = N / E;
    static constexpr int THREADS = 64 * W;
    static constexpr int ROWS = THREADS / T;
    static constexpr int L = LdsLen<N>::value;
    static_assert(ROWS >= 1 && THREADS % T == 0, "a row must fit in the workgroup");
};

template <int N, int E_, int CW_> struct ColCfgT {
    static constexpr int E = E_;
    static constexpr int T = N / E;
    static constexpr int CW = CW_;
    static constexpr int THREADS = CW * T;
    static constexpr int L = LdsLen<N>::value;
    static_assert(THREADS <= 1024 && THREADS % 64 == 0, "column workgroup size");
};

// Defaults (used by the THERMAL, tape and adjoint passes).
template <int N> using RowCfg = RowCfgT<N, FftPlan<N>::E, 1>;
template <int N> using ColCfg = ColCfgT<N, FftPlan<N>::E, (N == 512 ? 4 : 8)>;

}  // namespace shg
