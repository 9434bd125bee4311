// pass_cfg.h — thread layouts of the row and column passes.
#pragma once
#include "fft_lds.h"

namespace shg {

// Row pass: one wave per workgroup, T threads per row, ROWS rows per workgroup.
template <int N> struct RowCfg {
    static constexpr int E = FftPlan<N>::E;
    static constexpr int T = N / E;               // threads per row
    static constexpr int THREADS = 64;            // one wave per workgroup
    static constexpr int ROWS = THREADS / T;      // rows per workgroup
    static constexpr int L = LdsLen<N>::value;
};

// Column pass: CW adjacent columns per workgroup (one 64- or 128-byte segment per row),
// T threads per column.
template <int N> struct ColCfg {
    static constexpr int E = FftPlan<N>::E;
    static constexpr int T = N / E;               // threads per column
    static constexpr int CW = N == 512 ? 4 : 8;   // columns per workgroup
    static constexpr int THREADS = CW * T;
    static constexpr int L = LdsLen<N>::value;
};

}  // namespace shg
