// fft_lds.h — register/LDS Stockham FFTs for CDNA4 (gfx950), wave64.
//
// One transform of length N is owned by T = N/E threads; thread t holds the E elements
// x[t + j*T], j = 0..E-1, in registers, before and after the transform (natural order).
// A stage of radix R does E/R butterflies per thread in registers; between stages the
// values go once through LDS (autosort Stockham), so an N = 256 transform (16 x 16) costs
// one LDS round trip and N = 512 (8 x 8 x 8) two.  Every call transforms two fields at
// once (the fundamental and the second harmonic share positions, twiddles and index math).
//
// Sign convention: SIGN = +1 is FFTW_BACKWARD (e^{+2 pi i jk/N}), SIGN = -1 is
// FFTW_FORWARD; both unnormalised, as the reference's fftwf plans
// (src/headers/Solver.h:161-187).
#pragma once
#ifndef SHG_HOST_EMULATION
#include <hip/hip_runtime.h>
#endif

namespace shg {

__device__ __forceinline__ float2 c_add(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 c_sub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 c_mul(float2 a, float2 b) {
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// cos/sin of 2*pi*m/16, m = 0..7
__device__ constexpr float kCos16[8] = {1.0f, 0.92387953251128674f, 0.70710678118654752f, 0.38268343236508977f,
                                        0.0f, -0.38268343236508977f, -0.70710678118654752f, -0.92387953251128674f};
__device__ constexpr float kSin16[8] = {0.0f, 0.38268343236508977f, 0.70710678118654752f, 0.92387953251128674f,
                                        1.0f, 0.92387953251128674f, 0.70710678118654752f, 0.38268343236508977f};

// d * exp(SIGN * 2 pi i m / 16) with the trivial rotations spelled out.
template <int SIGN>
__device__ __forceinline__ float2 rot16(float2 d, int m) {
    if (m == 0) return d;
    if (m == 4) return SIGN > 0 ? make_float2(-d.y, d.x) : make_float2(d.y, -d.x);
    if (m == 2) {
        const float h = 0.70710678118654752f;
        return SIGN > 0 ? make_float2(h * (d.x - d.y), h * (d.y + d.x)) : make_float2(h * (d.x + d.y), h * (d.y - d.x));
    }
    if (m == 6) {
        const float h = 0.70710678118654752f;
        return SIGN > 0 ? make_float2(-h * (d.x + d.y), h * (d.x - d.y)) : make_float2(h * (d.y - d.x), -h * (d.x + d.y));
    }
    const float c = kCos16[m], s = SIGN > 0 ? kSin16[m] : -kSin16[m];
    return make_float2(c * d.x - s * d.y, c * d.y + s * d.x);
}

__host__ __device__ constexpr int bitrev(int k, int R) {
    int r = 0;
    for (int b = R >> 1; b > 0; b >>= 1, k >>= 1) r = (r << 1) | (k & 1);
    (void)r;
    return r;
}

// In-register DFT of size R (2..16), natural order in and out: radix-2 decimation in
// frequency with compile-time twiddles, then a compile-time bit-reversal rename.
template <int R, int SIGN>
__device__ __forceinline__ void dft(float2 (&x)[R]) {
#pragma unroll
    for (int len = R; len >= 2; len >>= 1) {
        const int half = len >> 1;
#pragma unroll
        for (int base = 0; base < R; base += len) {
#pragma unroll
            for (int j = 0; j < half; j++) {
                const float2 u = x[base + j], v = x[base + j + half];
                x[base + j] = c_add(u, v);
                x[base + j + half] = rot16<SIGN>(c_sub(u, v), j * (16 / len));
            }
        }
    }
    float2 y[R];
#pragma unroll
    for (int k = 0; k < R; k++) y[k] = x[bitrev(k, R)];
#pragma unroll
    for (int k = 0; k < R; k++) x[k] = y[k];
}

// LDS element index with one pad slot every 16 entries (breaks power-of-two strides).
__device__ __forceinline__ int lds_pad(int o) { return o + (o >> 4); }
template <int N> struct LdsLen { static constexpr int value = N + N / 16; };

// One Stockham stage, in registers.  Butterfly i = t + m*T takes x[i + r*N/R] (register
// m + r*M), twiddles it by w^(r*k) with k = i mod Ns and leaves y_r in register m + r*M.
// `tw` is the LDS table W_N^j = exp(-2 pi i j/N).
// TWO = false transforms `a` only (`b` is neither read nor written).
template <int N, int E, int R, int Ns, int SIGN, bool TWO = true>
__device__ __forceinline__ void stage_compute(float2 (&a)[E], float2 (&b)[E], int t, const float2 *tw) {
    constexpr int T = N / E;
    constexpr int M = E / R;
#pragma unroll
    for (int m = 0; m < M; m++) {
        const int i = t + m * T;
        const int k = i & (Ns - 1);
        float2 xa[R], xb[R];
#pragma unroll
        for (int r = 0; r < R; r++) {
            xa[r] = a[m + r * M];
            if constexpr (TWO) xb[r] = b[m + r * M];
        }
        if (Ns > 1) {
#pragma unroll
            for (int r = 1; r < R; r++) {
                const int e = (r * k * (N / (Ns * R))) & (N - 1);
                float2 w = tw[e];
                if (SIGN > 0) w.y = -w.y;
                xa[r] = c_mul(xa[r], w);
                if constexpr (TWO) xb[r] = c_mul(xb[r], w);
            }
        }
        dft<R, SIGN>(xa);
        if constexpr (TWO) dft<R, SIGN>(xb);
#pragma unroll
        for (int r = 0; r < R; r++) {
            a[m + r * M] = xa[r];
            if constexpr (TWO) b[m + r * M] = xb[r];
        }
    }
}

// Writes a stage's outputs to their autosort positions (i-k)*R + k + r*Ns.  After the
// last stage that position is t + (m + r*M)*T, i.e. the registers are already in order.
template <int N, int E, int R, int Ns, bool TWO = true, typename P>
__device__ __forceinline__ void stage_scatter(const float2 (&a)[E], const float2 (&b)[E], P la, P lb, int t, int st) {
    constexpr int T = N / E;
    constexpr int M = E / R;
#pragma unroll
    for (int m = 0; m < M; m++) {
        const int i = t + m * T;
        const int k = i & (Ns - 1);
        const int ob = (i - k) * R + k;
#pragma unroll
        for (int r = 0; r < R; r++) {
            const int o = lds_pad(ob + r * Ns) * st;
            la[o] = a[m + r * M];
            if constexpr (TWO) lb[o] = b[m + r * M];
        }
    }
}

template <int N, int E, bool TWO = true, typename P>
__device__ __forceinline__ void lds_gather(float2 (&a)[E], float2 (&b)[E], P la, P lb, int t, int st) {
    constexpr int T = N / E;
#pragma unroll
    for (int j = 0; j < E; j++) {
        const int o = lds_pad(t + j * T) * st;
        a[j] = la[o];
        if constexpr (TWO) b[j] = lb[o];
    }
}

// Per-length plans: elements per thread (the radix sequence is spelled out in fft2f).
template <int N> struct FftPlan;
template <> struct FftPlan<64>  { static constexpr int E = 8;  };   // 8 x 8
template <> struct FftPlan<128> { static constexpr int E = 16; };   // 16 x 8
template <> struct FftPlan<256> { static constexpr int E = 16; };   // 16 x 16
template <> struct FftPlan<512> { static constexpr int E = 8;  };   // 8 x 8 x 8

// Exchange between two stages: barrier (earlier readers of this LDS are done), scatter,
// barrier, gather.
template <int N, int E, int R, int Ns, bool TWO = true, typename P, typename Sync>
__device__ __forceinline__ void exchange(float2 (&a)[E], float2 (&b)[E], P la, P lb, int t, int st, Sync sync) {
    sync();
    stage_scatter<N, E, R, Ns, TWO>(a, b, la, lb, t, st);
    sync();
    lds_gather<N, E, TWO>(a, b, la, lb, t, st);
}

// Radix plan for length N with E elements per thread: radix E while it fits, then the
// remainder (N = E^k * r).  Stage s has stride Ns = product of the earlier radices.
__host__ __device__ constexpr int plan_radix(int N, int E, int s) {
    int ns = 1;
    for (int i = 0; i < s; i++) ns *= (N / ns >= E ? E : N / ns);
    return N / ns >= E ? E : N / ns;
}
__host__ __device__ constexpr int plan_stride(int N, int E, int s) {
    int ns = 1;
    for (int i = 0; i < s; i++) ns *= plan_radix(N, E, i);
    return ns;
}

template <int N, int E, int S, int SIGN, bool TWO = true, typename P, typename Sync>
__device__ __forceinline__ void fft_stages(float2 (&a)[E], float2 (&b)[E], P la, P lb, int t, int st,
                                           const float2 *tw, Sync sync) {
    constexpr int R = plan_radix(N, E, S), Ns = plan_stride(N, E, S);
    static_assert(R >= 2 && E % R == 0, "radix must divide the elements per thread");
    stage_compute<N, E, R, Ns, SIGN, TWO>(a, b, t, tw);
    if constexpr (Ns * R < N) {
        exchange<N, E, R, Ns, TWO>(a, b, la, lb, t, st, sync);
        fft_stages<N, E, S + 1, SIGN, TWO>(a, b, la, lb, t, st, tw, sync);
    }
}

// Full length-N transform of two fields held as described at the top of this file, with
// E elements per thread (E in {4, 8, 16}; T = N/E threads).  `sync` is the barrier of the
// caller's thread grouping; `la`/`lb` address the LDS of the two fields (element o of this
// transform at la[pad(o)*st]).
template <int N, int E, int SIGN, typename P, typename Sync>
__device__ __forceinline__ void fft2fe(float2 (&a)[E], float2 (&b)[E], P la, P lb, int t, int st, const float2 *tw,
                                       Sync sync) {
    static_assert(N >= 64 && N <= 512 && (N & (N - 1)) == 0, "supported lengths: 64..512, powers of two");
    fft_stages<N, E, 0, SIGN>(a, b, la, lb, t, st, tw, sync);
}

// The same for one field.
template <int N, int E, int SIGN, typename P, typename Sync>
__device__ __forceinline__ void fft1fe(float2 (&a)[E], P la, int t, int st, const float2 *tw, Sync sync) {
    static_assert(N >= 64 && N <= 512 && (N & (N - 1)) == 0, "supported lengths: 64..512, powers of two");
    fft_stages<N, E, 0, SIGN, false>(a, a, la, la, t, st, tw, sync);
}

// Ordering point for LDS exchanges inside one wavefront (a transform run by at most 64
// lanes of one wave): a wave's LDS operations complete in issue order, so only the
// compiler's reordering has to be fenced.
#ifndef SHG_HOST_EMULATION
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}
#endif

#ifndef SHG_HOST_EMULATION
// In-place transform of one field held in natural order in LDS (l[pad(o)], o < N) by the
// T = N/E lanes t of one wavefront: gather x[t + j T], transform, write back in order.
template <int N, int E, int SIGN>
__device__ __forceinline__ void fft_lds_field(float2 *l, int t, const float2 *tw) {
    constexpr int T = N / E;
    static_assert(T <= 64, "one wavefront per transform");
    float2 x[E];
#pragma unroll
    for (int j = 0; j < E; j++) x[j] = l[lds_pad(t + j * T)];
    fft1fe<N, E, SIGN>(x, l, t, 1, tw, [] { wave_sync(); });
    wave_sync();
#pragma unroll
    for (int j = 0; j < E; j++) l[lds_pad(t + j * T)] = x[j];
}
#endif

// The default plan of each length (FftPlan<N>::E elements per thread).
template <int N, int SIGN, typename P, typename Sync>
__device__ __forceinline__ void fft2f(float2 (&a)[FftPlan<N>::E], float2 (&b)[FftPlan<N>::E], P la, P lb, int t,
                                      int st, const float2 *tw, Sync sync) {
    fft2fe<N, FftPlan<N>::E, SIGN>(a, b, la, lb, t, st, tw, sync);
}

}  // namespace shg
