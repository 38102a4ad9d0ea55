// Compile-time specialised FFT building blocks for gfx950.
//
// RDft<M, INV>  : M-point DFT held entirely in registers (M <= ~25), unrolled with every
//                 index and twiddle a compile-time constant.  Composite M is split
//                 decimation-in-time, M = P * Q: Q-point DFTs of the P interleaved
//                 subsequences, twiddles w_M^(p k), then P-point DFTs.
// four_step     : N = N1 * N2 point transform of many lines in LDS with one exchange:
//                 X[k1 + N1 k2] = sum_j2 w_N2^(j2 k2) w_N^(j2 k1) sum_j1 x[j2 + N2 j1] w_N1^(j1 k1)
//                 thread (line, j2) does the N1-point DFTs, thread (line, k1) the N2-point.
// LDS layout is element-major, [element][line] with a padded line stride, so that the
// lanes of a wavefront (consecutive lines) touch consecutive words.
#pragma once
#include <hip/hip_runtime.h>

#include <type_traits>

#include "mdk_fft.h"

namespace rf {

// ---------------------------------------------------------------- constexpr trig
constexpr double kPi = 3.141592653589793238462643383279502884;

constexpr double wrap(double x) {
    while (x > kPi) x -= 2.0 * kPi;
    while (x < -kPi) x += 2.0 * kPi;
    return x;
}
constexpr double csin(double x) {
    x = wrap(x);
    double term = x, sum = x;
    for (int i = 1; i < 40; i++) {
        term *= -x * x / ((2.0 * i) * (2.0 * i + 1.0));
        sum += term;
    }
    return sum;
}
constexpr double ccos(double x) {
    x = wrap(x);
    double term = 1.0, sum = 1.0;
    for (int i = 1; i < 40; i++) {
        term *= -x * x / ((2.0 * i - 1.0) * (2.0 * i));
        sum += term;
    }
    return sum;
}

// forward twiddle exp(-2 pi i E / M)
template <int M, int E> struct Tw {
    static constexpr int e = ((E % M) + M) % M;
    static constexpr float re = (float)ccos(2.0 * kPi * e / M);
    static constexpr float im = (float)(-csin(2.0 * kPi * e / M));
};

template <int B, int E, class F> __device__ __forceinline__ void sfor(F &&f) {
    if constexpr (B < E) {
        f(std::integral_constant<int, B>{});
        sfor<B + 1, E>(f);
    }
}

template <int M, int E, bool INV> __device__ __forceinline__ float2 twiddle(float2 x) {
    constexpr int e = ((E % M) + M) % M;
    if constexpr (e == 0) {
        return x;
    } else if constexpr (4 * e == M) {  // -i (forward) / +i (inverse)
        return INV ? make_float2(-x.y, x.x) : make_float2(x.y, -x.x);
    } else if constexpr (2 * e == M) {
        return make_float2(-x.x, -x.y);
    } else if constexpr (4 * e == 3 * M) {  // +i (forward) / -i (inverse)
        return INV ? make_float2(x.y, -x.x) : make_float2(-x.y, x.x);
    } else {
        constexpr float wr = Tw<M, E>::re;
        constexpr float wi = INV ? -Tw<M, E>::im : Tw<M, E>::im;
        return make_float2(x.x * wr - x.y * wi, x.x * wi + x.y * wr);
    }
}

constexpr int pick_radix(int m) {
    if (m % 4 == 0 && m != 4 && m != 8) return 4;
    const int pr[] = {2, 3, 5, 7, 11, 13};
    for (int p : pr)
        if (m % p == 0) return p;
    return m;
}

constexpr bool is_base(int m) {
    return m == 1 || m == 2 || m == 3 || m == 4 || m == 5 || m == 7 || m == 8 || m == 11 || m == 13;
}

template <int M, bool INV, bool BASE = is_base(M)> struct RDft;

template <bool INV> struct RDft<1, INV, true> {
    static __device__ __forceinline__ void run(float2 *) {}
};

template <int M, bool INV> struct RDft<M, INV, true> {
    static __device__ __forceinline__ void run(float2 *v) { Dft<M, INV>::run(v); }
};

template <int M, bool INV> struct RDft<M, INV, false> {
    static constexpr int P = pick_radix(M);
    static constexpr int Q = M / P;
    static_assert(P * Q == M && P > 1 && Q > 1, "bad split");
    static __device__ __forceinline__ void run(float2 *v) {
        float2 y[M];
        sfor<0, P>([&](auto pc) {
            constexpr int p = decltype(pc)::value;
            float2 u[Q];
            sfor<0, Q>([&](auto qc) {
                constexpr int q = decltype(qc)::value;
                u[q] = v[p + P * q];
            });
            RDft<Q, INV>::run(u);
            sfor<0, Q>([&](auto kc) {
                constexpr int k = decltype(kc)::value;
                y[p * Q + k] = twiddle<M, p * k, INV>(u[k]);
            });
        });
        sfor<0, Q>([&](auto kc) {
            constexpr int k = decltype(kc)::value;
            float2 z[P];
            sfor<0, P>([&](auto pc) {
                constexpr int p = decltype(pc)::value;
                z[p] = y[p * Q + k];
            });
            RDft<P, INV>::run(z);
            sfor<0, P>([&](auto sc) {
                constexpr int s = decltype(sc)::value;
                v[k + Q * s] = z[s];
            });
        });
    }
};

// ---------------------------------------------------------------- line geometry
constexpr int split_n2(int n) {
    int best = 1;
    for (int d = 1; d * d <= n; d++)
        if (n % d == 0) best = d;
    return best;
}

template <int N> struct Geo {
    static constexpr int N2 = split_n2(N);  // <= sqrt(N)
    static constexpr int N1 = N / N2;       // >= sqrt(N)
    static constexpr int T = N1;            // threads per line
    static_assert(N1 <= 32, "line too long for the register-resident split");
};

// General strides: element e of line `line` is S[e * ES + line * LST]; the N-point
// twiddle w_N^m is tw[m * TWS] (TWS > 1 when tw is a table for a multiple of N).
// PAD: PAD spare elements after each group of N2 (the exchange then spans N + N1 PAD
// elements); it moves the groups of the k1 = 0, 1, ... threads onto other LDS banks.
template <int N, bool INV, int PAD = 0>
__device__ __forceinline__ void step1_g(float2 *v, float2 *S, const float2 *tw, int TWS, int ES, int LST, int line,
                                        int j2) {
    constexpr int N1 = Geo<N>::N1, N2 = Geo<N>::N2;
    RDft<N1, INV>::run(v);
    sfor<0, N1>([&](auto kc) {
        constexpr int k1 = decltype(kc)::value;
        float2 x = v[k1];
        if constexpr (k1 != 0) {
            float2 w = tw[j2 * k1 * TWS];
            if (INV) w.y = -w.y;
            x = make_float2(x.x * w.x - x.y * w.y, x.x * w.y + x.y * w.x);
        }
        S[(k1 * (N2 + PAD) + j2) * ES + line * LST] = x;
    });
}

template <int N, bool INV, int PAD = 0>
__device__ __forceinline__ void step2_g(float2 *y, const float2 *S, int ES, int LST, int line, int k1) {
    constexpr int N2 = Geo<N>::N2;
    sfor<0, N2>([&](auto jc) {
        constexpr int j2 = decltype(jc)::value;
        y[j2] = S[(k1 * (N2 + PAD) + j2) * ES + line * LST];
    });
    RDft<N2, INV>::run(y);
}

// Step 1 of the four-step transform for one line, from values already in registers:
// v[j1] = x[j2 + N2 j1]; writes w_N^(j2 k1) * DFT_N1(v)[k1] to S[(k1 N2 + j2) LS + line].
template <int N, bool INV>
__device__ __forceinline__ void step1_store(float2 *v, float2 *S, const float2 *tw, int LS, int line, int j2) {
    constexpr int N1 = Geo<N>::N1, N2 = Geo<N>::N2;
    RDft<N1, INV>::run(v);
    sfor<0, N1>([&](auto kc) {
        constexpr int k1 = decltype(kc)::value;
        float2 x = v[k1];
        if constexpr (k1 != 0) {
            float2 w = tw[j2 * k1];
            if (INV) w.y = -w.y;
            x = make_float2(x.x * w.x - x.y * w.y, x.x * w.y + x.y * w.x);
        }
        S[(k1 * N2 + j2) * LS + line] = x;
    });
}

// Step 2: reads S[(k1 N2 + j2) LS + line] for all j2 and returns y[k2] = X[k1 + N1 k2].
template <int N, bool INV>
__device__ __forceinline__ void step2_load(float2 *y, const float2 *S, int LS, int line, int k1) {
    constexpr int N2 = Geo<N>::N2;
    sfor<0, N2>([&](auto jc) {
        constexpr int j2 = decltype(jc)::value;
        y[j2] = S[(k1 * N2 + j2) * LS + line];
    });
    RDft<N2, INV>::run(y);
}

// In-place four-step transform of the L lines held element-major in S (stride LS).
// Thread (line = tid % L, u = tid / L); all threads of the block must call it.
template <int N, int L, bool INV>
__device__ __forceinline__ void four_step_inplace(float2 *S, const float2 *tw, int LS, int line, int u,
                                                  bool active) {
    constexpr int N1 = Geo<N>::N1, N2 = Geo<N>::N2;
    float2 v[N1];
    const bool a1 = active && u < N2;
    if (a1)
        sfor<0, N1>([&](auto jc) {
            constexpr int j1 = decltype(jc)::value;
            v[j1] = S[(u + N2 * j1) * LS + line];
        });
    __syncthreads();
    if (a1) step1_store<N, INV>(v, S, tw, LS, line, u);
    __syncthreads();
    float2 y[N2];
    const bool a2 = active && u < N1;
    if (a2) step2_load<N, INV>(y, S, LS, line, u);
    __syncthreads();
    if (a2)
        sfor<0, N2>([&](auto kc) {
            constexpr int k2 = decltype(kc)::value;
            S[(u + N1 * k2) * LS + line] = y[k2];
        });
    __syncthreads();
}

}  // namespace rf
