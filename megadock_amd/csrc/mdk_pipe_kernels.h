// FFT-side docking pipeline specialised at compile time per grid size: the kernels and
// the per-size driver run_pipe.  Same data flow and buffers as the generic kernels in
// mdk_dock.hip (see there), but every 1D transform is a four-step register FFT
// (mdk_regfft.h) with a single LDS exchange.  Included by mdk_pipe.hip (the tuned sizes,
// with every kernel variant) and by mdk_pipe_size.hip (one further size per object file,
// default variants only); everything here has internal linkage.
#pragma once
#include <hip/hip_runtime.h>

#include <cmath>

#include "mdk_internal.h"
#include "mdk_regfft.h"

namespace {

using rf::Geo;
using rf::sfor;

// Global accesses as a block-uniform base plus a 32-bit byte offset per lane, so that
// the address is formed by the load/store itself (scalar base + vector offset) instead
// of 64-bit vector arithmetic per access.
template <class T> __device__ __forceinline__ T ld_at(const T *base, unsigned byte_off) {
    return *reinterpret_cast<const T *>(reinterpret_cast<const char *>(base) + byte_off);
}
template <class T> __device__ __forceinline__ void st_at(T *base, unsigned byte_off, const T &v) {
    *reinterpret_cast<T *>(reinterpret_cast<char *>(base) + byte_off) = v;
}


// ------------------------------------------------------------------ forward z
// lines (b, x, y), x, y < ng: the ng ligand voxels of the row sit at z = nd .. nd+ng-1;
// their values come straight from the voxel masks and the charge grid into the step-1
// registers.  Thread (line, u) with u fastest, so a line's voxels go to neighbouring lanes.
template <int N, int L>
__global__ void __launch_bounds__(L *Geo<N>::T) kp_fwd_z(VoxGrid vg, float2 *__restrict__ a1,
                                                          const float2 *__restrict__ twg, int nlines) {
    constexpr int N1 = Geo<N>::N1, N2 = Geo<N>::N2, T = Geo<N>::T, LS = L + 1, NT = L * T;
    constexpr int NG = N / 2, ND = N / 4;  // ligand voxels per axis and their offset in the line
    __shared__ float2 S[N * LS];
    __shared__ float2 tw[N];
    const int tid = threadIdx.x, line = tid / T, u = tid - line * T;
    for (int q = tid; q < N; q += NT) tw[q] = twg[q];
    const int gl = blockIdx.x * L + line;
    const bool act = gl < nlines;
    const int b = gl / (NG * NG), row = gl - b * (NG * NG);
    __syncthreads();
    if (u < N2) {
        float2 v[N1];
        sfor<0, N1>([&](auto jc) {
            constexpr int j1 = decltype(jc)::value;
            const int e = u + N2 * j1;
            v[j1] = (act && e >= ND && e < ND + NG) ? vox_take(vg, b, row, e - ND) : make_float2(0.f, 0.f);
        });
        rf::step1_store<N, false>(v, S, tw, LS, line, u);
    }
    __syncthreads();
    if (u < N1) {
        float2 y[N2];
        rf::step2_load<N, false>(y, S, LS, line, u);
        float2 *dst = a1 + (size_t)blockIdx.x * L * N;  // uniform base, 32-bit lane offsets
        if (act)
            sfor<0, N2>([&](auto kc) {
                constexpr int k2 = decltype(kc)::value;
                st_at(dst, 8u * (unsigned)(line * N + u + N1 * k2), y[k2]);
            });
    }
}

// ------------------------------------------------------------------ forward y
// block (kz chunk, x, b); rows y < ng of a1 sit at y + nd; loads straight to registers
template <int N, int L>
__global__ void __launch_bounds__(L *Geo<N>::T) kp_fwd_y(const float2 *__restrict__ a1, float2 *__restrict__ a2,
                                                          const float2 *__restrict__ twg) {
    constexpr int N1 = Geo<N>::N1, N2 = Geo<N>::N2, T = Geo<N>::T, LS = L + 1, NT = L * T;
    constexpr int NG = N / 2, ND = N / 4;
    __shared__ float2 S[N * LS];
    __shared__ float2 tw[N];
    const int tid = threadIdx.x, line = tid % L, u = tid / L;
    for (int q = tid; q < N; q += NT) tw[q] = twg[q];
    const int kz = blockIdx.x * L + line, x = blockIdx.y, b = blockIdx.z;
    const bool col = kz < N;
    const float2 *src = a1 + ((size_t)b * NG + x) * NG * N;
    __syncthreads();
    if (u < N2) {
        float2 v[N1];
        sfor<0, N1>([&](auto jc) {
            constexpr int j1 = decltype(jc)::value;
            const int y = u + N2 * j1 - ND;
            v[j1] = (col && y >= 0 && y < NG) ? ld_at(src, 8u * (unsigned)(y * N + kz)) : make_float2(0.f, 0.f);
        });
        rf::step1_store<N, false>(v, S, tw, LS, line, u);
    }
    __syncthreads();
    if (u < N1) {
        float2 y[N2];
        rf::step2_load<N, false>(y, S, LS, line, u);
        float2 *dst = a2 + ((size_t)b * NG + x) * N * N;
        if (col)
            sfor<0, N2>([&](auto kc) {
                constexpr int k2 = decltype(kc)::value;
                st_at(dst, 8u * (unsigned)((u + N1 * k2) * N + kz), y[k2]);
            });
    }
}

// ------------------------------------------------------------------ forward z + y, fused
// Persistent workgroups walking the (b, x) planes, x < ng, of the ligand.  Per plane:
//   z-FFT of the ng rows y (voxel values straight from the masks and the charge grid into
//   the step-1 registers), kept in LDS as A[kz][y];
//   y-FFT of the nf columns kz in chunks of C, the ng non-zero rows at y + nd read from A,
//   the outputs written straight to a2[b][x][ky][kz].
// It replaces kp_fwd_z + kp_fwd_y and never writes a1.
// Where ng rows of T threads exceed the workgroup (WIDE: nf = 150 - 176), step 1 of the z
// transform takes N2 threads a row and step 2 two (row, k1) tasks a thread.
template <int N> constexpr bool fwd_zy_wide() { return N / 2 * Geo<N>::T > 1024; }
template <int N, int C> constexpr bool fwd_zy_fits() {
    constexpr int NG = N / 2, T = Geo<N>::T;
    return NG * Geo<N>::N2 <= 1024 && NG * Geo<N>::N1 <= 2048 && C * T <= 1024 && N % C == 0 &&
           (size_t)N * (NG + 1) * 8 + (size_t)N * (C + 1) * 8 + (size_t)N * 8 <= 150 * 1024;
}

// PF: the next plane's mask words and charges are loaded while the current plane is
// transformed (their zeroing of the charge grid follows when they are used).
template <int N, int C, bool PF = false>
__global__ void __launch_bounds__(1024) kp_fwd_zy(VoxGrid vg, float2 *__restrict__ a2, const float2 *__restrict__ twg,
                                                   int nplanes) {
    constexpr int NG = N / 2, ND = N / 4;
    constexpr int N1 = Geo<N>::N1, N2 = Geo<N>::N2, T = Geo<N>::T;
    // A's row stride: odd, so that the y step's lanes (one column kz each, stride LA) spread
    // over the LDS banks
    constexpr int LA = NG | 1, LE = C + 1;
    static_assert(fwd_zy_fits<N, C>(), "fused forward pass does not fit");
    // one spare row after each group of N2 in the z exchange, where it still fits
    constexpr int PZ = (size_t)(N + N1) * LA * 8 + (size_t)N * LE * 8 + (size_t)N * 8 <= 150 * 1024 ? 1 : 0;
    constexpr int RW = (NG + 31) / 32;
    static_assert(RW <= 4, "row of at most 128 voxels");
    __shared__ float2 A[(N + N1 * PZ) * LA];  // z-transformed plane, [kz][y]
    __shared__ float2 E[N * LE];  // y-exchange of one chunk of columns
    __shared__ float2 tw[N];
    for (int q = threadIdx.x; q < N; q += 1024) tw[q] = twg[q];
    // the z-row task of this thread: row y of a plane, step-1 index u (u fastest, so that a
    // row's voxels sit in neighbouring lanes)
    constexpr bool WIDE = fwd_zy_wide<N>();
    constexpr int ZT = WIDE ? N2 : T;  // threads a row in step 1
    const int zy = (int)threadIdx.x / ZT, zu = (int)threadIdx.x - zy * ZT;
    const bool zact = zy < NG && zu < N2;
    // a row's raw voxel data: its mask words and the charges of this thread's voxels
    uint32_t rd[RW], ro[RW];
    float rq[N1];
    auto load_row = [&](int p) {
#pragma unroll
        for (int w = 0; w < RW; w++) rd[w] = ro[w] = 0u;
#pragma unroll
        for (int j1 = 0; j1 < N1; j1++) rq[j1] = 0.0f;
        if (zact && p < nplanes) {
            const int b = p / NG, x = p - b * NG, row = x * NG + zy;
            const uint32_t *dr = vg.dmo + (size_t)b * 2 * vg.mw + (size_t)row * RW;
#pragma unroll
            for (int w = 0; w < RW; w++) {
                rd[w] = dr[w];
                ro[w] = dr[vg.mw + w];
            }
            const float *cr = vg.chg + (size_t)b * vg.ng3 + (size_t)row * NG;
#pragma unroll
            for (int j1 = 0; j1 < N1; j1++) {
                const int e = zu + N2 * j1;
                if (e >= ND && e < ND + NG) rq[j1] = cr[e - ND];
            }
        }
    };
    if (PF) load_row(blockIdx.x);
    for (int p = blockIdx.x; p < nplanes; p += gridDim.x) {  // uniform across the block
        const int b = p / NG, x = p - b * NG;
        // opaque thread index: keeps the per-thread addresses out of the plane loop
        int tid = threadIdx.x;
        asm volatile("" : "+v"(tid));
        if (!PF) load_row(p);
        __syncthreads();  // A, E (and tw) free
        // ---- z-FFT of row y: thread (y, u), the row's voxel values from the mask bits and
        // the charges into the step-1 registers; each non-zero charge read is cleared
        {
            const int y = tid / ZT, u = tid - y * ZT;
            const bool zrow = y < NG;
            if (zrow && u < N2) {
                float *cr = vg.chg + (size_t)b * vg.ng3 + (size_t)(x * NG + y) * NG;
                const float delta = vg.delta;
                float2 v[N1] = {};
#pragma unroll
                for (int j1 = 0; j1 < N1; j1++) {
                    const int e = u + N2 * j1;
                    if (e >= ND && e < ND + NG) {
                        const int k = e - ND, w = k >> 5;
                        const uint32_t d = w == 0 ? rd[0] : w == 1 ? rd[RW > 1 ? 1 : 0] : w == 2 ? rd[RW > 2 ? 2 : 0] : rd[RW > 3 ? 3 : 0];
                        const uint32_t o = w == 0 ? ro[0] : w == 1 ? ro[RW > 1 ? 1 : 0] : w == 2 ? ro[RW > 2 ? 2 : 0] : ro[RW > 3 ? 3 : 0];
                        const uint32_t bit = 1u << (k & 31);
                        const float q = rq[j1];
                        if (__float_as_uint(q) != 0u) cr[k] = 0.0f;  // keep the charge grid zero
                        v[j1] = make_float2((d & bit) ? delta : ((o & bit) ? 1.0f : 0.0f), q);
                    }
                }
                if (PF) load_row(p + gridDim.x);  // in flight while this plane is transformed
                rf::step1_g<N, false, PZ>(v, A, tw, 1, LA, 1, y, u);
            }
            __syncthreads();
            if constexpr (!WIDE) {
                float2 w[N2] = {};
                if (zrow && u < N1) rf::step2_g<N, false, PZ>(w, A, LA, 1, y, u);
                __syncthreads();
                if (zrow && u < N1)
                    sfor<0, N2>([&](auto kc) {
                        constexpr int k2 = decltype(kc)::value;
                        A[(u + N1 * k2) * LA + y] = w[k2];
                    });
            } else {  // tasks (row, k1) number tid and tid + 1024
                const int ya = tid / N1, ua = tid - ya * N1, yb = (tid + 1024) / N1, ub = tid + 1024 - yb * N1;
                float2 wa[N2] = {}, wb[N2] = {};
                if (ya < NG) rf::step2_g<N, false, PZ>(wa, A, LA, 1, ya, ua);
                if (yb < NG) rf::step2_g<N, false, PZ>(wb, A, LA, 1, yb, ub);
                __syncthreads();
                if (ya < NG)
                    sfor<0, N2>([&](auto kc) {
                        constexpr int k2 = decltype(kc)::value;
                        A[(ua + N1 * k2) * LA + ya] = wa[k2];
                    });
                if (yb < NG)
                    sfor<0, N2>([&](auto kc) {
                        constexpr int k2 = decltype(kc)::value;
                        A[(ub + N1 * k2) * LA + yb] = wb[k2];
                    });
            }
        }
        __syncthreads();
        // ---- y-FFT of the columns, C at a time: thread (column, u), column fastest
        float2 *dst = a2 + (size_t)p * N * N;  // a2[b][x]: [ky][kz]
        const int cl = tid % C, u = tid / C;
        for (int c0 = 0; c0 < N; c0 += C) {  // uniform
            const int kz = c0 + cl;
            if (u < N2) {
                float2 v[N1] = {};
                sfor<0, N1>([&](auto jc) {
                    constexpr int j1 = decltype(jc)::value;
                    const int yy = u + N2 * j1 - ND;
                    if (yy >= 0 && yy < NG) v[j1] = A[kz * LA + yy];
                });
                rf::step1_g<N, false>(v, E, tw, 1, LE, 1, cl, u);
            }
            __syncthreads();
            if (u < N1) {
                float2 w[N2];
                rf::step2_g<N, false>(w, E, LE, 1, cl, u);
                sfor<0, N2>([&](auto kc) {
                    constexpr int k2 = decltype(kc)::value;
                    st_at(dst, 8u * (unsigned)((u + N1 * k2) * N + kz), w[k2]);
                });
            }
            __syncthreads();  // E is rewritten by the next chunk
        }
    }
}

// ------------------------------------------------------------------ forward z + y, half planes
// The fused forward pass for sizes whose z-transformed plane does not fit one workgroup's
// LDS (nf = 256).  With M = nf/2 and W = exp(-2 pi i / nf), the kz = 2m + h outputs of a
// z row a (non-zero at z in [nd, nd + ng), nd = nf/4, ng = M) are the M-point DFT of
//   z'_h[j] = (a[j] + (-1)^h a[j + M]) W^(h j),   j < M,
// where one of the two terms is zero: z'_h[j] = r[j - nd] W^(hj) for j >= nd and
// (-1)^h r[j + nd] W^(hj) for j < nd (r: the row's voxels).  So per (rotation, x) plane and
// parity h the z outputs form an M x ng half-plane A_h[m][y] in LDS.  Each of its columns
// m then goes through the y transform the same way, split by the parity g of ky: both
// halves' step-1 inputs come from the same column values, the g = 1 step-1 results wait in
// registers while g = 0 uses the column's own storage as its exchange, and each step 2
// writes its outputs straight to HBM.  The output is d_a2 with the kz of each run of 16 in
// parity-split order: kz = 2m + h sits at 16 (m / 8) + 8 h + m % 8 of its row (half_slot),
// so the 16 columns kz = 16c .. 16c + 15 that a correlation workgroup reads still fill one
// 128-byte line (kp_x_corr reads it with PERM).
__device__ __forceinline__ int half_slot(int h, int m) { return ((m >> 3) << 4) + (h << 3) + (m & 7); }
__device__ __forceinline__ int half_slot_of_kz(int kz) { return half_slot(kz & 1, kz >> 1); }

template <int N> constexpr bool fwd_half_fits() {
    constexpr int M = N / 2;
    return N % 4 == 0 && (M & (M - 1)) == 0 && M * Geo<M>::N2 == 1024 && Geo<M>::N1 * 64 == 1024 &&
           (size_t)M * (M + 1) * 8 + (size_t)N * 8 <= 150 * 1024;
}

template <int N>
__global__ void __launch_bounds__(1024) kp_fwd_half(VoxGrid vg, float2 *__restrict__ a2, const float2 *__restrict__ twg,
                                                     int nplanes) {
    constexpr int NG = N / 2, ND = N / 4, M = N / 2;
    constexpr int N1 = Geo<M>::N1, N2 = Geo<M>::N2;  // M-point transforms: step 1 of N1 points
    constexpr int LA = NG + 1;                        // odd row stride of A[m][y]
    constexpr int RW = (NG + 31) / 32;
    static_assert(fwd_half_fits<N>(), "half-plane forward does not fit");
    static_assert(32 % N2 == 0 && ND % 32 == 0, "a z task's mask words are fixed at compile time");
    __shared__ float2 A[M * LA];
    __shared__ float2 tw[N];
    for (int q = threadIdx.x; q < N; q += 1024) tw[q] = twg[q];
    for (int p = blockIdx.x; p < nplanes; p += gridDim.x) {  // uniform across the block
        const int b = p / NG, x = p - b * NG;
        float2 *dst = a2 + (size_t)p * N * N;  // a2[b][x]: [ky][h M + m]
#pragma unroll 1
        for (int h = 0; h < 2; h++) {  // uniform
            int tid = threadIdx.x;
            asm volatile("" : "+v"(tid));  // the thread's addresses are made anew in each half
            __syncthreads();  // A free
            // ---- z: row y, step-1 index u; the row's voxel values from the mask bits and the
            // charges (each non-zero charge cleared at its second read, h = 1)
            {
                const int y = tid % NG, u = tid / NG;  // NG rows x N2 tasks = 1024
                const int row = x * NG + y;
                const uint32_t *dr = vg.dmo + (size_t)b * 2 * vg.mw + (size_t)row * RW;
                float *cr = vg.chg + (size_t)b * vg.ng3 + (size_t)row * NG;
                const float delta = vg.delta;
                uint32_t dw[RW], ow[RW];  // the row's mask words
#pragma unroll
                for (int w = 0; w < RW; w++) {
                    dw[w] = dr[w];
                    ow[w] = dr[vg.mw + w];
                }
                float2 v[N1];
                sfor<0, N1>([&](auto jc) {
                    constexpr int j1 = decltype(jc)::value;
                    // j = u + N2 j1 with u < N2: j < ND and the mask word of k are fixed by j1
                    constexpr bool lo = N2 * j1 < ND;
                    constexpr int kw = (N2 * j1 + (lo ? ND : -ND)) / 32;
                    const int j = u + N2 * j1;
                    const int k = lo ? j + ND : j - ND;  // the voxel r[k] behind z'_h[j]
                    const uint32_t bit = 1u << (k & 31);
                    const float q = cr[k];
                    if (h == 1 && __float_as_uint(q) != 0u) cr[k] = 0.0f;  // keep the charge grid zero
                    float2 val = make_float2((dw[kw] & bit) ? delta : ((ow[kw] & bit) ? 1.0f : 0.0f), q);
                    if (h == 1) {
                        if (lo) val = make_float2(-val.x, -val.y);
                        const float2 w = tw[j];  // W^j
                        val = make_float2(val.x * w.x - val.y * w.y, val.x * w.y + val.y * w.x);
                    }
                    v[j1] = val;
                });
                rf::step1_g<M, false>(v, A, tw, 2, LA, 1, y, u);
            }
            __syncthreads();
            // step 2 in two rounds of NG/2 rows, each read and then written back in place as
            // A[m][y] (m = k1 + N1 k2); rows of one round share no storage with the other's
#pragma unroll 1
            for (int r = 0; r < 2; r++) {  // uniform
                const int y = r * (NG / 2) + tid % (NG / 2), u = tid / (NG / 2);  // N1 tasks a row
                float2 w[N2];
                rf::step2_g<M, false>(w, A, LA, 1, y, u);
                __syncthreads();
                sfor<0, N2>([&](auto kc) {
                    constexpr int k2 = decltype(kc)::value;
                    A[(u + N1 * k2) * LA + y] = w[k2];
                });
                __syncthreads();
            }
            // ---- y: column m (its ng values A[m][0..ng)), step-1 index j2, both parities g of ky
            {
                const int m = tid % M, j2 = tid / M;  // M columns x N2 tasks = 1024
                // g = 1's step-1 DFT first, its results kept in registers until g = 0 is done;
                // then g = 0's values, read again (the column is still intact)
                float2 v1[N1];
                sfor<0, N1>([&](auto jc) {
                    constexpr int j1 = decltype(jc)::value;
                    const int j = j2 + N2 * j1;
                    const float2 a = A[m * LA + (j >= ND ? j - ND : j + ND)];
                    const float2 sa = j < ND ? make_float2(-a.x, -a.y) : a;
                    const float2 w = tw[j];
                    v1[j1] = make_float2(sa.x * w.x - sa.y * w.y, sa.x * w.y + sa.y * w.x);
                });
                rf::RDft<N1, false>::run(v1);
                {
                    float2 v0[N1];
                    sfor<0, N1>([&](auto jc) {
                        constexpr int j1 = decltype(jc)::value;
                        const int j = j2 + N2 * j1;
                        v0[j1] = A[m * LA + (j >= ND ? j - ND : j + ND)];
                    });
                    __syncthreads();  // every column read before the exchange overwrites it
                    rf::step1_g<M, false>(v0, A, tw, 2, 1, LA, m, j2);
                }
                __syncthreads();
#pragma unroll 1
                for (int r = 0; r < 2; r++) {  // uniform; step 2 of g = 0 in two rounds of M/2 columns
                    const int mc = r * (M / 2) + tid % (M / 2), k1 = tid / (M / 2);
                    float2 w[N2];
                    rf::step2_g<M, false>(w, A, 1, LA, mc, k1);
                    sfor<0, N2>([&](auto kc) {
                        constexpr int k2 = decltype(kc)::value;
                        const int ky = 2 * (k1 + N1 * k2);
                        st_at(dst, 8u * (unsigned)(ky * N + half_slot(h, mc)), w[k2]);
                    });
                }
                __syncthreads();  // g = 0 has read the exchange
                // the twiddled g = 1 step-1 results into the exchange
                sfor<0, N1>([&](auto kc) {
                    constexpr int k1 = decltype(kc)::value;
                    float2 xv = v1[k1];
                    if constexpr (k1 != 0) {
                        const float2 w = tw[2 * j2 * k1];
                        xv = make_float2(xv.x * w.x - xv.y * w.y, xv.x * w.y + xv.y * w.x);
                    }
                    A[m * LA + k1 * N2 + j2] = xv;
                });
                __syncthreads();
#pragma unroll 1
                for (int r = 0; r < 2; r++) {  // uniform
                    const int mc = r * (M / 2) + tid % (M / 2), k1 = tid / (M / 2);
                    float2 w[N2];
                    rf::step2_g<M, false>(w, A, 1, LA, mc, k1);
                    sfor<0, N2>([&](auto kc) {
                        constexpr int k2 = decltype(kc)::value;
                        const int ky = 2 * (k1 + N1 * k2) + 1;
                        st_at(dst, 8u * (unsigned)(ky * N + half_slot(h, mc)), w[k2]);
                    });
                }
            }
        }
    }
}

// ------------------------------------------------------------------ x + correlation
// block (kz chunk c, ky, rotation group).  Lines 0..L-1: (ky, kz = c L + l); lines
// L..2L-1: partner columns (-ky, -kz).  Forward x-FFT of all 2L lines, the Hermitian
// fold Q(k) = (conj(R(k)) L(k) + conj(conj(R(-k)) L(-k))) / 2, inverse x-FFT of Q.
// The receptor values of the tile are loaded once and reused for RB rotations.
// B_FAST: the rotation is the fastest grid index, so the workgroups that share a
// receptor tile run back to back and find it in L2 / the last-level cache.
// QA: the folded lines Q share the LDS of S (two more barriers, a third less LDS).
// SPEC: store the folded spectrum Q itself, cc[b][kx][ky][kz], and skip the inverse x-FFT
// (the spectral score evaluation reads it).
// FLAT: the block's L columns are entries c L .. c L + L - 1 of the flat list, ky slowest,
// of the columns kz >= KZ0 = L floor((nf/2 + 1) / L) (the ones that do not fill a whole
// chunk) over all ky; the chunked launch covers kz < KZ0, so no padding column is
// transformed.
// PERM: d_a2 holds each run of 16 kz in parity-split order (kp_fwd_half, half_slot).
template <int N, int L, int RB, bool CACHE_R, bool B_FAST = false, bool QA = false, bool SPEC = false,
          bool FLAT = false, bool PERM = false>
__global__ void __launch_bounds__(2 * L *Geo<N>::T)
    kp_x_corr(const float2 *__restrict__ a2, const float2 *__restrict__ rec, float2 *__restrict__ cc,
              const float2 *__restrict__ twg, int H, int Hp, int B) {
    constexpr int N1 = Geo<N>::N1, N2 = Geo<N>::N2, T = Geo<N>::T;
    constexpr int NG = N / 2, ND = N / 4;
    constexpr int L2 = 2 * L, LS2 = L2 + 1, LS = L + 1, NT = L2 * T;
    constexpr int PER = (L * N + NT - 1) / NT;  // fold elements per thread
    __shared__ float2 S[N * LS2];
    __shared__ float2 Qs[QA ? 1 : N * LS];
    __shared__ float2 tw[N];
    float2 *const Q = QA ? S : Qs;
    const int tid = threadIdx.x;
    for (int q = tid; q < N; q += NT) tw[q] = twg[q];
    const int c = B_FAST ? blockIdx.y : blockIdx.x, ky = B_FAST ? blockIdx.z : blockIdx.y;
    const int b0 = (B_FAST ? blockIdx.x : blockIdx.z) * RB;
    // column (cky, ckz) of line l < L, and whether it exists
    constexpr int KZ0 = FLAT ? (N / 2 + 1) / L * L : 0, TWN = FLAT ? (N / 2 + 1) % L : 1;
    static_assert(!FLAT || (TWN > 0 && !SPEC), "flat list of the remaining columns");
    auto col = [&](int l, int &cky, int &ckz) {
        if constexpr (FLAT) {
            const int f = c * L + l;
            cky = f / TWN;
            ckz = KZ0 + (f - cky * TWN);
        } else {
            cky = ky;
            ckz = c * L + l;
        }
    };
    // fold elements of this thread: column l (fixed), rows kx = fx0 + FX p, p < PER
    static_assert(NT % L == 0, "fold mapping");
    constexpr int FX = NT / L;
    const int fl = tid % L, fx0 = tid / L;
    int fky, fkz;
    col(fl, fky, fkz);
    const bool fcol = FLAT ? fky < N : fkz < H;
    const int fkyp = (N - fky) % N;
    const unsigned fa0 = 8u * (unsigned)((fx0 * N + fky) * N + fkz);                    // + 8 FX N^2 p
    const unsigned fd0 = 8u * (unsigned)(((N - fx0) * N + fkyp) * N + (N - fkz) % N);  // - 8 FX N^2 p
    // receptor values of this thread's fold elements (kept in registers when CACHE_R)
    auto rec_pair = [&](int p, float2 &a, float2 &d) {
        a = d = make_float2(0.f, 0.f);
        if (fx0 + FX * p < N && fcol) {
            a = ld_at(rec, fa0 + 8u * (unsigned)(FX * p * N * N));
            // partner row -kx: N - kx for kx > 0 and 0 for kx = 0
            const unsigned od = (fx0 + FX * p == 0) ? fd0 - 8u * (unsigned)(N * N * N) : fd0 - 8u * (unsigned)(FX * p * N * N);
            d = ld_at(rec, od);
        }
    };
    float2 r1[CACHE_R ? PER : 1], r2[CACHE_R ? PER : 1];
    if constexpr (CACHE_R) {
#pragma unroll
        for (int p = 0; p < PER; p++) rec_pair(p, r1[p], r2[p]);
    }
    const int line = tid % L2, u = tid / L2;
    int lky, kzl;
    col(line % L, lky, kzl);
    const bool valid_line = FLAT ? lky < N : kzl < H;
    const int kyl = line < L ? lky : (N - lky) % N;
    const int kzs = line < L ? kzl : (N - kzl) % N;
#pragma unroll 1
    for (int bb = 0; bb < RB; bb++) {
        const int b = b0 + bb;
        if (b >= B) break;  // uniform across the block
        const float2 *src = a2 + (size_t)b * NG * N * N;  // uniform base, 32-bit lane offsets
        __syncthreads();  // S, Q free (previous rotation done)
        if (u < N2) {
            float2 v[N1];
            const int kzp = PERM ? half_slot_of_kz(kzs) : kzs;  // kz's place in a row of d_a2
            const unsigned o0 = 8u * (unsigned)((u - ND) * N * N + kyl * N + kzp);  // used only when x is in range
            sfor<0, N1>([&](auto jc) {
                constexpr int j1 = decltype(jc)::value;
                const int x = u + N2 * j1 - ND;
                v[j1] = (valid_line && x >= 0 && x < NG) ? ld_at(src, o0 + 8u * N2 * j1 * N * N) : make_float2(0.f, 0.f);
            });
            rf::step1_store<N, false>(v, S, tw, LS2, line, u);
        }
        __syncthreads();
        float2 y[N2];
        if (u < N1) rf::step2_load<N, false>(y, S, LS2, line, u);
        __syncthreads();
        if (u < N1)
            sfor<0, N2>([&](auto kc) {
                constexpr int k2 = decltype(kc)::value;
                S[(u + N1 * k2) * LS2 + line] = y[k2];
            });
        __syncthreads();
        // Hermitian fold into Q
        float2 qv[PER];
#pragma unroll
        for (int p = 0; p < PER; p++) {
            const int kx = fx0 + FX * p;
            if (kx < N) {
                const int kxp = kx == 0 ? 0 : N - kx;
                const float2 l1 = S[kx * LS2 + fl], l2 = S[kxp * LS2 + L + fl];
                float2 a, d;
                if constexpr (CACHE_R) {
                    a = r1[p];
                    d = r2[p];
                } else {
                    rec_pair(p, a, d);
                }
                const float2 p1 = make_float2(a.x * l1.x + a.y * l1.y, a.x * l1.y - a.y * l1.x);
                const float2 p2 = make_float2(d.x * l2.x + d.y * l2.y, d.x * l2.y - d.y * l2.x);
                qv[p] = make_float2(0.5f * (p1.x + p2.x), 0.5f * (p1.y - p2.y));
            }
        }
        if constexpr (SPEC) {
            float2 *dst = cc + (size_t)b * N * N * Hp;
            if (fkz < Hp)
#pragma unroll
                for (int p = 0; p < PER; p++)
                    if (fx0 + FX * p < N) st_at(dst, 8u * (unsigned)(((fx0 + FX * p) * N + fky) * Hp + fkz), qv[p]);
            continue;
        }
        if (QA) __syncthreads();  // every read of S above is done before Q overwrites it
#pragma unroll
        for (int p = 0; p < PER; p++)
            if (fx0 + FX * p < N) Q[(fx0 + FX * p) * LS + fl] = qv[p];
        __syncthreads();
        // inverse x-FFT of the L lines of Q; thread (lq, uq) over the first L*T threads
        const int lq = tid % L, uq = tid / L;
        const bool act = tid < L * T;
        {
            float2 v[N1];
            if (act && uq < N2)
                sfor<0, N1>([&](auto jc) {
                    constexpr int j1 = decltype(jc)::value;
                    v[j1] = Q[(uq + N2 * j1) * LS + lq];
                });
            if (QA) __syncthreads();  // Q is read before step 1 overwrites it
            if (act && uq < N2) rf::step1_store<N, true>(v, S, tw, LS, lq, uq);
        }
        __syncthreads();
        if (act && uq < N1) {
            float2 z[N2];
            rf::step2_load<N, true>(z, S, LS, lq, uq);
            int sky, kz;
            col(lq, sky, kz);
            if (FLAT ? sky < N : kz < Hp) {
                float2 *dst = cc + (size_t)b * N * N * Hp;  // uniform base, 32-bit lane offsets
                const unsigned o0 = 8u * (unsigned)((uq * N + sky) * Hp + kz);
                sfor<0, N2>([&](auto kc) {
                    constexpr int k2 = decltype(kc)::value;
                    st_at(dst, o0 + 8u * (unsigned)(N1 * k2 * N * Hp), z[k2]);
                });
            }
        }
    }
}

// ------------------------------------------------------------------ inverse y, in place
template <int N, int L>
__global__ void __launch_bounds__(L *Geo<N>::T) kp_inv_y(float2 *__restrict__ cc, const float2 *__restrict__ twg, int Hp) {
    constexpr int N1 = Geo<N>::N1, N2 = Geo<N>::N2, T = Geo<N>::T, LS = L + 1, NT = L * T;
    __shared__ float2 S[N * LS];
    __shared__ float2 tw[N];
    const int tid = threadIdx.x, line = tid % L, u = tid / L;
    for (int q = tid; q < N; q += NT) tw[q] = twg[q];
    const int kz = blockIdx.x * L + line, x = blockIdx.y, b = blockIdx.z;
    float2 *base = cc + ((size_t)b * N + x) * N * Hp;
    __syncthreads();
    if (u < N2) {
        float2 v[N1];
        sfor<0, N1>([&](auto jc) {
            constexpr int j1 = decltype(jc)::value;
            v[j1] = base[(size_t)(u + N2 * j1) * Hp + kz];
        });
        rf::step1_store<N, true>(v, S, tw, LS, line, u);
    }
    __syncthreads();
    if (u < N1) {
        float2 y[N2];
        rf::step2_load<N, true>(y, S, LS, line, u);
        sfor<0, N2>([&](auto kc) {
            constexpr int k2 = decltype(kc)::value;
            base[(size_t)(u + N1 * k2) * Hp + kz] = y[k2];
        });
    }
}

// Block-wide top-t: each wavefront extracts the top-t of its lanes' lists, lane 0 of each
// wavefront leaves them in wl, and thread 0 merges the NW short lists and writes t
// candidates to dst.
template <int TT, int NW>
__device__ __forceinline__ void block_top(TopR<TT> &tl, int t, Cand *wl, Cand *dst) {
    float ov[TT];
    int oi[TT];
    wave_top<TT>(tl, t, ov, oi);
    const int tid = threadIdx.x;
    if ((tid & 63) == 0)
#pragma unroll
        for (int q = 0; q < TT; q++)
            if (q < t) wl[(tid >> 6) * TT + q] = Cand{ov[q], oi[q]};
    __syncthreads();
    if (tid == 0) {
        TopR<TT> m;
        m.init();
        for (int w = 0; w < NW; w++)
#pragma unroll
            for (int q = 0; q < TT; q++)
                if (q < t) m.push(wl[w * TT + q].key, wl[w * TT + q].idx);
#pragma unroll
        for (int q = 0; q < TT; q++)
            if (q < t) dst[q] = Cand{m.v[q], m.i[q]};
    }
}

// ------------------------------------------------------------------ inverse z + top-t
// block (x, b): the N rows r = x*N + y of plane x of rotation b, L rows at a time; a row's
// spectrum is Hermitian in kz and only kz <= N/2 is stored.  As in the fused inverse, a row
// goes through the half-length complex-to-real step: with M = N/2 and w = exp(+2 pi i / N),
//   Z[k] = (X[k] + conj(X[M-k])) + i w^k (X[k] - conj(X[M-k])),  k < M,
// whose M-point inverse gives z[m] = x[2m] + i x[2m+1], two real scores per output.  Every
// thread keeps a running top-TT (TT >= t, compile time) of its outputs over the plane's row
// groups; the lists are merged across each wavefront with lane shuffles and across
// wavefronts through LDS once per plane, leaving the plane's top-t as t candidates.  The
// test hook writes the real score rows to smap instead.
// The workgroup has L * N1z threads rounded up to whole wavefronts (the extra lanes only
// take part in the wavefront merges); a last row group that runs past the plane's N rows
// transforms zeros there and pushes nothing from them.
template <int N, int L> constexpr int inv_z_threads() { return (L * Geo<N / 2>::N1 + 63) / 64 * 64; }
template <int N, int L, int TT>
__global__ void __launch_bounds__((inv_z_threads<N, L>()))
    kp_inv_z_top(const float2 *__restrict__ cc, Cand *__restrict__ cand, float *__restrict__ smap,
                 const float2 *__restrict__ twg, int H, int Hp, int t, float nf3f) {
    constexpr int M = N / 2, HC = M + 1;
    constexpr int N1z = Geo<M>::N1, N2z = Geo<M>::N2, NT = inv_z_threads<N, L>(), NA = L * N1z;
    constexpr int LSR = HC % 2 ? HC : HC + 1;  // row stride in LDS (odd)
    constexpr int NROW = N * N, NW = NT / 64, RG = (N + L - 1) / L;
    static_assert(N2z <= N1z, "one task per thread and step");
    __shared__ float2 S[L * LSR];
    __shared__ float2 tw[N];
    __shared__ Cand wl[NW * TT];
    const int tid = threadIdx.x, line = tid % L, u = tid / L;
    const bool act = tid < NA;  // lanes beyond the row group's tasks only merge
    for (int q = tid; q < N; q += NT) tw[q] = twg[q];
    const int b = blockIdx.y;
    const float2 *src = cc + ((size_t)b * NROW) * Hp;
    // Only scores that can reach the wavefront's top-t are pushed: at least t scores are
    // >= tau (the t largest lane maxima of a row group, the largest such bound over the
    // plane so far, taken from its first four row groups and every fourth after; the t
    // scores behind it were pushed when it was raised), so anything whose key falls below
    // tau's loses.  Keys: raw values for t == 1, fl(value / nf^3) for t > 1 (TT > 1
    // exactly when t > 1).
    constexpr bool QUOT = TT > 1;
    TopR<TT> tl;
    tl.init();
    float run_tau = -INFINITY, cut = -INFINITY;
    for (int g = 0; g < RG; g++) {
        const int y0 = g * L, row0 = blockIdx.x * N + y0;
        const bool row_ok = act && y0 + line < N;
        __syncthreads();  // S is free: the previous row group has read it
        for (int q = tid; q < L * HC; q += NT) {
            const int l = q / HC, kz = q - l * HC;
            S[l * LSR + kz] = y0 + l < N ? src[(size_t)(row0 + l) * Hp + kz] : make_float2(0.f, 0.f);
        }
        __syncthreads();
        float2 v[N1z];
        if (act && u < N2z)
            sfor<0, N1z>([&](auto jc) {
                constexpr int j1 = decltype(jc)::value;
                const int k = u + N2z * j1;
                const float2 a = S[line * LSR + k], c = S[line * LSR + M - k], w = tw[k];
                const float2 s1 = make_float2(a.x + c.x, a.y - c.y);  // X[k] + conj(X[M-k])
                const float2 d1 = make_float2(a.x - c.x, a.y + c.y);  // X[k] - conj(X[M-k])
                // e = conj(w) d1; Z = s1 + i e
                const float2 e1 = make_float2(w.x * d1.x + w.y * d1.y, w.x * d1.y - w.y * d1.x);
                v[j1] = make_float2(s1.x - e1.y, s1.y + e1.x);
            });
        __syncthreads();  // every read of S is done before step 1 writes into it
        if (act && u < N2z) rf::step1_g<M, true>(v, S, tw, 2, 1, LSR, line, u);
        __syncthreads();
        float2 y[N2z];
        if (act) rf::step2_g<M, true>(y, S, 1, LSR, line, u);
        const int r = row0 + line;
        // z = 2m and 2m+1 of row r are y[k2].x and .y with m = u + N1z k2
        if (smap) {  // uniform
            if (row_ok)
                sfor<0, N2z>([&](auto kc) {
                    constexpr int k2 = decltype(kc)::value;
                    const size_t o = ((size_t)b * NROW + r) * N + 2 * (u + N1z * k2);
                    smap[o] = y[k2].x;
                    smap[o + 1] = y[k2].y;
                });
            continue;
        }
        if (g < 4 || (g & 3) == 0) {  // uniform
            float mx = -INFINITY;
            if (row_ok)
                sfor<0, N2z>([&](auto kc) {
                    constexpr int k2 = decltype(kc)::value;
                    mx = fmaxf(mx, fmaxf(y[k2].x, y[k2].y));
                });
            const float tau = wave_kth_max(mx, t);
            if (tau > run_tau) {  // uniform
                run_tau = tau;
                cut = QUOT ? quotient_cut(tau, nf3f) : tau;
            }
        }
        if (row_ok)
            sfor<0, N2z>([&](auto kc) {
                constexpr int k2 = decltype(kc)::value;
                const int z0 = 2 * (u + N1z * k2);
                const float v0 = y[k2].x, v1 = y[k2].y;
                if (v0 >= cut) tl.push(QUOT ? v0 / nf3f : v0, r * N + z0);
                if (v1 >= cut) tl.push(QUOT ? v1 / nf3f : v1, r * N + z0 + 1);
            });
    }
    if (smap) return;  // uniform
    block_top<TT, NW>(tl, t, wl, cand + ((size_t)b * gridDim.x + blockIdx.x) * t);
}

template <int N, int L, int TT> void launch_inv_z(const PipeArgs &a, hipStream_t s) {
    kp_inv_z_top<N, L, TT><<<dim3(N, a.B), inv_z_threads<N, L>(), 0, s>>>(a.cc, a.cand, a.smap, a.tw4, a.H, a.Hp,
                                                                          a.t, a.nf3f);
}

// the fused inverse keeps a whole (nf x (nf/2+1)) half-spectrum plane in LDS and gives
// every thread at most two tasks of each step of the y transform
constexpr int plane_stride(int n) { return ((n / 2 + 1) % 2) ? n / 2 + 1 : n / 2 + 2; }
template <int N> constexpr bool fused_inverse() {
    return (size_t)N * plane_stride(N) * 8 <= 150 * 1024 && (N / 2 + 1) * Geo<N>::N1 <= 2 * 1024 &&
           (N / 2 + 1) * Geo<N>::N2 <= 2 * 1024;
}
template <int N> constexpr bool pipe_size_ok() { return N % 2 == 0; }

// ------------------------------------------------------------------ fused inverse y + z + top-t
// Persistent workgroups, each walking the (b, x) planes p = b*nf + x of cc (nf rows ky,
// kz <= nf/2 columns) with stride gridDim.x.  Per plane, held in LDS:
//   y-IFFT of every kz column in place (step-1 inputs come straight from HBM into
//   registers; the next plane's are loaded while the current one is computed);
//   per row, the half-length complex-to-real step: with M = nf/2 and w = exp(+2 pi i / nf),
//     Z[k] = (X[k] + conj(X[M-k])) + i w^k (X[k] - conj(X[M-k])),  k < M,
//   whose M-point inverse gives z[m] = x[2m] + i x[2m+1] (x the real score row);
//   the M-point inverse of each row, whose outputs go straight into the top-t lists.
// The top-t is written as one candidate list per plane.  PF: prefetch the next plane's
// step-1 inputs (in flight during the current plane) or load them at the start.
// SEG: each workgroup walks a contiguous range of planes instead; the planes of one
// rotation inside the range form a segment whose lane lists, and whose running threshold,
// carry over from plane to plane.  The segment's top-t goes to the list of its first plane
// and the lists of its other planes are left empty, so the wavefront merge and the
// block merge run once per segment rather than once per plane.
template <int N, int TT, bool PF, bool SEG = false>
__global__ void __launch_bounds__(1024)
    kp_inv_yz_top(const float2 *__restrict__ cc, Cand *__restrict__ cand, float *__restrict__ smap,
                  const float2 *__restrict__ twg, int Hp, int t, float nf3f, int nplanes) {
    constexpr int NTH = 1024, NW = NTH / 64;
    constexpr int M = N / 2, H = M + 1;
    constexpr int LSP = plane_stride(N);  // row stride of the plane (odd)
    constexpr int N1y = Geo<N>::N1, N2y = Geo<N>::N2;
    constexpr int N1z = Geo<M>::N1, N2z = Geo<M>::N2;
    constexpr int KZ = (N * N1z + NTH - 1) / NTH;    // z step-2 tasks per thread
    constexpr int KZ1 = (N * N2z + NTH - 1) / NTH;   // z step-1 tasks per thread (u < N2z)
    constexpr int KY1 = (H * N2y + NTH - 1) / NTH, KY2 = (H * N1y + NTH - 1) / NTH;  // y tasks per thread
    static_assert(!PF || (KY1 == 1 && KY2 == 1), "the prefetch holds one y task");
    __shared__ float2 P[N * LSP];
    __shared__ float2 tw[N];
    __shared__ Cand wl[NW * TT];
    const int tid = threadIdx.x;
    for (int q = tid; q < N; q += NTH) tw[q] = twg[q];
    // y task of a thread: column kz = yl, step-1 index yu (< N2y), step-2 index yu (< N1y)
    // planes of this workgroup: p_first, p_first + p_step, ... below p_end
    int p_first = blockIdx.x, p_end = nplanes, p_step = gridDim.x;
    if constexpr (SEG) {
        const int chunk = (nplanes + gridDim.x - 1) / gridDim.x;
        p_first = blockIdx.x * chunk;
        p_end = min(nplanes, p_first + chunk);
        p_step = 1;
    }
    float2 vn[N1y];  // step-1 inputs of the plane to come
    auto load_plane = [&](int p, int yl, int yu) {
        if (yu < N2y && p < p_end) {
            const float2 *src = cc + (size_t)p * N * Hp;  // uniform base, 32-bit lane offsets
            const unsigned off = (unsigned)(yu * Hp + yl);
            sfor<0, N1y>([&](auto jc) {
                constexpr int j1 = decltype(jc)::value;
                vn[j1] = src[off + (unsigned)(N2y * j1 * Hp)];
            });
        }
    };
    constexpr bool QUOT = TT > 1;
    // SEG: the running lane lists, the running threshold and the segment's first plane
    TopR<TT> tl;
    tl.init();
    float run_tau = -INFINITY, cut = -INFINITY;
    int seg0 = p_first;
    if (PF) load_plane(p_first, tid % H, tid / H);
    for (int p = p_first; p < p_end; p += p_step) {  // uniform across the block
        const int b = p / N, x = p - b * N;
        // The thread's LDS addresses are recomputed in every iteration: an opaque copy of
        // the thread index keeps them from being hoisted out of the plane loop, where they
        // would all stay live and spill.
        int tid = threadIdx.x;
        asm volatile("" : "+v"(tid));
        const int yl = tid % H, yu = tid / H;
        if (!PF) load_plane(p, yl, yu);
        __syncthreads();  // P, tw and wl are free: the previous plane is done with them
        // ---- y-IFFT of the H columns (line = kz, element = ky): element stride LSP
        if (yu < N2y) {
            float2 v[N1y];
            sfor<0, N1y>([&](auto jc) { v[decltype(jc)::value] = vn[decltype(jc)::value]; });
            rf::step1_g<N, true>(v, P, tw, 1, LSP, 1, yl, yu);
        }
        // sizes with more step-1 tasks than threads: the further tasks tid + j NTH
#pragma unroll
        for (int j = 1; j < KY1; j++) {
            const int task = tid + j * NTH, ylj = task % H, yuj = task / H;
            if (yuj < N2y) {
                const float2 *src = cc + (size_t)p * N * Hp;  // uniform base, 32-bit lane offsets
                const unsigned off = (unsigned)(yuj * Hp + ylj);
                float2 v[N1y];
                sfor<0, N1y>([&](auto jc) {
                    constexpr int j1 = decltype(jc)::value;
                    v[j1] = src[off + (unsigned)(N2y * j1 * Hp)];
                });
                rf::step1_g<N, true>(v, P, tw, 1, LSP, 1, ylj, yuj);
            }
        }
        if (PF) load_plane(p + p_step, yl, yu);
        __syncthreads();
        {
            float2 y[KY2][N2y];
#pragma unroll
            for (int j = 0; j < KY2; j++) {
                const int task = tid + j * NTH, ylj = task % H, yuj = task / H;
                if (yuj < N1y) rf::step2_g<N, true>(y[j], P, LSP, 1, ylj, yuj);
            }
            __syncthreads();
#pragma unroll
            for (int j = 0; j < KY2; j++) {
                const int task = tid + j * NTH, ylj = task % H, yuj = task / H;
                if (yuj < N1y)
                    sfor<0, N2y>([&](auto kc) {
                        constexpr int k2 = decltype(kc)::value;
                        P[(yuj + N1y * k2) * LSP + ylj] = y[j][k2];
                    });
            }
        }
        __syncthreads();
        // ---- M-point inverse of each row (line = ky, element stride 1, line stride LSP),
        // its inputs formed by the complex-to-real pre-step as they are loaded:
        // Z[k] = (X[k] + conj(X[M-k])) + i w^k (X[k] - conj(X[M-k])), w^k = conj(tw[k])
        // zero-initialised so that no value of the arrays is carried around the plane loop
        float2 y[KZ][N2z] = {};
        {
            float2 v[KZ1][N1z] = {};
#pragma unroll
            for (int j = 0; j < KZ1; j++) {
                const int task = tid + j * NTH, line = task % N, u = task / N;
                if (u < N2z)
                    sfor<0, N1z>([&](auto jc) {
                        constexpr int j1 = decltype(jc)::value;
                        const int k = u + N2z * j1;
                        const float2 a = P[line * LSP + k], c = P[line * LSP + M - k], w = tw[k];
                        const float2 s1 = make_float2(a.x + c.x, a.y - c.y);  // X[k] + conj(X[M-k])
                        const float2 d1 = make_float2(a.x - c.x, a.y + c.y);  // X[k] - conj(X[M-k])
                        // e = conj(w) d1; Z = s1 + i e
                        const float2 e1 = make_float2(w.x * d1.x + w.y * d1.y, w.x * d1.y - w.y * d1.x);
                        v[j][j1] = make_float2(s1.x - e1.y, s1.y + e1.x);
                    });
            }
            __syncthreads();
#pragma unroll
            for (int j = 0; j < KZ1; j++) {
                const int task = tid + j * NTH, line = task % N, u = task / N;
                if (u < N2z) rf::step1_g<M, true>(v[j], P, tw, 2, 1, LSP, line, u);
            }
            __syncthreads();
#pragma unroll
            for (int j = 0; j < KZ; j++) {
                const int task = tid + j * NTH, line = task % N, u = task / N;
                if (u < N1z) rf::step2_g<M, true>(y[j], P, 1, LSP, line, u);
            }
        }
        // z = 2m and 2m+1 of row `line` are y[j][k2].x and .y with m = u + N1z k2
        if (smap) {  // uniform
#pragma unroll
            for (int j = 0; j < KZ; j++) {
                const int task = tid + j * NTH, line = task % N, u = task / N;
                if (u < N1z)
                    sfor<0, N2z>([&](auto kc) {
                        constexpr int k2 = decltype(kc)::value;
                        const size_t o = (size_t)b * N * N * N + (size_t)(x * N + line) * N + 2 * (u + N1z * k2);
                        smap[o] = y[j][k2].x;
                        smap[o + 1] = y[j][k2].y;
                    });
            }
            continue;
        }
        // Only scores that can reach the wavefront's top-t are pushed: at least t scores are
        // >= tau (the t largest lane maxima), so anything whose key falls below tau's loses.
        // Keys: raw values for t == 1, fl(value / nf^3) for t > 1 (TT > 1 exactly when t > 1).
        // SEG: tau is the largest such bound over the segment's planes so far, taken from
        // its first four planes and every eighth after; the t scores behind it were pushed
        // when it was raised and can only be displaced by better ones.
        const int k = p - seg0;
        if (!SEG || k < 4 || (k & 7) == 0) {  // uniform
            float mx = -INFINITY;
#pragma unroll
            for (int j = 0; j < KZ; j++) {
                if ((tid + j * NTH) / N < N1z)
                    sfor<0, N2z>([&](auto kc) {
                        constexpr int k2 = decltype(kc)::value;
                        mx = fmaxf(mx, fmaxf(y[j][k2].x, y[j][k2].y));
                    });
            }
            const float tau = wave_kth_max(mx, t);
            if (!SEG || tau > run_tau) {  // uniform
                run_tau = tau;
                cut = QUOT ? quotient_cut(tau, nf3f) : tau;
            }
        }
        if (!SEG) tl.init();
#pragma unroll
        for (int j = 0; j < KZ; j++) {
            const int task = tid + j * NTH, line = task % N, u = task / N;
            if (u < N1z) {
                const int base = (x * N + line) * N + 2 * u;
                sfor<0, N2z>([&](auto kc) {
                    constexpr int k2 = decltype(kc)::value;
                    const float v0 = y[j][k2].x, v1 = y[j][k2].y;
                    if (v0 >= cut) tl.push(QUOT ? v0 / nf3f : v0, base + 2 * N1z * k2);
                    if (v1 >= cut) tl.push(QUOT ? v1 / nf3f : v1, base + 2 * N1z * k2 + 1);
                });
            }
        }
        if constexpr (SEG) {
            if (p != seg0 && tid < t) cand[(size_t)p * t + tid] = Cand{-INFINITY, kNoIndex};  // empty list
            if (p + 1 == p_end || (p + 1) / N != b) {  // uniform: the segment ends with this plane
                block_top<TT, NW>(tl, t, wl, cand + (size_t)seg0 * t);
                tl.init();
                run_tau = cut = -INFINITY;
                seg0 = p + 1;
            }
        } else {
            block_top<TT, NW>(tl, t, wl, cand + (size_t)p * t);
        }
    }
}

// ------------------------------------------------------------------ half-plane inverse
// The fused inverse for sizes whose half-spectrum plane does not fit one workgroup's LDS
// (nf = 256): each (rotation, x) plane in two halves, its even and its odd rows y.  With
// M = nf/2 and w = exp(+2 pi i / nf), the rows y = 2m + h of the y-IFFT of a column X are
// the M-point IFFT of
//   X[k] + X[k + M]              (h = 0)
//   (X[k] - X[k + M]) w^k        (h = 1),   k < M,
// so a half needs an M x (nf/2 + 1) half-plane in LDS.  Each half then goes through the
// same steps as kp_inv_yz_top (complex-to-real pre-step and M-point z-IFFT per row, running
// top-t over a workgroup's contiguous planes), and the plane is read from HBM twice instead
// of being transformed in place there (kp_inv_y) and read again (kp_inv_z_top).
// The columns kz = 0 and kz = nf/2 are real after the y-IFFT (each z row of the score is
// real, so its spectrum is Hermitian and both of these entries are real), so they go
// through the y-IFFT together as one column, X(0) + i X(nf/2), and the pre-step takes them
// apart as the real and imaginary parts: M columns instead of M + 1, whole rounds.
// the input of row y = 2m + h's M-point y-IFFT at k from X[k] = a and X[k + M] = c;
// w = tw[k] = conj(w^k)
__device__ __forceinline__ float2 fold_half(float2 a, float2 c, int h, float2 w) {
    if (h == 0) return make_float2(a.x + c.x, a.y + c.y);
    const float2 d = make_float2(a.x - c.x, a.y - c.y);
    return make_float2(d.x * w.x + d.y * w.y, d.y * w.x - d.x * w.y);
}

template <int N> constexpr bool half_inverse_fits() {
    constexpr int M = N / 2, N1 = Geo<M>::N1, N2 = Geo<M>::N2;
    // whole columns and rows a round of 1024 threads
    return N % 4 == 0 && (size_t)M * plane_stride(N) * 8 + (size_t)N * 8 + 16 * 16 * 8 <= 150 * 1024 && N1 <= 32 &&
           1024 % N1 == 0 && 1024 % N2 == 0 && M % (1024 / N1) == 0 && M * N2 <= 1024;
}

// NTH threads a workgroup.  With 1024 the odd half reads the plane again.  With 512
// (KEEP: two waves per SIMD, so twice the registers) every thread takes two tasks of each
// step, and the plane is read once: the odd rows' folded y inputs, (X[k] - X[k + M]) w^k,
// are made from the same loads as the even rows' and wait in registers meanwhile.
template <int N, int TT, int NTH = 1024>
__global__ void __launch_bounds__(NTH)
    kp_inv_yz_half(const float2 *__restrict__ cc, Cand *__restrict__ cand, float *__restrict__ smap,
                   const float2 *__restrict__ twg, int Hp, int t, float nf3f, int nplanes) {
    constexpr int NW = NTH / 64;
    constexpr bool KEEP = NTH < 1024;
    constexpr int M = N / 2;
    constexpr int LSP = plane_stride(N);  // row stride of the half-plane (odd)
    constexpr int N1 = Geo<M>::N1, N2 = Geo<M>::N2;  // both transforms are M-point
    // the y transform runs in rounds of whole columns, one task per thread: step 1 over
    // CS1 = NTH / N2 columns a round, step 2 over CS2 = NTH / N1 columns a round (each round's
    // columns are read, then written back in place, so no thread holds two tasks' outputs);
    // the z transform's step 1 in ZR1 rounds of NTH / N2 rows, all read before any is
    // written, its step 2 over RS = NTH / N1 rows a round, its outputs pushed round by round
    constexpr int CS1 = NTH / N2, CS2 = NTH / N1, RS = NTH / N1;
    constexpr int HC = M;  // y columns: kz = 1 .. M - 1 and the packed pair kz = 0, M
    constexpr int R1 = (HC + CS1 - 1) / CS1, R2 = (HC + CS2 - 1) / CS2, RZ = M / RS;
    constexpr int ZR1 = M * N2 / NTH;
    static_assert(half_inverse_fits<N>(), "half-plane inverse does not fit");
    static_assert(NTH % N1 == 0 && NTH % N2 == 0 && M % RS == 0 && (M * N2) % NTH == 0 && HC % CS1 == 0,
                  "whole columns and rows a round");
    __shared__ float2 P[M * LSP];
    __shared__ float2 tw[N];
    __shared__ Cand wl[NW * TT];
    for (int q = threadIdx.x; q < N; q += NTH) tw[q] = twg[q];
    // a contiguous range of planes per workgroup; the planes of one rotation in it form a
    // segment whose lane lists and running threshold carry over (as kp_inv_yz_top, SEG)
    const int chunk = (nplanes + gridDim.x - 1) / gridDim.x;
    const int p_first = blockIdx.x * chunk, p_end = min(nplanes, p_first + chunk);
    constexpr bool QUOT = TT > 1;
    TopR<TT> tl;
    tl.init();
    float run_tau = -INFINITY, cut = -INFINITY;
    int seg0 = p_first, units = 0;  // units: z rounds of the segment done so far
    for (int p = p_first; p < p_end; p++) {  // uniform across the block
        const int b = p / N, x = p - b * N;
        const float2 *src = cc + (size_t)p * N * Hp;  // uniform base, 32-bit lane offsets
        float2 keep[KEEP ? R1 : 1][KEEP ? N1 : 1];
        for (int h = 0; h < 2; h++) {  // uniform
            // opaque thread index: keeps the per-thread addresses out of the loops
            int tid = threadIdx.x;
            asm volatile("" : "+v"(tid));
            __syncthreads();  // P, tw and wl are free
            // ---- M-point y-IFFT of the H columns (line = kz, element = m), folded inputs;
            // thread (column yl = r CS1 + tid % CS1, index yu = tid / CS1) in round r
            sfor<0, R1>([&](auto rc) {
                constexpr int r = decltype(rc)::value;
                const int yl = r * CS1 + tid % CS1, yu = tid / CS1;
                float2 v[N1];
                if (!KEEP || h == 0) {  // uniform
                    sfor<0, N1>([&](auto jc) {
                        constexpr int j1 = decltype(jc)::value;
                        const int k = yu + N2 * j1;
                        const float2 a = src[(unsigned)(k * Hp + yl)], c = src[(unsigned)((k + M) * Hp + yl)];
                        v[j1] = fold_half(a, c, h, tw[k]);
                        if constexpr (KEEP) keep[KEEP ? r : 0][KEEP ? j1 : 0] = fold_half(a, c, 1, tw[k]);
                    });
                    if (yl == 0) {  // column 0 carries X(0) + i X(M): add i times the kz = M fold
                        asm volatile("" ::: "memory");  // after the main loads: no more of them in flight at once
                        sfor<0, N1>([&](auto jc) {
                            constexpr int j1 = decltype(jc)::value;
                            const int k = yu + N2 * j1;
                            const float2 an = src[(unsigned)(k * Hp + M)], cn = src[(unsigned)((k + M) * Hp + M)];
                            const float2 q = fold_half(an, cn, h, tw[k]);
                            v[j1] = make_float2(v[j1].x - q.y, v[j1].y + q.x);
                            if constexpr (KEEP) {
                                float2 &kk = keep[KEEP ? r : 0][KEEP ? j1 : 0];
                                const float2 q1 = fold_half(an, cn, 1, tw[k]);
                                kk = make_float2(kk.x - q1.y, kk.y + q1.x);
                            }
                        });
                    }
                } else {  // the odd rows' inputs, folded from the even rows' loads
                    sfor<0, N1>([&](auto jc) {
                        constexpr int j1 = decltype(jc)::value;
                        v[j1] = keep[KEEP ? r : 0][KEEP ? j1 : 0];
                    });
                }
                rf::step1_g<M, true>(v, P, tw, 2, LSP, 1, yl, yu);
            });
            __syncthreads();
#pragma unroll 1
            for (int r = 0; r < R2; r++) {  // uniform
                const int yl = r * CS2 + tid % CS2, yu = tid / CS2;
                float2 y[N2];
                if (yl < HC) rf::step2_g<M, true>(y, P, LSP, 1, yl, yu);
                __syncthreads();  // the round's columns are read before they are written back
                if (yl < HC)
                    sfor<0, N2>([&](auto kc) {
                        constexpr int k2 = decltype(kc)::value;
                        P[(yu + N1 * k2) * LSP + yl] = y[k2];
                    });
                __syncthreads();
            }
            // ---- M-point inverse of each of the M rows with the complex-to-real pre-step:
            // Z[k] = (X[k] + conj(X[M-k])) + i w^k (X[k] - conj(X[M-k])), w^k = conj(tw[k])
            {
                float2 v[ZR1][N1];
                sfor<0, ZR1>([&](auto rc) {  // step 1: M rows x N2 tasks, ZR1 rounds
                    constexpr int r = decltype(rc)::value;
                    const int q = r * NTH + tid, line = q % M, u = q / M;
                    sfor<0, N1>([&](auto jc) {
                        constexpr int j1 = decltype(jc)::value;
                        const int k = u + N2 * j1;
                        float2 a = P[line * LSP + k], c = P[line * LSP + (M - k) % M], w = tw[k];
                        if (j1 == 0 && u == 0) {  // k = 0: X(0) and X(M) from the packed column
                            c = make_float2(a.y, 0.0f);
                            a = make_float2(a.x, 0.0f);
                        }
                        const float2 s1 = make_float2(a.x + c.x, a.y - c.y);
                        const float2 d1 = make_float2(a.x - c.x, a.y + c.y);
                        const float2 e1 = make_float2(w.x * d1.x + w.y * d1.y, w.x * d1.y - w.y * d1.x);
                        v[r][j1] = make_float2(s1.x - e1.y, s1.y + e1.x);
                    });
                });
                __syncthreads();
                sfor<0, ZR1>([&](auto rc) {
                    constexpr int r = decltype(rc)::value;
                    const int q = r * NTH + tid, line = q % M, u = q / M;
                    rf::step1_g<M, true>(v[r], P, tw, 2, 1, LSP, line, u);
                });
                __syncthreads();
            }
            // step 2 and the top-t, RS rows a round: row `line` of the half is y = 2 line + h;
            // z = 2m and 2m+1 with m = u + N1 k2
#pragma unroll 1
            for (int r = 0; r < RZ; r++, units++) {  // uniform
                const int line = r * RS + tid % RS, u = tid / RS;
                float2 y[N2];
                rf::step2_g<M, true>(y, P, 1, LSP, line, u);
                if (smap) {  // uniform
                    sfor<0, N2>([&](auto kc) {
                        constexpr int k2 = decltype(kc)::value;
                        const size_t o = (size_t)b * N * N * N + (size_t)(x * N + 2 * line + h) * N + 2 * (u + N1 * k2);
                        smap[o] = y[k2].x;
                        smap[o + 1] = y[k2].y;
                    });
                    continue;
                }
                // the running threshold of the segment, raised on its first four rounds and
                // every eighth after (the t scores behind it are pushed when it is raised)
                if (units < 4 || (units & 7) == 0) {  // uniform
                    float mx = -INFINITY;
                    sfor<0, N2>([&](auto kc) {
                        constexpr int k2 = decltype(kc)::value;
                        mx = fmaxf(mx, fmaxf(y[k2].x, y[k2].y));
                    });
                    const float tau = wave_kth_max(mx, t);
                    if (tau > run_tau) {  // uniform
                        run_tau = tau;
                        cut = QUOT ? quotient_cut(tau, nf3f) : tau;
                    }
                }
                const int base = (x * N + 2 * line + h) * N + 2 * u;
                sfor<0, N2>([&](auto kc) {
                    constexpr int k2 = decltype(kc)::value;
                    const float v0 = y[k2].x, v1 = y[k2].y;
                    if (v0 >= cut) tl.push(QUOT ? v0 / nf3f : v0, base + 2 * N1 * k2);
                    if (v1 >= cut) tl.push(QUOT ? v1 / nf3f : v1, base + 2 * N1 * k2 + 1);
                });
            }
        }
        if (smap) continue;  // uniform
        if (p != seg0 && (int)threadIdx.x < t) cand[(size_t)p * t + threadIdx.x] = Cand{-INFINITY, kNoIndex};  // empty list
        if (p + 1 == p_end || (p + 1) / N != b) {  // uniform: the segment ends with this plane
            block_top<TT, NW>(tl, t, wl, cand + (size_t)seg0 * t);
            tl.init();
            run_tau = cut = -INFINITY;
            seg0 = p + 1;
            units = 0;
        }
    }
}

template <int N, int TT> void launch_inv_half(const PipeArgs &a, hipStream_t s) {
    const int nplanes = a.B * N;
    const int grid = nplanes < a.cu_count ? nplanes : a.cu_count;  // one workgroup per CU
    if (a.inv_variant == 1)  // 512 threads, the plane read once
        kp_inv_yz_half<N, TT, 512><<<grid, 512, 0, s>>>(a.cc, a.cand, a.smap, a.tw4, a.Hp, a.t, a.nf3f, nplanes);
    else
        kp_inv_yz_half<N, TT><<<grid, 1024, 0, s>>>(a.cc, a.cand, a.smap, a.tw4, a.Hp, a.t, a.nf3f, nplanes);
}

// the fused inverse can prefetch the next plane (variants 1, 3) where a thread holds one
// y task of each step
template <int N> constexpr bool inv_prefetch_fits() {
    return (N / 2 + 1) * Geo<N>::N1 <= 1024 && (N / 2 + 1) * Geo<N>::N2 <= 1024;
}

// LEVEL 2: every variant; LEVEL 1: running lists with (3) or without (2) the prefetch
template <int N, int TT, int LEVEL> void launch_inv_yz(const PipeArgs &a, hipStream_t s) {
    const int nplanes = a.B * N;
    const int grid = nplanes < a.cu_count ? nplanes : a.cu_count;  // one 1024-thread workgroup per CU
    if constexpr (LEVEL < 2) {
        if (a.inv_variant == 3 && inv_prefetch_fits<N>())
            kp_inv_yz_top<N, TT, inv_prefetch_fits<N>(), true><<<grid, 1024, 0, s>>>(a.cc, a.cand, a.smap, a.tw4, a.Hp,
                                                                                   a.t, a.nf3f, nplanes);
        else
            kp_inv_yz_top<N, TT, false, true><<<grid, 1024, 0, s>>>(a.cc, a.cand, a.smap, a.tw4, a.Hp, a.t, a.nf3f,
                                                                  nplanes);
    } else switch (a.inv_variant) {
        case 1:
            kp_inv_yz_top<N, TT, true><<<grid, 1024, 0, s>>>(a.cc, a.cand, a.smap, a.tw4, a.Hp, a.t, a.nf3f, nplanes);
            break;
        case 2:
            kp_inv_yz_top<N, TT, false, true><<<grid, 1024, 0, s>>>(a.cc, a.cand, a.smap, a.tw4, a.Hp, a.t, a.nf3f,
                                                                  nplanes);
            break;
        case 3:
            kp_inv_yz_top<N, TT, true, true><<<grid, 1024, 0, s>>>(a.cc, a.cand, a.smap, a.tw4, a.Hp, a.t, a.nf3f,
                                                                 nplanes);
            break;
        default:
            kp_inv_yz_top<N, TT, false><<<grid, 1024, 0, s>>>(a.cc, a.cand, a.smap, a.tw4, a.Hp, a.t, a.nf3f, nplanes);
    }
}

// columns per chunk of the fused forward pass: the largest divisor of nf whose chunk fits
// the workgroup and LDS, or 0 where none does (then z and y run as two passes)
template <int N> constexpr int fwd_zy_chunk() {
    constexpr int NG = N / 2, T = Geo<N>::T;
    if (NG * Geo<N>::N2 > 1024 || NG * Geo<N>::N1 > 2048) return 0;
    for (int c = N; c >= 8; c--)
        if (N % c == 0 && c * T <= 1024 && (size_t)N * (NG + 1) * 8 + (size_t)N * (c + 1) * 8 + (size_t)N * 8 <= 150 * 1024)
            return c;
    return 0;
}

// LEVEL 2: every kernel variant (the sizes tuned most, 80, 96, 144, 256); LEVEL 1: the
// variants worth tuning on every size (forward 0 and 2, correlation 5, 6 and 7, inverse 2
// and 3).  A variant not built for the size runs as the nearest one built (correlation 6,
// inverse 2).
template <int N, int LEVEL> hipError_t run_pipe(const PipeArgs &a) {
    constexpr bool FULL = LEVEL >= 2;
    static_assert(pipe_size_ok<N>(), "no specialised pipeline for this size");
    constexpr int T = Geo<N>::T;
    constexpr int L = 16, LX = 8;
    const int B = a.B, ng = a.ng;
    if (ng != N / 2 || a.nd != N / 4) return hipErrorInvalidValue;  // the kernels assume nf = 2 ng, nd = ng / 2
    hipStream_t s = a.stream;
    constexpr int CZY = fwd_zy_chunk<N>();
    bool fused_fwd = false, perm = false;
    if constexpr (CZY == 0 && fwd_half_fits<N>()) {
        if (a.fwd_variant != 1) {  // fused z + y over half planes; d_a2 in parity-split kz order
            const int nplanes = B * ng;
            const int grid = nplanes < a.cu_count ? nplanes : a.cu_count;
            kp_fwd_half<N><<<grid, 1024, 0, s>>>(a.vox, a.a2, a.tw4, nplanes);
            if (a.ev) (void)hipEventRecord(a.ev[0], s);  // the z stage is fused into the y stage
            fused_fwd = perm = true;
        }
    }
    if constexpr (CZY > 0) {
        if (!FULL || a.fwd_variant == 0 || a.fwd_variant == 2) {  // fused z + y, the next plane prefetched (2: not)
            const int nplanes = B * ng;
            const int grid = nplanes < a.cu_count ? nplanes : a.cu_count;
            if (a.fwd_variant == 2)
                kp_fwd_zy<N, CZY, false><<<grid, 1024, 0, s>>>(a.vox, a.a2, a.tw4, nplanes);
            else
                kp_fwd_zy<N, CZY, true><<<grid, 1024, 0, s>>>(a.vox, a.a2, a.tw4, nplanes);
            if (a.ev) (void)hipEventRecord(a.ev[0], s);  // the z stage is fused into the y stage
            fused_fwd = true;
        }
    }
    if constexpr (FULL || CZY == 0) {
        if (!fused_fwd) {
            const int lines_z = B * ng * ng;
            kp_fwd_z<N, L><<<(lines_z + L - 1) / L, L * T, 0, s>>>(a.vox, a.a1, a.tw4, lines_z);
            if (a.ev) (void)hipEventRecord(a.ev[0], s);
            kp_fwd_y<N, L><<<dim3((N + L - 1) / L, ng, B), L * T, 0, s>>>(a.a1, a.a2, a.tw4);
        }
    }
    if (a.ev) (void)hipEventRecord(a.ev[1], s);
    if (a.spectrum_only) {
        if (perm) {
            if constexpr (fwd_half_fits<N>())
                kp_x_corr<N, LX, 1, false, true, true, true, false, true><<<dim3(B, a.Hp / LX, N), 2 * LX * T, 0, s>>>(
                    a.a2, a.rec, a.cc, a.tw4, a.H, a.Hp, B);
        } else {
            kp_x_corr<N, LX, 1, false, true, true, true><<<dim3(B, a.Hp / LX, N), 2 * LX * T, 0, s>>>(a.a2, a.rec, a.cc,
                                                                                                      a.tw4, a.H, a.Hp, B);
        }
        return hipGetLastError();
    }
    // correlation-pass variant: one rotation per workgroup with the rotation index fastest in
    // the grid, so consecutive workgroups share the receptor tile in L2, folded lines kept
    // in the exchange buffer's LDS, 16 columns per workgroup with the receptor values
    // loaded first (6, the default) or where the fold needs them (5), 8 columns (4); 8
    // columns with a separate fold buffer
    // (3); receptor tile cached in registers for 4 rotations (0); re-read from L2 for each
    // of 4 rotations (1); one rotation per workgroup, tile index fastest (2); as 6 without
    // the padding columns (7)
    constexpr int LC = 2 * LX, NCH = (N / 2 + 1) / LC, NFL = ((N / 2 + 1) % LC * N + LC - 1) / LC;
    if (perm) {  // the parity-split d_a2: variant 6's tiling, or 7's (whole chunks and a flat tail)
        if constexpr (fwd_half_fits<N>()) {
            if (a.variant == 7 && NFL > 0) {
                if constexpr (NCH > 0)
                    kp_x_corr<N, LC, 1, true, true, true, false, false, true><<<dim3(B, NCH, N), 2 * LC * T, 0, s>>>(
                        a.a2, a.rec, a.cc, a.tw4, a.H, a.Hp, B);
                if constexpr (NFL > 0)
                    kp_x_corr<N, LC, 1, true, true, true, false, true, true><<<dim3(B, NFL, 1), 2 * LC * T, 0, s>>>(
                        a.a2, a.rec, a.cc, a.tw4, a.H, a.Hp, B);
            } else {
                kp_x_corr<N, 2 * LX, 1, true, true, true, false, false, true><<<dim3(B, a.Hp / (2 * LX), N), 4 * LX * T, 0, s>>>(
                    a.a2, a.rec, a.cc, a.tw4, a.H, a.Hp, B);
            }
        }
    } else if constexpr (!FULL) {
        if (a.variant == 5) {
            kp_x_corr<N, 2 * LX, 1, false, true, true><<<dim3(B, a.Hp / (2 * LX), N), 4 * LX * T, 0, s>>>(
                a.a2, a.rec, a.cc, a.tw4, a.H, a.Hp, B);
        } else if (a.variant == 7 && NFL > 0) {
            if constexpr (NCH > 0)
                kp_x_corr<N, LC, 1, true, true, true><<<dim3(B, NCH, N), 2 * LC * T, 0, s>>>(a.a2, a.rec, a.cc, a.tw4,
                                                                                           a.H, a.Hp, B);
            if constexpr (NFL > 0)
                kp_x_corr<N, LC, 1, true, true, true, false, true><<<dim3(B, NFL, 1), 2 * LC * T, 0, s>>>(
                    a.a2, a.rec, a.cc, a.tw4, a.H, a.Hp, B);
        } else {
            kp_x_corr<N, 2 * LX, 1, true, true, true><<<dim3(B, a.Hp / (2 * LX), N), 4 * LX * T, 0, s>>>(
                a.a2, a.rec, a.cc, a.tw4, a.H, a.Hp, B);
        }
    } else switch (a.variant) {
        case 0:
            kp_x_corr<N, LX, 4, true><<<dim3(a.Hp / LX, N, (B + 3) / 4), 2 * LX * T, 0, s>>>(
                a.a2, a.rec, a.cc, a.tw4, a.H, a.Hp, B);
            break;
        case 1:
            kp_x_corr<N, LX, 4, false><<<dim3(a.Hp / LX, N, (B + 3) / 4), 2 * LX * T, 0, s>>>(
                a.a2, a.rec, a.cc, a.tw4, a.H, a.Hp, B);
            break;
        case 2:
            kp_x_corr<N, LX, 1, false><<<dim3(a.Hp / LX, N, B), 2 * LX * T, 0, s>>>(a.a2, a.rec, a.cc, a.tw4, a.H, a.Hp, B);
            break;
        case 3:
            kp_x_corr<N, LX, 1, false, true><<<dim3(B, a.Hp / LX, N), 2 * LX * T, 0, s>>>(a.a2, a.rec, a.cc, a.tw4, a.H, a.Hp, B);
            break;
        case 4:
            kp_x_corr<N, LX, 1, false, true, true><<<dim3(B, a.Hp / LX, N), 2 * LX * T, 0, s>>>(a.a2, a.rec, a.cc, a.tw4,
                                                                                              a.H, a.Hp, B);
            break;
        case 5:  // as 4 with 16 columns per workgroup
            kp_x_corr<N, 2 * LX, 1, false, true, true><<<dim3(B, a.Hp / (2 * LX), N), 4 * LX * T, 0, s>>>(
                a.a2, a.rec, a.cc, a.tw4, a.H, a.Hp, B);
            break;
        case 7: {  // as 6 over the whole chunks of 16 columns, the remaining columns as a flat list
            if constexpr (NCH > 0)
                kp_x_corr<N, LC, 1, true, true, true><<<dim3(B, NCH, N), 2 * LC * T, 0, s>>>(a.a2, a.rec, a.cc, a.tw4,
                                                                                           a.H, a.Hp, B);
            if constexpr (NFL > 0)
                kp_x_corr<N, LC, 1, true, true, true, false, true><<<dim3(B, NFL, 1), 2 * LC * T, 0, s>>>(
                    a.a2, a.rec, a.cc, a.tw4, a.H, a.Hp, B);
            break;
        }
        default:  // 6: as 5 with the receptor tile loaded into registers before the forward x-FFT,
                  // so that its latency hides behind the transform
            kp_x_corr<N, 2 * LX, 1, true, true, true><<<dim3(B, a.Hp / (2 * LX), N), 4 * LX * T, 0, s>>>(
                a.a2, a.rec, a.cc, a.tw4, a.H, a.Hp, B);
    }
    if (a.ev) (void)hipEventRecord(a.ev[2], s);
    if constexpr (fused_inverse<N>()) {
        if (a.ev) (void)hipEventRecord(a.ev[3], s);  // inverse y is fused into the next pass
        if (a.t <= 1) launch_inv_yz<N, 1, LEVEL>(a, s);
        else if (a.t <= 2) launch_inv_yz<N, 2, LEVEL>(a, s);
        else if (a.t <= 4) launch_inv_yz<N, 4, LEVEL>(a, s);
        else if (a.t <= 8) launch_inv_yz<N, 8, LEVEL>(a, s);
        else launch_inv_yz<N, 16, LEVEL>(a, s);
    } else {
        bool half = false;
        if constexpr (half_inverse_fits<N>()) half = !FULL || a.inv_variant != 0;
        if (half) {  // the half-plane fused inverse (default where it fits; variant 0: two passes)
            if constexpr (half_inverse_fits<N>()) {
                if (a.ev) (void)hipEventRecord(a.ev[3], s);  // inverse y is fused into the next pass
                if (a.t <= 1) launch_inv_half<N, 1>(a, s);
                else if (a.t <= 2) launch_inv_half<N, 2>(a, s);
                else if (a.t <= 4) launch_inv_half<N, 4>(a, s);
                else if (a.t <= 8) launch_inv_half<N, 8>(a, s);
                else launch_inv_half<N, 16>(a, s);
            }
        } else {
            kp_inv_y<N, L><<<dim3(a.Hp / L, N, B), L * T, 0, s>>>(a.cc, a.tw4, a.Hp);
            if (a.ev) (void)hipEventRecord(a.ev[3], s);
            if (a.t <= 1) launch_inv_z<N, L, 1>(a, s);
            else if (a.t <= 2) launch_inv_z<N, L, 2>(a, s);
            else if (a.t <= 4) launch_inv_z<N, L, 4>(a, s);
            else if (a.t <= 8) launch_inv_z<N, L, 8>(a, s);
            else launch_inv_z<N, L, 16>(a, s);
        }
    }
    if (a.ev) (void)hipEventRecord(a.ev[4], s);
    return hipGetLastError();
}

}  // namespace
