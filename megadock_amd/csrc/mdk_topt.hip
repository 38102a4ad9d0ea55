// Per-rotation top-t for t above the register lists of the fused inverse passes
// (t > MDK_MAXT).  The inverse passes write the batch's real score maps; this selects each
// rotation's t best entries with the order of the reference's insertion top-t for t > 1
// (src/fft_process.cpp:407-431): key fl(value / nf^3) descending, ties to the lower linear
// grid index, only keys above the starting value -99999 kept.
//
// Each map entry i of a rotation gets the 64-bit key
//   K = sortable(fl(v / nf^3)) << 32 | (2^32 - 1 - i)
// (sortable() maps float order to unsigned order after -0 -> +0, so equal keys tie and the
// lower index wins), all distinct, so the wanted entries are the t largest K:
//   1. radix select per rotation, 8 bits a level from the top, every workgroup of a level
//      adding its histogram of one chunk of one rotation; a level whose chosen bucket
//      holds exactly the entries still wanted ends the select early (the threshold is
//      then that bucket's lowest key);
//   2. compaction of the entries with K >= threshold (exactly min(t, nf^3) of them);
//   3. one workgroup per rotation sorts them (bitonic, in LDS up to 8192 keys, else in
//      global memory) and writes the rotation's t SortScore entries.
#include <hip/hip_runtime.h>

#include "mdk_internal.h"
#include "../../include/mdk_hip.h"

namespace {

constexpr int kChunk = 4096;     // map entries per histogram workgroup (16 per thread)
constexpr int kTSortLds = 8192;  // keys sorted in LDS

__device__ __forceinline__ unsigned long long topt_key(float v, float nf3f, unsigned i) {
    const float f = v / nf3f + 0.0f;  // the reference's key; -0 -> +0
    unsigned u = __float_as_uint(f);
    u ^= (u & 0x80000000u) ? 0xffffffffu : 0x80000000u;
    return ((unsigned long long)u << 32) | (0xffffffffu - i);
}

__device__ __forceinline__ float key_float(unsigned long long k) {
    const unsigned u = (unsigned)(k >> 32);
    return __uint_as_float((u & 0x80000000u) ? (u ^ 0x80000000u) : ~u);
}

__global__ void k_tsel_init(TopTState *st, unsigned rank) {
    TopTState &s = st[blockIdx.x];
    s.hist[threadIdx.x] = 0u;
    if (threadIdx.x == 0) {
        s.prefix = 0ull;
        s.mask = 0ull;
        s.rank = rank;
        s.count = 0u;
        s.done = 0u;
    }
}

// grid (chunks, B): histogram of digit `shift` over the entries matching the prefix
__global__ void __launch_bounds__(256) k_tsel_hist(const float *__restrict__ smap, unsigned n3, float nf3f,
                                                   TopTState *st, int shift) {
    __shared__ unsigned h[256];
    TopTState &s = st[blockIdx.y];
    if (s.done) return;  // uniform
    h[threadIdx.x] = 0u;
    __syncthreads();
    const unsigned long long prefix = s.prefix, mask = s.mask;
    const float *m = smap + (size_t)blockIdx.y * n3;
    const unsigned i0 = blockIdx.x * (unsigned)kChunk;
    const unsigned i1 = min(n3, i0 + (unsigned)kChunk);
    for (unsigned i = i0 + threadIdx.x; i < i1; i += 256u) {
        const unsigned long long k = topt_key(m[i], nf3f, i);
        if ((k & mask) == prefix) atomicAdd(&h[(unsigned)(k >> shift) & 255u], 1u);
    }
    __syncthreads();
    if (h[threadIdx.x]) atomicAdd(&s.hist[threadIdx.x], h[threadIdx.x]);
}

// one block per rotation: the digit of the wanted key at this level
__global__ void k_tsel_pick(TopTState *st, int shift) {
    TopTState &s = st[blockIdx.x];
    if (threadIdx.x == 0 && !s.done) {
        unsigned c = 0u;
        for (int d = 255; d >= 0; d--) {
            const unsigned h = s.hist[d];
            if (c + h >= s.rank) {
                s.prefix |= (unsigned long long)d << shift;
                s.mask |= 255ull << shift;
                s.rank -= c;
                if (s.rank == h) s.done = 1u;  // the whole bucket is wanted
                break;
            }
            c += h;
        }
    }
    __syncthreads();
    for (int d = threadIdx.x; d < 256; d += blockDim.x) s.hist[d] = 0u;
}

// grid (chunks, B): the entries at or above the threshold, in any order
__global__ void __launch_bounds__(256) k_tsel_compact(const float *__restrict__ smap, unsigned n3, float nf3f,
                                                      TopTState *st, unsigned long long *keys, unsigned P) {
    TopTState &s = st[blockIdx.y];
    const unsigned long long thr = s.prefix;
    const float *m = smap + (size_t)blockIdx.y * n3;
    unsigned long long *out = keys + (size_t)blockIdx.y * P;
    const unsigned i0 = blockIdx.x * (unsigned)kChunk;
    const unsigned i1 = min(n3, i0 + (unsigned)kChunk);
    for (unsigned i = i0 + threadIdx.x; i < i1; i += 256u) {
        const unsigned long long k = topt_key(m[i], nf3f, i);
        if (k >= thr) out[atomicAdd(&s.count, 1u)] = k;
    }
}

// one workgroup per rotation: descending bitonic sort of its P keys (padded with 0, which
// sorts last), then the rotation's t results
__global__ void __launch_bounds__(1024) k_tsel_sort_out(unsigned long long *keys, unsigned P, const TopTState *st,
                                                        int t, int nf, mdk_sortscore *__restrict__ out,
                                                        const int *__restrict__ rot_id, const int *__restrict__ boff) {
    __shared__ unsigned long long sl[kTSortLds];
    const int b = blockIdx.x;
    const unsigned n = st[b].count;
    unsigned long long *g = keys + (size_t)b * P;
    const bool lds = P <= (unsigned)kTSortLds;
    unsigned long long *s = lds ? sl : g;
    for (unsigned i = threadIdx.x; i < P; i += 1024u) s[i] = i < n ? g[i] : 0ull;
    __syncthreads();
    for (unsigned k = 2; k <= P; k <<= 1)
        for (unsigned j = k >> 1; j > 0; j >>= 1) {
            for (unsigned i = threadIdx.x; i < P; i += 1024u) {
                const unsigned l = i ^ j;
                if (l > i) {
                    const unsigned long long a = s[i], c = s[l];
                    const bool desc = (i & k) == 0;
                    if (desc ? (a < c) : (a > c)) {
                        s[i] = c;
                        s[l] = a;
                    }
                }
            }
            __syncthreads();
        }
    const int b0 = boff ? *boff : 0;
    const int rid = rot_id[b0 + b];
    mdk_sortscore *o = out + ((size_t)b0 + b) * t;
    const int nf2 = nf * nf;
    for (int q = threadIdx.x; q < t; q += 1024) {
        mdk_sortscore r;
        const unsigned long long k = (unsigned)q < n ? s[q] : 0ull;
        const float key = key_float(k);
        // sorted, so the kept entries are a prefix; the reference's slots start at -99999
        const bool keep = (unsigned)q < n && key > -99999.0f;
        const int idx = keep ? (int)(0xffffffffu - (unsigned)(k & 0xffffffffull)) : 0;
        r.score = keep ? key : -99999.0f;
        r.index[0] = rid;
        r.index[1] = idx / nf2;
        r.index[2] = (idx / nf) % nf;
        r.index[3] = idx % nf;
        o[q] = r;
    }
}

}  // namespace

unsigned topt_sort_size(int t, int nf) {
    const unsigned long long n3 = (unsigned long long)nf * nf * nf;
    const unsigned rank = (unsigned long long)t < n3 ? (unsigned)t : (unsigned)n3;
    unsigned P = 1;
    while (P < rank) P <<= 1;
    return P;
}

hipError_t launch_topt(const float *smap, int B, int nf, int t, TopTState *st, unsigned long long *keys,
                       mdk_sortscore *out, const int *rot_id, const int *boff, hipStream_t s) {
    if (B < 1) return hipSuccess;
    const unsigned n3 = (unsigned)nf * nf * nf;
    const float nf3f = (float)n3;
    const unsigned rank = (unsigned)t < n3 ? (unsigned)t : n3;
    const unsigned P = topt_sort_size(t, nf);
    const dim3 grid((n3 + kChunk - 1) / kChunk, B);
    k_tsel_init<<<B, 256, 0, s>>>(st, rank);
    for (int shift = 56; shift >= 0; shift -= 8) {
        k_tsel_hist<<<grid, 256, 0, s>>>(smap, n3, nf3f, st, shift);
        k_tsel_pick<<<B, 256, 0, s>>>(st, shift);
    }
    k_tsel_compact<<<grid, 256, 0, s>>>(smap, n3, nf3f, st, keys, P);
    k_tsel_sort_out<<<B, 1024, 0, s>>>(keys, P, st, t, nf, out, rot_id, boff);
    return hipGetLastError();
}
