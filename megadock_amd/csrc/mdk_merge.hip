// Device-side merge of the per-rotation results into the best N decoys (replaces the
// reference's host sort, src/fft_process.cpp:469-479, `sort_index` with
// `std::partial_sort` and `ascending_struct` :462-466).
//
// Order: score descending, then position in the results ascending, i.e. rotation
// ascending, then rank within the rotation (the reference leaves tie order to its
// library's heap; this order is deterministic).  Each result i gets the 64-bit key
//   K = sortable(score) << 32 | (2^32 - 1 - i)
// where sortable() maps float order to unsigned order (after -0 -> +0, so that equal
// scores tie), so the wanted order is K descending and all keys are distinct.
//   1. radix select, 8 bits per level from the top: the N-th largest key T;
//   2. compaction of the N keys >= T;
//   3. bitonic sort of those keys (in LDS up to 8192, else in global passes);
//   4. gather of the results in key order.
#include <hip/hip_runtime.h>

#include <mutex>

#include "mdk_internal.h"
#include "../../include/mdk_hip.h"

namespace {

struct SelState {
    unsigned long long prefix, mask;  // key bits fixed so far
    unsigned int rank;                // 1-based rank of the wanted key among those matching prefix
    unsigned int count;               // compaction cursor
    unsigned int hist[256];
};

__device__ __forceinline__ unsigned long long merge_key(const mdk_sortscore *r, unsigned i) {
    const float f = r[i].score + 0.0f;  // -0 -> +0
    unsigned u = __float_as_uint(f);
    u ^= (u & 0x80000000u) ? 0xffffffffu : 0x80000000u;
    return ((unsigned long long)u << 32) | (0xffffffffu - i);
}

__global__ void k_sel_init(SelState *st, unsigned rank) {
    const int q = threadIdx.x;
    st->hist[q] = 0u;
    if (q == 0) {
        st->prefix = 0ull;
        st->mask = 0ull;
        st->rank = rank;
        st->count = 0u;
    }
}

__global__ void __launch_bounds__(256) k_sel_hist(const mdk_sortscore *__restrict__ res, unsigned m, SelState *st,
                                                  int shift) {
    __shared__ unsigned h[256];
    h[threadIdx.x] = 0u;
    __syncthreads();
    const unsigned long long prefix = st->prefix, mask = st->mask;
    for (unsigned i = blockIdx.x * 256u + threadIdx.x; i < m; i += gridDim.x * 256u) {
        const unsigned long long k = merge_key(res, i);
        if ((k & mask) == prefix) atomicAdd(&h[(unsigned)(k >> shift) & 255u], 1u);
    }
    __syncthreads();
    if (h[threadIdx.x]) atomicAdd(&st->hist[threadIdx.x], h[threadIdx.x]);
}

// one thread: the digit of the wanted key at this level
__global__ void k_sel_pick(SelState *st, int shift) {
    unsigned c = 0u;
    for (int d = 255; d >= 0; d--) {
        const unsigned h = st->hist[d];
        if (c + h >= st->rank) {
            st->prefix |= (unsigned long long)d << shift;
            st->mask |= 255ull << shift;
            st->rank -= c;
            break;
        }
        c += h;
    }
    for (int d = 0; d < 256; d++) st->hist[d] = 0u;
}

__global__ void __launch_bounds__(256) k_sel_compact(const mdk_sortscore *__restrict__ res, unsigned m, SelState *st,
                                                     unsigned long long *keys) {
    const unsigned long long t = st->prefix;  // the N-th largest key
    for (unsigned i = blockIdx.x * 256u + threadIdx.x; i < m; i += gridDim.x * 256u) {
        const unsigned long long k = merge_key(res, i);
        if (k >= t) keys[atomicAdd(&st->count, 1u)] = k;
    }
}

// keys[n, P) = 0, which sorts after every result key
__global__ void k_sel_pad(unsigned long long *keys, unsigned n, unsigned P) {
    const unsigned i = n + blockIdx.x * 256u + threadIdx.x;
    if (i < P) keys[i] = 0ull;
}

constexpr int kSortLds = 8192;

// descending bitonic sort of P <= kSortLds keys (P a power of two) in one workgroup
__global__ void __launch_bounds__(1024) k_sort_lds(unsigned long long *keys, unsigned P) {
    __shared__ unsigned long long s[kSortLds];
    for (unsigned i = threadIdx.x; i < P; i += 1024u) s[i] = keys[i];
    __syncthreads();
    for (unsigned k = 2; k <= P; k <<= 1)
        for (unsigned j = k >> 1; j > 0; j >>= 1) {
            for (unsigned i = threadIdx.x; i < P; i += 1024u) {
                const unsigned l = i ^ j;
                if (l > i) {
                    const unsigned long long a = s[i], b = s[l];
                    const bool desc = (i & k) == 0;
                    if (desc ? (a < b) : (a > b)) {
                        s[i] = b;
                        s[l] = a;
                    }
                }
            }
            __syncthreads();
        }
    for (unsigned i = threadIdx.x; i < P; i += 1024u) keys[i] = s[i];
}

// one compare-exchange step (k, j) of the descending bitonic network over global memory
__global__ void __launch_bounds__(256) k_sort_step(unsigned long long *keys, unsigned P, unsigned k, unsigned j) {
    const unsigned i = blockIdx.x * 256u + threadIdx.x;
    if (i >= P) return;
    const unsigned l = i ^ j;
    if (l <= i) return;
    const unsigned long long a = keys[i], b = keys[l];
    const bool desc = (i & k) == 0;
    if (desc ? (a < b) : (a > b)) {
        keys[i] = b;
        keys[l] = a;
    }
}

__global__ void k_sel_gather(const unsigned long long *__restrict__ keys, unsigned n,
                             const mdk_sortscore *__restrict__ res, mdk_sortscore *out) {
    const unsigned q = blockIdx.x * 256u + threadIdx.x;
    if (q < n) out[q] = res[0xffffffffu - (unsigned)(keys[q] & 0xffffffffull)];
}

}  // namespace

size_t topn_scratch_bytes(int n) {
    unsigned P = 1;
    while (P < (unsigned)n) P <<= 1;
    return sizeof(SelState) + sizeof(unsigned long long) * P + 256;
}

hipError_t launch_topn(const mdk_sortscore *res, unsigned m, unsigned n, void *work, mdk_sortscore *out,
                       hipStream_t s) {
    if (n == 0 || m == 0) return hipSuccess;
    if (n > m) n = m;
    unsigned P = 1;
    while (P < n) P <<= 1;
    SelState *st = reinterpret_cast<SelState *>(work);
    unsigned long long *keys =
        reinterpret_cast<unsigned long long *>(reinterpret_cast<char *>(work) + ((sizeof(SelState) + 255) & ~size_t(255)));
    const unsigned blocks = (m + 255u) / 256u < 1024u ? (m + 255u) / 256u : 1024u;
    k_sel_init<<<1, 256, 0, s>>>(st, n);
    for (int shift = 56; shift >= 0; shift -= 8) {
        k_sel_hist<<<blocks, 256, 0, s>>>(res, m, st, shift);
        k_sel_pick<<<1, 1, 0, s>>>(st, shift);
    }
    k_sel_compact<<<blocks, 256, 0, s>>>(res, m, st, keys);
    if (P > n) k_sel_pad<<<(P - n + 255u) / 256u, 256, 0, s>>>(keys, n, P);
    if (P <= (unsigned)kSortLds) {
        k_sort_lds<<<1, 1024, 0, s>>>(keys, P);
    } else {
        for (unsigned k = 2; k <= P; k <<= 1)
            for (unsigned j = k >> 1; j > 0; j >>= 1) k_sort_step<<<(P + 255u) / 256u, 256, 0, s>>>(keys, P, k, j);
    }
    k_sel_gather<<<(n + 255u) / 256u, 256, 0, s>>>(keys, n, res, out);
    return hipGetLastError();
}

// ------------------------------------------------------------ allocation-free merge
// Fixed device scratch of mdk_topn_noalloc: the select state and keys of launch_topn,
// then the n merged results.  Every GPU the library runs on holds its own copy until the
// process ends.  Host code never uses the array's name as an address: the address of
// this device's copy comes from hipGetSymbolAddress, looked up once per device at
// context creation (topn_static_setup) and kept in g_static_ptr.
static __device__ __align__(256) unsigned char scratch[MDK_TOPN_STATIC_BYTES];
// the largest merge fits: state, keys of the power of two MDK_TOPN_STATIC_MAX, then results
static_assert((MDK_TOPN_STATIC_MAX & (MDK_TOPN_STATIC_MAX - 1)) == 0, "the key count is a power of two");
static_assert(((sizeof(SelState) + 8ull * MDK_TOPN_STATIC_MAX + 256 + 255) & ~255ull) +
                      sizeof(mdk_sortscore) * MDK_TOPN_STATIC_MAX <= MDK_TOPN_STATIC_BYTES,
              "the fixed merge array holds MDK_TOPN_STATIC_MAX results");

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_static_mu;
void *g_static_ptr[kMaxDevices] = {};
}  // namespace

size_t topn_static_out_offset(int n) { return (topn_scratch_bytes(n) + 255) & ~size_t(255); }

hipError_t topn_static_setup(int device) {
    if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_static_mu);
    if (g_static_ptr[device]) return hipSuccess;
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));  // this device is current
    if (e != hipSuccess) return e;
    if (!p) return hipErrorInvalidSymbol;
    g_static_ptr[device] = p;
    return hipSuccess;
}

void *topn_static_ptr(int device) {
    if (device < 0 || device >= kMaxDevices) return nullptr;
    std::lock_guard<std::mutex> lock(g_static_mu);
    return g_static_ptr[device];
}
