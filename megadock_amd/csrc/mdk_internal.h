// Internal interfaces shared by the HIP translation units of libmdk.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/mdk_hip.h"

// Ligand data resident on the device (one context).
struct LigandDev {
    int na;           // atoms
    int ng;           // voxels per axis
    float gw;         // voxel pitch (A)
    float delta;      // core value (-b)
    float pad;        // float(ng * gw / 2)
    float cx, cy, cz; // ligand centre
    const float *xyz;    // na*3 input coordinates
    const float *rc2;    // r^2 * 1.5
    const float *rs2;    // r^2 * 1.0
    const float *charge; // na
    const float *gc;     // ng grid coordinates
};

// Voxel masks are stored row by row: row (i, j) of ng voxels along k in row_words(ng)
// 32-bit words, bit k & 31 of word k >> 5.
__host__ __device__ inline int row_words(int ng) { return (ng + 31) / 32; }
inline int mask_words(int ng) { return ng * ng * row_words(ng); }
// per-voxel atom counts, two 16-bit counters per word (ligands have < 65536 atoms)
inline int count_words(int ng) { return (ng * ng * ng + 1) / 2; }
// scratch words per rotation: core + surface masks, delta + value-1 masks, counts
inline size_t vox_scratch_words(int ng) { return 4 * (size_t)mask_words(ng) + count_words(ng); }

// Voxelization scratch of a context, laid out for max_batch rotations:
//   cs  [b][2][mask_words]  core and surface bits (cleared per batch)
//   dmo [b][2][mask_words]  delta mask D and value-1 mask O (written per batch)
//   cnt [b][count_words]    atom counts per voxel, zero between batches
struct VoxScratch {
    uint32_t *cs, *dmo, *cnt;
};
inline VoxScratch vox_scratch(uint32_t *base, int max_batch, int ng) {
    const size_t mw = mask_words(ng);
    return VoxScratch{base, base + 2 * mw * max_batch, base + 4 * mw * max_batch};
}

// A batch of ligand grids as the FFT passes read them: shape value delta where D is set,
// 1 where O is set, 0 elsewhere; charge from the dense grid chg[b][ng^3], which is zero
// between batches (vox_take clears each voxel it reads).
struct VoxGrid {
    const uint32_t *dmo;
    float *chg;
    int ng, rw, mw, ng3;
    float delta;
};
__device__ __forceinline__ float2 vox_take(const VoxGrid &g, int b, int row, int k) {
    const uint32_t *d = g.dmo + (size_t)b * 2 * g.mw + row * g.rw + (k >> 5);
    const uint32_t bit = 1u << (k & 31), dw = d[0], ow = d[g.mw];
    const float sv = (dw & bit) ? g.delta : ((ow & bit) ? 1.0f : 0.0f);
    float *cp = g.chg + (size_t)b * g.ng3 + (size_t)row * g.ng + k;
    const float q = *cp;
    if (__float_as_uint(q) != 0u) *cp = 0.0f;
    return make_float2(sv, q);
}

// Ligand voxelization of B rotations (rot: B*9 rotation coefficients, row-major r11..r33)
// into the masks of vs and the charge grid chg (B * ng^3, zero on entry); vidx: B*na ints.
// boff (optional): device word with the batch's first row of rot (graph-captured batches)
hipError_t launch_voxelize(const LigandDev &lig, const float *rot, int B, const VoxScratch &vs, int *vidx,
                           float *chg, hipStream_t s, const int *boff = nullptr);
// (grid_r, grid_i) pairs lg[b*ng^3 + v] from a voxelized batch; clears its charge grid
hipError_t launch_vox_expand(const VoxGrid &vg, int B, float2 *lg, hipStream_t s);

#define MDK_MAXT 16  // largest t (results per rotation) the device top-t lists hold
#define MDK_MAXT_SELECT (1 << 20)  // largest t of all: above MDK_MAXT, the score maps go through launch_topt

// Per-rotation radix-select state of launch_topt (mdk_topt.hip)
struct TopTState {
    unsigned long long prefix, mask;  // key bits fixed so far
    unsigned rank;                    // entries still wanted among those matching prefix
    unsigned count;                   // compaction cursor
    unsigned done;                    // the threshold is prefix (its lower bits zero)
    unsigned pad;
    unsigned hist[256];
};
// keys sorted per rotation by launch_topt for t results per rotation (a power of two)
unsigned topt_sort_size(int t, int nf);
// Top-t (t > MDK_MAXT) of B score maps smap[b][nf^3] (unnormalised real scores) into
// out[(b0 + b) * t + q] as SortScore with rotation number rot_id[b0 + b], where b0 = *boff
// (0 without boff); st: B states, keys: B * topt_sort_size(t, nf) words.
hipError_t launch_topt(const float *smap, int B, int nf, int t, TopTState *st, unsigned long long *keys,
                       mdk_sortscore *out, const int *rot_id, const int *boff, hipStream_t s);

// A score candidate: its order key (see better()) and its linear grid index
// (x*nf + y)*nf + z.
struct Cand {
    float key;
    int idx;
};

// Candidate order of the reference's per-rotation top-t (src/fft_process.cpp:407-446):
// key descending, then linear index ascending, where the key is the raw value for t == 1
// (strict '<' from the first index) and fl(value / nf^3) for t > 1 (each value is
// normalised before the strict '>' insertion).
__device__ __forceinline__ bool better(float ka, int ia, float kb, int ib) {
    return ka > kb || (ka == kb && ia < ib);
}

constexpr int kNoIndex = 0x7fffffff;  // placeholder entries (-inf, kNoIndex) lose to every candidate

// top-TT list with a compile-time length, best first
template <int TT> struct TopR {
    float v[TT];
    int i[TT];
    __device__ __forceinline__ void init() {
#pragma unroll
        for (int q = 0; q < TT; q++) {
            v[q] = -INFINITY;
            i[q] = kNoIndex;
        }
    }
    __device__ __forceinline__ void push(float vc, int ic) {
        if (!better(vc, ic, v[TT - 1], i[TT - 1])) return;
#pragma unroll
        for (int q = 0; q < TT; q++) {
            if (better(vc, ic, v[q], i[q])) {
                const float tv = v[q];
                const int ti = i[q];
                v[q] = vc;
                i[q] = ic;
                vc = tv;
                ic = ti;
            }
        }
    }
    __device__ __forceinline__ void pop() {
#pragma unroll
        for (int q = 0; q + 1 < TT; q++) {
            v[q] = v[q + 1];
            i[q] = i[q + 1];
        }
        v[TT - 1] = -INFINITY;
        i[TT - 1] = kNoIndex;
    }
};

// Wavefront-wide reductions with DPP lane moves inside each row of 16 lanes (quad
// swaps, then row rotations by 4 and 8) and four lane reads across the rows; the result
// is the same in every lane.  Every lane of the wavefront must take part.
template <int CTRL> __device__ __forceinline__ int dpp_mov(int v) {
    return __builtin_amdgcn_update_dpp(v, v, CTRL, 0xf, 0xf, false);
}
__device__ __forceinline__ float wave_max(float v) {
    v = fmaxf(v, __int_as_float(dpp_mov<0xb1>(__float_as_int(v))));   // quad_perm [1,0,3,2]
    v = fmaxf(v, __int_as_float(dpp_mov<0x4e>(__float_as_int(v))));   // quad_perm [2,3,0,1]
    v = fmaxf(v, __int_as_float(dpp_mov<0x124>(__float_as_int(v))));  // row_ror:4
    v = fmaxf(v, __int_as_float(dpp_mov<0x128>(__float_as_int(v))));  // row_ror:8
    const int b = __float_as_int(v);
    return fmaxf(fmaxf(__int_as_float(__builtin_amdgcn_readlane(b, 0)), __int_as_float(__builtin_amdgcn_readlane(b, 16))),
                 fmaxf(__int_as_float(__builtin_amdgcn_readlane(b, 32)), __int_as_float(__builtin_amdgcn_readlane(b, 48))));
}
__device__ __forceinline__ int wave_min(int v) {
    v = min(v, dpp_mov<0xb1>(v));
    v = min(v, dpp_mov<0x4e>(v));
    v = min(v, dpp_mov<0x124>(v));
    v = min(v, dpp_mov<0x128>(v));
    return min(min(__builtin_amdgcn_readlane(v, 0), __builtin_amdgcn_readlane(v, 16)),
               min(__builtin_amdgcn_readlane(v, 32), __builtin_amdgcn_readlane(v, 48)));
}

// Top-t of the sorted lists held by the 64 lanes of a wavefront, in t rounds: the best
// head key (wave_max), the lowest index among the heads holding it, and a pop by the lane
// that owns that entry.  Indices are unique except for the (-inf, kNoIndex) placeholders.
// Every lane receives the result in (ov, oi).
template <int TT, int R = 0>
__device__ __forceinline__ void wave_top(TopR<TT> &L, int t, float *ov, int *oi) {
    if constexpr (R < TT) {  // one round per rank, unrolled at compile time
        if (R >= t) return;  // uniform
        const float m = wave_max(L.v[0]);
        const bool at = L.v[0] == m;
        const unsigned long long hs = __ballot(at);
        int wi;
        if (__popcll(hs) == 1)
            wi = __builtin_amdgcn_readlane(L.i[0], __ffsll((long long)hs) - 1);
        else
            wi = wave_min(at ? L.i[0] : kNoIndex);
        const bool own = at && L.i[0] == wi;
        // the owner's own key (m may differ from it in the sign of a zero)
        ov[R] = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(L.v[0]), __ffsll((long long)__ballot(own)) - 1));
        oi[R] = wi;
        if (own) L.pop();
        wave_top<TT, R + 1>(L, t, ov, oi);
    }
}

// The t-th largest of the lane values m across the wavefront (counting repeats; -inf if
// fewer than t lanes hold a value).  With m = each lane's largest score, at least t of the
// wavefront's scores are >= the result.  Every lane of the wavefront must call it.
__device__ __forceinline__ float wave_kth_max(float m, int t) {
    float tau = -INFINITY;
    const int lane = (int)(threadIdx.x & 63);
    for (int r = 0; r < t; r++) {  // t is uniform
        const float w = wave_max(m);
        tau = w;
        const unsigned long long holders = __ballot(m == w);
        if (holders && lane == __ffsll((long long)holders) - 1) m = -INFINITY;  // retire one holder
    }
    return tau;
}

// Largest float f such that every v < f has fl(v / c) < fl(tau / c), for c > 0 with c
// exact in 24 bits: f <= c * prev(fl(tau / c)) and the product is exact in double.
// Scores below f cannot tie with tau after the division, so they lose to it.
__device__ __forceinline__ float quotient_cut(float tau, float c) {
    const float qp = nextafterf(tau / c, -INFINITY);
    const double d = (double)c * (double)qp;
    float f = (float)d;
    if ((double)f > d) f = nextafterf(f, -INFINITY);
    return f;
}

// register top-t list with a run-time t <= MDK_MAXT (static indexing); (wv, wi) mirrors
// the t-th entry so that most candidates are rejected with one comparison
struct TopList {
    float k[MDK_MAXT];
    int i[MDK_MAXT];
    float wv;
    int wi;
    __device__ void init() {
#pragma unroll
        for (int q = 0; q < MDK_MAXT; q++) {
            k[q] = -INFINITY;
            i[q] = kNoIndex;
        }
        wv = -INFINITY;
        wi = kNoIndex;
    }
    __device__ void push(int t, float kc, int ic) {
        if (!better(kc, ic, wv, wi)) return;
#pragma unroll
        for (int q = 0; q < MDK_MAXT; q++) {
            if (q < t && better(kc, ic, k[q], i[q])) {
                const float tk = k[q];
                const int ti = i[q];
                k[q] = kc;
                i[q] = ic;
                kc = tk;
                ic = ti;
            }
            if (q == t - 1) {
                wv = k[q];
                wi = i[q];
            }
        }
    }
    // insert another list sorted best-first; stops at its first rejected entry
    __device__ void merge_sorted(int t, const Cand *other) {
        for (int q = 0; q < t; q++) {
            if (!better(other[q].key, other[q].idx, wv, wi)) break;
            push(t, other[q].key, other[q].idx);
        }
    }
    __device__ void store(int t, Cand *dst) const {
        for (int q = 0; q < t; q++) {
#pragma unroll
            for (int s = 0; s < MDK_MAXT; s++)
                if (s == q) dst[q] = Cand{k[s], i[s]};
        }
    }
};

// Arguments of the FFT-side pipeline for one batch (buffers as in mdk_dock.hip).
struct PipeArgs {
    int B, ng, nd, H, Hp, t;
    float nf3f;
    VoxGrid vox;        // voxelized ligands of the batch
    float2 *a1, *a2, *cc;
    const float2 *rec;
    Cand *cand;
    float *smap;
    const float2 *tw4;  // exp(-2 pi i m / nf), m < nf
    hipStream_t stream;
    hipEvent_t *ev;     // optional: 5 events, one after each pass
    int variant;        // kernel variant of the correlation pass (MDK_XCORR_VARIANT)
    int cu_count;       // compute units of the device (persistent grids)
    int inv_variant;    // fused inverse: 0 loads each plane at its start, 1 prefetches, 2 and 3 as 0 and 1
                        // with running top-t lists over contiguous planes (MDK_INV_VARIANT); at
                        // nf = 256: 0 two passes, 1 the half-plane inverse with 512 threads
                        // reading each plane once (the default), 2-3 with 1024 reading it twice
    bool spectrum_only = false;  // stop after the fold: cc holds Q[b][kx][ky][kz]
    int fwd_variant = 0;         // forward passes: 0 fused z + y where it fits, the next plane prefetched (nf = 256: over half planes), 1 separate, 2 fused without the prefetch (MDK_FWD_VARIANT)
};

// Compile-time specialised pipeline (mdk_pipe.hip) for the grid sizes it was built for.
bool pipe_available(int nf);
// default kernel variants of a size (correlation, inverse, forward); false without a
// specialised pipeline
bool pipe_defaults(int nf, int *xcorr, int *inv, int *fwd);
// whether a size's specialised pipeline was built with variant v of a pass (which: 0
// correlation, 1 inverse, 2 forward)
bool pipe_has_variant(int nf, int which, int v);
// top-t candidate lists the specialised pipeline leaves per rotation (each of t entries)
int pipe_cands_per_rotation(int nf);
hipError_t pipe_run(int nf, const PipeArgs &a);

// Device merge (mdk_merge.hip): scratch bytes of launch_topn for n results, and the merge of
// the best n of m results res[] into out[] (work: topn_scratch_bytes(n) bytes of device memory)
size_t topn_scratch_bytes(int n);
hipError_t launch_topn(const mdk_sortscore *res, unsigned m, unsigned n, void *work, mdk_sortscore *out,
                       hipStream_t s);
// The allocation-free merge (mdk_topn_noalloc) keeps its select state, keys and n results
// in a fixed device array of MDK_TOPN_STATIC_BYTES per GPU: n <= MDK_TOPN_STATIC_MAX fits.
#define MDK_TOPN_STATIC_BYTES (1u << 20)
#define MDK_TOPN_STATIC_MAX 32768
// byte offset of the results within that array for n results
size_t topn_static_out_offset(int n);
// look the array's device address up for `device` (which must be current); once per device
hipError_t topn_static_setup(int device);
// the address looked up by topn_static_setup, or nullptr
void *topn_static_ptr(int device);
