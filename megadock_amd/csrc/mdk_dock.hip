// Batched FFT docking pipeline for MI355X (gfx950) and the libmdk C ABI.
//
// Per batch of B rotations (nf = 2 ng points per axis, nd = (nf - ng) / 2 offset of the
// ligand inside the FFT grid, H = nf/2 + 1):
//   voxelize      ligand grids lg[b][ng^3]                           (mdk_vox.hip)
//   k_fwd_z       z-FFT of the ng^2 non-zero rows        -> a1[b][ng][ng][nf]
//   k_fwd_y       y-FFT of the ng*nf non-zero columns     -> a2[b][ng][nf][nf]
//   k_x_corr      x-FFT, conj(R)L, Hermitian fold, x-IFFT -> cc[b][nf][nf][Hp]
//   k_inv_y       y-IFFT in place on the kz <= nf/2 half
//   k_inv_z_top   z-IFFT of the Hermitian rows (real output) + per-row top-t
//   k_reduce_top  per-rotation top-t, written as SortScore
// The score is Re IFFT(conj(R) L) (src/fft_process.cpp:333-346, :386-388), which equals
// IFFT of the Hermitian part Q(k) = (P(k) + conj(P(-k))) / 2 of P = conj(R) L; only the
// kz <= nf/2 half of Q is formed, so the inverse transform moves half the data.
// The zero padding of the ligand (7/8 of the grid) is never read: the forward passes
// transform only the rows that hold ligand voxels.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "mdk_fft.h"
#include "mdk_internal.h"
#include "../../include/mdk_hip.h"

// real-space detail sums (mdk_detail.hip)
hipError_t launch_detail(const float2 *lg, const float2 *recgrid, const float *rpsc, const int *trans_dev, int B, int ng,
                         float *out_dev, hipStream_t s);

hipError_t launch_decoy_coords(const float *xyz, int na, const float *pre, const float *rot, const int *trans, int nd,
                               float spacing, const float *r, float *out, hipStream_t s);

// receptor grids on the device (mdk_recvox.hip)
size_t receptor_voxel_scratch_bytes(int ng);
hipError_t launch_receptor_voxel(const float *xyz, const float *radius, const float *charge, const float *ace,
                                 const float *gc, int na, int ng, float gw, float aceratio, float rec_core, float eratio,
                                 float gc0, float gc1, void *scratch, float2 *out, hipStream_t s);

// off-grid scores (mdk_grad.hip)
size_t grad_scratch_doubles(int max_batch);
hipError_t launch_spec_eval(const float2 *Q, const float *t_dev, int B, int nf, int H, int Hp, double *part,
                            float *out_dev, hipStream_t s);
hipError_t launch_companion(const LigandDev &lig, const float *wshape, const float2 *recgrid, const float *rot_dev,
                            const float *t_dev, int B, int nf, double *part, float *out_dev, hipStream_t s);

#define MDK_NSTAGE 7  // voxelize, fwd_z, fwd_y, x_corr, inv_y, inv_z_top, reduce

namespace {

thread_local std::string g_err;

int fail(const std::string &msg) {
    g_err = msg;
    return 1;
}

#define HIPCHK(expr)                                                                     \
    do {                                                                                 \
        hipError_t e_ = (expr);                                                          \
        if (e_ != hipSuccess)                                                            \
            return fail(std::string(#expr) + ": " + hipGetErrorString(e_));              \
    } while (0)

constexpr int WZ = 16;   // lines per workgroup, z passes
constexpr int WY = 16;   // lines per workgroup, y passes
constexpr int WX = 8;    // kz columns per workgroup in the x pass (plus 8 partner columns)
constexpr int NT = 256;  // threads per workgroup

__device__ __forceinline__ void load_tw(float2 *tw_s, const float2 *tw_g, int ntw) {
    for (int q = threadIdx.x; q < ntw; q += blockDim.x) tw_s[q] = tw_g[q];
}

// ---------------------------------------------------------------- forward z
// lines: (b, x, y) with x, y < ng; row z < ng of lg sits at z + nd of the FFT line.
__global__ void __launch_bounds__(NT) k_fwd_z(const float2 *__restrict__ lg, float2 *__restrict__ a1,
                                              FftPlan p, const float2 *__restrict__ twg, int ng,
                                              int nd, int nlines) {
    extern __shared__ float2 sm[];
    const int nf = p.n, LS = WZ + 1;
    float2 *tw = sm, *A = sm + p.ntw, *Bf = A + nf * LS;
    load_tw(tw, twg, p.ntw);
    const int line0 = blockIdx.x * WZ;
    for (int q = threadIdx.x; q < WZ * nf; q += NT) {
        const int l = q / nf, z = q % nf;
        float2 v = make_float2(0.f, 0.f);
        if (line0 + l < nlines && z >= nd && z < nd + ng) v = lg[(size_t)(line0 + l) * ng + (z - nd)];
        A[z * LS + l] = v;
    }
    __syncthreads();
    float2 *res = lds_fft<false>(A, Bf, tw, p, LS, threadIdx.x % WZ, threadIdx.x / WZ, NT / WZ, true);
    for (int q = threadIdx.x; q < WZ * nf; q += NT) {
        const int l = q / nf, kz = q % nf;
        if (line0 + l < nlines) a1[(size_t)(line0 + l) * nf + kz] = res[kz * LS + l];
    }
}

// ---------------------------------------------------------------- forward y
// block (kz chunk, x, b); rows y < ng of a1 sit at y + nd.
__global__ void __launch_bounds__(NT) k_fwd_y(const float2 *__restrict__ a1, float2 *__restrict__ a2,
                                              FftPlan p, const float2 *__restrict__ twg, int ng,
                                              int nd) {
    extern __shared__ float2 sm[];
    const int nf = p.n, LS = WY + 1;
    float2 *tw = sm, *A = sm + p.ntw, *Bf = A + nf * LS;
    load_tw(tw, twg, p.ntw);
    const int kz0 = blockIdx.x * WY, x = blockIdx.y, b = blockIdx.z;
    const float2 *src = a1 + ((size_t)b * ng + x) * ng * nf;
    for (int q = threadIdx.x; q < WY * nf; q += NT) {
        const int l = q % WY, y = q / WY;
        float2 v = make_float2(0.f, 0.f);
        if (kz0 + l < nf && y >= nd && y < nd + ng) v = src[(size_t)(y - nd) * nf + kz0 + l];
        A[y * LS + l] = v;
    }
    __syncthreads();
    float2 *res = lds_fft<false>(A, Bf, tw, p, LS, threadIdx.x % WY, threadIdx.x / WY, NT / WY, true);
    float2 *dst = a2 + ((size_t)b * ng + x) * nf * nf;
    for (int q = threadIdx.x; q < WY * nf; q += NT) {
        const int l = q % WY, ky = q / WY;
        if (kz0 + l < nf) dst[(size_t)ky * nf + kz0 + l] = res[ky * LS + l];
    }
}

// ---------------------------------------------------------------- x pass, receptor
// plain forward x-FFT of the receptor: rec[kx][ky][kz] (full spectrum)
__global__ void __launch_bounds__(NT) k_fwd_x_plain(const float2 *__restrict__ a2, float2 *__restrict__ rec,
                                                    FftPlan p, const float2 *__restrict__ twg, int ng,
                                                    int nd) {
    extern __shared__ float2 sm[];
    const int nf = p.n, LS = WY + 1;
    float2 *tw = sm, *A = sm + p.ntw, *Bf = A + nf * LS;
    load_tw(tw, twg, p.ntw);
    const int kz0 = blockIdx.x * WY, ky = blockIdx.y;
    for (int q = threadIdx.x; q < WY * nf; q += NT) {
        const int l = q % WY, x = q / WY;
        float2 v = make_float2(0.f, 0.f);
        if (kz0 + l < nf && x >= nd && x < nd + ng) v = a2[((size_t)(x - nd) * nf + ky) * nf + kz0 + l];
        A[x * LS + l] = v;
    }
    __syncthreads();
    float2 *res = lds_fft<false>(A, Bf, tw, p, LS, threadIdx.x % WY, threadIdx.x / WY, NT / WY, true);
    for (int q = threadIdx.x; q < WY * nf; q += NT) {
        const int l = q % WY, kx = q / WY;
        if (kz0 + l < nf) rec[((size_t)kx * nf + ky) * nf + kz0 + l] = res[kx * LS + l];
    }
}

// ---------------------------------------------------------------- x pass, docking
// block (kz chunk c, ky, b).  Lines 0..WX-1: (ky, kz = c*WX + l); lines WX..2WX-1: the
// partner columns (-ky, -kz).  Forward x-FFT of all 2*WX, then
//   Q(k) = (conj(R(k)) L(k) + conj(conj(R(-k)) L(-k))) / 2
// and an inverse x-FFT of the WX lines of Q.
__global__ void __launch_bounds__(NT) k_x_corr(const float2 *__restrict__ a2, const float2 *__restrict__ rec,
                                               float2 *__restrict__ cc, FftPlan p,
                                               const float2 *__restrict__ twg, int ng, int nd, int H,
                                               int Hp) {
    extern __shared__ float2 sm[];
    const int nf = p.n, LS = 2 * WX + 1;
    float2 *tw = sm, *A = sm + p.ntw, *Bf = A + nf * LS;
    load_tw(tw, twg, p.ntw);
    const int c = blockIdx.x, ky = blockIdx.y, b = blockIdx.z;
    const int kyp = (nf - ky) % nf;
    const float2 *src = a2 + (size_t)b * ng * nf * nf;
    for (int q = threadIdx.x; q < 2 * WX * nf; q += NT) {
        const int l = q % (2 * WX), x = q / (2 * WX);
        const int kz = c * WX + (l % WX);
        float2 v = make_float2(0.f, 0.f);
        if (kz < H && x >= nd && x < nd + ng) {
            if (l < WX) v = src[((size_t)(x - nd) * nf + ky) * nf + kz];
            else v = src[((size_t)(x - nd) * nf + kyp) * nf + (nf - kz) % nf];
        }
        A[x * LS + l] = v;
    }
    __syncthreads();
    float2 *L = lds_fft<false>(A, Bf, tw, p, LS, threadIdx.x % (2 * WX), threadIdx.x / (2 * WX),
                               NT / (2 * WX), true);
    float2 *Qb = (L == A) ? Bf : A;
    for (int q = threadIdx.x; q < WX * nf; q += NT) {
        const int l = q % WX, kx = q / WX;
        const int kz = c * WX + l;
        float2 Q = make_float2(0.f, 0.f);
        if (kz < H) {
            const int kxp = (nf - kx) % nf, kzp = (nf - kz) % nf;
            const float2 r1 = rec[((size_t)kx * nf + ky) * nf + kz];
            const float2 r2 = rec[((size_t)kxp * nf + kyp) * nf + kzp];
            const float2 l1 = L[kx * LS + l], l2 = L[kxp * LS + WX + l];
            // conj(r1) * l1 and conj(r2) * l2
            const float2 p1 = make_float2(r1.x * l1.x + r1.y * l1.y, r1.x * l1.y - r1.y * l1.x);
            const float2 p2 = make_float2(r2.x * l2.x + r2.y * l2.y, r2.x * l2.y - r2.y * l2.x);
            Q = make_float2(0.5f * (p1.x + p2.x), 0.5f * (p1.y - p2.y));
        }
        Qb[kx * LS + l] = Q;
    }
    __syncthreads();
    float2 *other = (Qb == A) ? Bf : A;
    float2 *res = lds_fft<true>(Qb, other, tw, p, LS, threadIdx.x % WX, threadIdx.x / WX, NT / WX, true);
    float2 *dst = cc + (size_t)b * nf * nf * Hp;
    for (int q = threadIdx.x; q < WX * nf; q += NT) {
        const int l = q % WX, x = q / WX;
        const int kz = c * WX + l;
        if (kz < Hp) dst[((size_t)x * nf + ky) * Hp + kz] = res[x * LS + l];
    }
}

// ---------------------------------------------------------------- inverse y (in place)
__global__ void __launch_bounds__(NT) k_inv_y(float2 *__restrict__ cc, FftPlan p,
                                              const float2 *__restrict__ twg, int Hp) {
    extern __shared__ float2 sm[];
    const int nf = p.n, LS = WY + 1;
    float2 *tw = sm, *A = sm + p.ntw, *Bf = A + nf * LS;
    load_tw(tw, twg, p.ntw);
    const int kz0 = blockIdx.x * WY, x = blockIdx.y, b = blockIdx.z;
    float2 *base = cc + ((size_t)b * nf + x) * nf * Hp;
    for (int q = threadIdx.x; q < WY * nf; q += NT) {
        const int l = q % WY, ky = q / WY;
        A[ky * LS + l] = base[(size_t)ky * Hp + kz0 + l];
    }
    __syncthreads();
    float2 *res = lds_fft<true>(A, Bf, tw, p, LS, threadIdx.x % WY, threadIdx.x / WY, NT / WY, true);
    for (int q = threadIdx.x; q < WY * nf; q += NT) {
        const int l = q % WY, y = q / WY;
        base[(size_t)y * Hp + kz0 + l] = res[y * LS + l];
    }
}

// ---------------------------------------------------------------- inverse z + top-t
// rows r = (b*nf + x)*nf + y; the row's spectrum is Hermitian, only kz <= nf/2 stored.
// Writes either the real score row (smap) or the row's top-t candidates.
__global__ void __launch_bounds__(NT) k_inv_z_top(const float2 *__restrict__ cc, Cand *__restrict__ cand,
                                                  float *__restrict__ smap, FftPlan p,
                                                  const float2 *__restrict__ twg, int H, int Hp,
                                                  int nrows, int t, float nf3f) {
    extern __shared__ float2 sm[];
    const int nf = p.n, LS = WZ + 1;
    float2 *tw = sm, *A = sm + p.ntw, *Bf = A + nf * LS;
    Cand *lists = reinterpret_cast<Cand *>(Bf + nf * LS);  // [WZ][NT/WZ][t]
    load_tw(tw, twg, p.ntw);
    const int row0 = blockIdx.x * WZ;
    for (int q = threadIdx.x; q < WZ * H; q += NT) {
        const int l = q / H, kz = q % H;
        float2 v = make_float2(0.f, 0.f);
        if (row0 + l < nrows) v = cc[(size_t)(row0 + l) * Hp + kz];
        A[kz * LS + l] = v;
        if (kz >= 1 && nf - kz > kz) A[(nf - kz) * LS + l] = make_float2(v.x, -v.y);
    }
    __syncthreads();
    float2 *res = lds_fft<true>(A, Bf, tw, p, LS, threadIdx.x % WZ, threadIdx.x / WZ, NT / WZ, true);
    const int l = threadIdx.x % WZ, jt = threadIdx.x / WZ, tpl = NT / WZ;
    const int r = row0 + l;
    if (smap) {
        if (r < nrows)
            for (int z = jt; z < nf; z += tpl) smap[(size_t)r * nf + z] = res[z * LS + l].x;
        return;
    }
    // per-thread top-t over its elements; key = value (t == 1) or value / nf^3 (t > 1)
    TopList tl;
    tl.init();
    if (r < nrows) {
        const int rr = r % (nf * nf);  // row within the rotation: x*nf + y
        for (int z = jt; z < nf; z += tpl) {
            const float v = res[z * LS + l].x;
            tl.push(t, (t == 1) ? v : v / nf3f, rr * nf + z);
        }
    }
    Cand *my = lists + ((size_t)l * tpl + jt) * t;
    for (int q = 0; q < t; q++) {
#pragma unroll
        for (int s = 0; s < MDK_MAXT; s++)
            if (s == q) my[q] = Cand{tl.k[s], tl.i[s]};
    }
    __syncthreads();
    if (jt == 0 && r < nrows) {
        TopList m;
        m.init();
        const Cand *all = lists + (size_t)l * tpl * t;
        for (int q = 0; q < tpl * t; q++) m.push(t, all[q].key, all[q].idx);
        for (int q = 0; q < t; q++) {
#pragma unroll
            for (int s = 0; s < MDK_MAXT; s++)
                if (s == q) cand[(size_t)r * t + q] = Cand{m.k[s], m.i[s]};
        }
    }
}

// ---------------------------------------------------------------- per-rotation top-t
// Merges a rotation's candidate lists into its top-t, then applies the reference's
// starting values: t == 1 starts from (0.0, index 0) and keeps a candidate only if its
// value is > 0; t > 1 starts from -99999 and keeps candidates whose normalised value is
// > -99999 (src/fft_process.cpp:407-446).  Scores are written normalised by nf^3.
// boff (optional): device word with the batch's first rotation, added to rot_id and out
__global__ void __launch_bounds__(NT) k_reduce_top(const Cand *__restrict__ cand, mdk_sortscore *__restrict__ out,
                                                   const int *__restrict__ rot_id, const int *__restrict__ boff,
                                                   int lists_per_rot, int t, int nf, float nf3f) {
    __shared__ Cand lists[NT * MDK_MAXT];
    const int b = blockIdx.x;
    const int b0 = boff ? *boff : 0;
    rot_id += b0;
    out += (size_t)b0 * t;
    const Cand *src = cand + (size_t)b * lists_per_rot * t;
    TopList tl;
    tl.init();
    const int n = lists_per_rot * t;
    for (int q = threadIdx.x; q < n; q += NT) tl.push(t, src[q].key, src[q].idx);
    tl.store(t, lists + threadIdx.x * t);
    __syncthreads();
    // tree merge of the NT lists (every thread reaches every barrier)
    for (int half = NT / 2; half >= 1; half >>= 1) {
        if (threadIdx.x < half) tl.merge_sorted(t, lists + (threadIdx.x + half) * t);
        __syncthreads();
        if (threadIdx.x < half) tl.store(t, lists + threadIdx.x * t);
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const int nf2 = nf * nf;
        bool keep = true;  // the lists are sorted, so the kept entries are a prefix
        for (int q = 0; q < t; q++) {
            const Cand c = lists[q];
            keep = keep && c.idx != kNoIndex && (t == 1 ? c.key > 0.0f : c.key > -99999.0f);
            const int idx = keep ? c.idx : 0;
            mdk_sortscore s;
            s.score = keep ? (t == 1 ? c.key / nf3f : c.key) : (t == 1 ? 0.0f : -99999.0f);
            s.index[0] = rot_id[b];
            s.index[1] = idx / nf2;
            s.index[2] = (idx / nf) % nf;
            s.index[3] = idx % nf;
            out[(size_t)b * t + q] = s;
        }
    }
}

// Batch cursor of a captured batch graph: ctl = {base, step, count}; every replay docks
// rotations from base + count * step and advances count.
__global__ void k_batch_begin(int *ctl, int *boff) {
    if (threadIdx.x == 0) {
        const int q = ctl[2];
        boff[0] = ctl[0] + q * ctl[1];
        ctl[2] = q + 1;
    }
}

__global__ void k_batch_reset(int *ctl, int base, int step) {
    if (threadIdx.x == 0) {
        ctl[0] = base;
        ctl[1] = step;
        ctl[2] = 0;
    }
}

__global__ void k_iota(int *p, int first, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = first + i;
}

// ---------------------------------------------------------------- 1D test kernel
__global__ void __launch_bounds__(NT) k_lines(const float2 *__restrict__ in, float2 *__restrict__ out,
                                              FftPlan p, const float2 *__restrict__ twg, int nlines,
                                              int inv) {
    extern __shared__ float2 sm[];
    const int n = p.n, LS = WZ + 1;
    float2 *tw = sm, *A = sm + p.ntw, *Bf = A + n * LS;
    load_tw(tw, twg, p.ntw);
    const int line0 = blockIdx.x * WZ;
    for (int q = threadIdx.x; q < WZ * n; q += NT) {
        const int l = q / n, e = q % n;
        A[e * LS + l] = (line0 + l < nlines) ? in[(size_t)(line0 + l) * n + e] : make_float2(0.f, 0.f);
    }
    __syncthreads();
    float2 *res = inv ? lds_fft<true>(A, Bf, tw, p, LS, threadIdx.x % WZ, threadIdx.x / WZ, NT / WZ, true)
                      : lds_fft<false>(A, Bf, tw, p, LS, threadIdx.x % WZ, threadIdx.x / WZ, NT / WZ, true);
    for (int q = threadIdx.x; q < WZ * n; q += NT) {
        const int l = q / n, e = q % n;
        if (line0 + l < nlines) out[(size_t)(line0 + l) * n + e] = res[e * LS + l];
    }
}

// ---------------------------------------------------------------- host-side plan
// dynamic LDS above the default 64 KiB window must be requested per kernel
template <typename K> hipError_t allow_lds(K kernel, size_t bytes) {
    if (bytes <= 64 * 1024) return hipSuccess;
    return hipFuncSetAttribute(reinterpret_cast<const void *>(kernel),
                               hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
}

bool make_plan(int n, FftPlan &p, std::vector<float2> &tw) {
    static const int pref[] = {8, 4, 2, 3, 5, 7, 11, 13};
    int m = n, ns = 0;
    int fac[MDK_MAX_STAGES];
    // radix 8 first, then 4, a trailing 2 if needed, then odd primes
    for (int a = 0; a < 8 && m > 1; a++)
        while (m % pref[a] == 0) {
            if (ns >= MDK_MAX_STAGES) return false;
            fac[ns++] = pref[a];
            m /= pref[a];
        }
    if (m != 1) return false;
    p.n = n;
    p.nst = ns;
    tw.clear();
    int Ns = 1;
    for (int s = 0; s < ns; s++) {
        const int R = fac[s];
        p.radix[s] = R;
        p.twoff[s] = (int)tw.size();
        for (int r = 1; r < R; r++)
            for (int k = 0; k < Ns; k++) {
                const double a = -2.0 * M_PI * (double)r * k / ((double)Ns * R);
                tw.push_back(make_float2((float)std::cos(a), (float)std::sin(a)));
            }
        Ns *= R;
    }
    if (tw.empty()) tw.push_back(make_float2(1.f, 0.f));
    p.ntw = (int)tw.size();
    return true;
}

}  // namespace

// ==================================================================== context
// Buffers of one batch in flight, and the stream its kernels run on.
struct BatchSlot {
    hipStream_t stream = nullptr;
    uint32_t *d_bits = nullptr;  // voxelization scratch, laid out by vs
    VoxScratch vs{};
    float *d_chg = nullptr;      // charge grids, max_batch * ng^3, zero between batches
    int *d_vidx = nullptr;
    float2 *d_lg = nullptr, *d_a1 = nullptr, *d_a2 = nullptr, *d_cc = nullptr;
    Cand *d_cand = nullptr;
    mdk_sortscore *d_bout = nullptr;  // max_batch * t
    // t > MDK_MAXT: the batch's score maps and the per-rotation select (launch_topt)
    float *d_smap = nullptr;              // max_batch * nf^3
    TopTState *d_tsel = nullptr;          // max_batch
    unsigned long long *d_tkeys = nullptr;  // max_batch * topt_sort_size(t, nf)
    hipEvent_t done = nullptr;        // recorded after a range's last batch on this stream
    // one full batch captured as a graph (MDK_GRAPH): cursor, batch offset, executable
    int *d_ctl = nullptr;             // {base, step, count}, then the batch offset
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

struct mdk_ctx {
    int device = 0, nf = 0, ng = 0, na = 0, t = 1, max_batch = 1;
    int nd = 0, H = 0, Hp = 0;
    hipStream_t stream = nullptr;  // slot[0].stream: set-up, single batches, copies
    FftPlan plan{};
    float2 *d_tw = nullptr;
    float2 *d_tw4 = nullptr;  // exp(-2 pi i m / nf): twiddles of the specialised pipeline
    bool use_pipe = false;    // compile-time specialised kernels exist for nf
    int xcorr_variant = 6;    // MDK_XCORR_VARIANT: correlation-pass variant (tuning)
    int cu_count = 256;       // compute units (persistent grids)
    int inv_variant = 2;      // MDK_INV_VARIANT: fused-inverse variant (tuning)
    int fwd_variant = 0;      // MDK_FWD_VARIANT: 0 fused forward z + y, prefetching (nf = 256: over half planes), 1 separate passes, 2 fused, not prefetching
    float2 *d_rec = nullptr;
    bool have_rec = false, have_lig = false;
    // ligand
    float *d_xyz = nullptr, *d_rc2 = nullptr, *d_rs2 = nullptr, *d_charge = nullptr, *d_gc = nullptr;
    LigandDev lig{};
    // rotations
    float *d_rot = nullptr;
    int nrot = 0;
    int *d_rotid = nullptr;  // iota table, rotation i has number i
    float *d_brot = nullptr; // staging for mdk_dock (max_batch * 9)
    int *d_bid = nullptr;    // staging ids
    // batch buffers: one set per stream; mdk_dock_range alternates batches between the
    // streams so that one batch's memory-bound passes overlap another's compute-bound ones
    static constexpr int kMaxSlots = 3;
    BatchSlot slot[kMaxSlots];
    int nslots = 1;                // batch slots and streams (MDK_STREAMS: 1 to 3)
    bool use_graph = false;        // MDK_GRAPH: full batches of mdk_dock_range replay a captured graph
    mdk_sortscore *d_results = nullptr; // nrot * t
    // off-grid scores and gradients (mdk_grad.hip)
    float2 *d_recgrid = nullptr;  // receptor grids (grid_r, grid_i), ng^3, real space
    float *d_wshape = nullptr;    // companion shape weight per ligand atom
    float *d_bt = nullptr;        // translations of a batch (max_batch * 3)
    float *d_gout = nullptr;      // results of a batch (max_batch * 7)
    double *d_gpart = nullptr;    // partial sums
    float *d_rpsc = nullptr;      // receptor shape grid without the ACE term (mdk_score_detail), ng^3
    int *d_btrans = nullptr;      // grid translations of a batch (max_batch * 3)
    float *d_dsum = nullptr;      // detail sums of a batch (max_batch * 9)
    bool have_recgrid = false;
    void *d_topn_scratch = nullptr;      // device merge (mdk_topn)
    mdk_sortscore *d_topn_out = nullptr;
    int topn_cap = 0;
    mdk_sortscore *h_topn = nullptr;  // pinned host copy of a started merge (mdk_topn_start)
    int h_topn_cap = 0, topn_pending = 0;
    hipEvent_t topn_done = nullptr;
    // pinned staging of the receptor grids and of the ligand's atom arrays: the set-up calls
    // copy from them on the context's stream, behind the work queued before; an event after
    // each copy guards the buffer's reuse
    float2 *h_rstage = nullptr;
    float *h_lstage = nullptr;
    size_t h_rstage_n = 0, h_lstage_n = 0;  // bytes
    hipEvent_t ev_rstage = nullptr, ev_lstage = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    std::vector<hipEvent_t> tev;  // mdk_dock_range_timed: MDK_NSTAGE + 1 events per batch
};

namespace {

size_t lds_bytes(const mdk_ctx *c, int lines) {
    return sizeof(float2) * ((size_t)c->plan.ntw + 2 * (size_t)c->nf * (lines + 1));
}

VoxGrid vox_grid(const mdk_ctx *c, const BatchSlot &S) {
    const int ng = c->ng;
    return VoxGrid{S.vs.dmo, S.d_chg, ng, row_words(ng), mask_words(ng), ng * ng * ng, c->lig.delta};
}

// Run the forward ligand passes + correlation + inverse passes for B rotations that
// launch_voxelize has voxelized; output top-t candidates (or the score map).
// ev (optional): events recorded after each of the five FFT-side stages
// spectrum_only: stop after the fold, leaving Q[b][kx][ky][kz] in S.d_cc (specialised sizes)
// from_lg: the ligand grids are already in S.d_lg (given by the host, mdk_dock_grid); the
// generic kernels read them there
int run_fft_pipeline(mdk_ctx *c, BatchSlot &S, int B, float *smap_dev, hipEvent_t *ev = nullptr,
                     bool spectrum_only = false, bool from_lg = false) {
    const int ng = c->ng, nf = c->nf, nd = c->nd;
    hipStream_t s = S.stream;
    if (c->use_pipe && !from_lg) {
        PipeArgs a;
        a.B = B;
        a.ng = ng;
        a.nd = nd;
        a.H = c->H;
        a.Hp = c->Hp;
        a.t = smap_dev ? 1 : c->t;  // the score maps need no top-t lists
        a.nf3f = (float)((size_t)nf * nf * nf);
        a.vox = vox_grid(c, S);
        a.a1 = S.d_a1;
        a.a2 = S.d_a2;
        a.cc = S.d_cc;
        a.rec = c->d_rec;
        a.cand = S.d_cand;
        a.smap = smap_dev;
        a.tw4 = c->d_tw4;
        a.stream = s;
        a.ev = ev;
        a.variant = c->xcorr_variant;
        a.cu_count = c->cu_count;
        a.inv_variant = c->inv_variant;
        a.fwd_variant = c->fwd_variant;
        a.spectrum_only = spectrum_only;
        HIPCHK(pipe_run(nf, a));
        return 0;
    }
    if (!from_lg) HIPCHK(launch_vox_expand(vox_grid(c, S), B, S.d_lg, s));
    const int lines_z = B * ng * ng;
    k_fwd_z<<<(lines_z + WZ - 1) / WZ, NT, lds_bytes(c, WZ), s>>>(S.d_lg, S.d_a1, c->plan, c->d_tw, ng, nd, lines_z);
    if (ev) HIPCHK(hipEventRecord(ev[0], s));
    k_fwd_y<<<dim3((nf + WY - 1) / WY, ng, B), NT, lds_bytes(c, WY), s>>>(S.d_a1, S.d_a2, c->plan, c->d_tw, ng, nd);
    if (ev) HIPCHK(hipEventRecord(ev[1], s));
    k_x_corr<<<dim3(c->Hp / WX, nf, B), NT, lds_bytes(c, 2 * WX), s>>>(S.d_a2, c->d_rec, S.d_cc, c->plan, c->d_tw, ng, nd, c->H, c->Hp);
    if (ev) HIPCHK(hipEventRecord(ev[2], s));
    k_inv_y<<<dim3(c->Hp / WY, nf, B), NT, lds_bytes(c, WY), s>>>(S.d_cc, c->plan, c->d_tw, c->Hp);
    if (ev) HIPCHK(hipEventRecord(ev[3], s));
    const int rows = B * nf * nf;
    const int tl = smap_dev ? 1 : c->t;  // the score maps need no top-t lists
    const size_t lds_z = lds_bytes(c, WZ) + sizeof(Cand) * (size_t)NT * tl;
    k_inv_z_top<<<(rows + WZ - 1) / WZ, NT, lds_z, s>>>(S.d_cc, S.d_cand, smap_dev, c->plan, c->d_tw, c->H, c->Hp, rows, tl,
                                                       (float)((size_t)nf * nf * nf));
    if (ev) HIPCHK(hipEventRecord(ev[4], s));
    HIPCHK(hipGetLastError());
    return 0;
}

// ev (optional): MDK_NSTAGE + 1 events bracketing the stages of one batch
// boff (optional): device word with the batch's first row, added to rot_dev, id_dev, out_dev
int dock_batch(mdk_ctx *c, BatchSlot &S, const float *rot_dev, const int *id_dev, int B, mdk_sortscore *out_dev,
               hipEvent_t *ev = nullptr, const int *boff = nullptr) {
    if (ev) HIPCHK(hipEventRecord(ev[0], S.stream));
    HIPCHK(launch_voxelize(c->lig, rot_dev, B, S.vs, S.d_vidx, S.d_chg, S.stream, boff));
    if (ev) HIPCHK(hipEventRecord(ev[1], S.stream));
    const int nf = c->nf;
    if (c->t > MDK_MAXT) {  // score maps, then the per-rotation select
        if (run_fft_pipeline(c, S, B, S.d_smap, ev ? ev + 2 : nullptr)) return 1;
        HIPCHK(launch_topt(S.d_smap, B, nf, c->t, S.d_tsel, S.d_tkeys, out_dev, id_dev, boff, S.stream));
        if (ev) HIPCHK(hipEventRecord(ev[7], S.stream));
        return 0;
    }
    if (run_fft_pipeline(c, S, B, nullptr, ev ? ev + 2 : nullptr)) return 1;
    const int lists = c->use_pipe ? pipe_cands_per_rotation(nf) : nf * nf;
    k_reduce_top<<<B, NT, 0, S.stream>>>(S.d_cand, out_dev, id_dev, boff, lists, c->t, nf, (float)((size_t)nf * nf * nf));
    if (ev) HIPCHK(hipEventRecord(ev[7], S.stream));
    HIPCHK(hipGetLastError());
    return 0;
}

// Captures one full batch (max_batch rotations) of mdk_dock_range on slot S as a graph
// whose replays take their first rotation from the slot's cursor (k_batch_begin).  The
// graph bakes in the ligand, the rotation table, the results array and the kernel
// variants, so set-up calls that change them drop it (drop_graphs).
int ensure_graph(mdk_ctx *c, BatchSlot &S) {
    if (S.exec) return 0;
    if (!S.d_ctl) HIPCHK(hipMalloc((void **)&S.d_ctl, 4 * sizeof(int)));
    HIPCHK(hipStreamBeginCapture(S.stream, hipStreamCaptureModeThreadLocal));
    k_batch_begin<<<1, 64, 0, S.stream>>>(S.d_ctl, S.d_ctl + 3);
    const int rc = dock_batch(c, S, c->d_rot, c->d_rotid, c->max_batch, c->d_results, nullptr, S.d_ctl + 3);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(S.stream, &g);
    if (rc || e != hipSuccess || !g) {
        if (g) (void)hipGraphDestroy(g);
        return rc ? rc : fail(std::string("graph capture failed: ") + hipGetErrorString(e));
    }
    S.graph = g;
    HIPCHK(hipGraphInstantiate(&S.exec, g, nullptr, nullptr, 0));
    return 0;
}

void drop_graphs(mdk_ctx *c) {
    for (int k = 0; k < mdk_ctx::kMaxSlots; k++) {
        BatchSlot &S = c->slot[k];
        if (S.exec || S.graph) {
            if (S.stream) (void)hipStreamSynchronize(S.stream);
            if (S.exec) (void)hipGraphExecDestroy(S.exec);
            if (S.graph) (void)hipGraphDestroy(S.graph);
            S.exec = nullptr;
            S.graph = nullptr;
        }
    }
}

// Set-up calls replace device data that queued batches read (ligand atoms, receptor
// spectrum and grids, rotation table, results array), so that a caller may change them
// right after an asynchronous mdk_dock_range: the rotation table and a given spectrum wait
// for every slot's stream first; the receptor grids and the ligand atoms are uploaded on
// the context's stream behind the queued batches instead (mdk_dock_range joins its other
// streams into it at its end).
int wait_slots(mdk_ctx *c) {
    for (int k = 0; k < c->nslots; k++) HIPCHK(hipStreamSynchronize(c->slot[k].stream));
    return 0;
}

template <typename T> int dmalloc(T **p, size_t n) {
    HIPCHK(hipMalloc((void **)p, sizeof(T) * (n ? n : 1)));
    return 0;
}

}  // namespace

// the copy of mdk_copy_probe: n4 16-byte words, grid-stride
__global__ void __launch_bounds__(256) k_copy_probe(const float4 *__restrict__ a, float4 *__restrict__ b, size_t n4) {
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n4; i += (size_t)gridDim.x * 256) b[i] = a[i];
}

// ==================================================================== C ABI
extern "C" {

int mdk_version(void) { return 1; }

const char *mdk_last_error(void) { return g_err.c_str(); }

int mdk_rotation_coef(const float *theta, int R, float *rotcoef) {
#pragma clang fp contract(off)
    if (!theta || !rotcoef || R < 0) return fail("mdk_rotation_coef: invalid argument");
    for (int q = 0; q < R; q++) {
        // src/docking.cpp:376-384: float cos/sin (std::cos(float) = cosf), float products
        const float t0 = theta[3 * q], t1 = theta[3 * q + 1], t2 = theta[3 * q + 2];
        float *r = rotcoef + 9 * (size_t)q;
        r[0] = std::cos(t0) * std::cos(t2) - std::sin(t0) * std::cos(t1) * std::sin(t2);
        r[3] = std::sin(t0) * std::cos(t2) + std::cos(t0) * std::cos(t1) * std::sin(t2);
        r[6] = std::sin(t1) * std::sin(t2);
        r[1] = -std::cos(t0) * std::sin(t2) - std::sin(t0) * std::cos(t1) * std::cos(t2);
        r[4] = -std::sin(t0) * std::sin(t2) + std::cos(t0) * std::cos(t1) * std::cos(t2);
        r[7] = std::sin(t1) * std::cos(t2);
        r[2] = std::sin(t0) * std::sin(t1);
        r[5] = -std::cos(t0) * std::sin(t1);
        r[8] = std::cos(t1);
    }
    return 0;
}

int mdk_device_count(int *count) {
    HIPCHK(hipGetDeviceCount(count));
    return 0;
}

int mdk_create(int device, int nf, int ng, int na, int t, int max_batch, mdk_ctx **out) {
    if (!out) return fail("mdk_create: null output pointer");
    *out = nullptr;
    if (nf < 4 || ng < 2 || ng > nf || na < 1 || t < 1 || t > MDK_MAXT_SELECT || max_batch < 1)
        return fail("mdk_create: invalid sizes (need 4 <= nf, 2 <= ng <= nf, na >= 1, 1 <= t <= 1048576, max_batch >= 1)");
    if (nf != 2 * ng) return fail("mdk_create: nf must be 2*ng (src/docking.cpp:23-25)");
    if (na > 65535) return fail("mdk_create: at most 65535 ligand atoms");
    mdk_ctx *c = new mdk_ctx();
    c->device = device;
    c->nf = nf;
    c->ng = ng;
    c->na = na;
    c->t = t;
    c->max_batch = max_batch;
    c->nd = (nf - ng) / 2;
    c->H = nf / 2 + 1;
    c->Hp = ((c->H + WY - 1) / WY) * WY;
    std::vector<float2> tw;
    if (!make_plan(nf, c->plan, tw)) {
        delete c;
        return fail("mdk_create: nf has a prime factor above 13");
    }
    auto bail = [&](int rc) {
        mdk_destroy(c);
        return rc;
    };
    if (hipSetDevice(device) != hipSuccess) { delete c; return fail("mdk_create: hipSetDevice failed"); }
    {
        // the fixed merge array of mdk_topn_noalloc: its address on this device, once per device
        const hipError_t e = topn_static_setup(device);
        if (e != hipSuccess) {
            delete c;
            return fail(std::string("mdk_create: hipGetSymbolAddress of the merge array failed: ") + hipGetErrorString(e));
        }
    }
    {
        // two batches in flight on two streams (one kernel's tail overlaps the next batch's
        // first kernels; 61.0k against 59.6k rotations/s at 144^3; a third stream gave
        // 61.5k-62.8k against 63.0k in full runs), unless the batches are small or
        // MDK_STREAMS asks for 1 or 3
        const char *ns = std::getenv("MDK_STREAMS");
        c->nslots = ns ? std::max(1, std::min(mdk_ctx::kMaxSlots, std::atoi(ns))) : (max_batch >= 32 ? 2 : 1);
        const char *gr = std::getenv("MDK_GRAPH");
        c->use_graph = gr && gr[0] == '1';
    }
    for (int k = 0; k < c->nslots; k++)
        if (hipStreamCreateWithFlags(&c->slot[k].stream, hipStreamNonBlocking) != hipSuccess ||
            hipEventCreateWithFlags(&c->slot[k].done, hipEventDisableTiming) != hipSuccess)
            return bail(fail("mdk_create: stream creation failed"));
    c->stream = c->slot[0].stream;
    const size_t nf3 = (size_t)nf * nf * nf, ng3 = (size_t)ng * ng * ng, B = max_batch;
    if (dmalloc(&c->d_tw, tw.size()) || dmalloc(&c->d_rec, nf3) || dmalloc(&c->d_xyz, 3 * (size_t)na) ||
        dmalloc(&c->d_rc2, na) || dmalloc(&c->d_rs2, na) || dmalloc(&c->d_charge, na) || dmalloc(&c->d_gc, ng) ||
        dmalloc(&c->d_brot, 9 * B) || dmalloc(&c->d_bid, B) || dmalloc(&c->d_recgrid, ng3) ||
        dmalloc(&c->d_wshape, na) || dmalloc(&c->d_bt, 3 * B) || dmalloc(&c->d_gout, 7 * B) ||
        dmalloc(&c->d_gpart, grad_scratch_doubles((int)B)) || dmalloc(&c->d_rpsc, ng3) || dmalloc(&c->d_btrans, 3 * B) ||
        dmalloc(&c->d_dsum, 9 * B))
        return bail(1);
    for (int k = 0; k < c->nslots; k++) {
        BatchSlot &S = c->slot[k];
        if (dmalloc(&S.d_bits, B * vox_scratch_words(ng)) || dmalloc(&S.d_vidx, B * na) || dmalloc(&S.d_lg, B * ng3) ||
            dmalloc(&S.d_chg, B * ng3) || dmalloc(&S.d_a1, B * ng * ng * nf) || dmalloc(&S.d_a2, B * ng * nf * nf) ||
            dmalloc(&S.d_cc, B * nf * nf * c->Hp) || dmalloc(&S.d_cand, B * nf * nf * std::min(t, MDK_MAXT)) ||
            dmalloc(&S.d_bout, B * t))
            return bail(1);
        if (t > MDK_MAXT && (dmalloc(&S.d_smap, B * nf3) || dmalloc(&S.d_tsel, B) ||
                             dmalloc(&S.d_tkeys, B * topt_sort_size(t, nf))))
            return bail(1);
        S.vs = vox_scratch(S.d_bits, B, ng);
        // the atom counts and the charge grids start, and are kept between batches, all-zero
        if (hipMemsetAsync(S.d_bits, 0, sizeof(uint32_t) * B * vox_scratch_words(ng), S.stream) != hipSuccess ||
            hipMemsetAsync(S.d_chg, 0, sizeof(float) * B * ng3, S.stream) != hipSuccess ||
            hipStreamSynchronize(S.stream) != hipSuccess)
            return bail(fail("mdk_create: scratch initialisation failed"));
    }
    if (hipMemcpy(c->d_tw, tw.data(), sizeof(float2) * tw.size(), hipMemcpyHostToDevice) != hipSuccess)
        return bail(fail("mdk_create: twiddle upload failed"));
    {
        std::vector<float2> tw4(nf);
        for (int m = 0; m < nf; m++) {
            const double a = -2.0 * M_PI * m / nf;
            tw4[m] = make_float2((float)std::cos(a), (float)std::sin(a));
        }
        if (dmalloc(&c->d_tw4, nf) ||
            hipMemcpy(c->d_tw4, tw4.data(), sizeof(float2) * nf, hipMemcpyHostToDevice) != hipSuccess)
            return bail(fail("mdk_create: twiddle upload failed"));
        const char *g = std::getenv("MDK_FORCE_GENERIC");
        c->use_pipe = pipe_available(nf) && !(g && g[0] == '1');
        // default variants per size (pipe_defaults: measured on MI355X), unless the
        // environment asks for others
        int dx = 6, di = 2, df = 0;
        (void)pipe_defaults(nf, &dx, &di, &df);
        const char *xv = std::getenv("MDK_XCORR_VARIANT");
        c->xcorr_variant = xv ? std::atoi(xv) : dx;
        const char *iv = std::getenv("MDK_INV_VARIANT");
        c->inv_variant = iv ? std::atoi(iv) : di;
        const char *fv = std::getenv("MDK_FWD_VARIANT");
        c->fwd_variant = fv ? std::atoi(fv) : df;
        int cus = 0;
        if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) == hipSuccess && cus > 0)
            c->cu_count = cus;
    }
    if (hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess)
        return bail(fail("mdk_create: event creation failed"));
    // the largest kernel LDS request must fit
    const int tl = std::min(t, MDK_MAXT);  // above, the lists are not used (launch_topt)
    const size_t lds_max = std::max(lds_bytes(c, 2 * WX), lds_bytes(c, WZ) + sizeof(Cand) * (size_t)NT * tl);
    if (lds_max > 160 * 1024) return bail(fail("mdk_create: nf too large for the LDS-resident passes"));
    if (allow_lds(k_fwd_z, lds_bytes(c, WZ)) != hipSuccess || allow_lds(k_fwd_y, lds_bytes(c, WY)) != hipSuccess ||
        allow_lds(k_fwd_x_plain, lds_bytes(c, WY)) != hipSuccess ||
        allow_lds(k_x_corr, lds_bytes(c, 2 * WX)) != hipSuccess || allow_lds(k_inv_y, lds_bytes(c, WY)) != hipSuccess ||
        allow_lds(k_inv_z_top, lds_bytes(c, WZ) + sizeof(Cand) * (size_t)NT * tl) != hipSuccess)
        return bail(fail("mdk_create: cannot raise the LDS limit"));
    *out = c;
    return 0;
}

int mdk_destroy(mdk_ctx *c) {
    if (!c) return 0;
    (void)hipSetDevice(c->device);
    for (int k = 0; k < mdk_ctx::kMaxSlots; k++)
        if (c->slot[k].stream) (void)hipStreamSynchronize(c->slot[k].stream);
    void *ptrs[] = {c->d_tw, c->d_tw4, c->d_rec, c->d_xyz, c->d_rc2, c->d_rs2, c->d_charge, c->d_gc, c->d_rot,
                    c->d_rotid, c->d_brot, c->d_bid, c->d_results, c->d_topn_scratch, c->d_topn_out,
                    c->d_recgrid, c->d_wshape, c->d_bt, c->d_gout, c->d_gpart, c->d_rpsc, c->d_btrans, c->d_dsum};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    if (c->h_topn) (void)hipHostFree(c->h_topn);
    if (c->topn_done) (void)hipEventDestroy(c->topn_done);
    if (c->h_rstage) (void)hipHostFree(c->h_rstage);
    if (c->h_lstage) (void)hipHostFree(c->h_lstage);
    if (c->ev_rstage) (void)hipEventDestroy(c->ev_rstage);
    if (c->ev_lstage) (void)hipEventDestroy(c->ev_lstage);
    drop_graphs(c);
    for (int k = 0; k < mdk_ctx::kMaxSlots; k++) {
        BatchSlot &S = c->slot[k];
        void *sp[] = {S.d_bits, S.d_chg, S.d_vidx, S.d_lg, S.d_a1, S.d_a2, S.d_cc, S.d_cand, S.d_bout, S.d_ctl,
                      S.d_smap, S.d_tsel, S.d_tkeys};
        for (void *p : sp)
            if (p) (void)hipFree(p);
        if (S.done) (void)hipEventDestroy(S.done);
        if (S.stream) (void)hipStreamDestroy(S.stream);
    }
    if (c->ev0) (void)hipEventDestroy(c->ev0);
    if (c->ev1) (void)hipEventDestroy(c->ev1);
    for (hipEvent_t e : c->tev) (void)hipEventDestroy(e);
    delete c;
    return 0;
}

// a pinned staging buffer of at least `bytes` whose previous copy has completed
int staging(void **buf, size_t &cap, size_t bytes, hipEvent_t &ev) {
    if (ev) HIPCHK(hipEventSynchronize(ev));  // its last copy is done (long since, as a rule)
    else HIPCHK(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
    if (bytes > cap) {
        if (*buf) HIPCHK(hipHostFree(*buf));
        *buf = nullptr;
        cap = 0;
        HIPCHK(hipHostMalloc(buf, bytes, hipHostMallocDefault));
        cap = bytes;
    }
    return 0;
}

// receptor spectrum from the receptor grids already in d_recgrid (src/fft_process.cpp:212-303);
// `wait`: return after it is done, else with it queued on the context's stream
int receptor_transform(mdk_ctx *c, bool wait = true) {
    const int ng = c->ng, nf = c->nf, nd = c->nd;
    const size_t ng3 = (size_t)ng * ng * ng;
    HIPCHK(hipMemcpyAsync(c->slot[0].d_lg, c->d_recgrid, sizeof(float2) * ng3, hipMemcpyDeviceToDevice, c->stream));
    const int lines_z = ng * ng;
    k_fwd_z<<<(lines_z + WZ - 1) / WZ, NT, lds_bytes(c, WZ), c->stream>>>(c->slot[0].d_lg, c->slot[0].d_a1, c->plan, c->d_tw, ng, nd, lines_z);
    k_fwd_y<<<dim3((nf + WY - 1) / WY, ng, 1), NT, lds_bytes(c, WY), c->stream>>>(c->slot[0].d_a1, c->slot[0].d_a2, c->plan, c->d_tw, ng, nd);
    k_fwd_x_plain<<<dim3((nf + WY - 1) / WY, nf), NT, lds_bytes(c, WY), c->stream>>>(c->slot[0].d_a2, c->d_rec, c->plan, c->d_tw, ng, nd);
    HIPCHK(hipGetLastError());
    if (wait) HIPCHK(hipStreamSynchronize(c->stream));
    c->have_rec = true;
    c->have_recgrid = true;
    return 0;
}

// Stream-ordered: the grids are copied to pinned staging at once (the caller may change them
// on return), then uploaded and transformed on the context's stream behind the batches
// queued before (mdk_dock_range joins its other streams into it at its end), without
// waiting for them.
int mdk_set_receptor(mdk_ctx *c, const float *grid_r, const float *grid_i) {
    if (!c || !grid_r || !grid_i) return fail("mdk_set_receptor: null argument");
    HIPCHK(hipSetDevice(c->device));
    const size_t ng3 = (size_t)c->ng * c->ng * c->ng;
    if (staging((void **)&c->h_rstage, c->h_rstage_n, sizeof(float2) * ng3, c->ev_rstage)) return 1;
    float2 *h = c->h_rstage;
    for (size_t i = 0; i < ng3; i++) h[i] = make_float2(grid_r[i], grid_i[i]);
    HIPCHK(hipMemcpyAsync(c->d_recgrid, h, sizeof(float2) * ng3, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipEventRecord(c->ev_rstage, c->stream));
    return receptor_transform(c, false);
}

int mdk_set_receptor_atoms(mdk_ctx *c, int na, const float *xyz, const float *radius, const float *charge,
                           const float *ace, const float *grid_coord, float gw, float aceratio, float rec_core,
                           float eratio) {
    if (!c || na < 0 || (na > 0 && (!xyz || !radius || !charge || !ace)) || !grid_coord)
        return fail("mdk_set_receptor_atoms: invalid argument");
    if (!(gw > 0.0f) || c->ng < 2) return fail("mdk_set_receptor_atoms: invalid grid");
    HIPCHK(hipSetDevice(c->device));
    if (wait_slots(c)) return 1;
    const int ng = c->ng;
    const size_t n = na > 0 ? (size_t)na : 1;
    float *d_xyz = nullptr, *d_rad = nullptr, *d_chg = nullptr, *d_ace = nullptr, *d_gc = nullptr;
    void *scratch = nullptr;
    int rc = 0;
    if (hipMalloc(&d_xyz, sizeof(float) * 3 * n) != hipSuccess || hipMalloc(&d_rad, sizeof(float) * n) != hipSuccess ||
        hipMalloc(&d_chg, sizeof(float) * n) != hipSuccess || hipMalloc(&d_ace, sizeof(float) * n) != hipSuccess ||
        hipMalloc(&d_gc, sizeof(float) * ng) != hipSuccess ||
        hipMalloc(&scratch, receptor_voxel_scratch_bytes(ng)) != hipSuccess)
        rc = fail("mdk_set_receptor_atoms: out of device memory");
    if (!rc && na > 0 &&
        (hipMemcpy(d_xyz, xyz, sizeof(float) * 3 * na, hipMemcpyHostToDevice) != hipSuccess ||
         hipMemcpy(d_rad, radius, sizeof(float) * na, hipMemcpyHostToDevice) != hipSuccess ||
         hipMemcpy(d_chg, charge, sizeof(float) * na, hipMemcpyHostToDevice) != hipSuccess ||
         hipMemcpy(d_ace, ace, sizeof(float) * na, hipMemcpyHostToDevice) != hipSuccess))
        rc = fail("mdk_set_receptor_atoms: upload failed");
    if (!rc && hipMemcpy(d_gc, grid_coord, sizeof(float) * ng, hipMemcpyHostToDevice) != hipSuccess)
        rc = fail("mdk_set_receptor_atoms: upload failed");
    if (!rc && launch_receptor_voxel(d_xyz, d_rad, d_chg, d_ace, d_gc, na, ng, gw, aceratio, rec_core, eratio,
                                     grid_coord[0], ng > 1 ? grid_coord[1] : grid_coord[0], scratch, c->d_recgrid,
                                     c->stream) != hipSuccess)
        rc = fail("mdk_set_receptor_atoms: kernel launch failed");
    if (!rc && hipStreamSynchronize(c->stream) != hipSuccess) rc = fail("mdk_set_receptor_atoms: voxelization failed");
    (void)hipFree(d_xyz);
    (void)hipFree(d_rad);
    (void)hipFree(d_chg);
    (void)hipFree(d_ace);
    (void)hipFree(d_gc);
    (void)hipFree(scratch);
    if (rc) return rc;
    return receptor_transform(c);
}

int mdk_get_receptor_grids(mdk_ctx *c, float *grid_r, float *grid_i) {
    if (!c || !grid_r || !grid_i) return fail("mdk_get_receptor_grids: null argument");
    if (!c->have_recgrid) return fail("mdk_get_receptor_grids: no receptor grids set");
    HIPCHK(hipSetDevice(c->device));
    const size_t ng3 = (size_t)c->ng * c->ng * c->ng;
    std::vector<float2> h(ng3);
    HIPCHK(hipMemcpyAsync(h.data(), c->d_recgrid, sizeof(float2) * ng3, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    for (size_t i = 0; i < ng3; i++) {
        grid_r[i] = h[i].x;
        grid_i[i] = h[i].y;
    }
    return 0;
}

int mdk_set_receptor_spectrum(mdk_ctx *c, const float *re, const float *im) {
    if (!c || !re || !im) return fail("mdk_set_receptor_spectrum: null argument");
    HIPCHK(hipSetDevice(c->device));
    if (wait_slots(c)) return 1;
    const size_t nf3 = (size_t)c->nf * c->nf * c->nf;
    std::vector<float2> h(nf3);
    for (size_t i = 0; i < nf3; i++) h[i] = make_float2(re[i], im[i]);
    HIPCHK(hipMemcpy(c->d_rec, h.data(), sizeof(float2) * nf3, hipMemcpyHostToDevice));
    c->have_rec = true;
    c->have_recgrid = false;  // no real-space grids behind this spectrum
    return 0;
}

int mdk_get_receptor_spectrum(mdk_ctx *c, float *re, float *im) {
    if (!c || !re || !im) return fail("mdk_get_receptor_spectrum: null argument");
    if (!c->have_rec) return fail("mdk_get_receptor_spectrum: no receptor set");
    HIPCHK(hipSetDevice(c->device));
    HIPCHK(hipStreamSynchronize(c->stream));  // a receptor transform may still be queued there
    const size_t nf3 = (size_t)c->nf * c->nf * c->nf;
    std::vector<float2> h(nf3);
    HIPCHK(hipMemcpy(h.data(), c->d_rec, sizeof(float2) * nf3, hipMemcpyDeviceToHost));
    for (size_t i = 0; i < nf3; i++) {
        re[i] = h[i].x;
        im[i] = h[i].y;
    }
    return 0;
}

int mdk_set_ligand(mdk_ctx *c, const float *xyz, const float *center, const float *radius,
                   const float *charge, const float *grid_coord, float gw, float delta) {
    if (!c) return fail("mdk_set_ligand: null context");
    return mdk_set_ligand_atoms(c, c->na, xyz, center, radius, charge, grid_coord, gw, delta);
}

int mdk_set_ligand_atoms(mdk_ctx *c, int na, const float *xyz, const float *center, const float *radius,
                         const float *charge, const float *grid_coord, float gw, float delta) {
    if (!c || !xyz || !center || !radius || !charge || !grid_coord) return fail("mdk_set_ligand: null argument");
    if (na < 1 || na > c->na) return fail("mdk_set_ligand: atom count outside 1..na_lig of the context");
    if (delta == 0.0f) return fail("mdk_set_ligand: core value 0 is not supported (the reference's last shape sweep is order-dependent then)");
    if (!(gw > 0.0f)) return fail("mdk_set_ligand: grid width must be positive");
    HIPCHK(hipSetDevice(c->device));
    drop_graphs(c);  // they hold the ligand's kernel arguments
    // stream-ordered, as mdk_set_receptor: staged at once, uploaded on the context's stream
    // behind the batches queued before
    const size_t n = (size_t)na, ng = (size_t)c->ng;
    if (staging((void **)&c->h_lstage, c->h_lstage_n, sizeof(float) * (7 * n + ng), c->ev_lstage)) return 1;
    float *h_xyz = c->h_lstage, *h_rc2 = h_xyz + 3 * n, *h_rs2 = h_rc2 + n, *h_chg = h_rs2 + n,
          *h_ws = h_chg + n, *h_gc = h_ws + n;
    std::memcpy(h_xyz, xyz, sizeof(float) * 3 * n);
    std::memcpy(h_chg, charge, sizeof(float) * n);
    std::memcpy(h_gc, grid_coord, sizeof(float) * ng);
    for (int i = 0; i < na; i++) {  // src/fft_process.cpp:784-787 / src/ligand.cpp:254-257
        h_rc2[i] = radius[i] * radius[i] * 1.5f;
        h_rs2[i] = radius[i] * radius[i] * 1.0f;
        // companion shape weight: the atom's volume in voxels
        h_ws[i] = (float)(4.0 / 3.0 * M_PI * (double)radius[i] * radius[i] * radius[i] / ((double)gw * gw * gw));
    }
    const struct {
        float *dst;
        const float *src;
        size_t n;
    } up[] = {{c->d_xyz, h_xyz, 3 * n}, {c->d_rc2, h_rc2, n}, {c->d_rs2, h_rs2, n},
              {c->d_charge, h_chg, n}, {c->d_wshape, h_ws, n}, {c->d_gc, h_gc, ng}};
    for (const auto &u : up)
        HIPCHK(hipMemcpyAsync(u.dst, u.src, sizeof(float) * u.n, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipEventRecord(c->ev_lstage, c->stream));
    LigandDev &L = c->lig;
    L.na = na;
    L.ng = c->ng;
    L.gw = gw;
    L.delta = delta;
    L.pad = (float)c->ng * gw / 2;  // src/ligand.cpp:210
    L.cx = center[0];
    L.cy = center[1];
    L.cz = center[2];
    L.xyz = c->d_xyz;
    L.rc2 = c->d_rc2;
    L.rs2 = c->d_rs2;
    L.charge = c->d_charge;
    L.gc = c->d_gc;
    c->have_lig = true;
    return 0;
}

int mdk_dock(mdk_ctx *c, const float *rotcoef, const int *rot_id, int B, mdk_sortscore *out) {
    if (!c || !rotcoef || !rot_id || !out) return fail("mdk_dock: null argument");
    if (!c->have_rec || !c->have_lig) return fail("mdk_dock: receptor and ligand must be set first");
    if (B < 1 || B > c->max_batch) return fail("mdk_dock: batch size out of range");
    HIPCHK(hipSetDevice(c->device));
    HIPCHK(hipMemcpyAsync(c->d_brot, rotcoef, sizeof(float) * 9 * B, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(c->d_bid, rot_id, sizeof(int) * B, hipMemcpyHostToDevice, c->stream));
    if (dock_batch(c, c->slot[0], c->d_brot, c->d_bid, B, c->slot[0].d_bout)) return 1;
    HIPCHK(hipMemcpyAsync(out, c->slot[0].d_bout, sizeof(mdk_sortscore) * B * c->t, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return 0;
}

int mdk_dock_grid(mdk_ctx *c, const float *grid_r, const float *grid_i, int rot_id, mdk_sortscore *out) {
    if (!c || !grid_r || !grid_i || !out) return fail("mdk_dock_grid: null argument");
    if (!c->have_rec) return fail("mdk_dock_grid: receptor must be set first");
    HIPCHK(hipSetDevice(c->device));
    BatchSlot &S = c->slot[0];
    const size_t ng3 = (size_t)c->ng * c->ng * c->ng;
    std::vector<float2> h(ng3);
    for (size_t i = 0; i < ng3; i++) h[i] = make_float2(grid_r[i], grid_i[i]);
    HIPCHK(hipMemcpyAsync(S.d_lg, h.data(), sizeof(float2) * ng3, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(c->d_bid, &rot_id, sizeof(int), hipMemcpyHostToDevice, c->stream));
    const int nf = c->nf;
    if (c->t > MDK_MAXT) {  // the score map, then the per-rotation select
        if (run_fft_pipeline(c, S, 1, S.d_smap, nullptr, false, true)) return 1;
        HIPCHK(launch_topt(S.d_smap, 1, nf, c->t, S.d_tsel, S.d_tkeys, S.d_bout, c->d_bid, nullptr, c->stream));
    } else {
        if (run_fft_pipeline(c, S, 1, nullptr, nullptr, false, true)) return 1;
        k_reduce_top<<<1, NT, 0, c->stream>>>(S.d_cand, S.d_bout, c->d_bid, nullptr, nf * nf, c->t, nf,
                                             (float)((size_t)nf * nf * nf));
        HIPCHK(hipGetLastError());
    }
    HIPCHK(hipMemcpyAsync(out, S.d_bout, sizeof(mdk_sortscore) * c->t, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));  // h is read by the upload until here
    return 0;
}

int mdk_set_rotations(mdk_ctx *c, const float *rotcoef, int R) {
    if (!c || !rotcoef || R < 1) return fail("mdk_set_rotations: invalid argument");
    HIPCHK(hipSetDevice(c->device));
    if (wait_slots(c)) return 1;
    drop_graphs(c);  // they hold the old table and results pointers
    if (c->d_rot) (void)hipFree(c->d_rot);
    if (c->d_rotid) (void)hipFree(c->d_rotid);
    if (c->d_results) (void)hipFree(c->d_results);
    c->d_rot = nullptr;
    c->d_rotid = nullptr;
    c->d_results = nullptr;
    if (dmalloc(&c->d_rot, 9 * (size_t)R) || dmalloc(&c->d_rotid, R) || dmalloc(&c->d_results, (size_t)R * c->t)) return 1;
    HIPCHK(hipMemcpy(c->d_rot, rotcoef, sizeof(float) * 9 * R, hipMemcpyHostToDevice));
    k_iota<<<(R + 255) / 256, 256, 0, c->stream>>>(c->d_rotid, 0, R);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(c->stream));
    c->nrot = R;
    return 0;
}

int mdk_dock_range(mdk_ctx *c, int first, int count) {
    if (!c) return fail("mdk_dock_range: null context");
    if (!c->have_rec || !c->have_lig || !c->d_rot) return fail("mdk_dock_range: receptor, ligand and rotations must be set first");
    if (first < 0 || count < 0 || first + count > c->nrot) return fail("mdk_dock_range: range outside the rotation table");
    HIPCHK(hipSetDevice(c->device));
    // the other streams start after everything queued on the main stream so far
    HIPCHK(hipEventRecord(c->slot[0].done, c->stream));
    for (int k = 1; k < c->nslots; k++) HIPCHK(hipStreamWaitEvent(c->slot[k].stream, c->slot[0].done, 0));
    int q = 0, b0 = first;
    const int nfull = count / c->max_batch;
    if (c->use_graph && nfull > 0) {  // full batches replay the captured graph of each slot
        const int used = std::min(c->nslots, nfull);
        for (int k = 0; k < used; k++) {
            BatchSlot &S = c->slot[k];
            if (ensure_graph(c, S)) return 1;
            k_batch_reset<<<1, 64, 0, S.stream>>>(S.d_ctl, first + k * c->max_batch, c->nslots * c->max_batch);
            HIPCHK(hipGetLastError());
        }
        for (; q < nfull; q++) HIPCHK(hipGraphLaunch(c->slot[q % c->nslots].exec, c->slot[q % c->nslots].stream));
        b0 = first + nfull * c->max_batch;
    }
    for (; b0 < first + count; b0 += c->max_batch, q++) {
        const int B = std::min(c->max_batch, first + count - b0);
        if (dock_batch(c, c->slot[q % c->nslots], c->d_rot + 9 * (size_t)b0, c->d_rotid + b0, B,
                       c->d_results + (size_t)b0 * c->t))
            return 1;
    }
    // and the main stream continues after all of them (results, merge, timing events)
    for (int k = 1; k < c->nslots; k++) {
        HIPCHK(hipEventRecord(c->slot[k].done, c->slot[k].stream));
        HIPCHK(hipStreamWaitEvent(c->stream, c->slot[k].done, 0));
    }
    return 0;
}

int mdk_sync(mdk_ctx *c) {
    if (!c) return fail("mdk_sync: null context");
    for (int k = 0; k < c->nslots; k++) HIPCHK(hipStreamSynchronize(c->slot[k].stream));
    return 0;
}

int mdk_get_results(mdk_ctx *c, int first, int count, mdk_sortscore *out) {
    if (!c || !out || !c->d_results) return fail("mdk_get_results: nothing to read");
    if (first < 0 || count < 0 || first + count > c->nrot) return fail("mdk_get_results: range outside the rotation table");
    HIPCHK(hipMemcpyAsync(out, c->d_results + (size_t)first * c->t, sizeof(mdk_sortscore) * count * c->t,
                          hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return 0;
}

// the merge on the context's stream, after every batch queued so far (mdk_dock_range
// joins its slots' streams into it), its n best copied to the pinned host buffer
int topn_enqueue(mdk_ctx *c, int nrot, int n, const char *who) {
    if (!c || !c->d_results) return fail(std::string(who) + ": nothing to merge");
    if (nrot < 1 || nrot > c->nrot || n < 1) return fail(std::string(who) + ": invalid count");
    if (c->topn_pending) return fail(std::string(who) + ": a started merge was not finished (mdk_topn_finish)");
    HIPCHK(hipSetDevice(c->device));
    // on the device: order (score desc, rotation asc, rank asc), mdk_merge.hip
    const size_t m = (size_t)nrot * c->t;
    const int keep = (int)std::min((size_t)n, m);
    if (keep > c->topn_cap) {
        if (c->d_topn_scratch) HIPCHK(hipFree(c->d_topn_scratch));
        if (c->d_topn_out) HIPCHK(hipFree(c->d_topn_out));
        c->d_topn_scratch = nullptr;
        c->d_topn_out = nullptr;
        c->topn_cap = 0;
        HIPCHK(hipMalloc(&c->d_topn_scratch, topn_scratch_bytes(keep)));
        HIPCHK(hipMalloc((void **)&c->d_topn_out, sizeof(mdk_sortscore) * keep));
        c->topn_cap = keep;
    }
    if (keep > c->h_topn_cap) {  // no merge of this context is in flight (topn_pending is 0)
        if (c->h_topn) HIPCHK(hipHostFree(c->h_topn));
        c->h_topn = nullptr;
        c->h_topn_cap = 0;
        HIPCHK(hipHostMalloc((void **)&c->h_topn, sizeof(mdk_sortscore) * keep, hipHostMallocDefault));
        c->h_topn_cap = keep;
    }
    if (!c->topn_done) HIPCHK(hipEventCreateWithFlags(&c->topn_done, hipEventDisableTiming));
    HIPCHK(launch_topn(c->d_results, (unsigned)m, (unsigned)keep, c->d_topn_scratch, c->d_topn_out, c->stream));
    HIPCHK(hipMemcpyAsync(c->h_topn, c->d_topn_out, sizeof(mdk_sortscore) * keep, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipEventRecord(c->topn_done, c->stream));
    c->topn_pending = keep;
    return 0;
}

int mdk_topn_start(mdk_ctx *c, int nrot, int n) { return topn_enqueue(c, nrot, n, "mdk_topn_start"); }

int mdk_topn_finish(mdk_ctx *c, mdk_sortscore *out) {
    if (!c || !out) return fail("mdk_topn_finish: null argument");
    if (!c->topn_pending) return fail("mdk_topn_finish: no merge was started (mdk_topn_start)");
    const int keep = c->topn_pending;
    c->topn_pending = 0;
    HIPCHK(hipSetDevice(c->device));
    HIPCHK(hipEventSynchronize(c->topn_done));
    std::memcpy(out, c->h_topn, sizeof(mdk_sortscore) * keep);
    return 0;
}

int mdk_topn(mdk_ctx *c, int nrot, int n, mdk_sortscore *out) {
    if (!out) return fail("mdk_topn: null output");
    if (topn_enqueue(c, nrot, n, "mdk_topn")) return 1;
    return mdk_topn_finish(c, out);
}

int mdk_topn_noalloc(mdk_ctx *c, int nrot, int n, mdk_sortscore *out) {
    if (!c || !out || !c->d_results) return fail("mdk_topn_noalloc: nothing to merge");
    if (nrot < 1 || nrot > c->nrot || n < 1) return fail("mdk_topn_noalloc: invalid count");
    const size_t m = (size_t)nrot * c->t;
    const int keep = (int)std::min((size_t)n, m);
    if (keep > MDK_TOPN_STATIC_MAX)
        return fail("mdk_topn_noalloc: " + std::to_string(keep) + " results exceed the fixed merge array's maximum of " +
                    std::to_string(MDK_TOPN_STATIC_MAX) + " (use mdk_topn)");
    if (m > 0xffffffffull) return fail("mdk_topn_noalloc: too many results to merge");
    HIPCHK(hipSetDevice(c->device));
    char *base = static_cast<char *>(topn_static_ptr(c->device));  // looked up by mdk_create
    if (!base) return fail("mdk_topn_noalloc: the merge array of this device was not looked up");
    mdk_sortscore *dout = reinterpret_cast<mdk_sortscore *>(base + topn_static_out_offset(keep));
    // the same kernels as mdk_topn; they initialise every word of the array they read
    HIPCHK(launch_topn(c->d_results, (unsigned)m, (unsigned)keep, base, dout, c->stream));
    HIPCHK(hipMemcpyAsync(out, dout, sizeof(mdk_sortscore) * keep, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return 0;
}

int mdk_time_range(mdk_ctx *c, int first, int count, int reps, float *ms) {
    if (!c || !ms || reps < 1) return fail("mdk_time_range: invalid argument");
    HIPCHK(hipSetDevice(c->device));
    HIPCHK(hipEventRecord(c->ev0, c->stream));
    for (int r = 0; r < reps; r++)
        if (mdk_dock_range(c, first, count)) return 1;
    HIPCHK(hipEventRecord(c->ev1, c->stream));
    HIPCHK(hipEventSynchronize(c->ev1));
    float t = 0.f;
    HIPCHK(hipEventElapsedTime(&t, c->ev0, c->ev1));
    *ms = t / reps;
    return 0;
}

int mdk_dock_range_timed(mdk_ctx *c, int first, int count, float *stage_ms, int *nstages, int *nfull) {
    if (!c || !stage_ms || !nstages || !nfull) return fail("mdk_dock_range_timed: null argument");
    if (!c->have_rec || !c->have_lig || !c->d_rot) return fail("mdk_dock_range_timed: receptor, ligand and rotations must be set first");
    if (first < 0 || count < 0 || first + count > c->nrot) return fail("mdk_dock_range_timed: range outside the rotation table");
    HIPCHK(hipSetDevice(c->device));
    const int nb = (count + c->max_batch - 1) / c->max_batch;
    const size_t need = (size_t)nb * (MDK_NSTAGE + 1);
    while (c->tev.size() < need) {  // the pool grows once to the largest range timed
        hipEvent_t e = nullptr;
        HIPCHK(hipEventCreate(&e));
        c->tev.push_back(e);
    }
    // as mdk_dock_range, always with eager launches, events around every stage of every batch
    HIPCHK(hipEventRecord(c->slot[0].done, c->stream));
    for (int k = 1; k < c->nslots; k++) HIPCHK(hipStreamWaitEvent(c->slot[k].stream, c->slot[0].done, 0));
    int q = 0;
    for (int b0 = first; b0 < first + count; b0 += c->max_batch, q++) {
        const int B = std::min(c->max_batch, first + count - b0);
        if (dock_batch(c, c->slot[q % c->nslots], c->d_rot + 9 * (size_t)b0, c->d_rotid + b0, B,
                       c->d_results + (size_t)b0 * c->t, &c->tev[(size_t)q * (MDK_NSTAGE + 1)]))
            return 1;
    }
    for (int k = 1; k < c->nslots; k++) {
        HIPCHK(hipEventRecord(c->slot[k].done, c->slot[k].stream));
        HIPCHK(hipStreamWaitEvent(c->stream, c->slot[k].done, 0));
    }
    for (int k = 0; k < c->nslots; k++) HIPCHK(hipStreamSynchronize(c->slot[k].stream));
    for (int k = 0; k < MDK_NSTAGE; k++) stage_ms[k] = 0.0f;
    int full = 0;
    for (int b = 0; b < q; b++) {
        if (b * c->max_batch + c->max_batch > count) break;  // the last, partial batch is left out
        const hipEvent_t *ev = &c->tev[(size_t)b * (MDK_NSTAGE + 1)];
        for (int k = 0; k < MDK_NSTAGE; k++) {
            float ms = 0.0f;
            HIPCHK(hipEventElapsedTime(&ms, ev[k], ev[k + 1]));
            stage_ms[k] += ms;
        }
        full++;
    }
    *nstages = MDK_NSTAGE;
    *nfull = full;
    return 0;
}

int mdk_profile_stages(mdk_ctx *c, int first, int count, float *ms, int *nstages) {
    if (!c || !ms || !nstages) return fail("mdk_profile_stages: null argument");
    if (!c->have_rec || !c->have_lig || !c->d_rot) return fail("mdk_profile_stages: receptor, ligand and rotations must be set first");
    if (first < 0 || count < 1 || count > c->max_batch || first + count > c->nrot)
        return fail("mdk_profile_stages: count must be 1..max_batch inside the rotation table");
    HIPCHK(hipSetDevice(c->device));
    hipEvent_t ev[MDK_NSTAGE + 1];
    int made = 0, rc = 0;
    for (; made < MDK_NSTAGE + 1; made++)
        if (hipEventCreate(&ev[made]) != hipSuccess) { rc = fail("mdk_profile_stages: event creation failed"); break; }
    if (!rc) rc = dock_batch(c, c->slot[0], c->d_rot + 9 * (size_t)first, c->d_rotid + first, count, c->d_results + (size_t)first * c->t, ev);
    if (!rc && hipEventSynchronize(ev[MDK_NSTAGE]) != hipSuccess) rc = fail("mdk_profile_stages: sync failed");
    for (int q = 0; !rc && q < MDK_NSTAGE; q++)
        if (hipEventElapsedTime(&ms[q], ev[q], ev[q + 1]) != hipSuccess) rc = fail("mdk_profile_stages: timing failed");
    for (int q = 0; q < made; q++) (void)hipEventDestroy(ev[q]);
    *nstages = MDK_NSTAGE;
    return rc;
}

int mdk_voxelize(mdk_ctx *c, const float *rotcoef, int B, float *grid_r, float *grid_i) {
    if (!c || !rotcoef || !grid_r || !grid_i) return fail("mdk_voxelize: null argument");
    if (!c->have_lig) return fail("mdk_voxelize: ligand not set");
    if (B < 1 || B > c->max_batch) return fail("mdk_voxelize: batch size out of range");
    HIPCHK(hipSetDevice(c->device));
    HIPCHK(hipMemcpyAsync(c->d_brot, rotcoef, sizeof(float) * 9 * B, hipMemcpyHostToDevice, c->stream));
    HIPCHK(launch_voxelize(c->lig, c->d_brot, B, c->slot[0].vs, c->slot[0].d_vidx, c->slot[0].d_chg, c->stream));
    HIPCHK(launch_vox_expand(vox_grid(c, c->slot[0]), B, c->slot[0].d_lg, c->stream));
    const size_t ng3 = (size_t)c->ng * c->ng * c->ng;
    std::vector<float2> h(B * ng3);
    HIPCHK(hipMemcpyAsync(h.data(), c->slot[0].d_lg, sizeof(float2) * h.size(), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    for (size_t i = 0; i < h.size(); i++) {
        grid_r[i] = h[i].x;
        grid_i[i] = h[i].y;
    }
    return 0;
}

int mdk_score_map(mdk_ctx *c, const float *rotcoef, float *out) {
    if (!c || !rotcoef || !out) return fail("mdk_score_map: null argument");
    if (!c->have_rec || !c->have_lig) return fail("mdk_score_map: receptor and ligand must be set first");
    HIPCHK(hipSetDevice(c->device));
    const size_t nf3 = (size_t)c->nf * c->nf * c->nf;
    float *d_map = nullptr;
    HIPCHK(hipMalloc(&d_map, sizeof(float) * nf3));
    int rc = 0;
    if (hipMemcpyAsync(c->d_brot, rotcoef, sizeof(float) * 9, hipMemcpyHostToDevice, c->stream) != hipSuccess ||
        launch_voxelize(c->lig, c->d_brot, 1, c->slot[0].vs, c->slot[0].d_vidx, c->slot[0].d_chg, c->stream) != hipSuccess)
        rc = fail("mdk_score_map: voxelization failed");
    if (!rc) rc = run_fft_pipeline(c, c->slot[0], 1, d_map);
    if (!rc && (hipMemcpyAsync(out, d_map, sizeof(float) * nf3, hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
                hipStreamSynchronize(c->stream) != hipSuccess))
        rc = fail("mdk_score_map: copy back failed");
    (void)hipFree(d_map);
    return rc;
}

int mdk_autotune(mdk_ctx *c, int first, int count, int *choice) {
    if (!c) return fail("mdk_autotune: null context");
    if (!c->have_rec || !c->have_lig || !c->d_rot) return fail("mdk_autotune: receptor, ligand and rotations must be set first");
    if (first < 0 || count < 1 || first + count > c->nrot) return fail("mdk_autotune: range outside the rotation table");
    HIPCHK(hipSetDevice(c->device));
    if (c->use_pipe) {
        // the variants built for this size: correlation x inverse first, then the forward
        // pass; each timed twice after one warm-up pass.  Every variant computes the same
        // results (same operations in the same order), except at nf = 256, where the two-pass
        // forms (inverse variant 0, forward variant 1), the half-plane fused passes and the
        // 512-thread half-plane inverse (variant 1) round differently
        auto time_it = [&](float &ms) {
            drop_graphs(c);
            return mdk_dock_range(c, first, count) || mdk_time_range(c, first, count, 2, &ms);
        };
        float best = 0.f;
        int bx = c->xcorr_variant, bi = c->inv_variant, bf = c->fwd_variant;
        for (int xv = 3; xv <= 7; xv++)  // 0..2 are far slower on every size measured
            for (int iv = 0; iv <= 3; iv++) {
                if (!pipe_has_variant(c->nf, 0, xv) || !pipe_has_variant(c->nf, 1, iv)) continue;
                c->xcorr_variant = xv;
                c->inv_variant = iv;
                float ms = 0.f;
                if (time_it(ms)) return 1;
                if (best == 0.f || ms < best) {
                    best = ms;
                    bx = xv;
                    bi = iv;
                }
            }
        c->xcorr_variant = bx;
        c->inv_variant = bi;
        for (int fv : {0, 1, 2}) {
            if (!pipe_has_variant(c->nf, 2, fv) || fv == bf) continue;
            const int keep = c->fwd_variant;
            c->fwd_variant = fv;
            float ms = 0.f;
            if (time_it(ms)) return 1;
            if (ms < best) {
                best = ms;
                bf = fv;
            }
            c->fwd_variant = keep;
        }
        c->fwd_variant = bf;
        drop_graphs(c);
    }
    if (choice) {
        choice[0] = c->xcorr_variant;
        choice[1] = c->inv_variant;
        choice[2] = c->fwd_variant;
    }
    return 0;
}

int mdk_score_spectral(mdk_ctx *c, const float *rotcoef, const float *t, int B, float *out) {
    if (!c || !rotcoef || !t || !out) return fail("mdk_score_spectral: null argument");
    if (!c->have_rec || !c->have_lig) return fail("mdk_score_spectral: receptor and ligand must be set first");
    if (B < 0) return fail("mdk_score_spectral: negative batch size");
    if (!c->use_pipe)
        return fail("mdk_score_spectral: needs an FFT size with specialised kernels (and MDK_FORCE_GENERIC unset)");
    if (c->H > 192) return fail("mdk_score_spectral: nf above 382 (three kz per lane in the evaluation)");
    HIPCHK(hipSetDevice(c->device));
    BatchSlot &S = c->slot[0];
    for (int q0 = 0; q0 < B; q0 += c->max_batch) {  // in batches of the context's size, in order
        const int nb = std::min(c->max_batch, B - q0);
        HIPCHK(hipMemcpyAsync(c->d_brot, rotcoef + 9 * (size_t)q0, sizeof(float) * 9 * nb, hipMemcpyHostToDevice,
                              c->stream));
        HIPCHK(hipMemcpyAsync(c->d_bt, t + 3 * (size_t)q0, sizeof(float) * 3 * nb, hipMemcpyHostToDevice, c->stream));
        HIPCHK(launch_voxelize(c->lig, c->d_brot, nb, S.vs, S.d_vidx, S.d_chg, c->stream));
        if (run_fft_pipeline(c, S, nb, nullptr, nullptr, true)) return 1;
        HIPCHK(launch_spec_eval(S.d_cc, c->d_bt, nb, c->nf, c->H, c->Hp, c->d_gpart, c->d_gout, c->stream));
        HIPCHK(hipMemcpyAsync(out + 4 * (size_t)q0, c->d_gout, sizeof(float) * 4 * nb, hipMemcpyDeviceToHost,
                              c->stream));
    }
    HIPCHK(hipStreamSynchronize(c->stream));
    return 0;
}

int mdk_score_companion(mdk_ctx *c, const float *rotcoef, const float *t, int B, float *out) {
    if (!c || !rotcoef || !t || !out) return fail("mdk_score_companion: null argument");
    if (!c->have_recgrid || !c->have_lig)
        return fail("mdk_score_companion: receptor grids (mdk_set_receptor) and ligand must be set first");
    if (B < 0) return fail("mdk_score_companion: negative batch size");
    HIPCHK(hipSetDevice(c->device));
    for (int q0 = 0; q0 < B; q0 += c->max_batch) {  // in batches of the context's size, in order
        const int nb = std::min(c->max_batch, B - q0);
        HIPCHK(hipMemcpyAsync(c->d_brot, rotcoef + 9 * (size_t)q0, sizeof(float) * 9 * nb, hipMemcpyHostToDevice,
                              c->stream));
        HIPCHK(hipMemcpyAsync(c->d_bt, t + 3 * (size_t)q0, sizeof(float) * 3 * nb, hipMemcpyHostToDevice, c->stream));
        HIPCHK(launch_companion(c->lig, c->d_wshape, c->d_recgrid, c->d_brot, c->d_bt, nb, c->nf, c->d_gpart,
                                c->d_gout, c->stream));
        HIPCHK(hipMemcpyAsync(out + 7 * (size_t)q0, c->d_gout, sizeof(float) * 7 * nb, hipMemcpyDeviceToHost,
                              c->stream));
    }
    HIPCHK(hipStreamSynchronize(c->stream));
    return 0;
}

int mdk_score_detail(mdk_ctx *c, const float *rec_rpsc, const float *rotcoef, const int *trans, int B, float *out) {
    if (!c || !rec_rpsc || !rotcoef || !trans || !out) return fail("mdk_score_detail: null argument");
    if (!c->have_recgrid || !c->have_lig)
        return fail("mdk_score_detail: receptor grids (mdk_set_receptor) and ligand must be set first");
    if (B < 0) return fail("mdk_score_detail: negative pose count");
    for (int q = 0; q < 3 * B; q++)
        if (trans[q] < -c->ng || trans[q] >= 3 * c->ng) return fail("mdk_score_detail: translation out of range");
    HIPCHK(hipSetDevice(c->device));
    const int ng = c->ng;
    const size_t ng3 = (size_t)ng * ng * ng;
    BatchSlot &S = c->slot[0];
    HIPCHK(hipMemcpyAsync(c->d_rpsc, rec_rpsc, sizeof(float) * ng3, hipMemcpyHostToDevice, c->stream));
    for (int q0 = 0; q0 < B; q0 += c->max_batch) {  // batches of the context's size
        const int nb = B - q0 < c->max_batch ? B - q0 : c->max_batch;
        HIPCHK(hipMemcpyAsync(c->d_brot, rotcoef + 9 * (size_t)q0, sizeof(float) * 9 * nb, hipMemcpyHostToDevice,
                              c->stream));
        HIPCHK(hipMemcpyAsync(c->d_btrans, trans + 3 * (size_t)q0, sizeof(int) * 3 * nb, hipMemcpyHostToDevice,
                              c->stream));
        HIPCHK(launch_voxelize(c->lig, c->d_brot, nb, S.vs, S.d_vidx, S.d_chg, c->stream));
        HIPCHK(launch_vox_expand(vox_grid(c, S), nb, S.d_lg, c->stream));
        HIPCHK(launch_detail(S.d_lg, c->d_recgrid, c->d_rpsc, c->d_btrans, nb, ng, c->d_dsum, c->stream));
        HIPCHK(hipMemcpyAsync(out + 9 * (size_t)q0, c->d_dsum, sizeof(float) * 9 * nb, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));  // the staging buffers are reused by the next batch
    }
    return 0;
}

int mdk_decoy_coords(int device, const float *xyz, int na, const float *rot_pre, const float *rotcoef, const int *trans,
                     int nd, float spacing, const float *offset, float *out) {
    if (!xyz || !rot_pre || !rotcoef || !trans || !offset || !out || na < 1 || nd < 0)
        return fail("mdk_decoy_coords: invalid argument");
    if (nd == 0) return 0;
    HIPCHK(hipSetDevice(device));
    float *d_xyz = nullptr, *d_pre = nullptr, *d_rot = nullptr, *d_out = nullptr;
    int *d_tr = nullptr;
    int rc = 0;
    const size_t no = 3 * (size_t)na * nd;
    if (hipMalloc(&d_xyz, sizeof(float) * 3 * na) != hipSuccess || hipMalloc(&d_pre, sizeof(float) * 9) != hipSuccess ||
        hipMalloc(&d_rot, sizeof(float) * 9 * nd) != hipSuccess || hipMalloc(&d_tr, sizeof(int) * 3 * nd) != hipSuccess ||
        hipMalloc(&d_out, sizeof(float) * no) != hipSuccess)
        rc = fail("mdk_decoy_coords: out of device memory");
    if (!rc && (hipMemcpy(d_xyz, xyz, sizeof(float) * 3 * na, hipMemcpyHostToDevice) != hipSuccess ||
                hipMemcpy(d_pre, rot_pre, sizeof(float) * 9, hipMemcpyHostToDevice) != hipSuccess ||
                hipMemcpy(d_rot, rotcoef, sizeof(float) * 9 * nd, hipMemcpyHostToDevice) != hipSuccess ||
                hipMemcpy(d_tr, trans, sizeof(int) * 3 * nd, hipMemcpyHostToDevice) != hipSuccess))
        rc = fail("mdk_decoy_coords: upload failed");
    if (!rc && launch_decoy_coords(d_xyz, na, d_pre, d_rot, d_tr, nd, spacing, offset, d_out, nullptr) != hipSuccess)
        rc = fail("mdk_decoy_coords: kernel launch failed");
    if (!rc && hipMemcpy(out, d_out, sizeof(float) * no, hipMemcpyDeviceToHost) != hipSuccess)
        rc = fail("mdk_decoy_coords: download failed");
    (void)hipFree(d_xyz);
    (void)hipFree(d_pre);
    (void)hipFree(d_rot);
    (void)hipFree(d_tr);
    (void)hipFree(d_out);
    return rc;
}

int mdk_copy_probe(int device, size_t bytes, int reps, float *ms) {
    if (!ms || bytes == 0 || reps < 1) return fail("mdk_copy_probe: invalid argument");
    if (bytes % 16) return fail("mdk_copy_probe: bytes must be a multiple of 16");
    HIPCHK(hipSetDevice(device));
    void *a = nullptr, *b = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = 0;
    if (hipMalloc(&a, bytes) != hipSuccess || hipMalloc(&b, bytes) != hipSuccess ||
        hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess)
        rc = fail("mdk_copy_probe: allocation failed");
    if (!rc && hipMemset(a, 1, bytes) != hipSuccess) rc = fail("mdk_copy_probe: memset failed");
    // one dispatch per copy, so that a profiler's counters of one dispatch see all its bytes
    const size_t n4 = bytes / 16;
    const unsigned blocks = (unsigned)std::min<size_t>((n4 + 255) / 256, 16384);
    if (!rc) {
        k_copy_probe<<<blocks, 256>>>(static_cast<const float4 *>(a), static_cast<float4 *>(b), n4);
        if (hipGetLastError() != hipSuccess || hipDeviceSynchronize() != hipSuccess) rc = fail("mdk_copy_probe: copy failed");
    }
    if (!rc) {
        (void)hipEventRecord(e0, nullptr);
        for (int r = 0; r < reps; r++)
            k_copy_probe<<<blocks, 256>>>(static_cast<const float4 *>(a), static_cast<float4 *>(b), n4);
        (void)hipEventRecord(e1, nullptr);
        if (hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(ms, e0, e1) != hipSuccess)
            rc = fail("mdk_copy_probe: timing failed");
        else
            *ms /= reps;
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (a) (void)hipFree(a);
    if (b) (void)hipFree(b);
    return rc;
}

int mdk_fft1d(int device, int n, int count, const float *in, float *out, int sign) {
    if (!in || !out || n < 2 || count < 1) return fail("mdk_fft1d: invalid argument");
    FftPlan p{};
    std::vector<float2> tw;
    if (!make_plan(n, p, tw)) return fail("mdk_fft1d: length has a prime factor above 13");
    const size_t lds = sizeof(float2) * ((size_t)p.ntw + 2 * (size_t)n * (WZ + 1));
    if (lds > 160 * 1024) return fail("mdk_fft1d: length too large");
    HIPCHK(hipSetDevice(device));
    float2 *d_in = nullptr, *d_out = nullptr, *d_tw = nullptr;
    const size_t bytes = sizeof(float2) * (size_t)n * count;
    int rc = 0;
    if (hipMalloc(&d_in, bytes) != hipSuccess || hipMalloc(&d_out, bytes) != hipSuccess ||
        hipMalloc(&d_tw, sizeof(float2) * tw.size()) != hipSuccess)
        rc = fail("mdk_fft1d: allocation failed");
    if (!rc && (hipMemcpy(d_in, in, bytes, hipMemcpyHostToDevice) != hipSuccess ||
                hipMemcpy(d_tw, tw.data(), sizeof(float2) * tw.size(), hipMemcpyHostToDevice) != hipSuccess))
        rc = fail("mdk_fft1d: upload failed");
    if (!rc && allow_lds(k_lines, lds) != hipSuccess) rc = fail("mdk_fft1d: cannot raise the LDS limit");
    if (!rc) {
        k_lines<<<(count + WZ - 1) / WZ, NT, lds>>>(d_in, d_out, p, d_tw, count, sign > 0 ? 1 : 0);
        if (hipGetLastError() != hipSuccess || hipDeviceSynchronize() != hipSuccess)
            rc = fail("mdk_fft1d: kernel failed");
    }
    if (!rc && hipMemcpy(out, d_out, bytes, hipMemcpyDeviceToHost) != hipSuccess) rc = fail("mdk_fft1d: download failed");
    if (d_in) (void)hipFree(d_in);
    if (d_out) (void)hipFree(d_out);
    if (d_tw) (void)hipFree(d_tw);
    return rc;
}

}  // extern "C"
