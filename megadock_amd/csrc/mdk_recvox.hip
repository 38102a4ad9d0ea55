// Receptor grids on the device, bit-exact with the reference's producer (Receptor::rpscace
// and Receptor::electro, src/receptor.cpp:20-346; the host restatement is
// host/receptor.cpp).  Compiled with -ffp-contract=off; division and square root are
// the correctly rounded ones.
//
//   core fill        one workgroup per atom over its box: voxels with d^2 < 1.5 r^2 get
//                    the core value (the same value from every atom, so order-free);
//   open space       one thread per ray, six rays per line of the grid, each until the
//                    first core voxel (rays write only "open", which never stops a ray);
//   two sweeps       the reference's in-place propagation in increasing linear order
//                    (src/receptor.cpp:159-189).  A voxel is open when the sweep reaches it
//                    if it was open, or was zero with an in-range predecessor (-z, -y, -x)
//                    that was open when reached: a recurrence over the planes i + j + k = s,
//                    run by one workgroup plane after plane; every zero neighbour of an
//                    in-range voxel open when reached ends the sweep open;
//   cavities         zero becomes core, open becomes zero;
//   surface, charge  one thread per voxel, summing over the atoms in index order (the
//                    atoms whose boxes meet the workgroup's tile, compacted in LDS chunk by
//                    chunk), exactly the reference's per-voxel sequence of roundings.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "mdk_internal.h"

namespace {

constexpr float kOpenSpace = -7777.0f;

struct RecAtoms {
    const float *xyz, *radius, *charge, *ace, *gc;
    int na, ng;
    float gw;
};

// the voxel holding coordinate p along one axis (src/receptor.cpp:73-83)
__device__ __forceinline__ int cell_of(float p, float gw, int ng) { return (int)(p / gw + ng / 2); }

__device__ __forceinline__ float sq_dist(float p, float g) {
    const float d = p - g;
    return d * d;
}

__global__ void k_rec_fill(float *g, size_t n, float v) {
    for (size_t q = blockIdx.x * (size_t)blockDim.x + threadIdx.x; q < n; q += (size_t)gridDim.x * blockDim.x) g[q] = v;
}

__global__ void __launch_bounds__(128) k_rec_core(RecAtoms A, float *gr, int range, float eps) {
    const int l = blockIdx.x, ng = A.ng;
    const float x = A.xyz[3 * l], y = A.xyz[3 * l + 1], z = A.xyz[3 * l + 2];
    const float r = A.radius[l], rc2 = r * r * 1.5f;
    int lo[3], n[3];
    const float p[3] = {x, y, z};
    for (int a = 0; a < 3; a++) {
        const int c = cell_of(p[a], A.gw, ng);
        lo[a] = max(c - range, 0);
        n[a] = max(min(c + range + 1, ng) - lo[a], 0);
    }
    for (int t = threadIdx.x; t < n[0] * n[1] * n[2]; t += blockDim.x) {
        const int i = lo[0] + t / (n[1] * n[2]), j = lo[1] + (t / n[2]) % n[1], k = lo[2] + t % n[2];
        const float ax = sq_dist(x, A.gc[i]), ay = sq_dist(y, A.gc[j]), az = sq_dist(z, A.gc[k]);
        if (ax > rc2 || ax + ay > rc2) continue;
        if (ax + ay + az < rc2) gr[((size_t)i * ng + j) * ng + k] = eps;
    }
}

// one thread per (direction, line): both walks of the line from its two faces
__global__ void k_rec_rays(float *gr, int ng, float eps) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= 3 * ng * ng) return;
    const int dir = t / (ng * ng), p = t % (ng * ng), u = p % ng, v = p / ng;
    const size_t n1 = ng, n2 = n1 * n1;
    auto idx = [&](int s) -> size_t {
        return dir == 0 ? u + n1 * v + n2 * s : dir == 1 ? s + n1 * u + n2 * v : v + n1 * s + n2 * u;
    };
    for (int s = 0; s < ng; s++) {
        float &g = gr[idx(s)];
        if (g == eps) break;
        g = kOpenSpace;
    }
    for (int s = ng - 1; s >= 0; s--) {
        float &g = gr[idx(s)];
        if (g == eps) break;
        g = kOpenSpace;
    }
}

// voxels the sweep visits (src/receptor.cpp:161-165)
__device__ __forceinline__ bool swept(int i, int j, int k, int ng) {
    return i >= 1 && i <= ng - 2 && j >= 2 && j <= ng - 2 && k >= 2 && k <= ng - 2;
}

// Open-when-reached flags of one sweep, plane s = i + j + k after plane: pr[v] = 1 when v
// is visited and open at that moment.  One workgroup; the previous plane's flags in LDS,
// indexed by (i, j).
constexpr int kSweepThreads = 1024;
__global__ void __launch_bounds__(kSweepThreads) k_rec_sweep_flags(const float *__restrict__ gr, uint8_t *pr, int ng) {
    extern __shared__ uint8_t plane[];  // [2][ng * ng]
    uint8_t *prev = plane, *cur = plane + ng * ng;
    for (int q = threadIdx.x; q < 2 * ng * ng; q += kSweepThreads) plane[q] = 0;
    __syncthreads();
    for (int s = 0; s <= 3 * (ng - 1); s++) {
        for (int q = threadIdx.x; q < ng * ng; q += kSweepThreads) {
            const int i = q / ng, j = q % ng, k = s - i - j;
            uint8_t f = 0;
            if (k >= 0 && k < ng) {
                const size_t v = ((size_t)i * ng + j) * ng + k;
                const float g = gr[v];
                bool open = g == kOpenSpace;
                if (!open && g == 0.0f)
                    open = (k > 0 && prev[q]) || (j > 0 && prev[q - 1]) || (i > 0 && prev[q - ng]);
                f = (open && swept(i, j, k, ng)) ? 1 : 0;
                pr[v] = f;
            }
            cur[q] = f;
        }
        __syncthreads();
        uint8_t *t = prev;
        prev = cur;
        cur = t;
    }
}

// end of a sweep: a zero voxel next to a voxel open when reached is open
__global__ void k_rec_sweep_apply(float *gr, const uint8_t *__restrict__ pr, int ng) {
    const size_t n1 = ng, n2 = n1 * n1, n3 = n2 * n1;
    const size_t v = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (v >= n3) return;
    if (gr[v] != 0.0f) return;
    const int i = (int)(v / n2), j = (int)((v / n1) % n1), k = (int)(v % n1);
    const bool m = (k > 0 && pr[v - 1]) || (k < ng - 1 && pr[v + 1]) || (j > 0 && pr[v - n1]) ||
                   (j < ng - 1 && pr[v + n1]) || (i > 0 && pr[v - n2]) || (i < ng - 1 && pr[v + n2]);
    if (m) gr[v] = kOpenSpace;
}

__global__ void k_rec_cavities(float *gr, size_t n3, float eps) {
    const size_t v = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (v >= n3) return;
    const float g = gr[v];
    if (g == 0.0f) gr[v] = eps;  // enclosed cavity
    else if (g == kOpenSpace) gr[v] = 0.0f;
}

// Surface score and electrostatics, one thread per voxel of a 4 x 8 x 8 tile.
constexpr int kTI = 4, kTJ = 8, kTK = 8, kSumThreads = kTI * kTJ * kTK;
struct TileAtom {
    float x, y, z, rj, q;
    double add;
    int lo[3], hi[3];
};

__device__ __forceinline__ void atom_box(const float *p, float gw, int ng, int range, int *lo, int *hi) {
    for (int a = 0; a < 3; a++) {
        const int c = cell_of(p[a], gw, ng);
        lo[a] = max(c - range, 0);
        hi[a] = min(c + range + 1, ng);
    }
}

__global__ void __launch_bounds__(kSumThreads)
    k_rec_sums(RecAtoms A, float *gr, float *gi, float eps, int surf_range, float aceratio, float eratio, float gwl,
               float charge_per_grid) {
    __shared__ TileAtom list[kSumThreads];
    __shared__ int wave_count[kSumThreads / 64];
    const int ng = A.ng, tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int t0[3] = {(int)blockIdx.z * kTI, (int)blockIdx.y * kTJ, (int)blockIdx.x * kTK};
    const int i = t0[0] + tid / (kTJ * kTK), j = t0[1] + (tid / kTK) % kTJ, k = t0[2] + tid % kTK;
    const bool mine = i < ng && j < ng && k < ng;
    const size_t v = mine ? ((size_t)i * ng + j) * ng + k : 0;
    const float gci = A.gc[mine ? i : 0], gcj = A.gc[mine ? j : 0], gck = A.gc[mine ? k : 0];
    const int t1[3] = {min(t0[0] + kTI, ng), min(t0[1] + kTJ, ng), min(t0[2] + kTK, ng)};
    // the atoms of the next chunk whose boxes meet this tile, in index order
    auto gather = [&](int base, bool elec) -> int {
        const int l = base + tid;
        bool keep = false;
        TileAtom a;
        if (l < A.na) {
            a.x = A.xyz[3 * l];
            a.y = A.xyz[3 * l + 1];
            a.z = A.xyz[3 * l + 2];
            const float p[3] = {a.x, a.y, a.z};
            if (!elec) {
                const float r = A.radius[l];
                a.rj = (r + 3.6f) * (r + 3.6f);
                a.add = 1.0 + (double)A.ace[l] * (-0.8) * (double)aceratio;
                atom_box(p, A.gw, ng, surf_range, a.lo, a.hi);
                keep = true;
            } else {
                a.q = A.charge[l];
                const float aq = fabsf(a.q);
                if (aq > 1.0e-6f) {
                    const int range = (double)aq < 0.07 ? 4 : ((double)aq < 0.21 ? 5 : (int)(aq / charge_per_grid));
                    atom_box(p, gwl, ng, range, a.lo, a.hi);
                    keep = true;
                }
            }
            if (keep)
                for (int d = 0; d < 3; d++) keep = keep && a.lo[d] < t1[d] && a.hi[d] > t0[d];
        }
        const unsigned long long bal = __ballot(keep);
        if (lane == 0) wave_count[wave] = __popcll(bal);
        __syncthreads();
        int off = 0, total = 0;
        for (int w = 0; w < kSumThreads / 64; w++) {
            if (w < wave) off += wave_count[w];
            total += wave_count[w];
        }
        if (keep) list[off + __popcll(bal & ((1ull << lane) - 1ull))] = a;
        __syncthreads();
        return total;
    };
    // surface: sum over atoms of 1 + ACE * (-0.8) * aceratio, in double, rounded per term
    float surf = 0.0f;
    for (int base = 0; base < A.na; base += kSumThreads) {
        const int n = gather(base, false);
        if (mine)
            for (int m = 0; m < n; m++) {
                const TileAtom &a = list[m];
                if (i < a.lo[0] || i >= a.hi[0] || j < a.lo[1] || j >= a.hi[1] || k < a.lo[2] || k >= a.hi[2]) continue;
                const float ax = sq_dist(a.x, gci), ay = sq_dist(a.y, gcj), az = sq_dist(a.z, gck);
                if (ax > a.rj || ax + ay > a.rj) continue;
                if (ax + ay + az < a.rj) surf = (float)((double)surf + a.add);
            }
        __syncthreads();  // the list is rewritten by the next chunk
    }
    float g = mine ? gr[v] : 0.0f;
    if (mine && g != eps) {
        g = surf;
        gr[v] = g;
    }
    // electrostatics on the voxels that are not core (src/receptor.cpp:258-346)
    const float beta = -2800.0f;
    float e = 0.0f;
    for (int base = 0; base < A.na; base += kSumThreads) {
        const int n = gather(base, true);
        if (mine && !(g < 0.0))
            for (int m = 0; m < n; m++) {
                const TileAtom &a = list[m];
                if (i < a.lo[0] || i >= a.hi[0] || j < a.lo[1] || j >= a.hi[1] || k < a.lo[2] || k >= a.hi[2]) continue;
                const float dij = sq_dist(a.x, gci) + sq_dist(a.y, gcj);
                const float d = sqrtf(dij + sq_dist(a.z, gck));
                const float er = d >= 8.0f ? 80.0f : (d > 6.0f ? 38.0f * d + -224.0f : 4.0f);
                e += eratio * beta * a.q / er / d;
            }
        __syncthreads();
    }
    if (mine) gi[v] = e;
}

__global__ void k_rec_pack(const float *__restrict__ gr, const float *__restrict__ gi, float2 *out, size_t n3) {
    const size_t v = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (v < n3) out[v] = make_float2(gr[v], gi[v]);
}

}  // namespace

size_t receptor_voxel_scratch_bytes(int ng) {
    const size_t n3 = (size_t)ng * ng * ng;
    return 2 * n3 * sizeof(float) + n3;  // grid_r, grid_i, sweep flags
}

// atoms (centred coordinates xyz, radius, charge, ace: device, na each; gc: device, ng)
hipError_t launch_receptor_voxel(const float *xyz, const float *radius, const float *charge, const float *ace,
                                 const float *gc, int na, int ng, float gw, float aceratio, float rec_core, float eratio,
                                 float gc0, float gc1, void *scratch, float2 *out, hipStream_t s) {
    const size_t n3 = (size_t)ng * ng * ng;
    float *gr = (float *)scratch, *gi = gr + n3;
    uint8_t *pr = (uint8_t *)(gi + n3);
    const RecAtoms A{xyz, radius, charge, ace, gc, na, ng, gw};
    const unsigned fill_blocks = (unsigned)((n3 + 255) / 256 < 65535 ? (n3 + 255) / 256 : 65535);
    k_rec_fill<<<fill_blocks, 256, 0, s>>>(gr, n3, 0.0f);
    const int core_range = (int)((2.4 + (double)gw - 0.01) / (double)gw);
    if (na > 0) k_rec_core<<<na, 128, 0, s>>>(A, gr, core_range, rec_core);
    k_rec_rays<<<(3 * ng * ng + 255) / 256, 256, 0, s>>>(gr, ng, rec_core);
    for (int sweep = 0; sweep < 2; sweep++) {
        k_rec_sweep_flags<<<1, kSweepThreads, 2 * (size_t)ng * ng, s>>>(gr, pr, ng);
        k_rec_sweep_apply<<<(unsigned)((n3 + 255) / 256), 256, 0, s>>>(gr, pr, ng);
    }
    k_rec_cavities<<<(unsigned)((n3 + 255) / 256), 256, 0, s>>>(gr, n3, rec_core);
    const int surf_range = (int)((12 + (double)gw - 0.01) / (double)gw);
    const float beta = -2800.0f, gwl = gc1 - gc0;
    const float charge_per_grid = -1 * gwl * 80.0f / (beta * eratio);
    const dim3 tiles((ng + kTK - 1) / kTK, (ng + kTJ - 1) / kTJ, (ng + kTI - 1) / kTI);
    k_rec_sums<<<tiles, kSumThreads, 0, s>>>(A, gr, gi, rec_core, surf_range, aceratio, eratio, gwl, charge_per_grid);
    k_rec_pack<<<(unsigned)((n3 + 255) / 256), 256, 0, s>>>(gr, gi, out, n3);
    return hipGetLastError();
}
