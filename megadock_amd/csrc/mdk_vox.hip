// Ligand voxelization on the device, bit-exact with the reference's CPU voxelizer
// (src/ligand.cpp:20-261, src/docking.cpp:368-396).
//
// Built with -ffp-contract=off: every float expression rounds after each operation in
// the same order as the reference's x86-64 build (no FMA), so rotated coordinates,
// squared distances and voxel indices are identical.
//
// The reference's sequential sweeps are restated as order-free rules:
//   core[v]   = some atom has (xd+yd)+zd < 1.5 r^2 in its 5^3 box      (ligand.cpp:68-101)
//   scraped   = interior(v) & core[v] & some 6-neighbour not core      (ligand.cpp:110-134;
//               the sweep tests ==0 on values it never turns into 0, so it does not cascade)
//   grid_r[v] = core & !scraped ? delta : (surf[v] ? 1 : 0)             (ligand.cpp:140-173)
//   delta != 1: interior delta voxels with a 0 neighbour become 1       (ligand.cpp:176-195;
//               again non-cascading because delta != 0 and 1 != 0)
//   grid_i[v] = charges of the atoms nearest to v, summed in atom order (ligand.cpp:212-235)
// delta == 0 makes the last rule order-dependent; the host rejects it.
//
// Representation handed to the FFT passes (VoxGrid): per rotation, a bit mask of the
// voxels whose shape value is delta (D) and of those whose value is 1 (O), each row
// (i, j) of ng voxels in its own rw = ceil(ng/32) words, and a dense float charge grid.
// The rules above become word-wide bit operations over rows; the charge grid and the
// per-voxel atom counts are kept all-zero between batches (their consumers clear what
// they read), so a batch writes only the voxels its atoms touch.
#pragma clang fp contract(off)

#include "mdk_internal.h"

namespace {

__device__ __forceinline__ void rotate_atom(const LigandDev &L, const float *r, int l, float &X,
                                            float &Y, float &Z) {
    const float x = L.xyz[3 * l] - L.cx;
    const float y = L.xyz[3 * l + 1] - L.cy;
    const float z = L.xyz[3 * l + 2] - L.cz;
    X = r[0] * x + r[1] * y + r[2] * z;
    Y = r[3] * x + r[4] * y + r[5] * z;
    Z = r[6] * x + r[7] * y + r[8] * z;
}

__global__ void k_clear(uint32_t *p, size_t n) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x)
        p[i] = 0u;
}

// set bits [k0, k1) of a row that starts at word `row` (k1 - k0 <= 5)
__device__ __forceinline__ void or_run(uint32_t *m, int row, int k0, int k1) {
    if (k1 <= k0) return;
    const int w0 = k0 >> 5, w1 = (k1 - 1) >> 5;
    if (w0 == w1) {
        atomicOr(&m[row + w0], ((1u << (k1 - k0)) - 1u) << (k0 & 31));
    } else {
        atomicOr(&m[row + w0], 0xffffffffu << (k0 & 31));
        atomicOr(&m[row + w1], 0xffffffffu >> (31 - ((k1 - 1) & 31)));
    }
}

// Atom l of rotation b: its core/surface bit runs into cb/sb (global or LDS masks of one
// rotation), its charge voxel into vidx and the voxel's atom count.
__device__ __forceinline__ void vox_atom(const LigandDev &L, const float *r, int l, int b, uint32_t *cb,
                                         uint32_t *sb, uint32_t *cnt, int *vidx, int cw) {
    float X, Y, Z;
    rotate_atom(L, r, l, X, Y, Z);
    const int ng = L.ng, rw = row_words(ng);
    const float h = (float)(ng / 2);
    const int i2 = (int)(X / L.gw + h);
    const int j2 = (int)(Y / L.gw + h);
    const int k2 = (int)(Z / L.gw + h);
    const int ia = max(i2 - 2, 0), ja = max(j2 - 2, 0), ka = max(k2 - 2, 0);
    const int ib = min(i2 + 3, ng), jb = min(j2 + 3, ng), kb = min(k2 + 3, ng);
    const float rc2 = L.rc2[l], rs2 = L.rs2[l];
    for (int i = ia; i < ib; i++) {
        const float dx = X - L.gc[i];
        const float xd = dx * dx;
        if (xd > rc2) continue;  // rs2 <= rc2, so neither sphere reaches this plane
        for (int j = ja; j < jb; j++) {
            const float dy = Y - L.gc[j];
            const float d2 = xd + dy * dy;
            if (d2 > rc2) continue;
            // the hits along k form one run per sphere (d3 grows away from the atom), but
            // each voxel is tested exactly as the reference does
            int c0 = kb, c1 = ka, s0 = kb, s1 = ka;
            for (int k = ka; k < kb; k++) {
                const float dz = Z - L.gc[k];
                const float d3 = d2 + dz * dz;
                if (d3 < rc2) { c0 = min(c0, k); c1 = k + 1; }
                if (d3 < rs2) { s0 = min(s0, k); s1 = k + 1; }
            }
            const int row = (i * ng + j) * rw;
            or_run(cb, row, c0, c1);
            or_run(sb, row, s0, s1);
        }
    }
    const int ci = max(0, min(ng - 1, (int)((X + L.pad) / L.gw)));
    const int cj = max(0, min(ng - 1, (int)((Y + L.pad) / L.gw)));
    const int ck = max(0, min(ng - 1, (int)((Z + L.pad) / L.gw)));
    const int v = (ci * ng + cj) * ng + ck;
    vidx[(size_t)b * L.na + l] = v;
    atomicAdd(&cnt[(size_t)b * cw + (v >> 1)], 1u << (16 * (v & 1)));
}

// one thread per (atom, rotation): core/surface bits, the charge voxel and its atom count
// boff (optional): device word holding the batch's first row of rot
__global__ void k_vox_atoms(LigandDev L, const float *__restrict__ rot, const int *__restrict__ boff, uint32_t *cs,
                            uint32_t *cnt, int *vidx, int mw, int cw) {
    const int l = blockIdx.x * blockDim.x + threadIdx.x;
    const int b = blockIdx.y;
    if (l >= L.na) return;
    const float *rb = rot + 9 * ((size_t)(boff ? *boff : 0) + b);
    float r[9];
#pragma unroll
    for (int q = 0; q < 9; q++) r[q] = rb[q];
    uint32_t *cb = cs + (size_t)b * 2 * mw;  // core bits
    vox_atom(L, r, l, b, cb, cb + mw, cnt, vidx, cw);
}

// bits k = 32 w + s of word w with lo <= k <= hi
__device__ __forceinline__ uint32_t range_bits(int w, int lo, int hi) {
    const int a = max(lo - 32 * w, 0), e = min(hi - 32 * w, 31);
    if (a > e) return 0u;
    const uint32_t upto = (e == 31) ? 0xffffffffu : ((1u << (e + 1)) - 1u);
    return upto & (0xffffffffu << a);
}

struct MaskRows {
    const uint32_t *m;
    int ng, rw;
    // word w of row (i, j); zero outside the grid
    __device__ __forceinline__ uint32_t at(int i, int j, int w) const {
        if (i < 0 || i >= ng || j < 0 || j >= ng || w < 0 || w >= rw) return 0u;
        return m[(i * ng + j) * rw + w];
    }
};

// core after the scrape step, word w of row (i, j)
__device__ __forceinline__ uint32_t core2_word(const MaskRows &C, int i, int j, int w) {
    const int ng = C.ng;
    const uint32_t c = C.at(i, j, w);
    if (i < 1 || i > ng - 2 || j < 1 || j > ng - 2 || c == 0u) return c;
    const uint32_t zm = (c << 1) | (C.at(i, j, w - 1) >> 31);  // bit k: voxel k-1 is core
    const uint32_t zp = (c >> 1) | (C.at(i, j, w + 1) << 31);  // bit k: voxel k+1 is core
    const uint32_t closed = zm & zp & C.at(i, j - 1, w) & C.at(i, j + 1, w) & C.at(i - 1, j, w) & C.at(i + 1, j, w);
    return c & ~(range_bits(w, 1, ng - 2) & ~closed);
}

// word q of the delta mask D and the value-1 mask O of one rotation from its core (C) and
// surface (S) bits
__device__ __forceinline__ void vox_mask_word(const MaskRows &C, const MaskRows &S, int q, int ng, float delta,
                                              uint32_t &dq, uint32_t &oq) {
    const int rw = C.rw;
    const int w = q % rw, row = q / rw, i = row / ng, j = row - i * ng;
    uint32_t d = core2_word(C, i, j, w);
    uint32_t o = ~d & S.at(i, j, w);
    if (delta != 1.0f && i >= 1 && i <= ng - 2 && j >= 1 && j <= ng - 2) {
        // interior delta voxels next to a voxel of value 0 (neither core nor surface) become 1
        auto zero = [&](int ii, int jj, int ww) { return ~core2_word(C, ii, jj, ww) & ~S.at(ii, jj, ww); };
        const uint32_t z0 = zero(i, j, w);
        const uint32_t zm = (z0 << 1) | (zero(i, j, w - 1) >> 31);
        const uint32_t zp = (z0 >> 1) | (zero(i, j, w + 1) << 31);
        const uint32_t nb = zm | zp | zero(i, j - 1, w) | zero(i, j + 1, w) | zero(i - 1, j, w) | zero(i + 1, j, w);
        const uint32_t to1 = d & range_bits(w, 1, ng - 2) & nb;
        d &= ~to1;
        o |= to1;
    }
    dq = d;
    oq = o;
}

// one thread per (row word, rotation): the delta mask D and the value-1 mask O
__global__ void k_vox_masks(const uint32_t *__restrict__ cs, uint32_t *dmo, int ng, int mw, float delta) {
    const int rw = row_words(ng);
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    const int b = blockIdx.y;
    if (q >= mw) return;
    const MaskRows C{cs + (size_t)b * 2 * mw, ng, rw}, S{cs + (size_t)b * 2 * mw + mw, ng, rw};
    uint32_t d, o;
    vox_mask_word(C, S, q, ng, delta, d, o);
    dmo[(size_t)b * 2 * mw + q] = d;
    dmo[(size_t)b * 2 * mw + mw + q] = o;
}

// One workgroup per rotation with its core and surface bits in LDS (grids whose masks
// fit): atom runs by LDS atomics, then the D and O words straight to dmo.  Replaces
// k_clear + k_vox_atoms + k_vox_masks where it fits.
constexpr int kVoxLdsThreads = 1024;
constexpr size_t kVoxLdsBytes = 132 * 1024;  // both masks of one rotation, at most
__global__ void __launch_bounds__(kVoxLdsThreads)
    k_vox_lds(LigandDev L, const float *__restrict__ rot, const int *__restrict__ boff, uint32_t *dmo, uint32_t *cnt,
              int *vidx, int mw, int cw) {
    extern __shared__ uint32_t msk[];  // [2][mw]: core, surface
    const int b = blockIdx.x, tid = threadIdx.x, ng = L.ng;
    for (int q = tid; q < 2 * mw; q += kVoxLdsThreads) msk[q] = 0u;
    const float *rb = rot + 9 * ((size_t)(boff ? *boff : 0) + b);
    float r[9];
#pragma unroll
    for (int q = 0; q < 9; q++) r[q] = rb[q];
    __syncthreads();
    for (int l = tid; l < L.na; l += kVoxLdsThreads) vox_atom(L, r, l, b, msk, msk + mw, cnt, vidx, cw);
    __syncthreads();
    const MaskRows C{msk, ng, row_words(ng)}, S{msk + mw, ng, row_words(ng)};
    uint32_t *db = dmo + (size_t)b * 2 * mw;
    for (int q = tid; q < mw; q += kVoxLdsThreads) {
        uint32_t d, o;
        vox_mask_word(C, S, q, ng, L.delta, d, o);
        db[q] = d;
        db[mw + q] = o;
    }
}

// Charges of the atoms nearest to each voxel, summed in atom order from +0.0f
// (src/ligand.cpp:216-234).  One workgroup per rotation.  With one or two atoms in a
// voxel the float sum does not depend on the order (0 + a = a and a + b = b + a
// exactly), so those are added atomically; atoms of voxels holding three or more are
// listed in LDS and each such voxel is summed in atom order by its lowest-index atom.
// Leaves the rotation's atom counts zero again.
constexpr int kChargeThreads = 1024;
constexpr int kCrowdCap = 2048;  // listed atoms of crowded voxels (beyond: global scan)

__global__ void __launch_bounds__(kChargeThreads)
    k_vox_charge(const int *__restrict__ vidx, uint32_t *cnt, const float *__restrict__ charge, float *chg,
                 int na, int ng3, int cw) {
    __shared__ int nlist;
    __shared__ int2 list[kCrowdCap];  // (voxel, atom)
    const int b = blockIdx.x, tid = threadIdx.x;
    const int *vi = vidx + (size_t)b * na;
    uint32_t *cb = cnt + (size_t)b * cw;
    float *g = chg + (size_t)b * ng3;
    if (tid == 0) nlist = 0;
    __syncthreads();
    for (int l = tid; l < na; l += kChargeThreads) {
        const int v = vi[l];
        const uint32_t n = (cb[v >> 1] >> (16 * (v & 1))) & 0xffffu;
        if (n <= 2) {
            atomicAdd(&g[v], charge[l]);
        } else {
            const int k = atomicAdd(&nlist, 1);
            if (k < kCrowdCap) list[k] = make_int2(v, l);
        }
    }
    __syncthreads();
    const int m = nlist;  // uniform
    if (m <= kCrowdCap) {
        for (int e = tid; e < m; e += kChargeThreads) {
            const int v = list[e].x, l = list[e].y;
            bool owner = true;
            for (int f = 0; f < m; f++)
                if (list[f].x == v && list[f].y < l) owner = false;
            if (!owner) continue;
            float sum = 0.0f;
            for (int cur = -1;;) {  // next atom of the voxel in index order
                int nxt = 0x7fffffff;
                for (int f = 0; f < m; f++)
                    if (list[f].x == v && list[f].y > cur && list[f].y < nxt) nxt = list[f].y;
                if (nxt == 0x7fffffff) break;
                sum += charge[nxt];
                cur = nxt;
            }
            g[v] = sum;
        }
    } else {
        for (int l = tid; l < na; l += kChargeThreads) {
            const int v = vi[l];
            const uint32_t n = (cb[v >> 1] >> (16 * (v & 1))) & 0xffffu;
            if (n <= 2) continue;
            bool owner = true;
            for (int q = 0; q < l; q++)
                if (vi[q] == v) owner = false;
            if (!owner) continue;
            float sum = 0.0f;
            for (int q = l; q < na; q++)
                if (vi[q] == v) sum += charge[q];
            g[v] = sum;
        }
    }
    __syncthreads();  // every count read above is done
    for (int l = tid; l < na; l += kChargeThreads) cb[vi[l] >> 1] = 0u;
}

// grids as (grid_r, grid_i) pairs; clears the charge grid it reads
__global__ void k_vox_expand(VoxGrid vg, float2 *lg, int ng3) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    const int b = blockIdx.y;
    if (v >= ng3) return;
    lg[(size_t)b * ng3 + v] = vox_take(vg, b, v / vg.ng, v % vg.ng);
}

}  // namespace

hipError_t launch_voxelize(const LigandDev &lig, const float *rot, int B, const VoxScratch &vs, int *vidx,
                           float *chg, hipStream_t s, const int *boff) {
    const int ng = lig.ng, mw = mask_words(ng), cw = count_words(ng);
    const int ng3 = ng * ng * ng;
    const size_t lds = 2 * sizeof(uint32_t) * (size_t)mw;
    if (lds <= kVoxLdsBytes) {  // one workgroup per rotation, bits in LDS
        k_vox_lds<<<B, kVoxLdsThreads, lds, s>>>(lig, rot, boff, vs.dmo, vs.cnt, vidx, mw, cw);
    } else {
        const size_t nclear = 2 * (size_t)B * mw;  // core and surface bits; counts are kept zero
        k_clear<<<dim3((unsigned)((nclear + 255) / 256 < 65535 ? (nclear + 255) / 256 : 65535)), 256, 0, s>>>(vs.cs,
                                                                                                       nclear);
        k_vox_atoms<<<dim3((lig.na + 127) / 128, B), 128, 0, s>>>(lig, rot, boff, vs.cs, vs.cnt, vidx, mw, cw);
        k_vox_masks<<<dim3((mw + 255) / 256, B), 256, 0, s>>>(vs.cs, vs.dmo, ng, mw, lig.delta);
    }
    k_vox_charge<<<B, kChargeThreads, 0, s>>>(vidx, vs.cnt, lig.charge, chg, lig.na, ng3, cw);
    return hipGetLastError();
}

hipError_t launch_vox_expand(const VoxGrid &vg, int B, float2 *lg, hipStream_t s) {
    const int ng3 = vg.ng * vg.ng * vg.ng;
    k_vox_expand<<<dim3((ng3 + 255) / 256, B), 256, 0, s>>>(vg, lg, ng3);
    return hipGetLastError();
}
