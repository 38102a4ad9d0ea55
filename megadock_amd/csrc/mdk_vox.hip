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

// set bits [k0, k1) of the row starting at voxel v0 (k1 - k0 <= 32)
__device__ __forceinline__ void or_run(uint32_t *m, int v0, int k0, int k1) {
    if (k1 <= k0) return;
    const int a = v0 + k0, e = v0 + k1;  // [a, e)
    const int w0 = a >> 5, w1 = (e - 1) >> 5;
    if (w0 == w1) {
        const uint32_t n = (uint32_t)(e - a);
        const uint32_t mask = (n == 32 ? 0xffffffffu : ((1u << n) - 1u)) << (a & 31);
        atomicOr(&m[w0], mask);
    } else {
        atomicOr(&m[w0], 0xffffffffu << (a & 31));
        atomicOr(&m[w1], 0xffffffffu >> (31 - ((e - 1) & 31)));
    }
}

// one thread per (atom, rotation): core/surface bits, the charge voxel and its atom count
__global__ void k_vox_atoms(LigandDev L, const float *__restrict__ rot, uint32_t *core,
                            uint32_t *surf, uint32_t *cnt, int *vidx, int nw, int cw) {
    const int l = blockIdx.x * blockDim.x + threadIdx.x;
    const int b = blockIdx.y;
    if (l >= L.na) return;
    float r[9];
#pragma unroll
    for (int q = 0; q < 9; q++) r[q] = rot[9 * b + q];
    float X, Y, Z;
    rotate_atom(L, r, l, X, Y, Z);
    const int ng = L.ng;
    const float h = (float)(ng / 2);
    const int i2 = (int)(X / L.gw + h);
    const int j2 = (int)(Y / L.gw + h);
    const int k2 = (int)(Z / L.gw + h);
    const int ia = max(i2 - 2, 0), ja = max(j2 - 2, 0), ka = max(k2 - 2, 0);
    const int ib = min(i2 + 3, ng), jb = min(j2 + 3, ng), kb = min(k2 + 3, ng);
    const float rc2 = L.rc2[l], rs2 = L.rs2[l];
    uint32_t *cb = core + (size_t)b * nw;
    uint32_t *sb = surf + (size_t)b * nw;
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
            const int v0 = (i * ng + j) * ng;
            or_run(cb, v0, c0, c1);
            or_run(sb, v0, s0, s1);
        }
    }
    const int ci = max(0, min(ng - 1, (int)((X + L.pad) / L.gw)));
    const int cj = max(0, min(ng - 1, (int)((Y + L.pad) / L.gw)));
    const int ck = max(0, min(ng - 1, (int)((Z + L.pad) / L.gw)));
    const int v = (ci * ng + cj) * ng + ck;
    vidx[(size_t)b * L.na + l] = v;
    atomicAdd(&cnt[(size_t)b * cw + (v >> 1)], 1u << (16 * (v & 1)));
}

__device__ __forceinline__ bool bit(const uint32_t *m, int v) { return (m[v >> 5] >> (v & 31)) & 1u; }

// core after the scrape step
__device__ __forceinline__ bool core2(const uint32_t *c, int ng, int i, int j, int k) {
    const int v = (i * ng + j) * ng + k;
    if (!bit(c, v)) return false;
    const bool interior = i >= 1 && i <= ng - 2 && j >= 1 && j <= ng - 2 && k >= 1 && k <= ng - 2;
    if (!interior) return true;
    const int ng2 = ng * ng;
    const bool open = !bit(c, v - 1) || !bit(c, v + 1) || !bit(c, v - ng) || !bit(c, v + ng) ||
                      !bit(c, v - ng2) || !bit(c, v + ng2);
    return !open;
}

__global__ void k_vox_grid(const uint32_t *core, const uint32_t *surf, float2 *lg, int ng, int nw,
                           float delta) {
    const int ng3 = ng * ng * ng;
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    const int b = blockIdx.y;
    if (v >= ng3) return;
    const uint32_t *c = core + (size_t)b * nw;
    const uint32_t *s = surf + (size_t)b * nw;
    const int i = v / (ng * ng), j = (v / ng) % ng, k = v % ng;
    const bool cc = core2(c, ng, i, j, k);
    float val = cc ? delta : (bit(s, v) ? 1.0f : 0.0f);
    if (cc && delta != 1.0f && i >= 1 && i <= ng - 2 && j >= 1 && j <= ng - 2 && k >= 1 &&
        k <= ng - 2) {
        const int di[6] = {-1, 1, 0, 0, 0, 0}, dj[6] = {0, 0, -1, 1, 0, 0}, dk[6] = {0, 0, 0, 0, -1, 1};
        bool zero_nb = false;
        for (int q = 0; q < 6; q++) {
            const int a = i + di[q], bb = j + dj[q], cq = k + dk[q];
            const int w = (a * ng + bb) * ng + cq;
            if (!core2(c, ng, a, bb, cq) && !bit(s, w)) zero_nb = true;
        }
        if (zero_nb) val = 1.0f;
    }
    lg[(size_t)b * ng3 + v] = make_float2(val, 0.0f);
}

// Charges of the atoms nearest to each voxel, summed in atom order from +0.0f
// (src/ligand.cpp:216-234).  With one or two atoms in a voxel the float sum does not
// depend on the order (0 + a = a and a + b = b + a exactly), so those are added
// atomically; a voxel holding three or more is summed in atom order by its
// lowest-index atom, which is then its only writer.
__global__ void k_vox_charge(const int *__restrict__ vidx, const uint32_t *__restrict__ cnt,
                             const float *__restrict__ charge, float2 *lg, int na, int ng3, int cw) {
    const int l = blockIdx.x * blockDim.x + threadIdx.x;
    const int b = blockIdx.y;
    if (l >= na) return;
    const int *vi = vidx + (size_t)b * na;
    const int v = vi[l];
    const uint32_t n = (cnt[(size_t)b * cw + (v >> 1)] >> (16 * (v & 1))) & 0xffffu;
    float *dst = &lg[(size_t)b * ng3 + v].y;
    if (n <= 2) {
        atomicAdd(dst, charge[l]);
        return;
    }
    for (int q = 0; q < l; q++)
        if (vi[q] == v) return;
    float sum = 0.0f;
    for (int q = l; q < na; q++)
        if (vi[q] == v) sum += charge[q];
    *dst = sum;
}

}  // namespace

hipError_t launch_voxelize(const LigandDev &lig, const float *rot, int B, uint32_t *bits, int *vidx,
                           float2 *lg, hipStream_t s) {
    const int nw = mask_words(lig.ng), cw = count_words(lig.ng);
    const size_t nclear = (size_t)B * (2 * (size_t)nw + cw);
    uint32_t *core = bits, *surf = bits + (size_t)B * nw, *cnt = bits + 2 * (size_t)B * nw;
    const int ng3 = lig.ng * lig.ng * lig.ng;
    k_clear<<<dim3((unsigned)((nclear + 255) / 256 < 65535 ? (nclear + 255) / 256 : 65535)), 256, 0, s>>>(bits, nclear);
    k_vox_atoms<<<dim3((lig.na + 127) / 128, B), 128, 0, s>>>(lig, rot, core, surf, cnt, vidx, nw, cw);
    k_vox_grid<<<dim3((ng3 + 255) / 256, B), 256, 0, s>>>(core, surf, lg, lig.ng, nw, lig.delta);
    k_vox_charge<<<dim3((lig.na + 127) / 128, B), 128, 0, s>>>(vidx, cnt, lig.charge, lg, lig.na, ng3, cw);
    return hipGetLastError();
}
