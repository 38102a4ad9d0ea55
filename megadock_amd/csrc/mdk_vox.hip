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

// one thread per (atom, rotation): core/surface bits and the charge voxel
__global__ void k_vox_atoms(LigandDev L, const float *__restrict__ rot, uint32_t *core,
                            uint32_t *surf, int *vidx, int nw) {
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
            for (int k = ka; k < kb; k++) {
                const float dz = Z - L.gc[k];
                const float d3 = d2 + dz * dz;
                const int v = (i * ng + j) * ng + k;
                if (d3 < rc2) atomicOr(&cb[v >> 5], 1u << (v & 31));
                if (d3 < rs2) atomicOr(&sb[v >> 5], 1u << (v & 31));
            }
        }
    }
    const int ci = max(0, min(ng - 1, (int)((X + L.pad) / L.gw)));
    const int cj = max(0, min(ng - 1, (int)((Y + L.pad) / L.gw)));
    const int ck = max(0, min(ng - 1, (int)((Z + L.pad) / L.gw)));
    vidx[(size_t)b * L.na + l] = (ci * ng + cj) * ng + ck;
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

// The first atom (lowest index) of each charge voxel sums that voxel's charges in
// atom order, starting from +0.0f, and is its only writer.
__global__ void k_vox_charge(const int *__restrict__ vidx, const float *__restrict__ charge,
                             float2 *lg, int na, int ng3) {
    const int l = blockIdx.x * blockDim.x + threadIdx.x;
    const int b = blockIdx.y;
    if (l >= na) return;
    const int *vi = vidx + (size_t)b * na;
    const int v = vi[l];
    for (int q = 0; q < l; q++)
        if (vi[q] == v) return;
    float sum = 0.0f;
    for (int q = l; q < na; q++)
        if (vi[q] == v) sum += charge[q];
    lg[(size_t)b * ng3 + v].y = sum;
}

}  // namespace

hipError_t launch_voxelize(const LigandDev &lig, const float *rot, int B, uint32_t *bits, int *vidx,
                           float2 *lg, hipStream_t s) {
    const int nw = mask_words(lig.ng);
    const size_t nbits = 2 * (size_t)B * nw;
    uint32_t *core = bits, *surf = bits + (size_t)B * nw;
    const int ng3 = lig.ng * lig.ng * lig.ng;
    k_clear<<<dim3((unsigned)((nbits + 255) / 256 < 65535 ? (nbits + 255) / 256 : 65535)), 256, 0, s>>>(bits, nbits);
    k_vox_atoms<<<dim3((lig.na + 127) / 128, B), 128, 0, s>>>(lig, rot, core, surf, vidx, nw);
    k_vox_grid<<<dim3((ng3 + 255) / 256, B), 256, 0, s>>>(core, surf, lg, lig.ng, nw, lig.delta);
    k_vox_charge<<<dim3((lig.na + 127) / 128, B), 128, 0, s>>>(vidx, lig.charge, lg, lig.na, ng3);
    return hipGetLastError();
}
