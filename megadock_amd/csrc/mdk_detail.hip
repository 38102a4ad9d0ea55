// Real-space re-scoring of poses: the sums of the reference's detailed output
// (Docking::output_detail, src/docking.cpp:556-868).  For a pose (rotation, grid
// translation t as a decoy's index[1..3]) the ligand grids L (shape, charge) are the
// device voxelizer's, and the receptor grids R (shape with the ACE term, charge) those
// of mdk_set_receptor; Rp is the receptor shape grid without the ACE term.  With
// s = (3 ng - t) mod 2 ng per axis, ligand voxel (l, m, n) meets receptor voxel
// (l + s - ng, ...) where that lies inside the grid, and per pose
//   rPSC = sum Rp * Lr,  ELEC = sum Ri * Li,  ACE = sum (Rr - Rp) * Lr,
// each split into its positive and negative products (the decoys' lines,
// src/docking.cpp:746-807) and each also unsplit (the lines of the poses around the input
// position, :646-654).  As in the reference the products are formed in float and each is
// added to its float sum in the order l, m, n ascending, so the sums are the reference's
// bit for bit:
// one wavefront per pose walks the rows (l, m) in order, its lanes form the products of
// 64 consecutive n at a time, and the non-zero ones are added one after another in lane
// order (a zero product is skipped by the reference and changes no sum).
#include <hip/hip_runtime.h>

#include "mdk_internal.h"

namespace {

// the products of `v` held by the lanes in `mask`, added to acc in lane order
__device__ __forceinline__ float add_in_lane_order(float acc, float v, unsigned long long mask) {
    while (mask) {  // uniform
        const int k = __ffsll((long long)mask) - 1;
        acc += __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), k));
        mask &= mask - 1;
    }
    return acc;
}

__global__ void __launch_bounds__(64)
    k_detail(const float2 *__restrict__ lg, const float2 *__restrict__ rec, const float *__restrict__ rpsc,
             const int *__restrict__ trans, int ng, float *__restrict__ out) {
    const int b = blockIdx.x, lane = threadIdx.x;
    const size_t ng3 = (size_t)ng * ng * ng;
    int s[3], lo[3], hi[3];
#pragma unroll
    for (int a = 0; a < 3; a++) {  // src/docking.cpp:720-744
        s[a] = (3 * ng - trans[3 * b + a]) % (2 * ng);
        lo[a] = s[a] < ng ? ng - s[a] : 0;
        hi[a] = s[a] < ng ? ng : 2 * ng - s[a];
    }
    float acc[9] = {};  // rPSC +/-, ELEC +/-, ACE +/-, then rPSC, ELEC, ACE unsplit
    const float2 *L = lg + (size_t)b * ng3;
    for (int l = lo[0]; l < hi[0]; l++)
        for (int m = lo[1]; m < hi[1]; m++) {
            const size_t lrow = ((size_t)l * ng + m) * ng;
            const size_t rrow = ((size_t)(l + s[0] - ng) * ng + (m + s[1] - ng)) * ng + (s[2] - ng);
            for (int n0 = lo[2]; n0 < hi[2]; n0 += 64) {
                const int n = n0 + lane;
                float p = 0.0f, e = 0.0f, c = 0.0f;
                if (n < hi[2]) {
                    const float2 lv = L[lrow + n], rv = rec[rrow + n];
                    const float rp = rpsc[rrow + n];
                    p = rp * lv.x;           // rPSC
                    e = rv.y * lv.y;         // ELEC
                    c = (rv.x - rp) * lv.x;  // ACE
                }
                const unsigned long long any = __ballot(p != 0.0f || e != 0.0f || c != 0.0f);
                if (!any) continue;  // uniform: nothing to add in this stretch of the row
                acc[0] = add_in_lane_order(acc[0], p, __ballot(p > 0.0f));
                acc[1] = add_in_lane_order(acc[1], p, __ballot(p < 0.0f));
                acc[2] = add_in_lane_order(acc[2], e, __ballot(e > 0.0f));
                acc[3] = add_in_lane_order(acc[3], e, __ballot(e < 0.0f));
                acc[4] = add_in_lane_order(acc[4], c, __ballot(c > 0.0f));
                acc[5] = add_in_lane_order(acc[5], c, __ballot(c < 0.0f));
                acc[6] = add_in_lane_order(acc[6], p, __ballot(p != 0.0f));
                acc[7] = add_in_lane_order(acc[7], e, __ballot(e != 0.0f));
                acc[8] = add_in_lane_order(acc[8], c, __ballot(c != 0.0f));
            }
        }
    if (lane < 9) {
        float v = acc[0];
#pragma unroll
        for (int k = 1; k < 9; k++)
            if (lane == k) v = acc[k];
        out[(size_t)b * 9 + lane] = v;
    }
}

}  // namespace

hipError_t launch_detail(const float2 *lg, const float2 *recgrid, const float *rpsc, const int *trans_dev, int B, int ng,
                         float *out_dev, hipStream_t s) {
    k_detail<<<B, 64, 0, s>>>(lg, recgrid, rpsc, trans_dev, ng, out_dev);
    return hipGetLastError();
}

// ------------------------------------------------------------------ decoy coordinates
// Atom a of decoy d (decoygen's createPDB, src/decoygen.cpp:65-103): the ligand atom
// relative to its centre, rotated by the .out file's receptor-line angles and then by the
// decoy's angles (rotation coefficients from the host, as decoygen's rotateAtom forms
// them), moved by minus its grid translation times the spacing and by the receptor centre.
namespace {
__global__ void k_decoy_coords(const float *__restrict__ xyz, int na, const float *__restrict__ pre,
                               const float *__restrict__ rot, const int *__restrict__ trans, float spacing, float r1,
                               float r2, float r3, float *__restrict__ out) {
    const int nb = (na + 255) / 256, d = blockIdx.x / nb, a = (blockIdx.x - d * nb) * blockDim.x + threadIdx.x;
    if (a >= na) return;
    const float x = xyz[3 * a], y = xyz[3 * a + 1], z = xyz[3 * a + 2];
    const float x1 = pre[0] * x + pre[1] * y + pre[2] * z;
    const float y1 = pre[3] * x + pre[4] * y + pre[5] * z;
    const float z1 = pre[6] * x + pre[7] * y + pre[8] * z;
    const float *R = rot + 9 * (size_t)d;
    const float x2 = R[0] * x1 + R[1] * y1 + R[2] * z1;
    const float y2 = R[3] * x1 + R[4] * y1 + R[5] * z1;
    const float z2 = R[6] * x1 + R[7] * y1 + R[8] * z1;
    float *o = out + 3 * ((size_t)d * na + a);
    o[0] = x2 - trans[3 * d] * spacing + r1;
    o[1] = y2 - trans[3 * d + 1] * spacing + r2;
    o[2] = z2 - trans[3 * d + 2] * spacing + r3;
}
}  // namespace

hipError_t launch_decoy_coords(const float *xyz, int na, const float *pre, const float *rot, const int *trans, int nd,
                               float spacing, const float *r, float *out, hipStream_t s) {
    const unsigned blocks = (unsigned)((na + 255) / 256) * (unsigned)nd;  // decoy-major, 256 atoms a block
    k_decoy_coords<<<blocks, 256, 0, s>>>(xyz, na, pre, rot, trans, spacing, r[0], r[1], r[2], out);
    return hipGetLastError();
}
