// Real-space re-scoring of poses: the sums of the reference's detailed output
// (Docking::output_detail, src/docking.cpp:556-868).  For a pose (rotation, grid
// translation t as a decoy's index[1..3]) the ligand grids L (shape, charge) are the
// device voxelizer's, and the receptor grids R (shape with the ACE term, charge) those
// of mdk_set_receptor; Rp is the receptor shape grid without the ACE term.  With
// s = (3 ng - t) mod 2 ng per axis, ligand voxel (l, m, n) meets receptor voxel
// (l + s - ng, ...) where that lies inside the grid, and per pose
//   rPSC = sum Rp * Lr,  ELEC = sum Ri * Li,  ACE = sum (Rr - Rp) * Lr,
// each split into its positive and negative products.  The products are formed in float
// as the reference forms them; the sums run in double, one workgroup per (pose, x slab)
// and then in slab order, so every call gives the same result.
#include <hip/hip_runtime.h>

#include "mdk_internal.h"

namespace {

constexpr int kDetailThreads = 256;

__global__ void __launch_bounds__(kDetailThreads)
    k_detail(const float2 *__restrict__ lg, const float2 *__restrict__ rec, const float *__restrict__ rpsc,
             const int *__restrict__ trans, int ng, double *__restrict__ part) {
    const int l = blockIdx.x, b = blockIdx.y, tid = threadIdx.x;
    const int ng3 = ng * ng * ng;
    int s[3], lo[3], hi[3];
#pragma unroll
    for (int a = 0; a < 3; a++) {  // src/docking.cpp:720-744
        s[a] = (3 * ng - trans[3 * b + a]) % (2 * ng);
        lo[a] = s[a] < ng ? ng - s[a] : 0;
        hi[a] = s[a] < ng ? ng : 2 * ng - s[a];
    }
    double acc[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};  // rPSC +/-, ELEC +/-, ACE +/-
    if (l >= lo[0] && l < hi[0]) {
        const int nm = hi[1] - lo[1], nn = hi[2] - lo[2];
        const float2 *L = lg + (size_t)b * ng3;
        for (int q = tid; q < nm * nn; q += kDetailThreads) {
            const int m = lo[1] + q / nn, n = lo[2] + q % nn;
            const int lz = (l * ng + m) * ng + n;
            const int rz = ((l + s[0] - ng) * ng + (m + s[1] - ng)) * ng + (n + s[2] - ng);
            const float2 lv = L[lz], rv = rec[rz];
            const float rp = rpsc[rz];
            const float p = rp * lv.x;               // rPSC
            const float e = rv.y * lv.y;             // ELEC
            const float c = (rv.x - rp) * lv.x;      // ACE
            if (p > 0.0f) acc[0] += p; else if (p < 0.0f) acc[1] += p;
            if (e > 0.0f) acc[2] += e; else if (e < 0.0f) acc[3] += e;
            if (c > 0.0f) acc[4] += c; else if (c < 0.0f) acc[5] += c;
        }
    }
    __shared__ double red[6][kDetailThreads];
#pragma unroll
    for (int k = 0; k < 6; k++) red[k][tid] = acc[k];
    __syncthreads();
    for (int h = kDetailThreads / 2; h >= 1; h >>= 1) {  // fixed tree: deterministic
        if (tid < h)
#pragma unroll
            for (int k = 0; k < 6; k++) red[k][tid] += red[k][tid + h];
        __syncthreads();
    }
    if (tid < 6) part[((size_t)b * ng + l) * 6 + tid] = red[tid][0];
}

__global__ void k_detail_finish(const double *__restrict__ part, int ng, double *__restrict__ out) {
    const int b = blockIdx.x, k = threadIdx.x;
    if (k >= 6) return;
    double v = 0.0;
    for (int l = 0; l < ng; l++) v += part[((size_t)b * ng + l) * 6 + k];
    out[(size_t)b * 6 + k] = v;
}

}  // namespace

size_t detail_scratch_doubles(int max_batch, int ng) { return (size_t)max_batch * ng * 6 + (size_t)max_batch * 6; }

hipError_t launch_detail(const float2 *lg, const float2 *recgrid, const float *rpsc, const int *trans_dev, int B, int ng,
                         double *scratch, double *out_dev, hipStream_t s) {
    k_detail<<<dim3(ng, B), kDetailThreads, 0, s>>>(lg, recgrid, rpsc, trans_dev, ng, scratch);
    k_detail_finish<<<B, 64, 0, s>>>(scratch, ng, out_dev);
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
