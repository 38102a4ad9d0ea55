// Scores off the grid, with their gradients, for refining docked poses.  The reference
// has no counterpart; these kernels extend its score (src/fft_process.cpp:333-446).
//
// 1. Spectral score at a continuous translation.  With Q the Hermitian fold of
//    conj(R^) L^ for one rotation (kz <= nf/2 stored), signed frequencies k' (k' = k for
//    k < nf/2, else k - nf) and phase phi = 2 pi k'.t / nf,
//        s(t) = 1/nf^3 sum_{kz <= nf/2} w(kz) Re[Q(k) e^{i phi}],  w = 1 at kz = 0, nf/2, else 2,
//    which equals the reported score value/nf^3 at every grid point t, and
//        ds/dt_a = 1/nf^3 sum w(kz) (-2 pi k'_a / nf) Im[Q(k) e^{i phi}].
// 2. Companion score of a pose with trilinear interpolation of the receptor grids: with
//    the ligand atom a rotated about the ligand centre, X_a = R (a - c), placed at
//    P_a = X_a - gw t' (t' the signed translation in voxels, as decoygen places poses),
//        s = sum_a [ w_a Rr(P_a) + q_a Ri(P_a) ],   w_a = (4/3) pi r_a^3 / gw^3,
//    where Rr, Ri are the receptor's shape and electrostatic grids (voxel i at
//    gw (i + 1/2 - ng/2), zero outside).  The atom adjoint is g_a = ds/dP_a; from it
//    ds/dt' = -gw sum_a g_a and, for a rotation R -> exp([w]x) R, ds/dw = sum_a X_a x g_a.
//    This spreads each atom trilinearly instead of over its voxel sphere: a smooth
//    companion of the hard voxelizer, recomputed per call, never stored.
#include <hip/hip_runtime.h>

#include <cmath>

#include "mdk_internal.h"

namespace {

constexpr int kSpecBlocks = 64;  // workgroups per rotation in the spectral evaluation
constexpr int kCompBlocks = 8;   // workgroups per pose in the companion score

__device__ __forceinline__ int signed_freq(int k, int nf) { return k < nf / 2 ? k : k - nf; }

// partial sums part[b][blk][4] = (value, d/dtx, d/dty, d/dtz) before scaling
__global__ void __launch_bounds__(256) k_spec_eval(const float2 *__restrict__ Q, const float *__restrict__ tq,
                                                   double *part, int nf, int H, int Hp) {
    extern __shared__ float2 ph[];  // phases e^{2 pi i k' t_a / nf}: [3][nf]
    const int b = blockIdx.y, tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    for (int k = tid; k < nf; k += 256) {
        const double f = 2.0 * M_PI * signed_freq(k, nf) / nf;
        for (int a = 0; a < 3; a++) {
            double sn, cs;
            sincos(f * tq[3 * b + a], &sn, &cs);
            ph[a * nf + k] = make_float2((float)cs, (float)sn);
        }
    }
    __syncthreads();
    const float2 *q = Q + (size_t)b * nf * nf * Hp;
    // lane owns kz = lane, lane + 64 and lane + 128 (< H; kz = 128 is the Nyquist plane of nf = 256)
    constexpr int JZ = 3;
    float2 ez[JZ];
    float wz[JZ], kzs[JZ];
    bool on[JZ];
    for (int j = 0; j < JZ; j++) {
        const int kz = lane + 64 * j;
        on[j] = kz < H;
        ez[j] = on[j] ? ph[2 * nf + kz] : make_float2(0.f, 0.f);
        wz[j] = (kz == 0 || 2 * kz == nf) ? 1.0f : 2.0f;
        kzs[j] = (float)signed_freq(kz, nf);
    }
    double v = 0.0, gx = 0.0, gy = 0.0, gz = 0.0;  // 8.5M terms a pose at 256^3: summed in double
    const int rows = nf * nf, wtot = kSpecBlocks * 4;
    for (int r = blockIdx.x * 4 + wv; r < rows; r += wtot) {
        const int kx = r / nf, ky = r - kx * nf;
        const float2 ex = ph[kx], ey = ph[nf + ky];
        const float2 er = make_float2(ex.x * ey.x - ex.y * ey.y, ex.x * ey.y + ex.y * ey.x);
        const float kxs = (float)signed_freq(kx, nf), kys = (float)signed_freq(ky, nf);
        for (int j = 0; j < JZ; j++) {
            if (!on[j]) continue;
            const float2 e = make_float2(er.x * ez[j].x - er.y * ez[j].y, er.x * ez[j].y + er.y * ez[j].x);
            const float2 x = q[(size_t)r * Hp + lane + 64 * j];
            const float zr = x.x * e.x - x.y * e.y, zi = x.x * e.y + x.y * e.x;
            v += wz[j] * zr;
            const float im = wz[j] * zi;
            gx += kxs * im;
            gy += kys * im;
            gz += kzs[j] * im;
        }
    }
    __shared__ double red[4][256];
    red[0][tid] = v;
    red[1][tid] = gx;
    red[2][tid] = gy;
    red[3][tid] = gz;
    __syncthreads();
    for (int h = 128; h >= 1; h >>= 1) {
        if (tid < h)
            for (int a = 0; a < 4; a++) red[a][tid] += red[a][tid + h];
        __syncthreads();
    }
    if (tid < 4) part[((size_t)b * kSpecBlocks + blockIdx.x) * 4 + tid] = red[tid][0];
}

// out[b][0..3] = (s, ds/dtx, ds/dty, ds/dtz), partial sums added in a fixed order
__global__ void k_spec_finish(const double *__restrict__ part, float *out, int B, int nf) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= B) return;
    double s[4] = {0, 0, 0, 0};
    for (int k = 0; k < kSpecBlocks; k++)
        for (int a = 0; a < 4; a++) s[a] += part[((size_t)b * kSpecBlocks + k) * 4 + a];
    const double n3 = (double)nf * nf * nf, g = -2.0 * M_PI / nf / n3;
    out[4 * b] = (float)(s[0] / n3);
    for (int a = 1; a < 4; a++) out[4 * b + a] = (float)(s[a] * g);
}

// trilinear value and gradient (per unit of u) of the two receptor grids at voxel
// coordinate u; corners outside the grid count as 0
__device__ __forceinline__ void trilinear(const float2 *__restrict__ rg, int ng, float ux, float uy, float uz,
                                          float2 &val, float2 &dx, float2 &dy, float2 &dz) {
    const float fx0 = floorf(ux), fy0 = floorf(uy), fz0 = floorf(uz);
    const int i0 = (int)fx0, j0 = (int)fy0, k0 = (int)fz0;
    const float fx = ux - fx0, fy = uy - fy0, fz = uz - fz0;
    val = dx = dy = dz = make_float2(0.f, 0.f);
    for (int c = 0; c < 8; c++) {
        const int di = c >> 2, dj = (c >> 1) & 1, dk = c & 1;
        const int i = i0 + di, j = j0 + dj, k = k0 + dk;
        if (i < 0 || i >= ng || j < 0 || j >= ng || k < 0 || k >= ng) continue;
        const float2 g = rg[((size_t)i * ng + j) * ng + k];
        const float wx = di ? fx : 1.f - fx, wy = dj ? fy : 1.f - fy, wz = dk ? fz : 1.f - fz;
        const float sx = di ? 1.f : -1.f, sy = dj ? 1.f : -1.f, sz = dk ? 1.f : -1.f;
        const float w = wx * wy * wz, ax = sx * wy * wz, ay = wx * sy * wz, az = wx * wy * sz;
        val.x += w * g.x;
        val.y += w * g.y;
        dx.x += ax * g.x;
        dx.y += ax * g.y;
        dy.x += ay * g.x;
        dy.y += ay * g.y;
        dz.x += az * g.x;
        dz.y += az * g.y;
    }
}

// part[b][blk][7] = (s, ds/dt'(3), ds/dw(3))
__global__ void __launch_bounds__(256) k_companion(LigandDev L, const float *__restrict__ wshape,
                                                   const float2 *__restrict__ rg, const float *__restrict__ rot,
                                                   const float *__restrict__ tq, int nf, double *part) {
    const int b = blockIdx.y, tid = threadIdx.x;
    float r[9];
    for (int q = 0; q < 9; q++) r[q] = rot[9 * b + q];
    float t3[3];
    for (int a = 0; a < 3; a++) {
        const float t = tq[3 * b + a];
        t3[a] = t < 0.5f * nf ? t : t - nf;  // signed translation (voxels)
    }
    const int ng = L.ng;
    const float gw = L.gw, h = 0.5f * ng - 0.5f;
    double acc[7] = {0, 0, 0, 0, 0, 0, 0};
    for (int a = blockIdx.x * 256 + tid; a < L.na; a += kCompBlocks * 256) {
        const float x = L.xyz[3 * a] - L.cx, y = L.xyz[3 * a + 1] - L.cy, z = L.xyz[3 * a + 2] - L.cz;
        const float X = r[0] * x + r[1] * y + r[2] * z;
        const float Y = r[3] * x + r[4] * y + r[5] * z;
        const float Z = r[6] * x + r[7] * y + r[8] * z;
        const float Px = X - gw * t3[0], Py = Y - gw * t3[1], Pz = Z - gw * t3[2];
        float2 val, dx, dy, dz;
        trilinear(rg, ng, Px / gw + h, Py / gw + h, Pz / gw + h, val, dx, dy, dz);
        const float ws = wshape[a], q = L.charge[a];
        const float s = ws * val.x + q * val.y;
        // g = ds/dP (per Angstrom)
        const float gx = (ws * dx.x + q * dx.y) / gw, gy = (ws * dy.x + q * dy.y) / gw, gz = (ws * dz.x + q * dz.y) / gw;
        acc[0] += s;
        acc[1] += -gw * gx;
        acc[2] += -gw * gy;
        acc[3] += -gw * gz;
        acc[4] += Y * gz - Z * gy;
        acc[5] += Z * gx - X * gz;
        acc[6] += X * gy - Y * gx;
    }
    __shared__ double red[7][256];
    for (int a = 0; a < 7; a++) red[a][tid] = acc[a];
    __syncthreads();
    for (int hh = 128; hh >= 1; hh >>= 1) {
        if (tid < hh)
            for (int a = 0; a < 7; a++) red[a][tid] += red[a][tid + hh];
        __syncthreads();
    }
    if (tid < 7) part[((size_t)b * kCompBlocks + blockIdx.x) * 7 + tid] = red[tid][0];
}

__global__ void k_comp_finish(const double *__restrict__ part, float *out, int B) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= B) return;
    for (int a = 0; a < 7; a++) {
        double s = 0;
        for (int k = 0; k < kCompBlocks; k++) s += part[((size_t)b * kCompBlocks + k) * 7 + a];
        out[7 * b + a] = (float)s;
    }
}

}  // namespace

size_t grad_scratch_doubles(int max_batch) {
    const int per = kSpecBlocks * 4 > kCompBlocks * 7 ? kSpecBlocks * 4 : kCompBlocks * 7;
    return (size_t)max_batch * per;
}

hipError_t launch_spec_eval(const float2 *Q, const float *t_dev, int B, int nf, int H, int Hp, double *part,
                            float *out_dev, hipStream_t s) {
    k_spec_eval<<<dim3(kSpecBlocks, B), 256, sizeof(float2) * 3 * nf, s>>>(Q, t_dev, part, nf, H, Hp);
    k_spec_finish<<<(B + 63) / 64, 64, 0, s>>>(part, out_dev, B, nf);
    return hipGetLastError();
}

hipError_t launch_companion(const LigandDev &lig, const float *wshape, const float2 *recgrid, const float *rot_dev,
                            const float *t_dev, int B, int nf, double *part, float *out_dev, hipStream_t s) {
    k_companion<<<dim3(kCompBlocks, B), 256, 0, s>>>(lig, wshape, recgrid, rot_dev, t_dev, nf, part);
    k_comp_finish<<<(B + 63) / 64, 64, 0, s>>>(part, out_dev, B);
    return hipGetLastError();
}
