// Internal interfaces shared by the HIP translation units of libmdk.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

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

// Ligand voxelization of B rotations into lg[b*ng^3 + v] = (grid_r, grid_i).
// rot: B*9 rotation coefficients (row-major r11..r33).  core/surf: 2 * B * nw words of
// bit masks (nw = ceil(ng^3 / 32)), vidx: B*na ints; all three are scratch.
hipError_t launch_voxelize(const LigandDev &lig, const float *rot, int B, uint32_t *bits,
                           int *vidx, float2 *lg, hipStream_t s);

inline int mask_words(int ng) { return (ng * ng * ng + 31) / 32; }
