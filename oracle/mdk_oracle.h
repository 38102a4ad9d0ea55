/*
 * CPU oracle for the FFT docking rotation loop -- TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C restatement of the reference's CPU (OpenMP + FFTW float) path.  Only the
 * tests, the smoke check and the CPU-baseline leg of the benchmark may load it, and
 * only as the checker: nothing in the product links or calls it.
 *
 * Every function names the reference lines it restates.  The FFT is this file's own
 * mixed-radix float FFT standing in for FFTW 3.3.8 (float build, FFTW_ESTIMATE),
 * which is not present in this image; its results are pinned against numpy's FFT in
 * float64, brute-force DFTs and closed forms by the tests, not against FFTW.
 */
#ifndef MDK_ORACLE_H
#define MDK_ORACLE_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Layout-identical to the reference's SortScore (src/fft_process.h:40-43). */
typedef struct {
    float score;
    int index[4]; /* rotation, tx, ty, tz */
} or_sortscore;

int or_version(void);

/* A2: ZXZ rotation coefficients r11..r33 (row-major r[0..8]) in float with
 * cosf/sinf, same expression order as src/docking.cpp:376-384. */
void or_rotation_coef(const float theta[3], float r[9]);

/* A2: x' = R (x - c) per atom, src/docking.cpp:386-393. */
void or_rotate_ligand(int na, const float *coord, const float center[3], const float r[9],
                      float *out);

/* A3: grid coordinates, src/docking.cpp:133-141. */
void or_grid_coord(int ng, float gw, float *out);

/* Protein centre (src/protein.cpp:67-111) and centring (src/protein.cpp:114-131). */
void or_center(int na, const float *coord, float center[3]);
void or_shift_center(int na, const float *coord, float *out);

/* A4 + A5: ligand shape and charge grids (src/ligand.cpp:20-239). grid_r, grid_i: ng^3. */
void or_ligand_voxel(int ng, float gw, int na, const float *xyz_rot, const float *radius,
                     const float *charge, const float *grid_coord, float delta,
                     float *grid_r, float *grid_i);

/* A14: receptor shape (rPSC + ACE) and electrostatics grids
 * (src/receptor.cpp:20-372). xyz: centred coordinates. */
void or_receptor_voxel(int ng, float gw, int na, const float *xyz, const float *radius,
                       const float *charge, const float *ace, const float *grid_coord,
                       float aceratio, float rec_core, float eratio,
                       float *grid_r, float *grid_i);

/* A7: embed ng^3 grids into the zero-padded nf^3 complex grid (src/fft_process.cpp:306-330). */
void or_embed(int ng, int nf, const float *grid_r, const float *grid_i, float *cplx);

/* A8/A10: unnormalised 3D complex FFT, sign -1 forward, +1 inverse. in/out nf^3
 * interleaved complex64; in == out allowed. */
int or_fft3d(int nf, const float *in, float *out, int sign);
/* 1D helper for tests: n-point FFT of `count` contiguous lines. */
int or_fft1d(int n, int count, const float *in, float *out, int sign);

/* A9: out = conj(R) * L with R in SoA form (src/fft_process.cpp:333-346). */
void or_convolution(size_t nf3, const float *rec_r, const float *rec_i, const float *lig,
                    float *out);

/* A11: per-rotation top-t of the real part (src/fft_process.cpp:399-458). */
void or_score_sort(int nf, int t, const float *fftout, int rot, or_sortscore *out);

/* A14 transform: embed + forward FFT + SoA split (src/fft_process.cpp:277-298). */
int or_receptor_fft(int ng, int nf, const float *grid_r, const float *grid_i, float *rec_r,
                    float *rec_i);

/* A12: order by score descending, ties by (rotation asc, position asc), keep the
 * first n (src/fft_process.cpp:462-498 sorts with std::partial_sort, whose tie order
 * is unspecified). */
void or_sort_topn(or_sortscore *top, int count, int n);

/* A13: the CPU rotation loop (src/docking.cpp:194-346): for each listed rotation,
 * rotate, voxelize, embed, FFT, product, inverse FFT, top-t.  OpenMP dynamic,1 over
 * rotations with `threads` threads.  rot_ids: the rotations to run (indices into
 * angles[3*R]); out: n_rot*t entries.  Returns 0 on success. */
int or_dock(int ng, int nf, float gw, int na, const float *coord, const float *center,
            const float *radius, const float *charge, float delta,
            const float *rec_r, const float *rec_i, const float *angles,
            const int *rot_ids, int n_rot, int t, int threads, or_sortscore *out);

/* A15 (-O): the float sums of Docking::output_detail for one pose at grid translation
 * trans (a decoy's index[1..3]), in the reference's order (src/docking.cpp:646-654,
 * :720-807): out = (rPSC+, rPSC-, ELEC+, ELEC-, ACE+, ACE-, rPSC, ELEC, ACE unsplit), cnt
 * (optional) the counts of the six split sums. */
void or_detail_sums_f32(int ng, const float *rec_r, const float *rec_i, const float *rec_rpsc, const float *lig_r,
                        const float *lig_i, const int trans[3], float out[9], int cnt[6]);

#ifdef __cplusplus
}
#endif
#endif
