/*
 * libmdk -- MI355X (gfx950) FFT docking rotation loop behind a C ABI.
 *
 * The entry points replace the GPU side of the reference's FFTProcess class
 * (src/fft_process.h:45-196); each declaration cites the reference interface it
 * stands in for.  All functions return 0 on success and a nonzero status otherwise;
 * mdk_last_error() then describes the failure.  No exception and no exit() crosses
 * this boundary: the host maps a nonzero status onto the reference's fatal path.
 *
 * Buffers passed in are host memory unless stated otherwise; they are copied, and the
 * device memory belongs to the context.  A context is used by one host thread at a
 * time (the reference binds OpenMP thread myid2 to device myid2,
 * src/fft_process.cpp:527).
 *
 * Ordering: mdk_dock_range queues work and returns; mdk_sync, mdk_get_results and
 * mdk_topn wait for it.  The set-up calls (mdk_set_receptor*, mdk_set_ligand*,
 * mdk_set_rotations) first wait for every queued batch of the context, so they may be
 * called right after mdk_dock_range without mdk_sync.
 */
#ifndef MDK_HIP_H
#define MDK_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

/* Layout-identical to SortScore (src/fft_process.h:40-43):
 * index = {rotation, tx, ty, tz}. */
typedef struct {
    float score;
    int index[4];
} mdk_sortscore;

typedef struct mdk_ctx mdk_ctx;

/* Library version (no reference counterpart). */
int mdk_version(void);

/* Description of the last failure in this thread (no reference counterpart; the
 * reference prints and exits, e.g. src/fft_process.cpp:538-541). */
const char *mdk_last_error(void);

/* Host helper: ZXZ rotation coefficients (R*9, row-major r11..r33) for Euler triples
 * theta (R*3), in float with cosf/sinf and the reference's expression order
 * (Docking::ligand_rotationz, src/docking.cpp:376-384). */
int mdk_rotation_coef(const float *theta, int R, float *rotcoef);

/* Number of visible GPUs (replaces cudaGetDeviceCount in src/main.cpp:87-109). */
int mdk_device_count(int *count);

/* Allocate device state for one receptor-ligand pair: grids of nf^3 points, ligand
 * voxel grids of ng^3, na_lig ligand atoms, t results per rotation (1..1048576; the
 * reference's -t, src/parameter.cpp:101-102) and rotation batches of up to max_batch.
 * Up to t = 16 the top-t lists are kept in registers by the inverse transform; above,
 * each batch's score maps (max_batch * nf^3 floats) go through a per-rotation radix select.
 * Replaces FFTProcess::alloc_array / alloc_fft (src/fft_process.cpp:29-209). */
int mdk_create(int device, int nf, int ng, int na_lig, int t, int max_batch, mdk_ctx **out);

/* Free everything (FFTProcess::fft_memory_free, src/fft_process.cpp:814-875). */
int mdk_destroy(mdk_ctx *ctx);

/* Receptor grids (ng^3 floats each, row-major x slowest): embedded in the nf^3 grid
 * and transformed on the device (FFTProcess::receptor_fft, src/fft_process.cpp:212-303).
 * Returns once the grids are copied to a staging buffer; the upload and the transform
 * are queued behind the context's queued batches, which see the previous receptor. */
int mdk_set_receptor(mdk_ctx *ctx, const float *grid_r, const float *grid_i);

/* Receptor spectrum given directly, SoA nf^3 floats each (the _FFT_rec_r/_FFT_rec_i
 * arrays of src/fft_process.cpp:257-260). */
int mdk_set_receptor_spectrum(mdk_ctx *ctx, const float *re, const float *im);

/* Receptor grids computed on the device from its atoms, bit-exact with the reference's
 * Receptor::rpscace + Receptor::electro (src/receptor.cpp:20-346): xyz (na*3) centred as
 * Protein::shift_center leaves them (src/protein.cpp:114-131), radius, charge and ACE
 * per atom (na), grid coordinates (ng, src/docking.cpp:139-141), voxel pitch gw, ACE
 * ratio (-d), core value (-a) and electrostatics ratio (-e); then transformed as by
 * mdk_set_receptor.  Replaces the host part of Docking::rec_init (src/docking.cpp:128-159). */
int mdk_set_receptor_atoms(mdk_ctx *ctx, int na, const float *xyz, const float *radius, const float *charge,
                           const float *ace, const float *grid_coord, float gw, float aceratio, float rec_core,
                           float eratio);

/* Decoy coordinates of decoygen (src/decoygen.cpp:65-103) for nd decoys at once, on the
 * device: xyz (na*3) the ligand atoms minus the ligand centre of the .out file, rot_pre
 * (9) the coefficients of the .out file's second-line angles, rotcoef (nd*9) those of the
 * decoys' angles (both as mdk_rotation_coef forms them, which is rotateAtom's
 * expression), trans (nd*3) the decoys' grid translations already mapped to t - nf for
 * t >= nf/2, spacing the grid width and offset (3) the receptor centre;
 * out (nd*na*3) = R_decoy (R_pre xyz) - trans * spacing + offset, in float as decoygen. */
int mdk_decoy_coords(int device, const float *xyz, int na, const float *rot_pre, const float *rotcoef, const int *trans,
                     int nd, float spacing, const float *offset, float *out);

/* Read the receptor grids back (ng^3 each): those of mdk_set_receptor or
 * mdk_set_receptor_atoms. */
int mdk_get_receptor_grids(mdk_ctx *ctx, float *grid_r, float *grid_i);

/* Read the receptor spectrum back (SoA, nf^3 each). */
int mdk_get_receptor_spectrum(mdk_ctx *ctx, float *re, float *im);

/* Ligand atoms: xyz (na*3, PDB coordinates), center (3), radius, charge (na), grid
 * coordinates (ng, src/docking.cpp:139-141), voxel pitch gw and core value (-b).
 * Replaces FFTProcess::ligand_data_transfer_gpu (src/fft_process.cpp:769-807). */
int mdk_set_ligand(mdk_ctx *ctx, const float *xyz, const float *center, const float *radius,
                   const float *charge, const float *grid_coord, float gw, float core_delta);

/* As mdk_set_ligand for a ligand of na atoms, 1 <= na <= the na_lig the context was
 * created with, so that one context serves the pairs of a job table that share an FFT
 * size.  No counterpart in the reference, which builds its FFTProcess per pair (the
 * table farm of src/mpidp.cpp runs one process per job).  Both ligand calls return once
 * the arrays are staged; the upload is queued behind the context's queued batches. */
int mdk_set_ligand_atoms(mdk_ctx *ctx, int na, const float *xyz, const float *center, const float *radius,
                         const float *charge, const float *grid_coord, float gw, float core_delta);

/* Dock B rotations given by host-computed coefficients rotcoef (B*9, row-major
 * r11..r33 as src/docking.cpp:376-384 computes them in float) with rotation numbers
 * rot_id (B); writes B*t results (the _Top entries of src/fft_process.cpp:640-643).
 * Replaces FFTProcess::cuda_fft + ligand_voxelization_on_gpu
 * (src/fft_process.cpp:502-765), batched. */
int mdk_dock(mdk_ctx *ctx, const float *rotcoef, const int *rot_id, int B, mdk_sortscore *out);

/* Dock one rotation whose ligand grids were voxelized by the caller: grid_r (shape) and
 * grid_i (charge), ng^3 floats each, row-major x slowest, as Ligand::rpscace and
 * Ligand::electro leave them (src/ligand.cpp:20-239); writes the rotation's t results
 * with index[0] = rot_id.  Replaces the CPU branch of the reference's loop,
 * FFTProcess::ligand_preparation + fft3d + convolution + score_sort
 * (src/fft_process.cpp:306-458), on the device. */
int mdk_dock_grid(mdk_ctx *ctx, const float *grid_r, const float *grid_i, int rot_id, mdk_sortscore *out);

/* Device-resident rotation table (R*9 coefficients, rotation i has number i), so that
 * mdk_dock_range runs without host traffic (the _Zangle table, src/parameter.h:87). */
int mdk_set_rotations(mdk_ctx *ctx, const float *rotcoef, int R);

/* Dock rotations [first, first+count) of the table, asynchronously, into the
 * device-resident result array (R*t entries, the reference's _Top vector). */
int mdk_dock_range(mdk_ctx *ctx, int first, int count);

/* Wait for queued work. */
int mdk_sync(mdk_ctx *ctx);

/* Copy results of rotations [first, first+count) to the host (count*t entries). */
int mdk_get_results(mdk_ctx *ctx, int first, int count, mdk_sortscore *out);

/* Merge the results of rotations [0, nrot) on the device and return the best n, ordered
 * by score descending, ties by rotation then rank ascending.  Replaces
 * FFTProcess::sort_index (src/fft_process.cpp:469-479), whose tie order is unspecified. */
int mdk_topn(mdk_ctx *ctx, int nrot, int n, mdk_sortscore *out);

/* mdk_topn in two calls: mdk_topn_start queues the merge behind the batches queued so far
 * and returns; mdk_topn_finish waits for it and writes the best n to out.  Between them
 * the caller may set up and queue other contexts' work, so the merge runs right after the
 * context's last batch instead of behind work queued later.  One merge per context may be
 * started at a time.  Extends the same reference step (src/fft_process.cpp:469-479). */
int mdk_topn_start(mdk_ctx *ctx, int nrot, int n);
int mdk_topn_finish(mdk_ctx *ctx, mdk_sortscore *out);

/* As mdk_topn, allocating nothing: no hipMalloc, hipFree or allocating library call on
 * any call.  No counterpart in the reference; it is for callers that cannot allocate
 * device memory (code under stream capture, tight serving loops).  The merge state,
 * keys and results live in one fixed array of 1 MiB per GPU (its address is looked up
 * by mdk_create on each device and kept until the process ends), so min(n, nrot * t)
 * may be at most 32768; larger requests fail with an error and mdk_topn serves them.
 * Every call on a GPU shares that GPU's array: calls that overlap on one GPU (from
 * several host threads or contexts) race and must be serialised by the caller. */
int mdk_topn_noalloc(mdk_ctx *ctx, int nrot, int n, mdk_sortscore *out);

/* Time mdk_dock_range(first, count) `reps` times with HIP events on the context's
 * stream; *ms = mean milliseconds per repetition. */
int mdk_time_range(mdk_ctx *ctx, int first, int count, int reps, float *ms);

/* Measurement hook: dock rotations [first, first+count) (count <= max_batch) as one
 * batch with HIP events between the stages (voxelize, forward z, forward y,
 * x + correlation, inverse y, inverse z + top-t, per-rotation reduce); ms[q] gets each
 * stage's milliseconds and *nstages their number (7). No reference counterpart. */
int mdk_profile_stages(mdk_ctx *ctx, int first, int count, float *ms, int *nstages);

/* Measurement hook: mdk_dock_range(first, count) with HIP events recorded on each
 * batch's stream around each of its stages (as mdk_profile_stages), then a wait for the
 * whole range; stage_ms[q] gets the summed milliseconds of stage q over the range's full
 * batches (*nfull of them; a last, partial batch is left out) and *nstages their number
 * (7).  Results are those of mdk_dock_range.  No reference counterpart. */
int mdk_dock_range_timed(mdk_ctx *ctx, int first, int count, float *stage_ms, int *nstages, int *nfull);

/* Test hook: ligand voxel grids for B rotations (B*ng^3 floats each). */
int mdk_voxelize(mdk_ctx *ctx, const float *rotcoef, int B, float *grid_r, float *grid_i);

/* Test hook: the full real score map (unnormalised, nf^3 floats, x slowest) of one
 * rotation, i.e. Re IFFT(conj(R) * FFT(ligand)). */
int mdk_score_map(mdk_ctx *ctx, const float *rotcoef, float *out);

/* Autotuner: times the kernel variants built for this context's FFT size (correlation-pass
 * tiling and receptor reuse and fused-inverse loading together, then the forward pass's
 * prefetch or split) on rotations [first, first+count) of the table and keeps the fastest
 * for this context; choice[0..2] receives the chosen (correlation, inverse, forward)
 * variants.  The variants give bitwise-identical results, except at nf = 256, where the
 * two-pass forms (inverse variant 0, forward variant 1), the half-plane fused passes and
 * the 512-thread half-plane inverse (variant 1) round differently.  The sizes 80, 96 and 144 have every variant, 256 correlation
 * 0-7 (6 with the fused forward), inverse 0-3 and forward 0-1, the other specialised sizes
 * correlation 5-7, inverse 2-3 and forward 0 and 2.  No reference counterpart. */
int mdk_autotune(mdk_ctx *ctx, int first, int count, int *choice);

/* Off-grid score of B poses with its translation gradient, for refining decoys.  No
 * reference counterpart (extends the score of src/fft_process.cpp:333-446).  Pose b is
 * rotation rotcoef[9b..9b+8] (row-major, as mdk_dock) and a continuous translation
 * t[3b..3b+2] in voxels, the units of a result's index[1..3].  out[4b..4b+3] =
 * (s, ds/dtx, ds/dty, ds/dtz) where s is the band-limited interpolation of the hard-voxel
 * score map (value / nf^3, equal to the reported score at grid points).  Needs an FFT
 * size with specialised kernels: every one from 64 to 320 of the reference's size table
 * (see megadock_amd/csrc/mdk_pipe.hip).  Any B >= 0; runs in batches of the context's
 * max_batch. */
int mdk_score_spectral(mdk_ctx *ctx, const float *rotcoef, const float *t, int B, float *out);

/* Companion score of B poses (same pose arguments) with the ligand atoms spread
 * trilinearly instead of voxelized: s = sum_a [w_a Rr(P_a) + q_a Ri(P_a)] with Rr, Ri the
 * receptor grids of mdk_set_receptor interpolated at the placed atom P_a = R(a - c) -
 * gw t' (t' the signed translation), w_a = (4/3) pi r_a^3 / gw^3.  out[7b..7b+6] =
 * (s, ds/dt' (3, per voxel), ds/dw (3, per radian of a rotation exp([w]x) applied after
 * R)).  No reference counterpart: the differentiable companion used to refine poses.
 * Any B >= 0; runs in batches of the context's max_batch. */
int mdk_score_companion(mdk_ctx *ctx, const float *rotcoef, const float *t, int B, float *out);

/* Real-space sums of the reference's detailed output (Docking::output_detail,
 * src/docking.cpp:556-868) for B poses (any B; run in batches of the context's
 * max_batch): rotation coefficients rotcoef (B*9, as mdk_dock)
 * and grid translations trans (B*3, a decoy's index[1..3], within -ng .. 3ng-1).  The
 * ligand is voxelized on the device; rec_rpsc (ng^3) is the receptor shape grid built
 * without the ACE term (aceratio 0), the other receptor grids are mdk_set_receptor's.
 * out (B*9): rPSC(+), rPSC(-), ELEC(+), ELEC(-), ACE(+), ACE(-) (the decoys' sums,
 * src/docking.cpp:746-807), then rPSC, ELEC, ACE unsplit (the sums of the poses around the
 * input position, :646-654); each the float sum of the reference's float products in its
 * order (l, m, n ascending), so equal to the reference's bit for bit.  Replaces the voxel
 * loops of src/docking.cpp:596-657 and :710-801. */
int mdk_score_detail(mdk_ctx *ctx, const float *rec_rpsc, const float *rotcoef, const int *trans, int B, float *out);

/* Measurement hook: device-to-device copy of `bytes` (a multiple of 16; read + write) by
 * one dispatch of the kernel k_copy_probe, `reps` times after one untimed copy; *ms = mean
 * milliseconds per copy.  Calibrates HBM counters (a profiler sees each copy's bytes in
 * one dispatch) and the achievable copy bandwidth.  No reference counterpart. */
int mdk_copy_probe(int device, size_t bytes, int reps, float *ms);

/* Test hook: `count` independent n-point complex FFTs (interleaved complex64), sign
 * -1 forward / +1 inverse, unnormalised, on device `device`. */
int mdk_fft1d(int device, int n, int count, const float *in, float *out, int sign);

#ifdef __cplusplus
}
#endif
#endif
