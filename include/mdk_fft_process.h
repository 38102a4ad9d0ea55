/*
 * FFTProcess with the reference's public interface (src/fft_process.h:45-196), running
 * the rotation loop on MI355X through the libmdk C ABI (mdk_hip.h).
 *
 * The class keeps the reference's constructor, method names, argument meanings and
 * error behaviour (a message, then exit(1)), so that Docking::rec_init / dockz /
 * output (src/docking.cpp:128-453) drive it unchanged:
 *   alloc_array          one libmdk context per host thread myid2 (GPU myid2 % GPUs)
 *   receptor_fft         the receptor grids, embedded and transformed on every GPU
 *   ligand_data_transfer_gpu  the ligand atoms and grid coordinates to every GPU
 *   rotation_index + cuda_fft  one rotation of thread myid2; rotations are gathered into
 *                        device batches and docked when a batch is full (and by
 *                        sort_index), their t results landing in _Top at rotation*t
 *   ligand_preparation + fft3d(-1) + convolution + fft3d(+1) + score_sort
 *                        one rotation from host voxel grids (the reference's CPU branch),
 *                        docked on the GPU by mdk_dock_grid at score_sort
 *   sort_index           the partial sort of _Top by score (src/fft_process.cpp:469-479)
 * _Top only holds a rotation's results after its batch ran: read top_score / top_index
 * after sort_index, as the reference's callers do.
 *
 * Host types: by default this header declares minimal CPUTime, Parallel, Parameter,
 * Protein, Receptor and Ligand classes holding the members FFTProcess reads, under the
 * reference's names.  Defining MDK_REFERENCE_TYPES before including it uses the
 * reference's own headers instead (cpu_time.h, parallel.h, parameter.h, receptor.h,
 * ligand.h, whose classes name FFTProcess a friend), so this file replaces
 * src/fft_process.h in the reference's source tree.
 */
#ifndef MDK_FFT_PROCESS_H
#define MDK_FFT_PROCESS_H

#include <cstddef>
#include <string>
#include <vector>

#include "mdk_hip.h"

#ifdef MDK_REFERENCE_TYPES
#include "cpu_time.h"
#include "ligand.h"
#include "parallel.h"
#include "parameter.h"
#include "receptor.h"
#else
// Timers FFTProcess adds to (src/cpu_time.h:31-60).
class CPUTime {
public:
    float t3_1_ligand_voxelization = 0, t3_2_fftprocess_ligand_fft = 0, t3_3_fftprocess_convolution = 0;
    float t3_4_fftprocess_fft_inverse = 0, t3_5_fftprocess_score_sort = 0, t3_6_fftprocess_sort_index = 0;
    float t6_data_transfer_rec = 0, t6_data_transfer_lig = 0, t6_data_transfer_in_loop = 0;
    long long m1_current_malloc_size = 0, m2_maximum_malloc_size = 0;
    void record_malloc(const int &size) {
        m1_current_malloc_size += (long long)size;
        if (m1_current_malloc_size > m2_maximum_malloc_size) m2_maximum_malloc_size = m1_current_malloc_size;
    }
    void record_free(const int &size) { m1_current_malloc_size -= (long long)size; }
};

// Host threads and GPUs of the run (src/parallel.h:25-55).
class Parallel {
public:
    explicit Parallel(const size_t &nproc2) : _Nproc2(nproc2) {}
    void nproc2(int i) { _Nproc2 = (size_t)i; }
    size_t nproc2() { return _Nproc2; }
    void num_gpu(int i) { _Num_gpu = i; }
    int num_gpu() { return _Num_gpu; }

private:
    size_t _Nproc2;
    int _Num_gpu = 0;
};

// The parameters FFTProcess reads (src/parameter.h:55-90), with the reference's defaults
// (src/parameter.cpp:266-308).
class Parameter {
public:
    int _Num_sort = 1;                     // -t
    int _Num_output = 2000;                // -N
    int _Num_rot_angles = 0;               // rotations of the set
    float grid_width = 1.2f;               // voxel pitch (A)
    float _rPSC_param_lig_core = 1.0f;     // -b
    float *_Zangle = nullptr;              // 3 * _Num_rot_angles ZXZ angles
};

// Atom data FFTProcess reads (src/pdb_entry.h:31-35, src/protein.h:30-38).
class Protein {
public:
    virtual ~Protein() {}
    int num_atoms() { return _Num_atoms; }
    int _Num_atoms = 0;
    float *_Coordinate = nullptr;  // 3 * _Num_atoms
    float *_Radius = nullptr, *_Charge = nullptr;
    float _Center[3] = {0, 0, 0};
};
class Receptor : public Protein {};
class Ligand : public Protein {};
#endif

typedef struct {  // src/fft_process.h:40-43
    float score;
    int index[4];
} SortScore;

class FFTProcess {
private:
    FFTProcess(FFTProcess &c);
    const FFTProcess &operator=(const FFTProcess &c);
    struct Thread;  // the libmdk context and pending rotations of one host thread myid2

    CPUTime *_cputime;
    Parallel *_parallel;
    Parameter *_parameter;
    Receptor *_receptor;
    Ligand *_ligand;
    std::vector<SortScore> _Top;
    int _Num_fft = 0;
    int *_Current_rot_angle_num = nullptr;
    std::vector<Thread *> _threads;
    std::vector<float> _rec_r, _rec_i;  // the receptor grids (ng^3), for contexts made later
    std::vector<float> _grid_coord;     // the ligand's grid coordinates (ng)
    bool _have_lig = false;
    Thread *thread(size_t myid2);
    void flush(size_t myid2);

protected:
    virtual void alloc_fft();

public:
    virtual void fft3d(const float &theta, size_t myid2);
    FFTProcess(CPUTime *pcputime, Parallel *pparallel, Parameter *pparameter, Receptor *preceptor, Ligand *pligand)
        : _cputime(pcputime), _parallel(pparallel), _parameter(pparameter), _receptor(preceptor), _ligand(pligand) {}
    virtual ~FFTProcess();
    virtual void alloc_array(const int &num_fft);
    virtual void receptor_fft(float *grid_r, float *grid_i);
    virtual void cuda_fft(float *grid_r, float *grid_i, float *grid_coord, float *atom_coord_rotated, float *theta,
                          size_t myid2);
    virtual void ligand_voxelization_on_gpu(float *theta, size_t myid2);
    virtual void ligand_data_transfer_gpu(float *grid_coord);
    virtual void ligand_preparation(float *grid_r, float *grid_i, size_t myid2);
    virtual void convolution(size_t myid2);
    virtual void score_sort(size_t myid2);
    virtual void fft_memory_free();
    virtual void top_score_clean();
    virtual int num_fft() { return _Num_fft; }
    virtual void num_fft(const int &i) { _Num_fft = i; }
    virtual void rotation_index(int i, const int &j) { _Current_rot_angle_num[i] = j; }
    virtual float top_score(const int &j) { return _Top[j].score; }
    virtual int top_index(const int &j, int k) { return _Top[j].index[k]; }
    virtual void sort_index(float *fwork, int *iwork);
};

#endif
