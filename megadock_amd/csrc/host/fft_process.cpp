// FFTProcess with the reference's public interface (src/fft_process.h:45-196) on the
// libmdk C ABI.  Built as libmdk_fftprocess.so against the minimal host types of
// mdk_fft_process.h; with -DMDK_REFERENCE_TYPES the same file compiles against the
// reference's own classes and replaces src/fft_process.cpp in its source tree.
//
// Per host thread myid2 (the reference's OpenMP thread, src/docking.cpp:238-307) the
// class keeps one libmdk context: on GPU myid2 for the GPU threads (myid2 < num_gpu, as
// src/fft_process.cpp:527 binds them), on GPU myid2 % GPUs for the others.  A GPU thread's
// rotations (cuda_fft) are gathered into batches of up to kBatch and docked together; the
// other threads' rotations arrive as host voxel grids (ligand_preparation) and are docked
// one by one at score_sort.  Errors end the process with a message, as the reference's do.
#include "../../../include/mdk_fft_process.h"

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "../../../include/mdk_hip.h"

namespace {

constexpr int kBatch = 128;  // rotations per device batch of a GPU thread

[[noreturn]] void fatal(const char *what) {
    std::fprintf(stderr, "[ERROR] %s: %s\n", what, mdk_last_error());
    std::exit(1);
}

bool by_score(const SortScore &s1, const SortScore &s2) { return s1.score > s2.score; }  // src/fft_process.cpp:462-466

}  // namespace

struct FFTProcess::Thread {
    mdk_ctx *ctx = nullptr;
    int batch = 1;
    bool gpu_thread = false;
    std::vector<float> rot;        // pending rotations: 9 coefficients each
    std::vector<int> ids;          // and their rotation numbers
    std::vector<mdk_sortscore> out;
    std::vector<float> lg_r, lg_i; // ligand grids of ligand_preparation (ng^3 each)
    bool prepared = false;
};

FFTProcess::~FFTProcess() {
    fft_memory_free();
    delete[] _Current_rot_angle_num;
}

// src/fft_process.cpp:29-76
void FFTProcess::alloc_array(const int &num_fft) {
    _Num_fft = num_fft;
    const int num_sort = _parameter->_Num_sort, no = _parameter->_Num_output;
    const size_t nproc2 = _parallel->nproc2();
    int num_toprank = _parameter->_Num_rot_angles * num_sort;
    if (no > num_toprank) num_toprank = no;
    alloc_fft();
    _Top.assign(num_toprank, SortScore{0.0f, {0, 0, 0, 0}});
    delete[] _Current_rot_angle_num;
    _Current_rot_angle_num = new int[nproc2]();
    _cputime->record_malloc((int)(sizeof(SortScore) * (size_t)num_toprank));
}

// src/fft_process.cpp:79-185: the device state of every GPU thread now; that of the other
// threads when they first dock
void FFTProcess::alloc_fft() {
    fft_memory_free();
    const size_t nproc2 = _parallel->nproc2();
    _threads.assign(nproc2, nullptr);
    const int num_gpu = _parallel->num_gpu();
    for (size_t t = 0; t < nproc2 && (int)t < num_gpu; t++) thread(t);
}

FFTProcess::Thread *FFTProcess::thread(size_t myid2) {
    if (myid2 >= _threads.size()) {
        std::fprintf(stderr, "[ERROR] FFTProcess: thread %zu outside the %zu threads of alloc_array\n", myid2,
                     _threads.size());
        std::exit(1);
    }
    if (_threads[myid2]) return _threads[myid2];
    int ndev = 0;
    if (mdk_device_count(&ndev) || ndev < 1) fatal("FFTProcess: no GPU visible (this build runs the loop on the GPU)");
    const int num_gpu = _parallel->num_gpu();
    Thread *T = new Thread();
    T->gpu_thread = (int)myid2 < num_gpu;
    T->batch = T->gpu_thread ? std::max(1, std::min(kBatch, _parameter->_Num_rot_angles)) : 1;
    const int device = (int)(myid2 % (size_t)ndev);
    if (mdk_create(device, _Num_fft, _Num_fft / 2, _ligand->num_atoms(), _parameter->_Num_sort, T->batch, &T->ctx))
        fatal("mdk_create");
    T->out.resize((size_t)T->batch * _parameter->_Num_sort);
    if (!_rec_r.empty() && mdk_set_receptor(T->ctx, _rec_r.data(), _rec_i.data())) fatal("mdk_set_receptor");
    if (_have_lig &&
        mdk_set_ligand(T->ctx, _ligand->_Coordinate, _ligand->_Center, _ligand->_Radius, _ligand->_Charge,
                       _grid_coord.data(), _parameter->grid_width, _parameter->_rPSC_param_lig_core))
        fatal("mdk_set_ligand");
    _threads[myid2] = T;
    return T;
}

// src/fft_process.cpp:212-303: embedded and transformed on the device of every context
void FFTProcess::receptor_fft(float *grid_r, float *grid_i) {
    const size_t ng3 = (size_t)(_Num_fft / 2) * (_Num_fft / 2) * (_Num_fft / 2);
    _rec_r.assign(grid_r, grid_r + ng3);
    _rec_i.assign(grid_i, grid_i + ng3);
    for (Thread *T : _threads)
        if (T && mdk_set_receptor(T->ctx, grid_r, grid_i)) fatal("mdk_set_receptor");
}

// src/fft_process.cpp:769-807
void FFTProcess::ligand_data_transfer_gpu(float *grid_coord) {
    _grid_coord.assign(grid_coord, grid_coord + _Num_fft / 2);
    _have_lig = true;
    for (Thread *T : _threads)
        if (T && mdk_set_ligand(T->ctx, _ligand->_Coordinate, _ligand->_Center, _ligand->_Radius, _ligand->_Charge,
                                grid_coord, _parameter->grid_width, _parameter->_rPSC_param_lig_core))
            fatal("mdk_set_ligand");
}

// src/fft_process.cpp:502-656: rotation theta of thread myid2, number
// _Current_rot_angle_num[myid2]; the work buffers of the reference's CPU voxelizer
// (grid_r, grid_i, grid_coord, atom_coord_rotated) are not needed, the device voxelizes
void FFTProcess::cuda_fft(float *, float *, float *, float *, float *theta, size_t myid2) {
    Thread *T = thread(myid2);
    float r[9];
    if (mdk_rotation_coef(theta, 1, r)) fatal("mdk_rotation_coef");
    T->rot.insert(T->rot.end(), r, r + 9);
    T->ids.push_back(_Current_rot_angle_num[myid2]);
    if ((int)T->ids.size() >= T->batch) flush(myid2);
}

// dock the thread's pending rotations; their results go to _Top[rotation * t + rank]
void FFTProcess::flush(size_t myid2) {
    Thread *T = _threads[myid2];
    const int B = (int)T->ids.size(), t = _parameter->_Num_sort;
    if (!T || B == 0) return;
    if (mdk_dock(T->ctx, T->rot.data(), T->ids.data(), B, T->out.data())) fatal("mdk_dock");
    for (int b = 0; b < B; b++)
        std::memcpy(&_Top[(size_t)T->ids[b] * t], &T->out[(size_t)b * t], sizeof(SortScore) * t);
    T->rot.clear();
    T->ids.clear();
}

// src/fft_process.cpp:660-765 runs inside cuda_fft here: the device voxelizes each
// rotation of a batch as the batch's first step, so a separate call has nothing to do
void FFTProcess::ligand_voxelization_on_gpu(float *, size_t) {}

// src/fft_process.cpp:306-330: the rotation's host voxel grids (shape, charge; ng^3 each)
void FFTProcess::ligand_preparation(float *grid_r, float *grid_i, size_t myid2) {
    Thread *T = thread(myid2);
    const size_t ng3 = (size_t)(_Num_fft / 2) * (_Num_fft / 2) * (_Num_fft / 2);
    T->lg_r.assign(grid_r, grid_r + ng3);
    T->lg_i.assign(grid_i, grid_i + ng3);
    T->prepared = true;
}

// src/fft_process.cpp:349-396 and :333-346: the forward transform, the product with the
// receptor spectrum and the inverse transform of the prepared grids run on the device as
// one call at score_sort, which needs their result
void FFTProcess::fft3d(const float &, size_t myid2) {
    if (!thread(myid2)->prepared) {
        std::fprintf(stderr, "[ERROR] FFTProcess::fft3d: no ligand grids prepared by thread %zu\n", myid2);
        std::exit(1);
    }
}

void FFTProcess::convolution(size_t myid2) { fft3d(1.0f, myid2); }

// src/fft_process.cpp:399-458: the rotation's t results into _Top
void FFTProcess::score_sort(size_t myid2) {
    Thread *T = thread(myid2);
    if (!T->prepared) {
        std::fprintf(stderr, "[ERROR] FFTProcess::score_sort: no ligand grids prepared by thread %zu\n", myid2);
        std::exit(1);
    }
    const int ang = _Current_rot_angle_num[myid2], t = _parameter->_Num_sort;
    if (mdk_dock_grid(T->ctx, T->lg_r.data(), T->lg_i.data(), ang, T->out.data())) fatal("mdk_dock_grid");
    std::memcpy(&_Top[(size_t)ang * t], T->out.data(), sizeof(SortScore) * t);
    T->prepared = false;
}

// src/fft_process.cpp:814-875
void FFTProcess::fft_memory_free() {
    for (Thread *&T : _threads) {
        if (T && T->ctx) mdk_destroy(T->ctx);
        delete T;
        T = nullptr;
    }
    _threads.clear();
}

// src/fft_process.cpp:482-498
void FFTProcess::top_score_clean() {
    int num_toprank = _parameter->_Num_rot_angles * _parameter->_Num_sort;
    if (_parameter->_Num_output > num_toprank) num_toprank = _parameter->_Num_output;
    for (int j = 0; j < num_toprank && j < (int)_Top.size(); j++) _Top[j].score = 0.0f;
}

// src/fft_process.cpp:469-479, after the GPU threads' last, partial batches
void FFTProcess::sort_index(float *, int *) {
    for (size_t t = 0; t < _threads.size(); t++)
        if (_threads[t]) flush(t);
    const int no = _parameter->_Num_output;
    const int nt = ((int)_Top.size() < no) ? (int)_Top.size() : no;
    std::partial_sort(_Top.begin(), _Top.begin() + nt, _Top.end(), by_score);
}
