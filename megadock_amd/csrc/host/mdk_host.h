// Host side of the drop-in: the reference's input layer, receptor producer, grid
// sizing, rotation loop driver and decoy output, in C++, calling the HIP path through
// the libmdk C ABI (include/mdk_hip.h).
//
// Classes keep the reference's names and roles:
//   Parameter  -- flags, defaults, CHARMM/ACE tables, rotation sets (src/parameter.cpp)
//   PDBEntry   -- PDB reader (src/pdb_entry.cpp)
//   Protein    -- per-atom radius/charge/ACE, edges, centre (src/protein.cpp)
//   Receptor   -- rPSC + ACE shape grid and electrostatics (src/receptor.cpp)
//   Ligand     -- atom data for the device voxelizer (src/ligand.cpp runs on the GPU)
//   FFTProcess -- per-rotation results, top-t merge, device context (src/fft_process.h)
//   Docking    -- rec_init / dockz / output (src/docking.cpp)
//   Control    -- grid sizing and the run (src/control.cpp)
#pragma once

#include <map>
#include <memory>
#include <string>
#include <vector>

#include "../../../include/mdk_hip.h"

namespace mdk {

struct SortScore {  // src/fft_process.h:40-43
    float score;
    int index[4];
};
static_assert(sizeof(SortScore) == sizeof(mdk_sortscore), "SortScore layout");

// Where the data tables live (directory holding rotations_*.txt and the .tsv tables).
std::string data_dir();

class Parameter {
public:
    // defaults of src/parameter.cpp:266-308
    std::string rec_file, lig_file, out_file;
    std::string detail_file, csv_file;  // -O outputs (src/parameter.cpp:81-92, :357-364)
    int detail_output_flag = 0;
    int score_func = 3;
    float grid_width = 1.2f;
    int num_grid = 0, num_fft = 0, num_fft_flag = 0;
    int num_sort = 1;
    float elec_ratio = 1.0f, ace_ratio = 1.0f;
    int num_output = 2000;
    int thread_limit = 9999999, gpu_limit = 9999999;
    float rec_core = -45.0f, lig_core = 1.0f;
    int rotation_set = 3600;
    int fft_base_set = 11;
    int progress = 0, load_balance = 0;
    int batch = 0;  // rotations per device batch (0: automatic)
    float grid_space_rec = 8.0f, grid_space_lig = 3.0f;
    std::vector<float> zangle;  // 3 * R
    int num_rot_angles = 0;

    // Parses the reference's command line (src/parameter.cpp:21-236).  Returns 0, or a
    // nonzero status with a message in `err` (the reference prints usage and exits).
    int initialize(int argc, char **argv, std::string &err);
    // Loads the tables (src/parameter.cpp:376-422).
    int load_tables(std::string &err);
    float atom_radius(const std::string &key) const;
    float atom_charge(const std::string &key) const;
    float atom_ace(const std::string &key) const;
    static int load_rotation_set(int n, std::vector<float> &out, std::string &err);
    static std::string usage();

private:
    std::map<std::string, float> charmm_r_, charmm_c_, ace_;
};

class PDBEntry {
public:
    explicit PDBEntry(const std::string &file) : file_(file) {}
    int pdb_read(std::string &err);  // src/pdb_entry.cpp:18-75
    static std::string erase_space(const std::string &s);  // src/pdb_entry.cpp:78-105
    int num_atoms() const { return (int)keys_.size(); }
    float coordinate(int i, int j) const { return xyz_[3 * i + j]; }
    const std::string &atom_type(int i) const { return keys_[i]; }
    const std::string &input_file() const { return file_; }
    const std::vector<float> &coords() const { return xyz_; }

protected:
    std::string file_;
    std::vector<float> xyz_;
    std::vector<std::string> keys_;
};

class Protein : public PDBEntry {
public:
    explicit Protein(const std::string &file) : PDBEntry(file) {}
    int initialize(const Parameter &p, std::string &err);  // src/protein.cpp:19-64
    float edge(int i, int j) const { return edge_[i][j]; }
    float center(int i) const { return center_[i]; }
    const float *center() const { return center_; }
    void shift_center(float *out) const;  // src/protein.cpp:114-131
    int num_grid() const { return num_grid_; }
    void num_grid(int n) { num_grid_ = n; }
    std::vector<float> radius, charge, ace;

protected:
    float edge_[3][2] = {{0, 0}, {0, 0}, {0, 0}};
    float center_[3] = {0, 0, 0};
    int num_grid_ = 0;
};

class Receptor : public Protein {
public:
    explicit Receptor(const std::string &file) : Protein(file) {}
    // Shape (rPSC + ACE) and electrostatics grids, ng^3 each, from centred coordinates
    // (src/receptor.cpp:20-372).
    // The per-voxel sums run over x-planes on `threads` host threads (0: the machine's
    // cores, at most 16), each voxel's terms in atom order, so the grids do not depend on it.
    void voxelize(int ng, float gw, const std::vector<float> &grid_coord, const float *xyz,
                  float aceratio, float rec_core, float eratio, float *grid_r, float *grid_i,
                  int threads = 0) const;
};

class Ligand : public Protein {
public:
    explicit Ligand(const std::string &file) : Protein(file) {}
};

// Grid coordinates (src/docking.cpp:133-141)
std::vector<float> grid_coord(int ng, float gw);

// Auto grid sizing (src/control.cpp:91-379): returns ng or -1.
int autogrid_receptor(const Protein &r, const Parameter &p);
int autogrid_ligand(const Protein &l, const Parameter &p);
// Forced -F sizing checks (src/control.cpp:382-465): 0 ok, nonzero when too small.
int check_grid(const Protein &r, const Protein &l, const Parameter &p, std::string &err);

// Per-rotation results and the device context (the FFTProcess role).
class FFTProcess {
public:
    FFTProcess(Parameter *p, Receptor *r, Ligand *l) : prm_(p), rec_(r), lig_(l) {}
    ~FFTProcess();
    int alloc_array(int num_fft, int device, std::string &err);      // src/fft_process.cpp:29-76
    int receptor_fft(const float *grid_r, const float *grid_i, std::string &err);  // :212-303
    int ligand_data_transfer_gpu(const std::vector<float> &grid_coord, std::string &err);  // :769-807
    // Docks rotations [first, first+count) of the parameter's rotation set.
    int dock_rotations(int first, int count, std::string &err);
    int sort_index(std::string &err);  // :469-479, merged on the device
    void top_score_clean();            // :482-498
    const SortScore *top() const { return top_.data(); }
    float top_score(int j) const { return top_[j].score; }
    int top_index(int j, int k) const { return top_[j].index[k]; }
    int num_fft() const { return num_fft_; }
    mdk_ctx *ctx() { return ctx_; }

private:
    Parameter *prm_;
    Receptor *rec_;
    Ligand *lig_;
    mdk_ctx *ctx_ = nullptr;
    int num_fft_ = 0;
    std::vector<SortScore> top_;
};

class Docking {
public:
    Docking(Parameter *p, Receptor *r, Ligand *l) : prm_(p), rec_(r), lig_(l) {}
    int initialize(int device, std::string &err);  // src/docking.cpp:18-32
    int rec_init(std::string &err);                // src/docking.cpp:128-159
    int dockz(std::string &err);                   // src/docking.cpp:194-346
    int output(std::string &err);                  // src/docking.cpp:399-453
    FFTProcess *fft_process() { return fft_.get(); }
    int num_grid() const { return num_grid_; }
    double t_receptor = 0, t_docking = 0;

private:
    Parameter *prm_;
    Receptor *rec_;
    Ligand *lig_;
    std::unique_ptr<FFTProcess> fft_;
    int num_grid_ = 0;
    std::vector<float> grid_coord_;
};

// Whole run as the reference's main(): parse flags, read PDBs, size grids, dock, write
// the .out file (src/main.cpp:38-140, src/control.cpp:18-88, :468-529).
int run(int argc, char **argv, std::string &err);

// The .out decoy file (src/docking.cpp:399-453) from the merged top list (num_output entries).
int write_decoys(const Parameter &prm, const Protein &rec, const Protein &lig, int num_grid, const SortScore *top,
                 std::string &err);

// The detailed output (-O): real-space score parts of the 7 poses around the input
// position and of the num_output decoys, in the .detail and .csv files
// (src/docking.cpp:456-898); the sums run on the device through mdk_score_detail.
int write_detail(const Parameter &prm, const Receptor &rec, const Ligand &lig, int num_grid, const SortScore *top,
                 mdk_ctx *ctx, std::string &err);

// Rotations per device batch for an FFT size: 8 GiB of batch buffers (36 GiB above nf =
// 192), at most 256 up to nf = 128, 128 up to nf = 192 and 192 above.
int auto_batch(int nf, int ng, int t);

// Device bytes of a context of that size and batch (mdk_create's buffers), an estimate
// for keeping several contexts alive within a memory budget.
double context_bytes(int nf, int ng, int t, int batch, int na);

// The job-table farm (src/mpidp.cpp), in process: see farm.cpp.
int run_table(int argc, char **argv, std::string &err);

}  // namespace mdk
