// Rotation-loop driver and decoy output of the drop-in (src/docking.cpp,
// src/fft_process.cpp, src/control.cpp:18-88, :468-529).  The loop itself runs on
// the GPU through the libmdk C ABI; this file prepares the receptor, hands the
// rotation table and ligand to the device, and writes the .out file.
#pragma clang fp contract(off)

#include <sys/time.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>

#include "mdk_host.h"

namespace mdk {

namespace {

double now() {
    struct timeval tv;
    gettimeofday(&tv, nullptr);
    return tv.tv_sec + 1e-6 * tv.tv_usec;
}

std::string abi_error(const char *what) { return std::string(what) + ": " + mdk_last_error(); }

const float PI_F = (float)(4.0 * std::atan(1.0));  // src/constant.h:24

}  // namespace

namespace {
// device bytes of one rotation's batch buffers (d_a1, d_a2, d_cc, d_cand; for t > 16 the
// score map and its select keys)
double batch_bytes_per_rotation(int nf, int ng, int t) {
    const double H = nf / 2 + 1;
    double per_rot = 8.0 * ((double)ng * ng * nf + (double)ng * nf * nf + (double)nf * nf * (H + 16) +
                            (double)nf * nf * std::min(t, 16));
    if (t > 16) {  // the score map of each rotation and its select keys (mdk_create)
        double P = 1;
        while (P < t) P *= 2;
        per_rot += 4.0 * nf * nf * nf + 8.0 * P;
    }
    return per_rot;
}
}  // namespace

double context_bytes(int nf, int ng, int t, int batch, int na) {
    // two batch slots (with their ligand grids, charge grids and atom indices), the receptor
    // spectrum and grids, and room for the rest
    const double ng3 = (double)ng * ng * ng;
    return 2.0 * batch * (batch_bytes_per_rotation(nf, ng, t) + 12.0 * ng3 + 4.0 * na) + 8.0 * nf * nf * nf +
           12.0 * ng3 + 64.0 * 1024 * 1024;
}

int auto_batch(int nf, int ng, int t) {
    const double per_rot = batch_bytes_per_rotation(nf, ng, t);
    // 8 GiB of batch buffers, at most 256 rotations a batch up to nf = 128 (184k against 175k
    // rotations/s at 96^3) and 128 up to nf = 192 (144^3: 61.8k for either, steadier at 128);
    // above, 36 GiB and at most 192 (256^3: 6.85k at 192 against 6.56k at 48 rotations)
    const double budget = (nf > 192 ? 36.0 : 8.0) * 1024 * 1024 * 1024;
    const int b = (int)(budget / per_rot);
    return std::max(1, std::min(nf <= 128 ? 256 : nf <= 192 ? 128 : 192, b));
}

// ------------------------------------------------------------------ FFTProcess
FFTProcess::~FFTProcess() {
    if (ctx_) mdk_destroy(ctx_);
}

int FFTProcess::alloc_array(int num_fft, int device, std::string &err) {
    num_fft_ = num_fft;
    const int ng = num_fft / 2;
    int batch = prm_->batch > 0 ? prm_->batch : auto_batch(num_fft, ng, prm_->num_sort);
    batch = std::min(batch, std::max(1, prm_->num_rot_angles));
    if (mdk_create(device, num_fft, ng, lig_->num_atoms(), prm_->num_sort, batch, &ctx_)) {
        err = abi_error("mdk_create");
        return 1;
    }
    const int toprank = std::max(prm_->num_rot_angles * prm_->num_sort, prm_->num_output);
    top_.assign(toprank, SortScore{0.0f, {0, 0, 0, 0}});
    return 0;
}

int FFTProcess::receptor_fft(const float *grid_r, const float *grid_i, std::string &err) {
    if (mdk_set_receptor(ctx_, grid_r, grid_i)) {
        err = abi_error("mdk_set_receptor");
        return 1;
    }
    return 0;
}

int FFTProcess::ligand_data_transfer_gpu(const std::vector<float> &gc, std::string &err) {
    std::vector<float> rot(9 * (size_t)prm_->num_rot_angles);
    if (mdk_rotation_coef(prm_->zangle.data(), prm_->num_rot_angles, rot.data()) ||
        mdk_set_rotations(ctx_, rot.data(), prm_->num_rot_angles)) {
        err = abi_error("rotation table");
        return 1;
    }
    if (mdk_set_ligand(ctx_, lig_->coords().data(), lig_->center(), lig_->radius.data(), lig_->charge.data(),
                       gc.data(), prm_->grid_width, prm_->lig_core)) {
        err = abi_error("mdk_set_ligand");
        return 1;
    }
    return 0;
}

int FFTProcess::dock_rotations(int first, int count, std::string &err) {
    if (mdk_dock_range(ctx_, first, count)) {
        err = abi_error("mdk_dock_range");
        return 1;
    }
    return 0;
}

void FFTProcess::top_score_clean() {
    for (auto &s : top_) s.score = 0.0f;
}

int FFTProcess::sort_index(std::string &err) {
    const int no = std::min((int)top_.size(), prm_->num_output);
    std::vector<mdk_sortscore> best(no);
    if (mdk_topn(ctx_, prm_->num_rot_angles, no, best.data())) {
        err = abi_error("mdk_topn");
        return 1;
    }
    std::memcpy(top_.data(), best.data(), sizeof(SortScore) * no);
    return 0;
}

// ------------------------------------------------------------------ Docking
int Docking::initialize(int device, std::string &err) {
    num_grid_ = std::max(rec_->num_grid(), lig_->num_grid());
    prm_->num_grid = num_grid_;
    prm_->num_fft = 2 * num_grid_;
    fft_.reset(new FFTProcess(prm_, rec_, lig_));
    return fft_->alloc_array(2 * num_grid_, device, err);
}

// The receptor grids are built on the device (mdk_set_receptor_atoms, bit-exact with
// Receptor::voxelize) and transformed there.
int Docking::rec_init(std::string &err) {
    const double t0 = now();
    const int ng = num_grid_;
    grid_coord_ = grid_coord(ng, prm_->grid_width);
    std::vector<float> xyz(3 * (size_t)rec_->num_atoms());
    rec_->shift_center(xyz.data());
    if (mdk_set_receptor_atoms(fft_->ctx(), rec_->num_atoms(), xyz.data(), rec_->radius.data(), rec_->charge.data(),
                               rec_->ace.data(), grid_coord_.data(), prm_->grid_width, prm_->ace_ratio, prm_->rec_core,
                               prm_->elec_ratio)) {
        err = abi_error("mdk_set_receptor_atoms");
        return 1;
    }
    t_receptor = now() - t0;
    return 0;
}

int Docking::dockz(std::string &err) {
    const double t0 = now();
    fft_->top_score_clean();
    std::printf("\nLigand = %s\nTarget receptors:\n %s\n\n", prm_->lig_file.c_str(), prm_->rec_file.c_str());
    if (fft_->ligand_data_transfer_gpu(grid_coord_, err)) return 1;
    const int R = prm_->num_rot_angles;
    const int chunk = prm_->progress ? std::max(R / 10, 1) : R;
    for (int a = 0; a < R; a += chunk) {
        const int n = std::min(chunk, R - a);
        if (fft_->dock_rotations(a, n, err)) return 1;
        if (prm_->progress) {
            if (mdk_sync(fft_->ctx())) {
                err = abi_error("mdk_sync");
                return 1;
            }
            std::printf("   >Ligand rotation = %5d / %5d\n", a + n, R);
        }
    }
    if (fft_->sort_index(err)) return 1;
    t_docking = now() - t0;
    return 0;
}

int write_decoys(const Parameter &prm, const Protein &rec, const Protein &lig, int num_grid, const SortScore *top,
                 std::string &err) {
    FILE *f = std::fopen(prm.out_file.c_str(), "w");
    if (!f) {
        err = "[ERROR] Output file [" + prm.out_file + "] not open!!";
        return 1;
    }
    std::fprintf(f, "%d\t%.2f\n", 2 * num_grid, prm.grid_width);  // src/docking.cpp:413-414
    std::fprintf(f, "\t%f\t%f\t%f\n", 0.0, 0.0, 0.0);             // receptor angle
    std::fprintf(f, "%s\t%f\t%f\t%f\n", prm.rec_file.c_str(), rec.center(0), rec.center(1), rec.center(2));
    std::fprintf(f, "%s\t%f\t%f\t%f\n", prm.lig_file.c_str(), lig.center(0), lig.center(1), lig.center(2));
    for (int rank = 0; rank < prm.num_output; rank++) {  // src/docking.cpp:434-448
        const int n = top[rank].index[0];
        float th[3];
        for (int k = 0; k < 3; k++) {
            th[k] = prm.zangle[3 * n + k];
            if (th[k] >= PI_F) th[k] = (float)(th[k] - PI_F * 2.0);
        }
        std::fprintf(f, "%f\t%f\t%f\t%d\t%d\t%d\t%.2f\n", th[0], th[1], th[2], top[rank].index[1],
                     top[rank].index[2], top[rank].index[3], top[rank].score);
    }
    std::fclose(f);
    return 0;
}

int Docking::output(std::string &err) {
    return write_decoys(*prm_, *rec_, *lig_, num_grid_, fft_->top(), err);
}

namespace {

// CA / backbone classes of src/protein.cpp:50-60: 1 for "CA ", 2 for "C " and "N ", else 0
int atom_class(const std::string &key) {
    if (key.compare(0, 3, "CA ") == 0) return 1;
    if (key.compare(0, 2, "C ") == 0 || key.compare(0, 2, "N ") == 0) return 2;
    return 0;
}

struct DetailRow {
    float total, rpsc, rpsc_plus, rpsc_minus, elec, ace, rmsd_lca;
    int near_native;
};

}  // namespace

int write_detail(const Parameter &prm, const Receptor &rec, const Ligand &lig, int num_grid, const SortScore *top,
                 mdk_ctx *ctx, std::string &err) {
    const int ng1 = num_grid, nf1 = 2 * num_grid;
    const size_t ng3 = (size_t)ng1 * ng1 * ng1;
    const float gw = prm.grid_width;
    const int natom = lig.num_atoms(), nout = prm.num_output;
    // receptor shape grid without the ACE term (src/docking.cpp:568-571)
    std::vector<float> rpsc(ng3), scratch(ng3);
    {
        std::vector<float> xyz(3 * (size_t)rec.num_atoms());
        rec.shift_center(xyz.data());
        rec.voxelize(ng1, gw, grid_coord(ng1, gw), xyz.data(), 0.0f, prm.rec_core, prm.elec_ratio, rpsc.data(),
                     scratch.data());
    }
    float coordr[3], coordl[3], diff_dis[3];
    int cor[3];
    for (int i = 0; i < 3; i++) {  // src/docking.cpp:541-546
        coordr[i] = rec.center(i);
        coordl[i] = lig.center(i);
        diff_dis[i] = coordr[i] - coordl[i];
        cor[i] = (int)std::round(diff_dis[i] / gw);
        if (cor[i] < 0) cor[i] += nf1;
    }
    // poses: the 7 around the input position (ligand unrotated, src/docking.cpp:596-615),
    // then the decoys
    const int np = 7 + nout;
    std::vector<int> trans(3 * (size_t)np);
    std::vector<float> theta(3 * (size_t)np, 0.0f), rot(9 * (size_t)np);
    int c[3] = {cor[0], cor[1], cor[2]};
    for (int rank = 0; rank < 7; rank++) {
        if (rank == 1) c[0]--;
        if (rank == 2) c[0] += 2;
        if (rank == 3) { c[0]--; c[1]--; }
        if (rank == 4) c[1] += 2;
        if (rank == 5) { c[1]--; c[2]--; }
        if (rank == 6) c[2] += 2;
        for (int k = 0; k < 3; k++) trans[3 * rank + k] = c[k];
    }
    for (int rank = 0; rank < nout; rank++) {
        const int a = top[rank].index[0];
        for (int k = 0; k < 3; k++) {
            theta[3 * (7 + rank) + k] = prm.zangle[3 * a + k];
            trans[3 * (7 + rank) + k] = top[rank].index[k + 1];
        }
    }
    if (mdk_rotation_coef(theta.data(), np, rot.data())) {
        err = std::string("mdk_rotation_coef: ") + mdk_last_error();
        return 1;
    }
    std::vector<float> sums(9 * (size_t)np);  // the reference's float sums, in its order
    if (mdk_score_detail(ctx, rpsc.data(), rot.data(), trans.data(), np, sums.data())) {
        err = std::string("mdk_score_detail: ") + mdk_last_error();
        return 1;
    }
    // RMSD of the translated (not rotated) ligand, as src/docking.cpp:663-693 and :829-851
    std::vector<int> cls(natom);
    int natom_lca = 0, natom_lbb = 0;
    for (int i = 0; i < natom; i++) {
        cls[i] = atom_class(lig.atom_type(i));
        if (cls[i] == 1) natom_lca++;
        if (cls[i] > 0) natom_lbb++;
    }
    std::vector<DetailRow> row(np);
    for (int p = 0; p < np; p++) {
        const float *d = &sums[9 * (size_t)p];
        DetailRow &r = row[p];
        r.rpsc_plus = d[0];
        r.rpsc_minus = d[1];
        // the poses around the input position print the unsplit sums (src/docking.cpp:646-655),
        // the decoys the sums of their split parts (:803-807)
        const bool around = p < 7;
        r.rpsc = around ? d[6] : r.rpsc_plus + r.rpsc_minus;
        const float elec = around ? d[7] : d[2] + d[3], ace = around ? d[8] : d[4] + d[5];
        r.elec = elec;
        r.ace = ace;
        r.total = r.rpsc + r.elec + r.ace;
        int tem[3];
        for (int k = 0; k < 3; k++) tem[k] = (3 * ng1 - trans[3 * p + k]) % (2 * ng1);
        float lca = 0.0f, lbb = 0.0f;
        for (int i = 0; i < natom; i++)
            for (int k = 0; k < 3; k++)
                if (cls[i] > 0) {
                    const float orig = lig.coordinate(i, k);
                    const float moved = orig + gw * (tem[k] - ng1) + diff_dis[k];
                    if (cls[i] == 1) lca += std::pow(moved - orig, 2.0);
                    lbb += std::pow(moved - orig, 2.0);
                }
        r.rmsd_lca = std::sqrt(lca / natom_lca);
        r.near_native = r.rmsd_lca < 5.0f ? 1 : 0;
    }
    FILE *fd = std::fopen(prm.detail_file.c_str(), "w");
    if (!fd) {
        err = "[ERROR] Output_detail file [" + prm.detail_file + "] not open!!";
        return 1;
    }
    FILE *fc = std::fopen(prm.csv_file.c_str(), "w");
    if (!fc) {
        std::fclose(fd);
        err = "[ERROR] Output_csv file [" + prm.csv_file + "] not open!!";
        return 1;
    }
    const char *columns = "Rank   Angle               Trans         total       rPSC     rPSC(+)  rPSC(-)  ELEC       RecACE     LCARMSD NearNat";
    std::fprintf(fd, "%d\t%.2f\n", nf1, gw);                 // src/docking.cpp:513-535
    std::fprintf(fd, "\t%f\t%f\t%f\n", 0.0, 0.0, 0.0);       // receptor angle
    std::fprintf(fd, "%s\t%f\t%f\t%f\n", prm.rec_file.c_str(), coordr[0], coordr[1], coordr[2]);
    std::fprintf(fd, "%s\t%f\t%f\t%f\n", prm.lig_file.c_str(), coordl[0], coordl[1], coordl[2]);
    std::fprintf(fd, "\nElec score ratio: %6.2f\nACE  score ratio: %6.2f\n", prm.elec_ratio, prm.ace_ratio);
    std::fprintf(fd, "\ncorrect position: angle = (%5.2f,%5.2f,%5.2f), trans = (%3d,%3d,%3d) \n\n", 0.0, 0.0, 0.0,
                 cor[0], cor[1], cor[2]);
    std::fprintf(fd, " *** score around correct position *** \n\n%s\n", columns);
    auto line = [&](int p, int rank, const float *th) {
        const DetailRow &r = row[p];
        std::fprintf(fd, "%5d: (%5.2f,%5.2f,%5.2f) (%3d,%3d,%3d) %11.2f %8.0f %8.0f %8.0f %10.2f %10.2f %7.3f %7d\n",
                     rank, th[0], th[1], th[2], trans[3 * p], trans[3 * p + 1], trans[3 * p + 2], r.total, r.rpsc,
                     r.rpsc_plus, r.rpsc_minus, r.elec, r.ace, r.rmsd_lca, r.near_native);
    };
    const float zero[3] = {0.0f, 0.0f, 0.0f};
    for (int rank = 0; rank < 7; rank++) line(rank, rank + 1, zero);
    std::fprintf(fd, "\n *** docking results *** \n\n%s\n", columns);
    std::fprintf(fc, "Rank,Anglex,Angley,Anglez,Transi,Transj,Transk,total,rPSC,rPSC_gain,rPSC_penalty,ELEC ,RecACE,LCARMSD,NearNat\n");
    for (int rank = 0; rank < nout; rank++) {
        const int p = 7 + rank;
        float th[3];
        for (int k = 0; k < 3; k++) {  // src/docking.cpp:819-823
            th[k] = theta[3 * p + k];
            if (th[k] >= PI_F) th[k] -= PI_F * 2.0;
        }
        line(p, rank + 1, th);
        const DetailRow &r = row[p];
        std::fprintf(fc, "%d, %f,%f,%f,%d,%d,%d,%f,%f,%f,%f,%f,%f,%f,%d\n", rank + 1, th[0], th[1], th[2], trans[3 * p],
                     trans[3 * p + 1], trans[3 * p + 2], r.total, r.rpsc, r.rpsc_plus, r.rpsc_minus, r.elec, r.ace,
                     r.rmsd_lca, r.near_native);
    }
    std::fclose(fd);
    std::fclose(fc);
    return 0;
}

// ------------------------------------------------------------------ run
int run(int argc, char **argv, std::string &err) {
    const double t0 = now();
    Parameter prm;
    const int rc = prm.initialize(argc, argv, err);
    if (rc) return rc;
    int ndev = 0;
    if (mdk_device_count(&ndev) || ndev < 1) {
        err = "no GPU visible: this build runs the rotation loop on the GPU";
        return 1;
    }
    if (prm.gpu_limit == 0) {
        err = "-G 0 asks for the CPU loop, which this build does not provide";
        return 1;
    }
    std::printf("# Using %3d GPU\n", 1);
    Receptor rec(prm.rec_file);
    Ligand lig(prm.lig_file);
    if (rec.initialize(prm, err) || lig.initialize(prm, err)) return 1;
    if (!prm.num_fft_flag) {
        const int ngr = autogrid_receptor(rec, prm), ngl = autogrid_ligand(lig, prm);
        rec.num_grid(ngr);
        lig.num_grid(ngl);
        std::printf("\nReceptor = %s\nNumber of grid = %d\nFFT N = %d\n", prm.rec_file.c_str(), ngr, 2 * ngr);
        std::printf("\nLigand = %s\nNumber of grid = %d\nFFT N = %d\n", prm.lig_file.c_str(), ngl, 2 * ngl);
    } else {
        if (check_grid(rec, lig, prm, err)) return 1;
        rec.num_grid(prm.num_grid);
        lig.num_grid(prm.num_grid);
    }
    Docking dock(&prm, &rec, &lig);
    if (dock.initialize(0, err)) return 1;
    const double t_init = now() - t0;
    std::printf("\n---------- Start docking calculations\n");
    if (dock.rec_init(err) || dock.dockz(err) || dock.output(err)) return 1;
    if (prm.detail_output_flag &&  // src/control.cpp:486-490
        write_detail(prm, rec, lig, dock.num_grid(), dock.fft_process()->top(), dock.fft_process()->ctx(), err))
        return 1;
    std::printf("\nElapsed time  = %8.2f sec\n", now() - t0);
    std::printf(" Initialization = %8.2f sec\n Receptor process = %8.2f sec\n Docking = %8.2f sec (%d rotations, %.1f rotations/s)\n",
                t_init, dock.t_receptor, dock.t_docking, prm.num_rot_angles, prm.num_rot_angles / dock.t_docking);
    return 0;
}

}  // namespace mdk
