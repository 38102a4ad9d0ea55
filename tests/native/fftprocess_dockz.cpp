// Test driver: FFTProcess (include/mdk_fft_process.h) driven the way the reference's
// Docking::rec_init and Docking::dockz drive theirs (src/docking.cpp:128-159, :194-346):
// alloc_array, receptor_fft, top_score_clean, ligand_data_transfer_gpu, then an OpenMP
// loop over the rotations with dynamic scheduling in which thread myid2 < num_gpu calls
// rotation_index + cuda_fft and the other threads rotation_index + ligand_preparation +
// fft3d(-1) + convolution + fft3d(+1) + score_sort on host voxel grids, then sort_index.
// Prints _Top after sort_index, all ROTATIONS * T entries, as "E j score i0 i1 i2 i3": the
// first N are the sorted top, the rest every other rotation result in no particular order.
//
//   fftprocess_dockz REC.pdb LIG.pdb ROTATIONS T N NPROC2 NUM_GPU
#include <omp.h>

#include <cstdio>
#include <cstdlib>
#include <vector>

#include "mdk_fft_process.h"
#include "mdk_hip.h"
#include "mdk_host.h"

static void check(int rc, const char *what) {
    if (rc) {
        std::fprintf(stderr, "%s failed: %s %s\n", what, mdk_last_error(), mdkh_last_error());
        std::exit(1);
    }
}

int main(int argc, char **argv) {
    if (argc != 8) {
        std::fprintf(stderr, "usage: %s REC.pdb LIG.pdb ROTATIONS T N NPROC2 NUM_GPU\n", argv[0]);
        return 2;
    }
    const int nrot_set = std::atoi(argv[3]), nproc2 = std::atoi(argv[6]), num_gpu = std::atoi(argv[7]);
    mdkh_protein *rec = nullptr, *lig = nullptr;
    check(mdkh_protein_load(argv[1], &rec), "mdkh_protein_load");
    check(mdkh_protein_load(argv[2], &lig), "mdkh_protein_load");
    int ng = 0, na_l = 0, nrot = 0;
    const float gw = 1.2f;
    check(mdkh_autogrid(rec, lig, gw, 11, &ng), "mdkh_autogrid");
    check(mdkh_protein_size(lig, &na_l), "mdkh_protein_size");
    const int nf = 2 * ng;
    std::vector<float> xyz(3 * (size_t)na_l), radius(na_l), charge(na_l);
    Ligand L;
    L._Num_atoms = na_l;
    check(mdkh_protein_atoms(lig, xyz.data(), radius.data(), charge.data(), nullptr, L._Center), "mdkh_protein_atoms");
    L._Coordinate = xyz.data();
    L._Radius = radius.data();
    L._Charge = charge.data();
    Receptor R;
    check(mdkh_rotation_set(nrot_set, nullptr, &nrot), "mdkh_rotation_set");
    std::vector<float> zangle(3 * (size_t)nrot);
    check(mdkh_rotation_set(nrot_set, zangle.data(), &nrot), "mdkh_rotation_set");
    Parameter P;
    P._Num_sort = std::atoi(argv[4]);
    P._Num_output = std::atoi(argv[5]);
    P._Num_rot_angles = nrot;
    P.grid_width = gw;
    P._Zangle = zangle.data();
    CPUTime T;
    Parallel par((size_t)nproc2);
    par.num_gpu(num_gpu);

    const size_t ng3 = (size_t)ng * ng * ng;
    std::vector<float> gc(ng), grid_r(ng3), grid_i(ng3);
    check(mdkh_grid_coord(ng, gw, gc.data()), "mdkh_grid_coord");
    check(mdkh_receptor_voxel(rec, ng, gw, 1.0f, -45.0f, 1.0f, grid_r.data(), grid_i.data()), "mdkh_receptor_voxel");

    // the host voxel grids of the threads without a GPU: the role of the reference's
    // Docking::ligand_rotationz + create_voxel (src/docking.cpp:162-191, :368-396), here
    // through the device voxelizer's test hook (bit-exact with Ligand::rpscace + electro)
    mdk_ctx *vox = nullptr;
    check(mdk_create(0, nf, ng, na_l, 1, 1, &vox), "mdk_create");
    check(mdk_set_ligand(vox, xyz.data(), L._Center, radius.data(), charge.data(), gc.data(), gw, 1.0f), "mdk_set_ligand");

    FFTProcess fft(&T, &par, &P, &R, &L);
    fft.alloc_array(nf);                       // Docking::maxsize_voxel (src/docking.cpp:72)
    fft.receptor_fft(grid_r.data(), grid_i.data());  // Docking::rec_init (:152)
    fft.top_score_clean();                     // Docking::dockz (:210)
    if (num_gpu > 0) fft.ligand_data_transfer_gpu(gc.data());  // (:233)
    std::vector<float> work(2 * ng3 * (size_t)nproc2);
#pragma omp parallel num_threads(nproc2)
    {
        const int myid2 = omp_get_thread_num();
#pragma omp for schedule(dynamic, 1)
        for (int ang = 0; ang < nrot; ang++) {
            fft.rotation_index(myid2, ang);
            float theta[3] = {zangle[3 * ang], zangle[3 * ang + 1], zangle[3 * ang + 2]};
            float *wr = &work[2 * ng3 * myid2], *wi = wr + ng3;
            if (myid2 < num_gpu) {
                fft.cuda_fft(wr, wi, gc.data(), nullptr, theta, myid2);
            } else {
                float coef[9];
#pragma omp critical(voxelize)
                {
                    check(mdk_rotation_coef(theta, 1, coef), "mdk_rotation_coef");
                    check(mdk_voxelize(vox, coef, 1, wr, wi), "mdk_voxelize");
                }
                fft.ligand_preparation(wr, wi, myid2);
                fft.fft3d(-1.0f, myid2);
                fft.convolution(myid2);
                fft.fft3d(1.0f, myid2);
                fft.score_sort(myid2);
            }
        }
    }
    fft.sort_index(nullptr, nullptr);          // (:319)
    for (int j = 0; j < nrot * P._Num_sort; j++)
        std::printf("E %d %.9g %d %d %d %d\n", j, fft.top_score(j), fft.top_index(j, 0), fft.top_index(j, 1),
                    fft.top_index(j, 2), fft.top_index(j, 3));
    mdk_destroy(vox);
    mdkh_protein_free(rec);
    mdkh_protein_free(lig);
    return 0;
}
