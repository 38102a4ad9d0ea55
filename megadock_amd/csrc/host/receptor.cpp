// Receptor grids on the host, as the reference computes them once per pair
// (src/receptor.cpp:20-372).  Compiled with -ffp-contract=off so every float
// operation rounds exactly as in the reference's x86-64 build.
//
// Shape grid: core voxels (d^2 < 1.5 r^2) get rec_core; open space is found by six
// ray sweeps and two in-place propagation sweeps whose order matters
// (src/receptor.cpp:106-189) and is kept; enclosed cavities become core; every other
// voxel gets sum over atoms within (r + 3.6)^2 of 1 + ACE * (-0.8) * aceratio, added
// in double and rounded to float at each step (src/receptor.cpp:209-249).
// Electrostatics: FTDock field with a distance-dependent dielectric on non-core voxels
// (src/receptor.cpp:258-346).
#pragma clang fp contract(off)

#include <algorithm>
#include <cmath>
#include <cstring>
#include <thread>

#include "mdk_host.h"

namespace mdk {

namespace {

struct Box {
    int lo[3], hi[3];  // half-open
};

// search box of +-range voxels around the voxel holding (x, y, z) (src/receptor.cpp:73-83)
Box atom_box(const float *p, float gw, int ng, int range) {
    Box b;
    for (int a = 0; a < 3; a++) {
        const int c = (int)(p[a] / gw + ng / 2);
        b.lo[a] = std::max(c - range, 0);
        b.hi[a] = std::min(c + range + 1, ng);
    }
    return b;
}

// fn(i0, i1) over x-planes [i0, i1) split among up to `threads` threads.  Every voxel
// belongs to one plane, so per-voxel sums keep their atom order whatever the split.
template <class F> void over_planes(int ng, int threads, F &&fn) {
    if (threads <= 0) threads = (int)std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    threads = std::min(threads, ng);
    if (threads <= 1) {
        fn(0, ng);
        return;
    }
    std::vector<std::thread> pool;
    for (int t = 0; t < threads; t++) pool.emplace_back([&, t] { fn(ng * t / threads, ng * (t + 1) / threads); });
    for (auto &th : pool) th.join();
}

}  // namespace

void Receptor::voxelize(int ng, float gw, const std::vector<float> &gc, const float *xyz, float aceratio,
                        float rec_core, float eratio, float *grid_r, float *grid_i, int threads) const {
    const int na = num_atoms();
    const size_t n1 = ng, n2 = n1 * n1, n3 = n2 * n1;
    const float eps = rec_core, open = -7777.0f;
    std::fill(grid_r, grid_r + n3, 0.0f);
    std::fill(grid_i, grid_i + n3, 0.0f);

    // squared per-axis distances atom -> grid line (src/protein.cpp:152-163)
    std::vector<float> d2x(na * n1), d2y(na * n1), d2z(na * n1);
    for (int l = 0; l < na; l++)
        for (size_t i = 0; i < n1; i++) {
            float dx = xyz[3 * l] - gc[i], dy = xyz[3 * l + 1] - gc[i], dz = xyz[3 * l + 2] - gc[i];
            d2x[l * n1 + i] = dx * dx;
            d2y[l * n1 + i] = dy * dy;
            d2z[l * n1 + i] = dz * dz;
        }
    auto at = [&](size_t i, size_t j, size_t k) -> float & { return grid_r[i * n2 + j * n1 + k]; };

    // core fill
    const int core_range = (int)((2.4 + (double)gw - 0.01) / (double)gw);
    over_planes(ng, threads, [&](int i0, int i1) {
        for (int l = 0; l < na; l++) {
            const float r = radius[l], rc2 = r * r * 1.5f;
            const Box b = atom_box(&xyz[3 * l], gw, ng, core_range);
            const float *ax = &d2x[l * n1], *ay = &d2y[l * n1], *az = &d2z[l * n1];
            for (int i = std::max(b.lo[0], i0); i < std::min(b.hi[0], i1); i++) {
                if (ax[i] > rc2) continue;
                for (int j = b.lo[1]; j < b.hi[1]; j++) {
                    if (ax[i] + ay[j] > rc2) continue;
                    for (int k = b.lo[2]; k < b.hi[2]; k++)
                        if (ax[i] + ay[j] + az[k] < rc2) at(i, j, k) = eps;
                }
            }
        }
    });

    // open space: rays from each face until the first core voxel
    for (size_t p = 0; p < n2; p++) {
        const size_t u = p % n1, v = p / n1;
        auto ray = [&](auto idx) {
            for (size_t s = 0; s < n1; s++) {
                float &g = grid_r[idx(s)];
                if (g == eps) break;
                g = open;
            }
            for (size_t s = n1; s-- > 0;) {
                float &g = grid_r[idx(s)];
                if (g == eps) break;
                g = open;
            }
        };
        ray([&](size_t s) { return u + n1 * v + n2 * s; });
        ray([&](size_t s) { return s + n1 * u + n2 * v; });
        ray([&](size_t s) { return v + n1 * s + n2 * u; });
    }
    // two in-place propagation sweeps in increasing linear order (order-dependent)
    for (int sweep = 0; sweep < 2; sweep++)
        for (size_t q = n2; q < n3 - n2; q++) {
            const int j = (int)((q / n1) % n1), k = (int)(q % n1);
            if (j < 2 || j > ng - 2 || k < 2 || k > ng - 2) continue;
            if (grid_r[q] != open) continue;
            const size_t nb[6] = {q + 1, q - 1, q + n1, q - n1, q + n2, q - n2};
            for (size_t w : nb)
                if (grid_r[w] == 0) grid_r[w] = open;
        }
    for (size_t q = 0; q < n3; q++)  // enclosed cavities become core
        if (grid_r[q] == 0) grid_r[q] = eps;
    for (size_t q = 0; q < n3; q++)
        if (grid_r[q] == open) grid_r[q] = 0;

    // surface score: rPSC + receptor ACE
    std::vector<float> surf(n3, 0.0f);
    const int surf_range = (int)((12 + (double)gw - 0.01) / (double)gw);
    over_planes(ng, threads, [&](int i0, int i1) {
        for (int l = 0; l < na; l++) {
            const float rj = (radius[l] + 3.6f) * (radius[l] + 3.6f);
            const double add = 1.0 + (double)ace[l] * (-0.8) * (double)aceratio;
            const Box b = atom_box(&xyz[3 * l], gw, ng, surf_range);
            const float *ax = &d2x[l * n1], *ay = &d2y[l * n1], *az = &d2z[l * n1];
            for (int i = std::max(b.lo[0], i0); i < std::min(b.hi[0], i1); i++) {
                if (ax[i] > rj) continue;
                for (int j = b.lo[1]; j < b.hi[1]; j++) {
                    if (ax[i] + ay[j] > rj) continue;
                    for (int k = b.lo[2]; k < b.hi[2]; k++)
                        if (ax[i] + ay[j] + az[k] < rj) {
                            float &s = surf[i * n2 + j * n1 + k];
                            s = (float)((double)s + add);
                        }
                }
            }
        }
    });
    for (size_t q = 0; q < n3; q++)
        if (grid_r[q] != eps) grid_r[q] = surf[q];

    // electrostatics (beta = -2800, src/docking.cpp:176)
    const float beta = -2800.0f;
    const float gwl = gc[1] - gc[0];
    const float charge_per_grid = -1 * gwl * 80.0f / (beta * eratio);
    over_planes(ng, threads, [&](int i0, int i1) {
        for (int l = 0; l < na; l++) {
            const float q = charge[l], aq = std::fabs(q);
            if (aq <= 1.0e-6f) continue;
            const int range = (double)aq < 0.07 ? 4 : ((double)aq < 0.21 ? 5 : (int)(aq / charge_per_grid));
            const Box b = atom_box(&xyz[3 * l], gwl, ng, range);
            const float *ax = &d2x[l * n1], *ay = &d2y[l * n1], *az = &d2z[l * n1];
            for (int i = std::max(b.lo[0], i0); i < std::min(b.hi[0], i1); i++)
                for (int j = b.lo[1]; j < b.hi[1]; j++) {
                    const float dij = ax[i] + ay[j];
                    for (int k = b.lo[2]; k < b.hi[2]; k++) {
                        const size_t w = i * n2 + j * n1 + k;
                        if (grid_r[w] < 0.0) continue;
                        const float d = std::sqrt(dij + az[k]);
                        const float er = d >= 8.0f ? 80.0f : (d > 6.0f ? 38.0f * d + -224.0f : 4.0f);
                        grid_i[w] += eratio * beta * q / er / d;
                    }
                }
        }
    });
}

}  // namespace mdk
