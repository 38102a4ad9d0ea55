// Specialised pipeline: the tuned sizes (every kernel variant, for the autotuner) and
// the dispatch to the further sizes built one per object file from mdk_pipe_size.hip.
#include "mdk_pipe_kernels.h"

// further sizes (mdk_pipe_size.hip, the Makefile's PIPE_SIZES)
#define MDK_PIPE_SIZES(X) \
    X(64) X(66) X(70) X(72) X(84) X(88) X(90) X(98) X(100) X(108) X(110) X(112) X(120) X(126) X(128) X(132) \
    X(140) X(150) X(154) X(160) X(162) X(168) X(176) X(180) X(192) X(196) X(198) X(200) X(210) X(216) X(220) \
    X(224) X(240) X(242) X(250) X(252) X(264) X(270) X(280) X(288) X(294) X(300) X(308) X(320)
#define MDK_DECLARE(n) hipError_t pipe_run_n##n(const PipeArgs &a);
MDK_PIPE_SIZES(MDK_DECLARE)
#undef MDK_DECLARE

// Default kernel variants per size: correlation, inverse, forward, as `python -m
// megadock_amd.tune` chose them on MI355X (mdk_autotune over the variants built for each
// size, 2048 rotations of a synthetic pair scaled to the grid).  A size not listed runs
// correlation 6, the inverse with the next plane prefetched (3) up to nf = 128 where it can,
// else 2, and the prefetching forward.
namespace {
struct Defaults {
    int nf, xcorr, inv, fwd;
};
constexpr Defaults kDefaults[] = {
    {64, 7, 3, 0}, {66, 7, 3, 2}, {70, 7, 3, 0}, {72, 7, 3, 0}, {80, 6, 3, 0}, {84, 6, 3, 0},
    {88, 6, 3, 0}, {90, 6, 2, 0}, {96, 7, 3, 0}, {98, 7, 3, 0}, {100, 6, 3, 0}, {108, 5, 3, 0},
    {110, 7, 3, 0}, {112, 6, 3, 0}, {120, 5, 3, 0}, {126, 6, 3, 0}, {128, 5, 2, 0}, {132, 7, 3, 0},
    {140, 6, 3, 0}, {144, 6, 3, 0}, {150, 5, 2, 0}, {154, 6, 2, 0}, {160, 7, 2, 0}, {162, 5, 2, 0},
    {168, 7, 2, 0}, {176, 5, 2, 0}, {180, 6, 2, 2}, {192, 6, 2, 2}, {256, 6, 1, 0},
};
template <int N> constexpr int further_inv_default() { return N <= 128 && inv_prefetch_fits<N>() ? 3 : 2; }
}  // namespace

bool pipe_defaults(int nf, int *xcorr, int *inv, int *fwd) {
    for (const Defaults &d : kDefaults)
        if (d.nf == nf) {
            *xcorr = d.xcorr;
            *inv = d.inv;
            *fwd = d.fwd;
            return true;
        }
    switch (nf) {
#define MDK_CASE(n)                              \
    case n:                                      \
        *xcorr = 6;                              \
        *inv = further_inv_default<n>();         \
        *fwd = 0;                                \
        return true;
        MDK_PIPE_SIZES(MDK_CASE)
#undef MDK_CASE
        default: return false;
    }
}

bool pipe_has_variant(int nf, int which, int v) {
    if (!pipe_available(nf)) return false;
    const bool all = nf == 80 || nf == 96 || nf == 144 || nf == 256;
    switch (which) {
        case 0: return all ? (v >= 0 && v <= 7) : (v >= 5 && v <= 7);  // correlation
        case 1:                                                         // inverse
            if (all) return v >= 0 && v <= 3;
            if (v == 2) return true;
            if (v != 3) return false;
            switch (nf) {
#define MDK_CASE(n) case n: return inv_prefetch_fits<n>();
                MDK_PIPE_SIZES(MDK_CASE)
#undef MDK_CASE
                default: return false;
            }
        case 2:  // forward; at nf = 256: 0 the half-plane fused pass, 1 the two passes
            if (nf == 256) return v == 0 || v == 1;
            return v == 0 || v == 2 || (all && v == 1);
        default: return false;
    }
}

int pipe_cands_per_rotation(int nf) {
    if (!pipe_available(nf)) return 0;
    return nf;  // one list per (b, x) plane
}

bool pipe_available(int nf) {
    switch (nf) {
        case 80: case 96: case 144: case 256: return true;
#define MDK_CASE(n) case n:
        MDK_PIPE_SIZES(MDK_CASE)
#undef MDK_CASE
        return true;
        default: return false;
    }
}

hipError_t pipe_run(int nf, const PipeArgs &a) {
    switch (nf) {
        case 80: return run_pipe<80, 2>(a);
        case 96: return run_pipe<96, 2>(a);
        case 144: return run_pipe<144, 2>(a);
        case 256: return run_pipe<256, 2>(a);
#define MDK_CASE(n) case n: return pipe_run_n##n(a);
        MDK_PIPE_SIZES(MDK_CASE)
#undef MDK_CASE
        default: return hipErrorInvalidValue;
    }
}
