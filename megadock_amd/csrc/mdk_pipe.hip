// Specialised pipeline: the tuned sizes (every kernel variant, for the autotuner) and
// the dispatch to the further sizes built one per object file from mdk_pipe_size.hip.
#include "mdk_pipe_kernels.h"

// further sizes (mdk_pipe_size.hip, the Makefile's PIPE_SIZES)
#define MDK_PIPE_SIZES(X) \
    X(64) X(66) X(70) X(72) X(84) X(88) X(90) X(98) X(100) X(108) X(110) X(112) X(120) X(126) X(128) X(132) \
    X(140) X(150) X(154) X(160) X(162) X(168) X(176) X(180) X(192)
#define MDK_DECLARE(n) hipError_t pipe_run_n##n(const PipeArgs &a);
MDK_PIPE_SIZES(MDK_DECLARE)
#undef MDK_DECLARE

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
        case 80: return run_pipe<80, true>(a);
        case 96: return run_pipe<96, true>(a);
        case 144: return run_pipe<144, true>(a);
        case 256: return run_pipe<256, true>(a);
#define MDK_CASE(n) case n: return pipe_run_n##n(a);
        MDK_PIPE_SIZES(MDK_CASE)
#undef MDK_CASE
        default: return hipErrorInvalidValue;
    }
}
