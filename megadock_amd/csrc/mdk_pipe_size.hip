// One further size of the specialised pipeline, with the kernel variants worth tuning on
// every size (run_pipe's LEVEL 1): compiled once per size with -DMDK_PIPE_N=<nf> (the
// Makefile's PIPE_SIZES).
#include "mdk_pipe_kernels.h"

#ifndef MDK_PIPE_N
#error "MDK_PIPE_N (the FFT size) must be defined"
#endif
#define MDK_CAT2(a, b) a##b
#define MDK_CAT(a, b) MDK_CAT2(a, b)

hipError_t MDK_CAT(pipe_run_n, MDK_PIPE_N)(const PipeArgs &a) { return run_pipe<MDK_PIPE_N, 1>(a); }
