// Reference point for measurement only: the vendor FFT (hipFFT on rocFFT) timed on the
// same grid size as the rotation loop.  Not part of libmdk; bench.py loads it on its own.
#include <hip/hip_runtime.h>
#include <hipfft/hipfft.h>

extern "C" {

// Milliseconds for one forward plus one inverse in-place 3D C2C FFT of `batch` nf^3
// complex64 grids, averaged over `reps`.  Returns 0 on success.
int mdkref_fft_time(int device, int nf, int batch, int reps, float *ms) {
    if (!ms || nf < 2 || batch < 1 || reps < 1) return 1;
    if (hipSetDevice(device) != hipSuccess) return 2;
    const size_t n = (size_t)nf * nf * nf * batch;
    hipfftComplex *d = nullptr;
    if (hipMalloc(&d, n * sizeof(hipfftComplex)) != hipSuccess) return 3;
    int rc = 0;
    hipfftHandle plan = 0;
    hipStream_t s = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int dims[3] = {nf, nf, nf};
    const int dist = nf * nf * nf;
    if (hipMemset(d, 0, n * sizeof(hipfftComplex)) != hipSuccess || hipStreamCreate(&s) != hipSuccess ||
        hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess ||
        hipfftPlanMany(&plan, 3, dims, nullptr, 1, dist, nullptr, 1, dist, HIPFFT_C2C, batch) != HIPFFT_SUCCESS ||
        hipfftSetStream(plan, s) != HIPFFT_SUCCESS) {
        rc = 4;
    } else {
        for (int w = 0; w < 2 && !rc; w++)
            if (hipfftExecC2C(plan, d, d, HIPFFT_FORWARD) != HIPFFT_SUCCESS ||
                hipfftExecC2C(plan, d, d, HIPFFT_BACKWARD) != HIPFFT_SUCCESS)
                rc = 5;
        if (!rc && hipEventRecord(e0, s) != hipSuccess) rc = 6;
        for (int r = 0; r < reps && !rc; r++)
            if (hipfftExecC2C(plan, d, d, HIPFFT_FORWARD) != HIPFFT_SUCCESS ||
                hipfftExecC2C(plan, d, d, HIPFFT_BACKWARD) != HIPFFT_SUCCESS)
                rc = 5;
        float t = 0.f;
        if (!rc && (hipEventRecord(e1, s) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                    hipEventElapsedTime(&t, e0, e1) != hipSuccess))
            rc = 6;
        *ms = t / reps;
    }
    if (plan) hipfftDestroy(plan);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (s) (void)hipStreamDestroy(s);
    (void)hipFree(d);
    return rc;
}

}  // extern "C"
