"""CPU: FFTProcess (include/mdk_fft_process.h, megadock_amd/csrc/host/fft_process.cpp)
keeps the reference's class interface: compiled with the reference's own CPUTime,
Parallel, Parameter, Receptor and Ligand headers (whose classes name FFTProcess a friend),
it replaces src/fft_process.h / src/fft_process.cpp in the reference's tree; and the
library built on the minimal host types exports every public method."""
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_SRC = "/root/reference/src"


@pytest.mark.skipif(not os.path.isdir(REF_SRC), reason="reference sources not present")
def test_facade_compiles_against_the_reference_classes():
    r = subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "tests", "native"), "refcheck", "REF=" + REF_SRC],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def test_facade_library_exports_the_reference_methods():
    lib = os.path.join(ROOT, "megadock_amd", "lib", "libmdk_fftprocess.so")
    assert os.path.exists(lib), "built by __graft_entry__.build()"
    syms = subprocess.run(["nm", "-DC", "--defined-only", lib], capture_output=True, text=True).stdout
    methods = ["FFTProcess::~FFTProcess", "FFTProcess::alloc_array(int const&)",
               "FFTProcess::alloc_fft()", "FFTProcess::receptor_fft(float*, float*)",
               "FFTProcess::cuda_fft(float*, float*, float*, float*, float*, unsigned long)",
               "FFTProcess::ligand_voxelization_on_gpu(float*, unsigned long)",
               "FFTProcess::ligand_data_transfer_gpu(float*)",
               "FFTProcess::ligand_preparation(float*, float*, unsigned long)",
               "FFTProcess::fft3d(float const&, unsigned long)", "FFTProcess::convolution(unsigned long)",
               "FFTProcess::score_sort(unsigned long)", "FFTProcess::fft_memory_free()",
               "FFTProcess::top_score_clean()", "FFTProcess::sort_index(float*, int*)"]
    for m in methods:
        assert m in syms, m
    # the inline accessors of the reference are declared in the header with the same signatures
    hdr = open(os.path.join(ROOT, "include", "mdk_fft_process.h")).read()
    for decl in [r"virtual int num_fft\(\)", r"virtual void num_fft\(const int &i\)",
                 r"virtual void rotation_index\(int i, const int &j\)", r"virtual float top_score\(const int &j\)",
                 r"virtual int top_index\(const int &j, int k\)",
                 r"FFTProcess\(CPUTime \*pcputime, Parallel \*pparallel, Parameter \*pparameter, Receptor \*preceptor, "
                 r"Ligand \*pligand\)"]:
        assert re.search(decl, hdr), decl
