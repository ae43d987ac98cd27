// Host build of the GPU's std::sort emulation (mspai_amd/csrc/introsort.h),
// compiled by tests/test_introsort.py to compare it with the host std::sort.
#include "introsort.h"

extern "C" void emul_sort_rho(int n, int* idx, double* rho)
{
    mspai::RhoPair* v = new mspai::RhoPair[n > 0 ? n : 1];
    for (int i = 0; i < n; ++i) v[i] = {idx[i], rho[i]};
    int stack[3 * mspai::introsort_detail::kStackFrames];
    mspai::std_sort_rho(v, v + n, stack);
    for (int i = 0; i < n; ++i) {
        idx[i] = v[i].idx;
        rho[i] = v[i].rho;
    }
    delete[] v;
}
