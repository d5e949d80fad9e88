// Test library (not part of the product): compares lem::recip_f32_to_f64, the
// float64 reciprocal used by the numba hit predicate, with the IEEE quotient
// 1.0 / (double)a over a range of float32 bit patterns.  Built in-tree next to
// this file by the package build; loaded by tests/test_gpu_reciprocal.py.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../lyceanem-python_amd/csrc/lem_device.h"

__global__ __launch_bounds__(256) void k_recip_check(uint64_t start, uint64_t n,
                                                     unsigned long long *__restrict__ bad,
                                                     unsigned int *__restrict__ first_bad)
{
    const uint64_t stride = (uint64_t)gridDim.x * 256;
    unsigned long long nbad = 0;
    unsigned int fb = 0xffffffffu;
    for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += stride) {
        const unsigned int bits = (unsigned int)(start + i);
        const float a = __uint_as_float(bits);
        if (a == 0.0f || __builtin_isinf(a))
            continue;  // zero is never passed; +-inf is covered by the predicate test
        const double ref = 1.0 / (double)a;
        const double got = lem::recip_f32_to_f64(a);
        const bool same = __builtin_isnan(ref) ? __builtin_isnan(got)
                                               : __double_as_longlong(ref) == __double_as_longlong(got);
        if (!same) {
            ++nbad;
            fb = bits < fb ? bits : fb;
        }
    }
    if (nbad) {
        atomicAdd(bad, nbad);
        atomicMin(first_bad, fb);
    }
}

// Checks the bit patterns [start, start + n) (n <= 2^32 - start) on `stream`;
// *bad (device) accumulates mismatches, *first_bad (device) the lowest one.
extern "C" int lem_test_recip_check(uint64_t start, uint64_t n, unsigned long long *bad,
                                    unsigned int *first_bad, void *stream)
{
    if (start + n > (1ull << 32) || !bad || !first_bad)
        return -1;
    hipLaunchKernelGGL(k_recip_check, dim3(4096), dim3(256), 0,
                       reinterpret_cast<hipStream_t>(stream), start, n, bad, first_bad);
    return hipGetLastError() == hipSuccess ? 0 : -2;
}
