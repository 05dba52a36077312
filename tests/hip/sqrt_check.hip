// sqrt_check.hip -- TEST INFRASTRUCTURE: the library's own square roots (swe::sqrt_rn for every
// argument, swe::sqrt_normal for finite arguments >= 2^-767, from csrc/swe_kernels.h) against the
// device sqrt(), bit for bit, on hashed doubles of every class.  Prints one line per class and exits
// with status 3 when any result differs.
//
//   sqrt_check [log2 of the count per class, default 28]
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <cstdlib>

#include "swe_kernels.h"

namespace {

__device__ __forceinline__ uint64_t mix64(uint64_t z) {
    z += 0x9e3779b97f4a7c15ull;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

// class 0: exponents -767..1023 (sqrt_normal); 1: [1, 2^20) (sqrt_normal); 2: mantissas near all
// ones or zeros (sqrt_normal); 3: any bit pattern (sqrt_rn); 4: tiny and denormal (sqrt_rn)
__device__ __forceinline__ double arg_of(int cls, uint64_t r) {
    const uint64_t man = r & 0xfffffffffffffull;
    switch (cls) {
        case 0: return __longlong_as_double((long long)(((r >> 52) % 1791 + 256) << 52 | man));
        case 1: return __longlong_as_double((long long)(((r >> 52) % 20 + 1023) << 52 | man));
        case 2: {
            const uint64_t m = (r & 1) ? (0xfffffffffffffull - ((r >> 1) & 0xffff)) : ((r >> 1) & 0xffff);
            return __longlong_as_double((long long)(((r >> 40) % 1791 + 256) << 52 | m));
        }
        case 3: return __longlong_as_double((long long)r);
        default: return __longlong_as_double((long long)(((r >> 52) % 260) << 52 | man));
    }
}

__global__ void k_check(uint64_t n, uint64_t seed, int cls, unsigned long long* bad, double* example) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    unsigned long long nb = 0;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const double x = arg_of(cls, mix64(i ^ seed));
        const double a = sqrt(x);
        const double b = (cls >= 3) ? swe::sqrt_rn(x) : swe::sqrt_normal(x);
        if (__double_as_longlong(a) != __double_as_longlong(b)) {
            ++nb;
            *example = x;
        }
    }
    if (nb) atomicAdd(bad, nb);
}

}  // namespace

int main(int argc, char** argv) {
    const int lg = (argc > 1) ? std::atoi(argv[1]) : 28;
    if (lg < 10 || lg > 34) {
        std::fprintf(stderr, "count exponent %d out of 10..34\n", lg);
        return 2;
    }
    const uint64_t n = 1ull << lg;
    unsigned long long* bad = nullptr;
    double* ex = nullptr;
    if (hipMalloc((void**)&bad, sizeof *bad) != hipSuccess || hipMalloc((void**)&ex, sizeof *ex) != hipSuccess) return 1;
    int rc = 0;
    for (int cls = 0; cls < 5; ++cls) {
        unsigned long long h = 0;
        double e = 0.0;
        if (hipMemset(bad, 0, sizeof *bad) != hipSuccess) return 1;
        hipLaunchKernelGGL(k_check, dim3(8192), dim3(256), 0, 0, n, 0x5eed0000ull + (uint64_t)cls, cls, bad, ex);
        if (hipMemcpy(&h, bad, sizeof h, hipMemcpyDeviceToHost) != hipSuccess) return 1;
        if (hipMemcpy(&e, ex, sizeof e, hipMemcpyDeviceToHost) != hipSuccess) return 1;
        std::printf("class %d: %llu of %llu differ", cls, h, (unsigned long long)n);
        if (h) {
            std::printf(" (for instance x = %.17g)", e);
            rc = 3;
        }
        std::printf("\n");
    }
    (void)hipFree(bad);
    (void)hipFree(ex);
    return rc;
}
