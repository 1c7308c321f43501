// Exhaustive check of the short correctly rounded forms the projections use
// (pbd_math.hpp): sqrt_rn(x) against sqrtf(x) for every float x in
// [2^-96, 2^96], and rcp_rn(d) against the IEEE quotient 1.0f / d for every
// float d in [2^-48, 2^48] -- the ranges in which project_distance takes
// them.  Prints one line per function and exits 0 when every result is
// bit-identical, 1 otherwise.
#include <hip/hip_runtime.h>

#include <cstdio>

#include "../positionbaseddynamicsold_amd/csrc/pbd_math.hpp"

namespace {

constexpr unsigned kMantissas = 1u << 23;

// one float per thread: exponent e (biased) with every mantissa
__global__ void k_check(int biased_lo, int nexp, unsigned long long* bad, unsigned* first_bad, int which) {
    const unsigned long long i = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (unsigned long long)nexp * kMantissas) return;
    const unsigned bits = ((unsigned)(biased_lo + (int)(i / kMantissas)) << 23) | (unsigned)(i % kMantissas);
    const float x = __uint_as_float(bits);
    float want, got;
    if (which == 0) {
        want = sqrtf(x);
        got = pbd::sqrt_rn(x);
    } else {
        want = 1.0f / x;
        got = pbd::rcp_rn(x);
    }
    if (__float_as_uint(want) != __float_as_uint(got)) {
        atomicAdd(bad, 1ull);
        atomicMin(first_bad, bits);
    }
}

bool ok(hipError_t e, const char* what) {
    if (e == hipSuccess) return true;
    std::fprintf(stderr, "%s: %s\n", what, hipGetErrorString(e));
    return false;
}

// every float with unbiased exponent in [lo, hi)
bool check(const char* name, int which, int lo, int hi, unsigned long long* d_bad, unsigned* d_first) {
    const int nexp = hi - lo;
    const unsigned long long n = (unsigned long long)nexp * kMantissas;
    unsigned long long bad = 0;
    unsigned first = 0xffffffffu;
    if (!ok(hipMemcpy(d_bad, &bad, sizeof(bad), hipMemcpyHostToDevice), "hipMemcpy") ||
        !ok(hipMemcpy(d_first, &first, sizeof(first), hipMemcpyHostToDevice), "hipMemcpy"))
        return false;
    const unsigned threads = 256;
    const unsigned blocks = (unsigned)((n + threads - 1) / threads);
    hipLaunchKernelGGL(k_check, dim3(blocks), dim3(threads), 0, 0, lo + 127, nexp, d_bad, d_first, which);
    if (!ok(hipGetLastError(), "launch") || !ok(hipDeviceSynchronize(), "kernel")) return false;
    if (!ok(hipMemcpy(&bad, d_bad, sizeof(bad), hipMemcpyDeviceToHost), "hipMemcpy") ||
        !ok(hipMemcpy(&first, d_first, sizeof(first), hipMemcpyDeviceToHost), "hipMemcpy"))
        return false;
    std::printf("%s: %llu floats in [2^%d, 2^%d), %llu differ", name, n, lo, hi, bad);
    if (bad) std::printf(" (first 0x%08x)", first);
    std::printf("\n");
    return bad == 0;
}

}  // namespace

int main() {
    unsigned long long* d_bad = nullptr;
    unsigned* d_first = nullptr;
    if (!ok(hipMalloc(&d_bad, sizeof(unsigned long long)), "hipMalloc") ||
        !ok(hipMalloc(&d_first, sizeof(unsigned)), "hipMalloc"))
        return 1;
    // [2^-96, 2^96] and [2^-48, 2^48]: the upper ends are powers of two,
    // checked with the exponent above the half-open range's last one
    bool good = check("sqrt_rn", 0, -96, 96, d_bad, d_first) && check("sqrt_rn", 0, 96, 97, d_bad, d_first);
    good = check("rcp_rn", 1, -48, 49, d_bad, d_first) && good;
    (void)hipFree(d_bad);
    (void)hipFree(d_first);
    std::printf(good ? "all equal\n" : "MISMATCH\n");
    return good ? 0 : 1;
}
