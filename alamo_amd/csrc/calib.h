// Attainable-bandwidth calibration of the roofline report (no reference counterpart): a streaming
// kernel with the byte pattern of the isotropic FP64 Fapply per node -- five reads (three
// displacement planes, lambda, mu) and three writes -- one node per work-item, no stencil, so its
// rate is what the memory system gives that traffic on the device at hand.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace alamo {

__global__ void __launch_bounds__(256) k_stream_5r3w(int64_t n, const double* __restrict__ a, double* __restrict__ o) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    const double l = a[3 * n + q], m = a[4 * n + q];
    o[q] = a[q] * l + m;
    o[n + q] = a[n + q] * l + m;
    o[2 * n + q] = a[2 * n + q] * l + m;
}

}  // namespace alamo
