// Batched factor + solve: the 256-thread shape, the work-queue reset and the
// dispatch over workgroup shapes (the kernel is in k_batched_body.hpp).
#include "k_batched_body.hpp"

namespace mfk {

hipError_t batched_launch_256(int N, hipStream_t s, const BatchArgs &a) { return launch_bt<256>(N, s, a); }

__global__ void k_counter_reset(unsigned int *counter) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *counter = 0u;
}

hipError_t batched_solve(int N, hipStream_t s, const BatchArgs &a) {
    if (a.threads != 64 && a.threads != 128 && a.threads != 256) return hipErrorInvalidValue;
    if (N < 1 || N > 4) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_counter_reset, dim3(1), dim3(64), 0, s, a.counter);
    if (a.count == 0) return hipGetLastError();
    if (a.threads == 64) return batched_launch_64(N, s, a);
    if (a.threads == 128) return batched_launch_128(N, s, a);
    return batched_launch_256(N, s, a);
}

int batched_max_order() { return BT_MAX; }
int batched_max_rhs() { return RHS_MAX; }

}  // namespace mfk
