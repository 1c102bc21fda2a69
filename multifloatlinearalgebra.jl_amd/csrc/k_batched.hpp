// Host-side interface of the batched persistent factor + solve (k_batched.hip).
#pragma once
#include "mf_common.hpp"
#include "../../include/mfla_hip.h"

namespace mfk {

using BatchEntry = mfla_batch_entry;
enum { BATCH_CHOLESKY = MFLA_BATCH_CHOLESKY, BATCH_LDLT = MFLA_BATCH_LDLT };

struct BatchArgs {
    const BatchEntry *entries;  // device
    int64_t count;
    double *A, *B, *dsub;       // device pools (offsets from entries)
    int64_t *pivots;
    uint8_t *blocks;
    int64_t *info;              // device, one per system
    double *origmax;            // device, limbs per system
    unsigned int *counter;      // device work-queue counter
    double alpha[4];            // Bunch-Kaufman alpha (host limbs)
    int64_t chol_block;
    int exact_backward;
    int64_t workgroups;         // persistent workgroups
    int threads;                // workgroup shape: 64, 128 or 256 threads (systems need n <= threads)
};

hipError_t batched_solve(int N, hipStream_t s, const BatchArgs &a);
// one launch of a workgroup shape (no counter reset; batched_solve dispatches)
hipError_t batched_launch_64(int N, hipStream_t s, const BatchArgs &a);
hipError_t batched_launch_128(int N, hipStream_t s, const BatchArgs &a);
hipError_t batched_launch_256(int N, hipStream_t s, const BatchArgs &a);
int batched_max_order();
int batched_max_rhs();

}  // namespace mfk
