// Host-side interface of the triangular-solve kernels (k_trsm.hip).
#pragma once
#include "mf_common.hpp"

namespace mfk {

struct TrsmArgs {
    int64_t n;        // unknowns per right-hand side
    int64_t nrhs;     // right-hand sides
    int forward;      // 1: dependencies on q < p; 0: on q > p
    int unit;         // unit diagonal (no division)
    int exact_order;  // backward only: 1 = MFLA's ascending-q chain, 0 = blocked
    int swap_product; // 1: each term is X(q)*CA(p,q) (MFLA's right-side solves,
                      // trsm.jl:136-149, 159-172); 0: CA(p,q)*X(q) (left side)
    int64_t status_limit;  // 0: stop on any recorded failure; > 0: only on
                           // failures at (one-based) columns <= status_limit
    // Look-ahead (blocked orders): with a second stream and three events, the
    // update of the rows beyond the next diagonal block runs on `side` while
    // the next block is solved; null = everything on the caller's stream.
    hipStream_t side;
    hipEvent_t ev_main, ev_side, ev_side2;
    // Look-ahead only: two device words (a counter the side stream advances
    // after each update, an error word); non-null: each diagonal kernel
    // waits on the counter for the update that last reaches its rows instead
    // of the stream waiting on an event.
    unsigned *far_flag;
    // With far_flag, forward: each diagonal kernel also applies its block's
    // terms to the next block's rows as they are solved (the same operations
    // and order as applying them after).
    int stream_next;
    // With far_flag: a device word each diagonal kernel marks at its start
    // (block b's far update waits for block b + 1's mark on the side stream).
    unsigned *start_sig;
    // The caller's stream and `side` have hardware queues of their own (so
    // a kernel on one may wait for work enqueued after it on the other).
    int dedicated;
};

// *flag = value once the work enqueued before it on stream s is done (a
// device-scope release; kernels on other streams wait with wait_flag).
hipError_t flag_launch(hipStream_t s, unsigned *flag, unsigned value);

// Solves in place on the solution view X(p, c) with coefficient view CA(p, q)
// and diagonal view D(p, p).
hipError_t trsm_solve(int N, hipStream_t s, const TrsmArgs &a, const DV &X, const DV &CA, const DV &D, const Status *st);

}  // namespace mfk
