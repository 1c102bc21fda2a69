// Shared device-side definitions: strided limb-planar matrix views and the
// launch helpers used by every kernel translation unit.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include "mf_arith.hpp"

namespace mfk {

// A strided view of a limb-planar matrix: element (i, j), limb l lives at
// p[l*ls + i*rs + j*cs]. A column-major plane has rs = 1, cs = ld; the
// transposed view of the same storage swaps rs and cs.
struct DV {
    double *p;
    int64_t rs, cs, ls;
};

template <int N> __device__ __forceinline__ mf::V<N> ld(const DV &v, int64_t i, int64_t j) {
    mf::V<N> r;
    const double *q = v.p + i * v.rs + j * v.cs;
#pragma unroll
    for (int l = 0; l < N; ++l) r.x[l] = q[l * v.ls];
    return r;
}
template <int N> __device__ __forceinline__ void stv(const DV &v, int64_t i, int64_t j, const mf::V<N> &x) {
    double *q = v.p + i * v.rs + j * v.cs;
#pragma unroll
    for (int l = 0; l < N; ++l) q[l * v.ls] = x.x[l];
}

// Scalars travel by value as plain limb arrays.
template <int N> struct S { double x[N]; };
template <int N> __device__ __forceinline__ mf::V<N> sv(const S<N> &s) {
    mf::V<N> r;
#pragma unroll
    for (int l = 0; l < N; ++l) r.x[l] = s.x[l];
    return r;
}

// Device status word shared by the kernels of one factorization: once a
// kernel records a nonzero status, every later kernel of the same call sees it
// and returns without touching memory, so the matrix is left exactly as the
// failing step left it.
struct Status {
    int64_t info;   // MFLA status code (0, -1, k > 0)
    int64_t aux;    // kernel-private scratch (pivot index, flags)
    unsigned long long sync;        // grid-barrier arrivals of the multi-workgroup kernels of a call
    unsigned long long sync_base;   // the arrivals before the next such launch
};

// A blocked-LDLT panel that stops by design (the host resumes the
// factorization after reading the code) records
// info = kPanelStopBase - 2 * pfirst - truncated: truncated = 0 when it
// extended past its last requested column by a 2x2 pivot, 1 when it reached
// a column that needs the row-max pass while the previous panel's trailing
// update was still running. The code names the panel's first column, so a
// kernel that reads the word alone knows which panel stopped.
constexpr int64_t kPanelStopBase = -16;
__host__ __device__ inline int64_t panel_stop_code(int64_t pfirst, bool truncated) { return kPanelStopBase - 2 * pfirst - (truncated ? 1 : 0); }
__host__ __device__ inline bool is_panel_stop(int64_t info) { return info <= kPanelStopBase; }
__host__ __device__ inline int64_t panel_stop_first(int64_t info) { return (kPanelStopBase - info) >> 1; }
__host__ __device__ inline bool panel_stop_truncated(int64_t info) { return ((kPanelStopBase - info) & 1) != 0; }

// Whether a kernel gated by the status returns at once. limit == 0: on any
// recorded status. limit > 0 (an update of the panel ending at one-based
// column `limit`, which may run beside later panels): only on failures at
// columns <= limit or a stop of a panel that starts at or before it. A code
// recorded by a later panel while the update runs leaves it running
// whichever value the load returns.
__device__ inline bool status_skips(const Status *st, int64_t limit) {
    const int64_t info = st->info;
    if (info == 0) return false;
    if (limit == 0) return true;
    if (info > 0) return info <= limit;
    if (is_panel_stop(info)) return panel_stop_first(info) + 1 <= limit;
    return true;
}

// Waits until *flag >= value (a device counter another stream's kernel
// advances, flag_launch); false after about 2^24 polls.
__device__ inline bool wait_flag(const unsigned *flag, unsigned value) {
    unsigned spins = 0;
    while (__hip_atomic_load(flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < value) {
        __builtin_amdgcn_s_sleep(4);
        if (++spins > (1u << 24)) return false;
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    return true;
}

inline int64_t cdiv(int64_t a, int64_t b) { return (a + b - 1) / b; }

}  // namespace mfk
