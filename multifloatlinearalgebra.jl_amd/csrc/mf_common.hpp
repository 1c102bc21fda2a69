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

inline int64_t cdiv(int64_t a, int64_t b) { return (a + b - 1) / b; }

}  // namespace mfk
