// C ABI of the library (include/mfla_hip.h): argument validation with MFLA's
// error semantics, and the host drivers of the blocked factorizations and
// solves. Every numerical step runs on the GPU; the host only sequences
// kernels, and reads back status words where MFLA returns them.
#include <hip/hip_runtime.h>
#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/mfla_hip.h"
#include "k_adjoint.hpp"
#include "k_batched.hpp"
#include "k_blas2.hpp"
#include "k_blas3.hpp"
#include "k_f64.hpp"
#include "k_factor.hpp"
#include "k_level3x.hpp"
#include "k_ozaki.hpp"
#include "k_qr.hpp"
#include "k_refine.hpp"
#include "k_trsm.hpp"
#include "mf_arith.hpp"

using mfk::DV;
using mfk::Status;

struct mfla_handle_s {
    int device = 0;
    hipStream_t stream = nullptr;
    Status *status = nullptr;   // device status word of the current call
    int *flags = nullptr;       // small device flag array
    void *scratch = nullptr;    // grows on demand (non-hot paths only)
    size_t scratch_bytes = 0;
    std::mutex lock;            // calls on one handle are serialised
    hipStream_t side = nullptr; // second stream of the look-ahead factorization
    hipEvent_t ev_join = nullptr, ev_panel = nullptr, ev_update = nullptr, ev_update2 = nullptr;
    // CU-partitioned streams of ldlt!'s look-ahead (created on first use):
    // panels on la_cus CUs, trailing updates on the others
    hipStream_t la_panel = nullptr, la_update = nullptr;
    int la_cus = 0;
    // pinned host staging for the small device-to-host reads a call waits on
    // (status, per-iteration column maxima): one DMA transfer each instead of
    // a copy through the runtime's pageable staging; null = pageable path
    unsigned char *pinned = nullptr;
    static constexpr size_t kPinnedBytes = 256 * 1024;
};

struct mfla_workspace_s {
    mfla_handle h = nullptr;
    int limbs = 0;
    int64_t factor_capacity = 0;
    int64_t block_capacity = 0;
    double *weighted = nullptr;  // three parts of limbs * factor_capacity * (block_capacity + 1)
};

// Device scratch of the no-allocation entry point (mfla_sytrf_noalloc): a
// file-scope device array, one copy per GPU, held until the process ends.
// Host code never passes the array's name to a kernel (on the host it is a
// stand-in, not a device address): each device's address is resolved once,
// with that device current, by hipGetSymbolAddress and kept in
// g_scratch_addr.
static __device__ __align__(256) double scratch[MFLA_NOALLOC_SCRATCH_BYTES / sizeof(double)];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_lock;
double *g_scratch_addr[kMaxDevices] = {};
hipError_t g_scratch_status[kMaxDevices] = {};
bool g_scratch_resolved[kMaxDevices] = {};

// Resolves the current device's copy of `scratch` (once per device).
void resolve_scratch(int device) {
    if (device < 0 || device >= kMaxDevices) return;
    std::lock_guard<std::mutex> g(g_scratch_lock);
    if (g_scratch_resolved[device]) return;
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    g_scratch_status[device] = e;
    g_scratch_addr[device] = e == hipSuccess ? static_cast<double *>(p) : nullptr;
    if (e != hipSuccess) (void)hipGetLastError();
    g_scratch_resolved[device] = true;
}

}  // namespace

namespace {

thread_local std::string g_err;

int fail(int code, const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

#define HIPCHK(expr)                                                                         \
    do {                                                                                     \
        hipError_t e_ = (expr);                                                              \
        if (e_ != hipSuccess) return fail(MFLA_ERR_HIP, "HIP error %s at %s:%d", hipGetErrorString(e_), __FILE__, __LINE__); \
    } while (0)

bool limbs_ok(int N) { return N >= 1 && N <= 4; }
int check_limbs(int N) { return limbs_ok(N) ? MFLA_OK : fail(MFLA_ERR_ARGUMENT, "only 1-4 limb MultiFloats are supported"); }

DV dvn(const mfla_matrix &m) { return DV{m.data, 1, m.ld, m.limb_stride}; }        // (i, j) -> M[i, j]
DV dvt(const mfla_matrix &m) { return DV{m.data, m.ld, 1, m.limb_stride}; }        // (i, j) -> M[j, i]
DV sub(const DV &v, int64_t i, int64_t j) { return DV{v.p + i * v.rs + j * v.cs, v.rs, v.cs, v.ls}; }

// Base.mightalias for limb-planar storage: overlapping address extents.
bool extents_overlap(const mfla_matrix &a, const mfla_matrix &b, int N) {
    if (a.rows * a.cols == 0 || b.rows * b.cols == 0) return false;
    auto ext = [N](const mfla_matrix &m, const double *&lo, const double *&hi) {
        lo = m.data;
        hi = m.data + (int64_t)(N - 1) * m.limb_stride + (m.cols - 1) * m.ld + m.rows;
    };
    const double *alo, *ahi, *blo, *bhi;
    ext(a, alo, ahi);
    ext(b, blo, bhi);
    return alo < bhi && blo < ahi;
}
bool same_matrix(const mfla_matrix &a, const mfla_matrix &b) {
    return a.data == b.data && a.rows == b.rows && a.cols == b.cols && a.ld == b.ld && a.limb_stride == b.limb_stride;
}
int check_desc(const mfla_matrix &m, const char *name) {
    if (m.rows < 0 || m.cols < 0) return fail(MFLA_ERR_DIMENSION, "%s has negative dimensions", name);
    if (m.rows * m.cols > 0 && (m.data == nullptr || m.ld < m.rows)) return fail(MFLA_ERR_ARGUMENT, "%s has an invalid descriptor", name);
    return MFLA_OK;
}

template <int N> mf::V<N> host_v(const double *p) { mf::V<N> v; for (int l = 0; l < N; ++l) v.x[l] = p[l]; return v; }
bool scalar_is_zero(const double *s, int N) { for (int l = 0; l < N; ++l) if (s[l] != 0.0) return false; return true; }
bool scalar_is_one(const double *s, int N) { if (s[0] != 1.0) return false; for (int l = 1; l < N; ++l) if (s[l] != 0.0) return false; return true; }

const mfla_config &cfg_or_default(const mfla_config *c, mfla_config &tmp) {
    if (c) return *c;
    mfla_config_default(&tmp);
    return tmp;
}

void *grow_scratch(mfla_handle h, size_t bytes) {
    if (h->scratch_bytes < bytes) {
        if (h->scratch) { (void)hipStreamSynchronize(h->stream); (void)hipFree(h->scratch); }
        h->scratch = nullptr;
        h->scratch_bytes = 0;
        if (hipMalloc(&h->scratch, bytes) != hipSuccess) return nullptr;
        h->scratch_bytes = bytes;
    }
    return h->scratch;
}

int64_t read_info(mfla_handle h, int64_t *aux = nullptr, hipStream_t on = nullptr) {
    Status local{};
    Status *st = h->pinned ? reinterpret_cast<Status *>(h->pinned) : &local;
    hipStream_t q = on ? on : h->stream;
    if (hipMemcpyAsync(st, h->status, sizeof(Status), hipMemcpyDeviceToHost, q) != hipSuccess) return INT64_MIN;
    if (hipStreamSynchronize(q) != hipSuccess) return INT64_MIN;
    if (aux) *aux = st->aux;
    return st->info;
}

// Device-to-host reads of `count` (pointer, bytes) pairs on stream s, then a
// wait: through the pinned staging when they fit, else directly.
hipError_t read_back(mfla_handle h, hipStream_t s, int count, void *const *dst, const void *const *src, const size_t *bytes) {
    size_t total = 0;
    for (int i = 0; i < count; ++i) total += (bytes[i] + 255) & ~(size_t)255;
    const bool staged = h->pinned && total + 256 <= mfla_handle_s::kPinnedBytes;
    size_t off = 256;                              // the first 256 bytes hold the status copy
    for (int i = 0; i < count; ++i) {
        void *to = staged ? h->pinned + off : dst[i];
        hipError_t e = hipMemcpyAsync(to, src[i], bytes[i], hipMemcpyDeviceToHost, s);
        if (e != hipSuccess) return e;
        off += (bytes[i] + 255) & ~(size_t)255;
    }
    hipError_t e = hipStreamSynchronize(s);
    if (e != hipSuccess || !staged) return e;
    off = 256;
    for (int i = 0; i < count; ++i) {
        std::memcpy(dst[i], h->pinned + off, bytes[i]);
        off += (bytes[i] + 255) & ~(size_t)255;
    }
    return hipSuccess;
}

const double kOne[4] = {1.0, 0.0, 0.0, 0.0};
const double kMinusOne[4] = {-1.0, 0.0, 0.0, 0.0};
const double kZero[4] = {0.0, 0.0, 0.0, 0.0};

// GEMM body shared by gemm!, LU's trailing update and the matrix residual.
// x3 auto/fused route: MFLA's four-row vector blocks use mulacc_x3 and the
// m mod 4 remainder rows the standard network (gemm.jl:411-452).
// stream: null = the handle's stream; status_limit as in TileArgs.
hipError_t gemm_launch(mfla_handle h, int N, int strategy, int tile, int64_t m, int64_t n, int64_t k, const DV &A, const DV &B,
                       const DV &C, const double *alpha, const double *beta, const Status *st, hipStream_t stream = nullptr,
                       int64_t status_limit = 0) {
    const bool fused = (N == 3) && (strategy == MFLA_GEMM_AUTO || strategy == MFLA_GEMM_FUSED);
    hipStream_t s = stream ? stream : h->stream;
    mfk::TileArgs t{};
    t.mode = mfk::MODE_AB; t.lower = 0; t.beta_zero = scalar_is_zero(beta, N) ? 1 : 0; t.tile = tile; t.n = n; t.K = k;
    t.status_limit = status_limit;
    if (fused) {
        const int64_t m4 = (m / 4) * 4;
        t.net = mf::NET_FUSED3; t.row_lo = 0; t.rows = m4;
        hipError_t e = mfk::tile_launch_n(N, s, t, A, B, C, alpha, beta, st);
        if (e != hipSuccess) return e;
        t.net = mf::NET_STD; t.row_lo = m4; t.rows = m - m4;
        return mfk::tile_launch_n(N, s, t, A, B, C, alpha, beta, st);
    }
    t.net = mf::NET_STD; t.row_lo = 0; t.rows = m;
    return mfk::tile_launch_n(N, s, t, A, B, C, alpha, beta, st);
}

// Lower-triangle update C = alpha*a*b + beta*C over i >= j (syrk!/gemmt!).
hipError_t lower_launch(mfla_handle h, int N, int tile, int64_t n, int64_t k, const DV &A, const DV &B, const DV &C,
                        const double *alpha, const double *beta, const Status *st) {
    mfk::TileArgs t{};
    t.mode = mfk::MODE_AB; t.net = mf::NET_STD; t.lower = 1; t.beta_zero = scalar_is_zero(beta, N) ? 1 : 0; t.tile = tile;
    t.row_lo = 0; t.rows = n; t.n = n; t.K = k;
    return mfk::tile_launch_n(N, h->stream, t, A, B, C, alpha, beta, st);
}

// Second stream (default priority) and events for look-ahead factorizations.
hipError_t ensure_side(mfla_handle h) {
    if (h->side) return hipSuccess;
    hipError_t e = hipStreamCreateWithFlags(&h->side, hipStreamNonBlocking);
    if (e != hipSuccess) { h->side = nullptr; return e; }
    if ((e = hipEventCreateWithFlags(&h->ev_join, hipEventDisableTiming)) != hipSuccess) return e;
    if ((e = hipEventCreateWithFlags(&h->ev_panel, hipEventDisableTiming)) != hipSuccess) return e;
    if ((e = hipEventCreateWithFlags(&h->ev_update2, hipEventDisableTiming)) != hipSuccess) return e;
    return hipEventCreateWithFlags(&h->ev_update, hipEventDisableTiming);
}

// The CU-partitioned streams of ldlt!'s look-ahead (default priority):
// panel_cus CUs for the panels, the device's other CUs for the trailing
// updates, so a panel's latency-bound waves never share a SIMD with the
// update's. Mask bit C x + 8 y + x (XCD x < 8, y < panel_cus / 8, C = CUs /
// 8) gives each of the eight XCDs panel_cus / 8 of them whether the driver
// maps mask bits to XCDs in blocks of C or in turn, so every XCD runs work
// of both streams (a kernel's workgroups are dealt to the XCDs in turn);
// panel_cus <= C keeps every bit inside its XCD's block. Fails (and the
// caller runs without look-ahead) on a device without eight XCDs of at
// least 16 CUs.
// CUs of ldlt! look-ahead's panel stream (four per XCD; the panels use at
// most 32 workgroups there, and the first part of each update runs beside them)
constexpr int kLdltPanelCUs = 32;
hipError_t ensure_la_streams(mfla_handle h, int panel_cus) {
    if (h->la_panel && h->la_cus == panel_cus) return hipSuccess;
    int cus = 0;
    hipError_t e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, h->device);
    if (e != hipSuccess) return e;
    const int C = cus / 8;
    if (cus % 8 != 0 || C < 16 || panel_cus % 8 != 0 || panel_cus < 8 || panel_cus > C || cus > 1024) return hipErrorNotSupported;
    for (hipStream_t *q : {&h->la_panel, &h->la_update})
        if (*q) { (void)hipStreamSynchronize(*q); (void)hipStreamDestroy(*q); *q = nullptr; }
    h->la_cus = 0;
    uint32_t pm[32] = {}, um[32] = {};
    for (int x = 0; x < 8; ++x)
        for (int y = 0; y < panel_cus / 8; ++y) {
            const int bit = C * x + 8 * y + x;
            pm[bit / 32] |= 1u << (bit % 32);
        }
    const uint32_t words = (uint32_t)((cus + 31) / 32);
    for (uint32_t w = 0; w < words; ++w) {
        const int valid = std::min(32, cus - (int)w * 32);
        const uint32_t all = valid == 32 ? 0xffffffffu : ((1u << valid) - 1);
        um[w] = all & ~pm[w];
    }
    if ((e = hipExtStreamCreateWithCUMask(&h->la_panel, words, pm)) != hipSuccess) { h->la_panel = nullptr; return e; }
    if ((e = hipExtStreamCreateWithCUMask(&h->la_update, words, um)) != hipSuccess) {
        (void)hipStreamDestroy(h->la_panel);
        h->la_panel = h->la_update = nullptr;
        return e;
    }
    h->la_cus = panel_cus;
    return hipSuccess;
}

// Ozaki-split gemm: C = alpha*A*B + beta*C with A (m x k) and B (k x n)
// sliced into digits on the INT8 matrix cores (scratch from the handle).
int ozaki_gemm(mfla_handle h, int N, int64_t m, int64_t n, int64_t k, const DV &A, const DV &B, const DV &C, const double *alpha,
               const double *beta) {
    const int64_t al = mfk::oz_row_align(), ld = (k + al - 1) / al * al;
    if (ld > 8192) return fail(MFLA_ERR_ARGUMENT, "gemm_strategy=:ozaki supports reduction lengths up to 8192");
    const int S = mfk::oz_digits(N), U = S;
    const int64_t mx = std::max(m, n);
    size_t off = 0;
    auto take = [&](size_t bytes) { size_t o = off; off += (bytes + 255) / 256 * 256; return o; };
    const size_t o_ea = take(sizeof(int) * m), o_eb = take(sizeof(int) * n), o_part = take(sizeof(double) * mfk::oz_chunks() * mx);
    const size_t o_da = take((size_t)S * m * ld), o_db = take((size_t)S * n * ld), o_pl = take(sizeof(int32_t) * (size_t)(U + 1) * m * n);
    char *base = (char *)grow_scratch(h, off);
    if (!base) return fail(MFLA_ERR_HIP, "scratch allocation failed (%zu bytes)", off);
    int *ea = (int *)(base + o_ea), *eb = (int *)(base + o_eb);
    double *part = (double *)(base + o_part);
    int8_t *da = (int8_t *)(base + o_da), *db = (int8_t *)(base + o_db);
    int32_t *pl = (int32_t *)(base + o_pl);
    hipStream_t s = h->stream;
    const DV Bt{B.p, B.cs, B.rs, B.ls};                 // Bt(j, kk) = B(kk, j)
    HIPCHK(mfk::oz_exponents(s, m, k, A, 0, ea, part));
    HIPCHK(mfk::oz_exponents(s, n, k, Bt, 0, eb, part));
    HIPCHK(mfk::oz_split(N, s, m, k, A, 0, ea, S, da, ld, (int64_t)m * ld));
    HIPCHK(mfk::oz_split(N, s, n, k, Bt, 0, eb, S, db, ld, (int64_t)n * ld));
    HIPCHK(mfk::oz_gemm(s, m, n, ld, da, ld, (int64_t)m * ld, S, db, ld, (int64_t)n * ld, S, U, pl));
    HIPCHK(mfk::oz_combine(N, s, m, n, U, pl, ea, eb, C, alpha, beta, scalar_is_zero(beta, N) ? 1 : 0, 0));
    return MFLA_OK;
}

// Kernel-side symmetric-access codes: 0 general, 1 lower, 2 upper.
int kernel_uplo(int uplo) { return uplo == MFLA_GENERAL ? 0 : (uplo == MFLA_LOWER ? 1 : 2); }

bool trsm_exact(const mfla_config &c, int64_t n) {
    if (c.trsm_order == MFLA_TRSM_ORDER_EXACT) return true;
    if (c.trsm_order == MFLA_TRSM_ORDER_BLOCKED) return false;
    return n <= 256;
}

#define HIPRET_(expr)                       \
    do {                                    \
        hipError_t e_ = (expr);             \
        if (e_ != hipSuccess) return e_;    \
    } while (0)

// Right-hand sides up to which a Float64x4 trsm_core look-ahead runs on the
// CU-partitioned streams, up to which a forward one streams the next block's
// rows, and up to which the diagonal kernels wait on a device counter.
constexpr int64_t kTrsmSplitMaxRhs = 64;
constexpr int64_t kTrsmStreamMaxRhs = 64;
constexpr int64_t kTrsmWaitMaxRhs = 256;

// trsm! core (trsm.jl:247-301) after argument checks; alpha scaling included.
hipError_t trsm_core(mfla_handle h, int N, int side, int uplo, int trans, int diag, const mfla_matrix &B, const mfla_matrix &A,
                     const double *alpha, const mfla_config &c, const Status *st) {
    const int64_t n = A.rows;
    if (alpha && !scalar_is_one(alpha, N)) {
        hipError_t e = mfk::scale(N, h->stream, B.rows, B.cols, dvn(B), alpha, st);
        if (e != hipSuccess) return e;
    }
    const bool transposed = trans == MFLA_TRANS;
    const bool eff_lower = ((uplo == MFLA_LOWER) == !transposed);
    mfk::TrsmArgs a{};
    a.n = n;
    a.unit = diag == MFLA_UNIT;
    a.exact_order = trsm_exact(c, n) ? 1 : 0;
    a.swap_product = side == MFLA_RIGHT ? 1 : 0;   // B[i,k]*coeff(k,j) (trsm.jl:136-172)
    DV X, CA;
    if (side == MFLA_LEFT) {
        X = dvn(B); a.nrhs = B.cols;
        CA = transposed ? dvt(A) : dvn(A);
        a.forward = eff_lower;
    } else {
        X = dvt(B); a.nrhs = B.rows;
        CA = transposed ? dvn(A) : dvt(A);
        a.forward = !eff_lower;
    }
    // Four limbs, few right-hand sides: with look-ahead the diagonal blocks
    // run on the handle's 32-CU stream and the far updates on the stream over
    // the other CUs, so neither shares a SIMD with the other (x4 4096 x 64:
    // 14.1 -> 13.5 ms per triangle). With 256 right-hand sides (x4 19.9 ->
    // 24.8 ms) or fewer limbs the shared placement is faster.
    hipStream_t ms = h->stream;
    if (n > 3 * 64 && (a.forward || !a.exact_order) && ensure_side(h) == hipSuccess) {
        const bool split = h->la_panel && h->la_update && N == 4 && a.nrhs <= kTrsmSplitMaxRhs;
        a.side = split ? h->la_update : h->side; a.ev_main = h->ev_panel; a.ev_side = h->ev_update; a.ev_side2 = h->ev_update2;
        // Forward, few right-hand sides: the next block's rows streamed inside
        // each diagonal kernel (device words 1 and 2 of the handle's flag
        // array). Each kernel then ends no earlier than the side stream's
        // update from the block before, one block of slack instead of two:
        // a gain where the diagonal chain is the bound (4096 x 64: x4 13.4 ->
        // 12.6 ms, x2 4.3 -> 3.7 ms), a loss where the updates are (256
        // right-hand sides and more; x3, whose updates cost more beside its
        // chain: 7.5 -> 7.8 ms).
        a.stream_next = a.forward && a.nrhs <= kTrsmStreamMaxRhs && N != 3;
        // Up to 256 right-hand sides each diagonal kernel waits for the side
        // stream's updates on the same device words instead of the stream
        // waiting on events (4096 x 256: x4 forward 19.9 -> 19.6 ms, backward
        // 20.3 -> 19.7 ms; 64: x4 backward 14.0 -> 13.5 ms). With 1024 the
        // waiting workgroups hold slots the updates need (x4 61.5 -> 64.7 ms).
        if (a.stream_next || a.nrhs <= kTrsmWaitMaxRhs) {
            a.far_flag = reinterpret_cast<unsigned *>(h->flags) + 1;   // words 1-2
            a.start_sig = reinterpret_cast<unsigned *>(h->flags) + 5;  // word 5
            a.dedicated = split ? 1 : 0;                                // CU-masked streams: queues of their own
        }
        HIPRET_(hipEventRecord(h->ev_join, h->stream));
        HIPRET_(hipStreamWaitEvent(a.side, h->ev_join, 0));
        if (split) {
            ms = h->la_panel;
            HIPRET_(hipStreamWaitEvent(ms, h->ev_join, 0));
        }
    }
    HIPRET_(mfk::trsm_solve(N, ms, a, X, CA, dvn(A), st));
    if (ms != h->stream) {
        HIPRET_(hipEventRecord(h->ev_join, ms));
        HIPRET_(hipStreamWaitEvent(h->stream, h->ev_join, 0));
    }
    return hipSuccess;
}

template <int N> void host_max_reduce(const double *colmax, int64_t count, double *out) {
    mf::V<N> mx = mf::zero<N>();
    for (int64_t c = 0; c < count; ++c) {
        mf::V<N> v;
        for (int l = 0; l < N; ++l) v.x[l] = colmax[l * count + c];
        mx = mf::vmax(mx, v);
    }
    for (int l = 0; l < N; ++l) out[l] = mx.x[l];
}
void host_max(int N, const double *colmax, int64_t count, double *out) {
    switch (N) {
        case 1: host_max_reduce<1>(colmax, count, out); break;
        case 2: host_max_reduce<2>(colmax, count, out); break;
        case 3: host_max_reduce<3>(colmax, count, out); break;
        default: host_max_reduce<4>(colmax, count, out); break;
    }
}

// original_maximum (types.jl:317-331): max |A| over all (region 0), lower
// (region 1) or upper (region 2) entries.
int original_maximum(mfla_handle h, int N, const mfla_matrix &A, int lower_only, double *out) {
    const int64_t cols = A.cols;
    for (int l = 0; l < N; ++l) out[l] = 0.0;
    if (A.rows * cols == 0) return MFLA_OK;
    double *buf = (double *)grow_scratch(h, sizeof(double) * N * cols);
    if (!buf) return fail(MFLA_ERR_HIP, "scratch allocation failed");
    HIPCHK(mfk::col_absmax_region(N, h->stream, dvn(A), A.rows, cols, lower_only, DV{buf, 1, cols, cols}));
    std::vector<double> host((size_t)N * cols);
    {
        void *dst_[1] = {host.data()};
        const void *src_[1] = {buf};
        const size_t by_[1] = {sizeof(double) * N * (size_t)cols};
        HIPCHK(read_back(h, h->stream, 1, dst_, src_, by_));
    }
    host_max(N, host.data(), cols, out);
    return MFLA_OK;
}

template <int N> void ldlt_alpha_t(double *out) {
    // alpha = (1 + sqrt(17)) / 8 (ldlt.jl:691)
    mf::V<N> a = mf::div(mf::add(mf::from<N>(1.0), mf::sqrt(mf::from<N>(17.0))), mf::from<N>(8.0));
    for (int l = 0; l < N; ++l) out[l] = a.x[l];
}
void ldlt_alpha(int N, double *out) {
    switch (N) {
        case 1: ldlt_alpha_t<1>(out); break;
        case 2: ldlt_alpha_t<2>(out); break;
        case 3: ldlt_alpha_t<3>(out); break;
        default: ldlt_alpha_t<4>(out); break;
    }
}

template <int N>
void backward_error_host(const double *an, const double *xn, const double *bn, const double *rn, bool finite, double *out) {
    // src/residual.jl:178-200
    auto put = [&](const mf::V<N> &v) { for (int l = 0; l < N; ++l) out[l] = v.x[l]; };
    if (!finite) { put(mf::from<N>(__builtin_nan(""))); return; }
    mf::V<N> A = host_v<N>(an), X = host_v<N>(xn), B = host_v<N>(bn), R = host_v<N>(rn);
    mf::V<N> ps = mf::vmax(A, B);
    if (mf::is_zero(ps)) { put(mf::is_zero(R) ? mf::zero<N>() : mf::from<N>(__builtin_inf())); return; }
    mf::V<N> ss = mf::vmax(X, mf::from<N>(1.0));
    mf::V<N> den = mf::add(mf::mul(mf::div(A, ps), mf::div(X, ss)), mf::div(mf::div(B, ps), ss));
    if (mf::is_zero(den)) { put(mf::is_zero(R) ? mf::zero<N>() : mf::from<N>(__builtin_inf())); return; }
    mf::V<N> sr = mf::div(mf::div(R, ps), ss);
    put(mf::div(sr, den));
}

}  // namespace

// ======================================================================
extern "C" {

int mfla_version(void) { return 1; }
const char *mfla_last_error(void) { return g_err.c_str(); }

void mfla_config_default(mfla_config *c) {
    std::memset(c, 0, sizeof(*c));
    c->reduction_tile = 64;
    c->column_tile = 16;
    c->cholesky_block = 16;
    c->lu_block = 16;
    c->ldlt_block = 0;
    c->ldlt_strategy = MFLA_LDLT_AUTO;
    c->gemm_strategy = MFLA_GEMM_AUTO;
    c->ldlt_blocked_crossover = 512;
    c->thread_count = 1;
    c->gemm_panel_columns = 0;
    c->gemm_micro_columns = 0;
    c->trsm_order = MFLA_TRSM_ORDER_AUTO;
    c->gemm_tile = -1;
    c->gemm_packed_crossover = INT64_MAX;
}

int mfla_handle_create(mfla_handle *out) {
    if (!out) return fail(MFLA_ERR_ARGUMENT, "null handle pointer");
    auto *h = new mfla_handle_s();
    if (hipGetDevice(&h->device) != hipSuccess) { delete h; return fail(MFLA_ERR_HIP, "no HIP device"); }
    resolve_scratch(h->device);
    if (hipMalloc(&h->status, sizeof(Status)) != hipSuccess) { delete h; return fail(MFLA_ERR_HIP, "status allocation failed"); }
    if (hipMalloc(&h->flags, 64 * sizeof(int)) != hipSuccess) { (void)hipFree(h->status); delete h; return fail(MFLA_ERR_HIP, "flag allocation failed"); }
    void *pin = nullptr;
    if (hipHostMalloc(&pin, mfla_handle_s::kPinnedBytes, hipHostMallocDefault) == hipSuccess) h->pinned = static_cast<unsigned char *>(pin);
    else (void)hipGetLastError();                  // pageable reads instead
    // the look-ahead streams now, so that no call pays for their queues
    // (a device where they cannot be made runs without look-ahead)
    if (ensure_side(h) != hipSuccess || ensure_la_streams(h, kLdltPanelCUs) != hipSuccess) (void)hipGetLastError();
    *out = h;
    return MFLA_OK;
}

int mfla_handle_destroy(mfla_handle h) {
    if (!h) return MFLA_OK;
    (void)hipStreamSynchronize(h->stream);
    (void)hipFree(h->status);
    (void)hipFree(h->flags);
    if (h->pinned) (void)hipHostFree(h->pinned);
    if (h->scratch) (void)hipFree(h->scratch);
    if (h->side) {
        (void)hipStreamSynchronize(h->side);
        (void)hipStreamDestroy(h->side);
        (void)hipEventDestroy(h->ev_join);
        (void)hipEventDestroy(h->ev_panel);
        (void)hipEventDestroy(h->ev_update);
        (void)hipEventDestroy(h->ev_update2);
    }
    for (hipStream_t q : {h->la_panel, h->la_update})
        if (q) { (void)hipStreamSynchronize(q); (void)hipStreamDestroy(q); }
    delete h;
    return MFLA_OK;
}

int mfla_handle_set_stream(mfla_handle h, void *stream) {
    if (!h) return fail(MFLA_ERR_ARGUMENT, "null handle");
    std::lock_guard<std::mutex> g(h->lock);
    h->stream = (hipStream_t)stream;
    return MFLA_OK;
}

int mfla_synchronize(mfla_handle h) {
    HIPCHK(hipStreamSynchronize(h->stream));
    return MFLA_OK;
}

// ------------------------------------------------------------ workspace
int mfla_workspace_create(mfla_handle h, int limbs, int64_t factor_capacity, int64_t ldlt_block_capacity, mfla_workspace *out) {
    if (check_limbs(limbs)) return MFLA_ERR_ARGUMENT;
    if (factor_capacity < 0) return fail(MFLA_ERR_ARGUMENT, "factor_capacity must be nonnegative");
    if (ldlt_block_capacity < 0) return fail(MFLA_ERR_ARGUMENT, "ldlt_block_capacity must be nonnegative");
    auto *w = new mfla_workspace_s();
    w->h = h;
    w->limbs = limbs;
    *out = w;
    int rc = mfla_workspace_ensure(w, factor_capacity, ldlt_block_capacity);
    if (rc != MFLA_OK) { delete w; *out = nullptr; }
    return rc;
}

int mfla_workspace_destroy(mfla_workspace w) {
    if (!w) return MFLA_OK;
    if (w->weighted) { (void)hipStreamSynchronize(w->h->stream); (void)hipFree(w->weighted); }
    delete w;
    return MFLA_OK;
}

int mfla_workspace_capacity(mfla_workspace w, int64_t *out2) {
    out2[0] = w->factor_capacity;
    out2[1] = w->block_capacity;
    return MFLA_OK;
}

int mfla_workspace_ensure(mfla_workspace w, int64_t factor_capacity, int64_t ldlt_block_capacity) {
    if (factor_capacity < 0 || ldlt_block_capacity < 0) return fail(MFLA_ERR_ARGUMENT, "workspace capacities must be nonnegative");
    const int64_t f = std::max(w->factor_capacity, factor_capacity);
    const int64_t b = std::max(w->block_capacity, ldlt_block_capacity);
    if (f == w->factor_capacity && b == w->block_capacity) return MFLA_OK;
    // three parts: ldlt!'s look-ahead keeps the two previous panels'
    // weighted rows while the next panel runs
    const size_t bytes = 3 * sizeof(double) * (size_t)w->limbs * (size_t)f * (size_t)(b > 0 ? b + 1 : 0);
    double *p = nullptr;
    if (bytes > 0 && hipMalloc(&p, bytes) != hipSuccess) return fail(MFLA_ERR_HIP, "workspace allocation failed");
    if (w->weighted) { (void)hipStreamSynchronize(w->h->stream); (void)hipFree(w->weighted); }
    w->weighted = p;
    w->factor_capacity = f;
    w->block_capacity = b;
    return MFLA_OK;
}

// ------------------------------------------------------------ arithmetic
int mfla_arith(mfla_handle h, int N, int op, int64_t count, const double *x, const double *y, const double *z, double *out) {
    if (check_limbs(N)) return MFLA_ERR_ARGUMENT;
    if (op < 0 || op > 12 || (op == 6 && N != 3)) return fail(MFLA_ERR_ARGUMENT, "unsupported arithmetic op %d for %d limbs", op, N);
    HIPCHK(mfk::arith(N, h->stream, op, count, x, y, z, out));
    return MFLA_OK;
}

int mfla_host_arith(int N, int op, int64_t count, const double *x, const double *y, const double *z, double *out) {
    if (check_limbs(N)) return MFLA_ERR_ARGUMENT;
    if (op < 0 || op > 12 || (op == 6 && N != 3)) return fail(MFLA_ERR_ARGUMENT, "unsupported arithmetic op %d for %d limbs", op, N);
    return mfk::host_arith(N, op, count, x, y, z, out) == 0 ? MFLA_OK : fail(MFLA_ERR_ARGUMENT, "arithmetic failed");
}

// types.jl:298-331 helpers (_all_finite, _lower_triangle_finite,
// _maximum_abs, _lower_maximum_abs) and the upper-triangle scan of
// diagnostics.jl:74-81. region: 0 all, 1 lower (i >= j), 2 upper (i <= j).
int mfla_absmax(mfla_handle h, int N, int region, mfla_matrix A, double *out) {
    if (check_limbs(N)) return MFLA_ERR_ARGUMENT;
    if (check_desc(A, "A")) return MFLA_ERR_ARGUMENT;
    if (region < 0 || region > 2) return fail(MFLA_ERR_ARGUMENT, "region must be 0, 1 or 2");
    std::lock_guard<std::mutex> g(h->lock);
    return original_maximum(h, N, A, region, out);
}

int mfla_all_finite(mfla_handle h, int N, int region, mfla_matrix A, int *out) {
    if (check_limbs(N)) return MFLA_ERR_ARGUMENT;
    if (check_desc(A, "A")) return MFLA_ERR_ARGUMENT;
    if (region < 0 || region > 2) return fail(MFLA_ERR_ARGUMENT, "region must be 0, 1 or 2");
    std::lock_guard<std::mutex> g(h->lock);
    HIPCHK(mfk::status_reset(h->stream, h->status));
    HIPCHK(mfk::finite_check(N, h->stream, A.rows, A.cols, dvn(A), region, h->status));
    const int64_t r = read_info(h);
    if (r == INT64_MIN) return fail(MFLA_ERR_HIP, "status read failed");
    *out = r == 0 ? 1 : 0;
    return MFLA_OK;
}

int mfla_copy(mfla_handle h, int N, mfla_matrix dst, mfla_matrix src) {
    if (check_limbs(N)) return MFLA_ERR_ARGUMENT;
    if (dst.rows != src.rows || dst.cols != src.cols) return fail(MFLA_ERR_DIMENSION, "copyto! dimensions differ");
    if (check_desc(dst, "dst") || check_desc(src, "src")) return MFLA_ERR_ARGUMENT;
    std::lock_guard<std::mutex> g(h->lock);
    HIPCHK(mfk::copy(N, h->stream, dst.rows, dst.cols, dvn(dst), dvn(src)));
    return MFLA_OK;
}

// ------------------------------------------------------------ conversions
int mfla_from_interleaved(mfla_handle h, int limbs, mfla_matrix dst, const double *src, int64_t src_ld) {
    if (check_limbs(limbs)) return MFLA_ERR_ARGUMENT;
    HIPCHK(mfk::aos_to_planar(limbs, h->stream, dst.rows, dst.cols, dvn(dst), src, src_ld));
    return MFLA_OK;
}

int mfla_to_interleaved(mfla_handle h, int limbs, double *dst, int64_t dst_ld, mfla_matrix src) {
    if (check_limbs(limbs)) return MFLA_ERR_ARGUMENT;
    HIPCHK(mfk::planar_to_aos(limbs, h->stream, src.rows, src.cols, dst, dst_ld, dvn(src)));
    return MFLA_OK;
}

// ------------------------------------------------------------ BLAS 3
int mfla_gemm_plan(int N, int64_t m, int64_t k, int64_t n, const mfla_config *config, mfla_gemm_plan_t *out) {
    mfla_config tmp;
    const mfla_config &c = cfg_or_default(config, tmp);
    if (!out) return fail(MFLA_ERR_ARGUMENT, "null plan pointer");
    if (check_limbs(N)) return MFLA_ERR_ARGUMENT;
    if (c.gemm_strategy < 0 || c.gemm_strategy > 4)
        return fail(MFLA_ERR_ARGUMENT, "gemm_strategy must be :auto, :direct, :packed, :fused, or :ozaki");
    // gemm_plan (gemm.jl:44-136): MFLA's CPU decomposition
    mfla_gemm_plan_t p{};
    const int64_t threads = c.thread_count;
    const int64_t base = N <= 1 ? 48 : N == 2 ? 32 : N == 3 ? 24 : 16;                       // gemm.jl:1-7
    const int64_t dpanel = threads >= 8 ? std::max<int64_t>(12, base * 3 / 4) : base;
    const int64_t panel = std::max<int64_t>(c.gemm_panel_columns > 0 ? c.gemm_panel_columns : dpanel, 1);
    const int64_t jobs = (std::max<int64_t>(n, 0) + panel - 1) / panel;
    const int64_t workers = std::max<int64_t>(1, std::min<int64_t>(threads, std::max<int64_t>(jobs, 1)));  // types.jl:249-250
    p.panel_columns = panel;
    p.workers = workers;
    if (c.gemm_strategy == MFLA_GEMM_FUSED) {
        if (N != 3) return fail(MFLA_ERR_ARGUMENT, "gemm_strategy=:fused is only available for Float64x3");
        p.strategy = MFLA_GEMM_FUSED; p.reason = MFLA_PLAN_FORCED_FUSED; p.micro_columns = 2;
    } else {
        const int64_t req = c.gemm_micro_columns > 0 ? c.gemm_micro_columns : (N <= 3 ? 4 : 2);   // gemm.jl:9-13
        if (req != 1 && req != 2 && req != 4) return fail(MFLA_ERR_ARGUMENT, "gemm_micro_columns must be 1, 2, or 4");
        const int64_t mc = std::min<int64_t>(req, panel);
        p.micro_columns = mc >= 4 ? 4 : mc >= 2 ? 2 : 1;
        if (c.gemm_strategy == MFLA_GEMM_DIRECT) {
            p.strategy = MFLA_GEMM_DIRECT; p.reason = MFLA_PLAN_FORCED_DIRECT;
        } else if (c.gemm_strategy == MFLA_GEMM_PACKED) {
            p.strategy = MFLA_GEMM_PACKED; p.reason = MFLA_PLAN_FORCED_PACKED;
            p.packed_elements_per_worker = std::max<int64_t>(k, 0) * panel;
        } else if (c.gemm_strategy == MFLA_GEMM_OZAKI) {
            p.strategy = MFLA_GEMM_OZAKI; p.reason = MFLA_PLAN_FORCED_OZAKI;
        } else {
            // volume-based crossover on near-square shapes (gemm.jl:110-127, :15-18)
            const int64_t crossover = std::max<int64_t>(c.gemm_packed_crossover, 1);
            const double work = (double)std::max<int64_t>(m, 0) * (double)std::max<int64_t>(k, 0) * (double)std::max<int64_t>(n, 0);
            const int64_t min_red = std::max<int64_t>(32, crossover / 2);
            const int64_t lo = std::min(m, std::min(k, n)), hi = std::max(m, std::max(k, n));
            const bool near_square = lo > 0 && hi <= 4 * lo;
            int reason;
            if (!near_square) reason = MFLA_PLAN_AUTO_OUTSIDE_CALIBRATED_SHAPE;
            else if (k < min_red) reason = MFLA_PLAN_AUTO_REDUCTION_TOO_SMALL;
            else if (work < (double)crossover * (double)crossover * (double)crossover) reason = MFLA_PLAN_AUTO_BELOW_CROSSOVER;
            else reason = MFLA_PLAN_AUTO_ABOVE_CROSSOVER;
            const bool packed = reason == MFLA_PLAN_AUTO_ABOVE_CROSSOVER;
            p.strategy = packed ? MFLA_GEMM_PACKED : (N == 3 ? MFLA_GEMM_FUSED : MFLA_GEMM_DIRECT);
            p.reason = packed ? reason : (N == 3 ? MFLA_PLAN_AUTO_FUSED_DIRECT : reason);
            p.packed_elements_per_worker = packed ? std::max<int64_t>(k, 0) * panel : 0;
        }
    }
    // the device launch gemm! makes (gemm_launch): the x3 mulacc route splits
    // off the m mod 4 remainder rows onto the standard network
    const bool fused = N == 3 && p.strategy == MFLA_GEMM_FUSED;
    const int64_t mm = std::max<int64_t>(m, 0), nn = std::max<int64_t>(n, 0);
    const int64_t main_rows = fused ? (mm / 4) * 4 : mm;
    int tr = 0, tc = 0;
    p.network = fused ? 1 : 0;
    if (p.strategy == MFLA_GEMM_OZAKI) {
        p.tile = -1; p.launches = 0; p.workgroups = 0;
    } else {
        p.tile = mfk::tile_geometry(N, main_rows, nn, k, false, c.gemm_tile, &tr, &tc);
        p.tile_rows = tr; p.tile_cols = tc;
        p.launches = (main_rows > 0 && nn > 0) ? 1 : 0;
        p.workgroups = mfk::cdiv(main_rows, tr) * mfk::cdiv(nn, tc);
        p.remainder_rows = mm - main_rows;
        if (p.remainder_rows > 0 && nn > 0) {
            int rr = 0, rc = 0;
            mfk::tile_geometry(N, p.remainder_rows, nn, k, false, c.gemm_tile, &rr, &rc);
            p.launches += 1;
            p.workgroups += mfk::cdiv(p.remainder_rows, rr) * mfk::cdiv(nn, rc);
        }
    }
    *out = p;
    return MFLA_OK;
}

int mfla_gemm(mfla_handle h, int N, mfla_matrix C, mfla_matrix A, mfla_matrix B, const double *alpha, const double *beta,
              const mfla_config *config) {
    mfla_config tmp;
    const mfla_config &c = cfg_or_default(config, tmp);
    const int64_t m = A.rows, k = A.cols, n = B.cols;
    if (B.rows != k) return fail(MFLA_ERR_DIMENSION, "gemm! inner dimensions differ");
    if (C.rows != m || C.cols != n) return fail(MFLA_ERR_DIMENSION, "gemm! output dimensions differ");
    if (check_limbs(N)) return MFLA_ERR_ARGUMENT;
    if (check_desc(A, "A") || check_desc(B, "B") || check_desc(C, "C")) return MFLA_ERR_ARGUMENT;
    if (extents_overlap(C, A, N) || extents_overlap(C, B, N)) return fail(MFLA_ERR_ALIAS, "gemm! does not support output/input aliasing");
    if (c.gemm_strategy < 0 || c.gemm_strategy > 4)
        return fail(MFLA_ERR_ARGUMENT, "gemm_strategy must be :auto, :direct, :packed, :fused, or :ozaki");
    if (c.gemm_strategy == MFLA_GEMM_FUSED && N != 3) return fail(MFLA_ERR_ARGUMENT, "gemm_strategy=:fused is only available for Float64x3");
    std::lock_guard<std::mutex> g(h->lock);
    if (c.gemm_strategy == MFLA_GEMM_OZAKI && m * n > 0 && k > 0)
        return ozaki_gemm(h, N, m, n, k, dvn(A), dvn(B), dvn(C), alpha ? alpha : kOne, beta ? beta : kZero);
    HIPCHK(gemm_launch(h, N, c.gemm_strategy, c.gemm_tile, m, n, k, dvn(A), dvn(B), dvn(C), alpha ? alpha : kOne, beta ? beta : kZero,
                       nullptr));
    return MFLA_OK;
}

int mfla_syrk(mfla_handle h, int N, int uplo, mfla_matrix C, mfla_matrix P, const double *alpha, const double *beta,
              const mfla_config *config) {
    mfla_config tmp;
    const mfla_config &c = cfg_or_default(config, tmp);
    if (uplo != MFLA_LOWER) return fail(MFLA_ERR_ARGUMENT, "syrk! currently supports only uplo=:lower or :L");
    const int64_t cols = P.cols;
    if (C.rows != cols || C.cols != cols) return fail(MFLA_ERR_DIMENSION, "syrk! output dimensions differ");
    if (check_limbs(N)) return MFLA_ERR_ARGUMENT;
    if (check_desc(C, "C") || check_desc(P, "P")) return MFLA_ERR_ARGUMENT;
    if (extents_overlap(C, P, N)) return fail(MFLA_ERR_ALIAS, "syrk! does not support output/input aliasing");
    std::lock_guard<std::mutex> g(h->lock);
    HIPCHK(lower_launch(h, N, c.gemm_tile, cols, P.rows, dvt(P), dvn(P), dvn(C), alpha ? alpha : kOne, beta ? beta : kZero, nullptr));
    return MFLA_OK;
}

int mfla_gemmt(mfla_handle h, int N, mfla_matrix C, mfla_matrix L, mfla_matrix R, const double *alpha, const double *beta,
               const mfla_config *config) {
    mfla_config tmp;
    const mfla_config &c = cfg_or_default(config, tmp);
    const int64_t rows = L.rows, red = L.cols;
    if (R.rows != rows || R.cols != red) return fail(MFLA_ERR_DIMENSION, "gemmt! input dimensions differ");
    if (C.rows != rows || C.cols != rows) return fail(MFLA_ERR_DIMENSION, "gemmt! output dimensions differ");
    if (check_limbs(N)) return MFLA_ERR_ARGUMENT;
    if (check_desc(C, "C") || check_desc(L, "L") || check_desc(R, "R")) return MFLA_ERR_ARGUMENT;
    if (extents_overlap(C, L, N) || extents_overlap(C, R, N)) return fail(MFLA_ERR_ALIAS, "gemmt! does not support output/input aliasing");
    std::lock_guard<std::mutex> g(h->lock);
    HIPCHK(lower_launch(h, N, c.gemm_tile, rows, red, dvn(L), dvt(R), dvn(C), alpha ? alpha : kOne, beta ? beta : kZero, nullptr));
    return MFLA_OK;
}

static int check_tr_symbols(int side, int uplo, int trans, int diag) {
    if (side != MFLA_LEFT && side != MFLA_RIGHT) return fail(MFLA_ERR_ARGUMENT, "side must be :left or :right");
    if (uplo != MFLA_LOWER && uplo != MFLA_UPPER) return fail(MFLA_ERR_ARGUMENT, "uplo must be :lower or :upper");
    if (trans != MFLA_NOTRANS && trans != MFLA_TRANS) return fail(MFLA_ERR_ARGUMENT, "trans must be :N or :T");
    if (diag != MFLA_NONUNIT && diag != MFLA_UNIT) return fail(MFLA_ERR_ARGUMENT, "diag must be :unit or :nonunit");
    return MFLA_OK;
}

int mfla_trsm(mfla_handle h, int N, int side, int uplo, int trans, int diag, mfla_matrix B, mfla_matrix A, const double *alpha,
              const mfla_config *config) {
    mfla_config tmp;
    const mfla_config &c = cfg_or_default(config, tmp);
    if (A.rows != A.cols) return fail(MFLA_ERR_DIMENSION, "trsm! requires a square triangular matrix");
    if (check_tr_symbols(side, uplo, trans, diag)) return MFLA_ERR_ARGUMENT;
    if (side == MFLA_LEFT && B.rows != A.rows) return fail(MFLA_ERR_DIMENSION, "trsm! left dimensions differ");
    if (side == MFLA_RIGHT && B.cols != A.rows) return fail(MFLA_ERR_DIMENSION, "trsm! right dimensions differ");
    if (check_limbs(N)) return MFLA_ERR_ARGUMENT;
    if (check_desc(A, "A") || check_desc(B, "B")) return MFLA_ERR_ARGUMENT;
    if (extents_overlap(B, A, N)) return fail(MFLA_ERR_ALIAS, "trsm! does not support output/input aliasing");
    std::lock_guard<std::mutex> g(h->lock);
    HIPCHK(trsm_core(h, N, side, uplo, trans, diag, B, A, alpha ? alpha : kOne, c, nullptr));
    return MFLA_OK;
}

// ------------------------------------------------------------ BLAS 2
int mfla_trsv(mfla_handle h, int N, int uplo, int trans, int diag, mfla_matrix x, mfla_matrix A, const double *alpha,
              const mfla_config *config) {
    mfla_config tmp;
    const mfla_config &c = cfg_or_default(config, tmp);
    if (A.rows != A.cols) return fail(MFLA_ERR_DIMENSION, "trsv! requires a square triangular matrix");
    if (x.rows != A.rows || x.cols != 1) return fail(MFLA_ERR_DIMENSION, "trsv! right-hand side length differs");
    if (check_tr_symbols(MFLA_LEFT, uplo, trans, diag)) return MFLA_ERR_ARGUMENT;
    if (check_limbs(N)) return MFLA_ERR_ARGUMENT;
    if (check_desc(A, "A") || check_desc(x, "x")) return MFLA_ERR_ARGUMENT;
    if (extents_overlap(x, A, N)) return fail(MFLA_ERR_ALIAS, "trsv! does not support output/input aliasing");
    std::lock_guard<std::mutex> g(h->lock);
    HIPCHK(trsm_core(h, N, MFLA_LEFT, uplo, trans, diag, x, A, alpha ? alpha : kOne, c, nullptr));
    return MFLA_OK;
}

int mfla_gemv(mfla_handle h, int N, int trans, mfla_matrix y, mfla_matrix A, mfla_matrix x, const double *alpha, const double *beta,
              const mfla_config *) {
    const int64_t m = A.rows, n = A.cols;
    if (trans != MFLA_NOTRANS && trans != MFLA_TRANS) return fail(MFLA_ERR_ARGUMENT, "trans must be :N or :T");
    if (check_limbs(N)) return MFLA_ERR_ARGUMENT;
    if (check_desc(A, "A") || check_desc(x, "x") || check_desc(y, "y")) return MFLA_ERR_ARGUMENT;
    if (extents_overlap(y, A, N) || extents_overlap(y, x, N)) return fail(MFLA_ERR_ALIAS, "gemv! does not support output/input aliasing");
    if (x.rows != (trans ? m : n) || x.cols != 1) return fail(MFLA_ERR_DIMENSION, "gemv! input length differs");
    if (y.rows != (trans ? n : m) || y.cols != 1) return fail(MFLA_ERR_DIMENSION, "gemv! output length differs");
    const double *al = alpha ? alpha : kOne, *be = beta ? beta : kZero;
    std::lock_guard<std::mutex> g(h->lock);
    HIPCHK(mfk::gemv(N, h->stream, trans, m, n, dvn(A), dvn(x), dvn(y), al, be, scalar_is_zero(be, N) ? 1 : 0, nullptr));
    return MFLA_OK;
}

int mfla_symv(mfla_handle h, int N, int uplo, mfla_matrix y, mfla_matrix A, mfla_matrix x, const double *alpha, const double *beta,
              const mfla_config *) {
    const int64_t n = A.rows;
    if (A.rows != A.cols) return fail(MFLA_ERR_DIMENSION, "symv! requires a square matrix");
    if (x.rows != n || x.cols != 1) return fail(MFLA_ERR_DIMENSION, "symv! input length differs");
    if (y.rows != n || y.cols != 1) return fail(MFLA_ERR_DIMENSION, "symv! output length differs");
    if (uplo != MFLA_LOWER && uplo != MFLA_UPPER) return fail(MFLA_ERR_ARGUMENT, "uplo must be :lower or :upper");
    if (check_limbs(N)) return MFLA_ERR_ARGUMENT;
    if (check_desc(A, "A") || check_desc(x, "x") || check_desc(y, "y")) return MFLA_ERR_ARGUMENT;
    if (extents_overlap(y, A, N) || extents_overlap(y, x, N)) return fail(MFLA_ERR_ALIAS, "symv! does not support output/input aliasing");
    const double *al = alpha ? alpha : kOne, *be = beta ? beta : kZero;
    std::lock_guard<std::mutex> g(h->lock);
    HIPCHK(mfk::symv(N, h->stream, uplo == MFLA_UPPER, n, 1, dvn(A), dvn(x), dvn(y), al, be, scalar_is_zero(be, N) ? 1 : 0));
    return MFLA_OK;
}

// ------------------------------------------------------ factorizations
int mfla_potrf(mfla_handle h, int N, mfla_matrix A, int64_t *info, const mfla_config *config) {
    mfla_config tmp;
    const mfla_config &c = cfg_or_default(config, tmp);
    const int64_t n = A.rows;
    if (A.rows != A.cols) return fail(MFLA_ERR_DIMENSION, "cholesky! requires a square matrix");
    if (check_limbs(N)) return MFLA_ERR_ARGUMENT;
    if (check_desc(A, "A")) return MFLA_ERR_ARGUMENT;
    std::lock_guard<std::mutex> g(h->lock);
    hipStream_t s = h->stream;
    Status *st = h->status;
    HIPCHK(mfk::status_reset(s, st));
    HIPCHK(mfk::finite_check(N, s, n, n, dvn(A), 1, st));
    const int64_t block = std::max<int64_t>(c.cholesky_block, 1);
    const DV a = dvn(A);
    // Right-looking blocked Cholesky (cholesky.jl:68-107) with look-ahead:
    // after panel j's trsm, the next block column is updated on the main
    // stream (so panel j+1 can start) while the rest of the trailing SYRK
    // runs on a second stream. Every element still receives the panel
    // updates in ascending panel order with the same per-update arithmetic,
    // so the factor is bitwise identical to the sequential order.
    const bool lookahead = n > 4 * block;
    hipStream_t s1 = s;
    if (lookahead) {
        HIPCHK(ensure_side(h));
        s1 = h->side;
        HIPCHK(hipEventRecord(h->ev_join, s));
        HIPCHK(hipStreamWaitEvent(s1, h->ev_join, 0));
    }
    for (int64_t first = 0; first < n; first += block) {
        const int64_t last = std::min(first + block, n) - 1, nb = last - first + 1;
        HIPCHK(mfk::chol_panel(N, s, a, first, last, st));
        if (last + 1 >= n) break;
        const int64_t rest = n - last - 1;
        // trsm!(A21, L11; side=:right, uplo=:lower, trans=:T) -> forward over panel columns
        mfk::TrsmArgs ta{};
        ta.n = nb; ta.nrhs = rest; ta.forward = 1; ta.unit = 0; ta.exact_order = 1; ta.swap_product = 1;
        const DV A21 = sub(a, last + 1, first), L11 = sub(a, first, first);
        const DV A21t{A21.p, A21.cs, A21.rs, A21.ls};
        HIPCHK(mfk::trsm_solve(N, s, ta, A21t, L11, L11, st));
        if (!lookahead) {
            // syrk!(A22, transpose(A21), -1, 1)
            HIPCHK(lower_launch(h, N, c.gemm_tile, rest, nb, A21, A21t, sub(a, last + 1, last + 1), kMinusOne, kOne, st));
            continue;
        }
        HIPCHK(hipEventRecord(h->ev_panel, s));
        // next block column (all rows below its diagonal), after the side
        // stream's update from the previous panel reached it
        const int64_t nb2 = std::min(block, rest);
        if (first > 0) HIPCHK(hipStreamWaitEvent(s, h->ev_update, 0));
        mfk::TileArgs t{};
        t.mode = mfk::MODE_AB; t.net = mf::NET_STD; t.lower = 1; t.beta_zero = 0; t.tile = c.gemm_tile;
        t.row_lo = 0; t.rows = rest; t.n = nb2; t.K = nb;
        HIPCHK(mfk::tile_launch_n(N, s, t, A21, A21t, sub(a, last + 1, last + 1), kMinusOne, kOne, st));
        if (rest > nb2) {
            // the remaining trailing matrix on the side stream; it skips its
            // work only for failures in panels up to this one (a failure in
            // a later panel happens after this update in MFLA's order)
            HIPCHK(hipStreamWaitEvent(s1, h->ev_panel, 0));
            const int64_t rest2 = rest - nb2;
            const DV B21 = sub(A21, nb2, 0);
            mfk::TileArgs u{};
            u.mode = mfk::MODE_AB; u.net = mf::NET_STD; u.lower = 1; u.beta_zero = 0; u.tile = c.gemm_tile;
            u.row_lo = 0; u.rows = rest2; u.n = rest2; u.K = nb; u.status_limit = last + 1;
            HIPCHK(mfk::tile_launch_n(N, s1, u, B21, DV{B21.p, B21.cs, B21.rs, B21.ls}, sub(a, last + 1 + nb2, last + 1 + nb2),
                                      kMinusOne, kOne, st));
            HIPCHK(hipEventRecord(h->ev_update, s1));
        }
    }
    if (lookahead) {
        HIPCHK(hipEventRecord(h->ev_join, s1));
        HIPCHK(hipStreamWaitEvent(s, h->ev_join, 0));
    }
    const int64_t r = read_info(h);
    if (r == INT64_MIN) return fail(MFLA_ERR_HIP, "status read failed");
    *info = r;
    return MFLA_OK;
}

int mfla_getrf(mfla_handle h, int N, mfla_matrix A, int64_t *ipiv, double *origmax, int64_t *info, const mfla_config *config,
               mfla_workspace) {
    mfla_config tmp;
    const mfla_config &c = cfg_or_default(config, tmp);
    const int64_t m = A.rows, n = A.cols, kmax = std::min(m, n);
    if (check_limbs(N)) return MFLA_ERR_ARGUMENT;
    if (check_desc(A, "A")) return MFLA_ERR_ARGUMENT;
    if (c.gemm_strategy < 0 || c.gemm_strategy > 3) return fail(MFLA_ERR_ARGUMENT, "gemm_strategy must be :auto, :direct, :packed, or :fused");
    std::lock_guard<std::mutex> g(h->lock);
    hipStream_t s = h->stream;
    Status *st = h->status;
    HIPCHK(mfk::status_reset(s, st));
    HIPCHK(mfk::iota1(s, ipiv, kmax));
    HIPCHK(mfk::finite_check(N, s, m, n, dvn(A), 0, st));
    double om[4] = {0, 0, 0, 0};
    if (int rc = original_maximum(h, N, A, 0, om)) return rc;
    const int64_t block = std::max<int64_t>(c.lu_block, 1);
    const DV a = dvn(A);
    // Blocked right-looking LU (lu.jl:128-195). Row interchanges reach the
    // columns outside the current panel once the panel is done (lu_laswp),
    // before anything reads them. With look-ahead, the trsm!/gemm! update of
    // the next panel's columns runs on the main stream and the rest of the
    // trailing update on a second stream while the next panel is factored;
    // every element receives the same operations in the same order as in
    // MFLA's sequence, so factors and pivots are bitwise identical.
    const bool lookahead = n - std::min<int64_t>(block, n) > 4 * block && kmax > block;
    hipStream_t s1 = s;
    if (lookahead) {
        HIPCHK(ensure_side(h));
        s1 = h->side;
        HIPCHK(hipEventRecord(h->ev_join, s));
        HIPCHK(hipStreamWaitEvent(s1, h->ev_join, 0));
    }
    bool side_pending = false;
    // trsm!(U12, L11; side=:left, uplo=:lower, trans=:N, diag=:unit) and
    // gemm!(A22, A21, U12, -1, 1) on the columns [c0, c0 + cols)
    auto update = [&](hipStream_t st_, int64_t first, int64_t last, int64_t c0, int64_t cols, int64_t limit) -> hipError_t {
        const int64_t nb = last - first + 1;
        mfk::TrsmArgs ta{};
        ta.n = nb; ta.nrhs = cols; ta.forward = 1; ta.unit = 1; ta.exact_order = 1; ta.status_limit = limit;
        const DV U12 = sub(a, first, c0), L11 = sub(a, first, first);
        hipError_t e = mfk::trsm_solve(N, st_, ta, U12, L11, L11, st);
        if (e != hipSuccess || last + 1 >= m) return e;
        return gemm_launch(h, N, c.gemm_strategy, c.gemm_tile, m - last - 1, cols, nb, sub(a, last + 1, first), U12, sub(a, last + 1, c0),
                           kMinusOne, kOne, st, st_, limit);
    };
    for (int64_t first = 0; first < kmax; first += block) {
        const int64_t last = std::min(first + block, kmax) - 1;
        for (int64_t k = first; k <= last; ++k) {
            HIPCHK(mfk::lu_pivot_swap(N, s, a, k, m, first, last + 1, ipiv, st));
            HIPCHK(mfk::lu_scale_update(N, s, a, k, last, m, st));
        }
        if (side_pending) {   // the previous trailing update owns the columns right of the next panel
            HIPCHK(hipStreamWaitEvent(s, h->ev_update, 0));
            side_pending = false;
        }
        HIPCHK(mfk::lu_laswp(N, s, a, first, last - first + 1, ipiv, 0, first, last + 1, n - last - 1));
        if (last + 1 >= n) continue;
        const int64_t rest = n - last - 1, nxt = std::min(block, rest);
        if (!lookahead || rest <= nxt || last + 1 >= kmax) {
            HIPCHK(update(s, first, last, last + 1, rest, 0));
            continue;
        }
        HIPCHK(hipEventRecord(h->ev_panel, s));
        HIPCHK(update(s, first, last, last + 1, nxt, 0));
        // the rest on the side stream; it skips its work only for failures in
        // panels up to this one (a later failure comes after this update in
        // MFLA's order)
        HIPCHK(hipStreamWaitEvent(s1, h->ev_panel, 0));
        HIPCHK(update(s1, first, last, last + 1 + nxt, rest - nxt, last + 1));
        HIPCHK(hipEventRecord(h->ev_update, s1));
        side_pending = true;
    }
    if (side_pending) HIPCHK(hipStreamWaitEvent(s, h->ev_update, 0));
    if (lookahead) {
        HIPCHK(hipEventRecord(h->ev_join, s1));
        HIPCHK(hipStreamWaitEvent(s, h->ev_join, 0));
    }
    const int64_t r = read_info(h);
    if (r == INT64_MIN) return fail(MFLA_ERR_HIP, "status read failed");
    *info = r;
    for (int l = 0; l < N; ++l) origmax[l] = (r == -1) ? 0.0 : om[l];
    return MFLA_OK;
}

// ldlt_plan (ldlt.jl:329-350); panel widths 16 / 12 / 8 (ldlt_tuning.jl:82-88)
static int ldlt_plan_core(int N, int64_t n, const mfla_config &c, mfla_ldlt_plan_t *p) {
    if (c.ldlt_strategy < 0 || c.ldlt_strategy > 2) return fail(MFLA_ERR_ARGUMENT, "ldlt_strategy must be :auto, :unblocked, or :blocked");
    const int64_t b = std::max<int64_t>(c.ldlt_block > 0 ? c.ldlt_block : (N <= 2 ? 16 : (N == 3 ? 12 : 8)), 1);
    if (c.ldlt_strategy == MFLA_LDLT_UNBLOCKED) {
        *p = {MFLA_LDLT_UNBLOCKED, MFLA_PLAN_FORCED_UNBLOCKED, 1};
    } else if (c.ldlt_strategy == MFLA_LDLT_BLOCKED) {
        *p = {MFLA_LDLT_BLOCKED, MFLA_PLAN_FORCED_BLOCKED, std::max<int64_t>(b, 2)};
    } else if (b > 1 && n >= std::max<int64_t>(c.ldlt_blocked_crossover, 1)) {
        *p = {MFLA_LDLT_BLOCKED, MFLA_PLAN_AUTO_ABOVE_CROSSOVER, b};
    } else {
        *p = {MFLA_LDLT_UNBLOCKED, MFLA_PLAN_AUTO_BELOW_CROSSOVER, 1};
    }
    return MFLA_OK;
}

int mfla_ldlt_plan(int N, int64_t n, const mfla_config *config, mfla_ldlt_plan_t *out) {
    mfla_config tmp;
    const mfla_config &c = cfg_or_default(config, tmp);
    if (!out) return fail(MFLA_ERR_ARGUMENT, "null plan pointer");
    if (check_limbs(N)) return MFLA_ERR_ARGUMENT;
    return ldlt_plan_core(N, n, c, out);
}

}  // extern "C"

// Layout of the no-allocation ldlt!'s device scratch (in doubles): the
// column maxima (limbs * n) at 0, their fold (limbs) at *fold_off, the
// blocked route's weighted panel (limbs * n * (block + 1)) at *panel_off;
// returns the total.
static size_t noalloc_layout(int N, int64_t n, const mfla_ldlt_plan_t &plan, size_t *fold_off, size_t *panel_off) {
    const size_t colmax_len = (size_t)N * (size_t)n;
    *fold_off = (colmax_len + 31) / 32 * 32;
    *panel_off = *fold_off + 32;
    const size_t panel_len = plan.strategy == MFLA_LDLT_BLOCKED ? (size_t)N * (size_t)n * (size_t)(plan.block_size + 1) : 0;
    return *panel_off + panel_len;
}

// Rows below a blocked-LDLT panel per workgroup of its multi-workgroup
// kernel. The environment variable MFLA_LDLT_PANEL_WORKGROUPS caps the
// workgroups (0: the 512-thread single-workgroup kernel); every choice gives
// the same bits.
constexpr int64_t kLdltPanelRows = 192;
constexpr int64_t kGridTimeoutInfo = -2;
constexpr int64_t kLdltSpecMaxDepth = 32;
static int64_t ldlt_panel_workgroup_cap() {
    const char *v = std::getenv("MFLA_LDLT_PANEL_WORKGROUPS");
    if (!v || !*v) return 64;
    const long long g = std::atoll(v);
    return g < 0 ? 0 : (g > 64 ? 64 : g);
}
// The environment variable MFLA_LDLT_LOOKAHEAD: 0 never, 1 whenever the
// call can, otherwise from kLdltLookaheadMinN on (the bits never change).
constexpr int64_t kLdltLookaheadMinN = 1536;
static int ldlt_lookahead_mode() {
    const char *v = std::getenv("MFLA_LDLT_LOOKAHEAD");
    if (v && v[0] == '0' && !v[1]) return 0;
    if (v && v[0] == '1' && !v[1]) return 1;
    return 2;
}

// ldlt! shared by mfla_sytrf and mfla_sytrf_noalloc. static_buf == nullptr:
// the blocked route's weighted panel comes from the workspace (or is
// allocated for the call) and original_maximum folds the column maxima on
// the host. Otherwise static_buf (static_doubles long) holds the column
// maxima (limbs * n), their fold (limbs) and the weighted panel
// (limbs * n * (block + 1)), and nothing is allocated.
static int sytrf_run(mfla_handle h, int N, const mfla_matrix &A, const mfla_matrix &dsub, int64_t *pivots, uint8_t *blocks,
                     double *origmax, int64_t *info, const mfla_config &c, mfla_workspace ws, double *static_buf,
                     size_t static_doubles) {
    const int64_t n = A.rows;
    if (A.rows != A.cols) return fail(MFLA_ERR_DIMENSION, "ldlt! requires a square matrix");
    if (check_limbs(N)) return MFLA_ERR_ARGUMENT;
    if (check_desc(A, "A") || check_desc(dsub, "dsub")) return MFLA_ERR_ARGUMENT;
    if (dsub.rows != n || dsub.cols != 1) return fail(MFLA_ERR_DIMENSION, "ldlt! dsub length differs");
    mfla_ldlt_plan_t plan;
    if (ldlt_plan_core(N, n, c, &plan) != MFLA_OK) return MFLA_ERR_ARGUMENT;
    const bool blocked = plan.strategy == MFLA_LDLT_BLOCKED;
    const int64_t b = plan.block_size;
    size_t fold_off = 0, panel_off = 0;
    const size_t need = noalloc_layout(N, n, plan, &fold_off, &panel_off);
    if (static_buf && need > static_doubles)
        return fail(MFLA_ERR_WORKSPACE, "ldlt! without allocation needs %zu bytes of device scratch (limbs * n * (block + 2) "
                    "doubles plus 256 bytes), the library holds %zu", sizeof(double) * need, sizeof(double) * static_doubles);

    std::lock_guard<std::mutex> g(h->lock);
    hipStream_t s = h->stream;
    Status *st = h->status;
    HIPCHK(mfk::status_reset(s, st));
    HIPCHK(mfk::iota1(s, pivots, n));
    HIPCHK(mfk::fill_u8(s, blocks, n, 0));
    HIPCHK(mfk::zero_fill(N, s, n, 1, dvn(dsub)));
    HIPCHK(mfk::finite_check(N, s, n, n, dvn(A), 1, st));
    HIPCHK(mfk::mirror_lower(N, s, dvn(A), n, st));
    double om[4] = {0, 0, 0, 0};
    if (!static_buf) {
        if (int rc = original_maximum(h, N, A, 1, om)) return rc;
    } else if (n > 0) {
        // original_maximum without allocation: the column maxima and their
        // ordered fold on the device, N doubles read back
        HIPCHK(mfk::col_absmax_region(N, s, dvn(A), n, n, 1, DV{static_buf, 1, n, n}));
        HIPCHK(mfk::absmax_fold(N, s, DV{static_buf, 1, n, n}, n, static_buf + fold_off));
        void *dst_[1] = {om};
        const void *src_[1] = {static_buf + fold_off};
        const size_t by_[1] = {sizeof(double) * (size_t)N};
        HIPCHK(read_back(h, s, 1, dst_, src_, by_));
    }
    double alpha[4];
    ldlt_alpha(N, alpha);
    const DV a = dvn(A), d = dvn(dsub);
    if (!blocked) {
        HIPCHK(mfk::ldlt_unblocked(N, s, a, n, d, pivots, blocks, alpha, st));
    } else {
        // one part: the weighted panel (limbs * n * (block + 1) doubles),
        // which the multi-workgroup panels also use for their keep sets and
        // partials; the look-ahead rotates three parts
        const size_t half = (size_t)N * (size_t)n * (size_t)(b + 1);
        double *W = nullptr;
        size_t avail = 0;
        bool own = false;
        if (static_buf) {
            W = static_buf + panel_off;
            avail = static_doubles - panel_off;
        } else if (ws && ws->limbs == N && ws->factor_capacity >= n && ws->block_capacity >= b) {
            W = ws->weighted;
            avail = 3 * (size_t)N * (size_t)ws->factor_capacity * (size_t)(ws->block_capacity + 1);
        }
        // the panel kernel runs over one workgroup per kLdltPanelRows rows
        // below the panel (at most the MFLA_LDLT_PANEL_WORKGROUPS cap) when
        // the weighted-panel buffer holds its keep sets and partials
        const bool grid_fits = mfk::ldlt_panel_grid_doubles(N, n) <= half;
        const int64_t gcap = ldlt_panel_workgroup_cap();
        const bool use_grid = grid_fits && gcap > 0;     // its panels also write the weighted rows
        auto workgroups = [&](int64_t pf) { return std::max<int64_t>(1, std::min(gcap, mfk::cdiv(n - pf - 1, kLdltPanelRows))); };
        // Look-ahead: while it splits, the panels run on a stream confined to
        // a few CUs and each panel's trailing update is cut in three: the
        // next panel's columns (the panel stream, before the next panel), the
        // columns of the panel after it and the rest (a second stream over
        // the other CUs, beside the next panel). The next panel's columns then
        // wait only for the second part of the previous update, done long
        // before. Every element still receives the panels' updates in
        // ascending order with the same arithmetic, so the bits do not change.
        // The next panel must not read past its columns meanwhile: it runs in
        // truncation mode and stops at its first column that needs the
        // row-max pass (a truncation stop); the host then restores the panel's
        // columns and runs it again after the updates, and splits no update
        // for the next la_hold panels, which run with their updates on the
        // call's stream as without look-ahead. The weighted panels rotate over
        // three buffers (panel j + 3 reuses panel j's after panel j's last
        // part). Not in the no-allocation call (creating the streams allocates).
        const int lam = ldlt_lookahead_mode();
        const int64_t g0 = use_grid ? workgroups(0) : 0;
        bool la = use_grid && !static_buf && lam != 0 && (lam == 1 || n >= kLdltLookaheadMinN) && g0 <= 32 && n > 2 * b;
        if (la) {
            if (ensure_side(h) != hipSuccess || ensure_la_streams(h, kLdltPanelCUs) != hipSuccess) { (void)hipGetLastError(); la = false; }
        }
        if (!W) {
            const size_t len = (la ? 3 : 1) * half;
            if (hipMalloc(&W, sizeof(double) * len) != hipSuccess) return fail(MFLA_ERR_HIP, "weighted panel allocation failed");
            own = true;
            avail = len;
        }
        if (avail < 3 * half) la = false;
        hipStream_t ps = s, us = s;                      // while splitting: panels, trailing updates
        // `to` waits for the work launched on `from` so far
        auto order = [&](hipStream_t from, hipStream_t to, hipEvent_t ev) -> hipError_t {
            if (from == to) return hipSuccess;
            const hipError_t e = hipEventRecord(ev, from);
            return e == hipSuccess ? hipStreamWaitEvent(to, ev, 0) : e;
        };
        // work launched on ps / us since the call's stream last waited for it
        bool ps_open = false, us_open = false;
        // the call's stream after every kernel launched so far
        auto gather = [&]() -> hipError_t {
            hipError_t e = hipSuccess;
            if (ps_open && (e = order(ps, s, h->ev_join)) != hipSuccess) return e;
            if (us_open && (e = order(us, s, h->ev_update2)) != hipSuccess) return e;
            ps_open = us_open = false;
            return e;
        };
        // Device words 3 and 4 of the handle's flag array: a counter the
        // update stream advances after each split update's second part, and
        // an error word. A narrow first part (k_cols) waits on the counter
        // for the previous panel's second part instead of its stream waiting
        // on an event (~11 us each even when satisfied); that part was
        // enqueued before it, on the other stream, so the wait ends.
        unsigned *la_ctr = reinterpret_cast<unsigned *>(h->flags) + 3;
        unsigned la_seq = 0;
        if (la) {
            ps = h->la_panel;
            us = h->la_update;
            HIPCHK(mfk::flag_launch(s, la_ctr, 0));
            HIPCHK(order(s, ps, h->ev_join));
            HIPCHK(order(s, us, h->ev_update2));
        }
        // The panels work on the lower triangle only (their interchanges move
        // lower-triangle elements, as MFLA's full swap does on a symmetric
        // block) and write row k's upper part when column k freezes, so the
        // trailing block is never mirrored. update: panel [pf, pl]'s trailing
        // update of columns [c0, c1), rows from c0 down, on stream q; limit:
        // the status limit (nonzero for the parts that may run beside later
        // panels: skipped only when this panel or an earlier one stopped).
        auto narrow = [&](int64_t c0, int64_t c1) { return c1 - c0 <= 8 && c.gemm_tile < 0; };
        auto update = [&](hipStream_t q, int64_t pf, int64_t pl, const double *Wh, int64_t c0, int64_t c1, int64_t limit,
                          unsigned wait_value = 0) -> hipError_t {
            if (c0 >= c1) return hipSuccess;
            const int64_t tfirst = pl + 1, tcount = n - tfirst;
            const DV Wn{const_cast<double *>(Wh) + (c0 - tfirst), tcount, 1, tcount * (b + 1)};
            if (narrow(c0, c1))
                return mfk::cols_launch_n(N, q, n - c0, c1 - c0, pl - pf + 1, true, sub(a, c0, pf), Wn, sub(a, c0, c0), kMinusOne, kOne, st,
                                          limit, wait_value ? la_ctr : nullptr, wait_value, la_ctr + 1);
            mfk::TileArgs t{};
            t.mode = mfk::MODE_AB; t.net = mf::NET_STD; t.lower = 1; t.beta_zero = 0; t.tile = c.gemm_tile;
            t.row_lo = 0; t.rows = n - c0; t.n = c1 - c0; t.K = pl - pf + 1; t.status_limit = limit;
            const DV Wt{const_cast<double *>(Wh) + (c0 - tfirst), tcount, 1, tcount * (b + 1)};
            return mfk::tile_launch_n(N, q, t, sub(a, c0, pf), Wt, sub(a, c0, c0), kMinusOne, kOne, st);
        };
        // the whole update (after the weighted panel when the panels do not
        // write it)
        auto trailing = [&](hipStream_t q, int64_t pf, int64_t pl, double *Wh) -> hipError_t {
            if (pl >= n - 1) return hipSuccess;
            const int64_t tfirst = pl + 1, tcount = n - tfirst;
            hipError_t e = hipSuccess;
            if (!use_grid && (e = mfk::ldlt_weighted(N, q, a, pf, pl, tfirst, tcount, d, blocks, DV{Wh, 1, tcount, tcount * (b + 1)}, st)) !=
                                 hipSuccess)
                return e;
            return update(q, pf, pl, Wh, tfirst, n, 0);
        };
        // The host launches the panels and their trailing updates ahead,
        // assuming each panel ends at its requested last column, and reads the
        // status after `depth` panels: a panel that extended (a 2x2 pivot at
        // its last column) or stopped recorded a stop code (is_panel_stop),
        // so the kernels launched after it did nothing; the loop then resumes
        // after it (or at it). The depth doubles after a read without a stop
        // (up to kLdltSpecMaxDepth) and drops to one after a stop.
        struct Launched { int64_t pfirst, reqlast; double *Wh; };
        std::vector<Launched> pending;
        int64_t depth = 4, pfirst = 0, hidx = 0;
        bool prev_split = false;                         // the previous update's later parts may still run
        int64_t la_hold = 0, la_backoff = 8, la_stops = 0;
        int rc = MFLA_OK;
        while (pfirst < n) {
            const int64_t reqlast = std::min(pfirst + b, n) - 1;
            double *Wh = la ? W + (size_t)(hidx % 3) * half : W;
            // beside the previous update's later parts: on the panel stream,
            // in truncation mode
            const hipStream_t pq = prev_split ? ps : s;
            if (use_grid) {
                HIPCHK(mfk::ldlt_panel_grid(N, pq, a, n, pfirst, reqlast, d, pivots, blocks, alpha, Wh, 1, prev_split ? 1 : 0, b + 1, st,
                                            (int)workgroups(pfirst)));
            } else {
                HIPCHK(mfk::ldlt_panel(N, pq, a, n, pfirst, reqlast, d, pivots, blocks, alpha, Wh, 1, st));
            }
            pending.push_back({pfirst, reqlast, Wh});
            const bool split = la && la_hold == 0 && reqlast + 1 + b < n;
            if (la_hold > 0) --la_hold;
            const int64_t t0c = reqlast + 1, t1c = std::min(t0c + b, n), t2c = std::min(t0c + 2 * b, n);
            if (!split) {
                // after the previous update's later parts; the next panel on
                // the call's stream
                const hipStream_t tq = prev_split ? us : s;
                HIPCHK(order(pq, tq, h->ev_panel));
                HIPCHK(trailing(tq, pfirst, reqlast, Wh));
                HIPCHK(order(tq, s, h->ev_update));
                us_open = us_open || tq == us;
            } else {
                // first part on the panel stream (after the previous update's
                // second part), the others on the update stream
                HIPCHK(order(pq, ps, h->ev_join));
                HIPCHK(order(pq, us, h->ev_panel));
                const bool dev_wait = prev_split && narrow(t0c, t1c);
                if (prev_split && !dev_wait) HIPCHK(hipStreamWaitEvent(ps, h->ev_update, 0));
                HIPCHK(update(ps, pfirst, reqlast, Wh, t0c, t1c, 0, dev_wait ? la_seq : 0));
                HIPCHK(update(us, pfirst, reqlast, Wh, t1c, t2c, reqlast + 1));
                HIPCHK(hipEventRecord(h->ev_update, us));
                HIPCHK(mfk::flag_launch(us, la_ctr, ++la_seq));
                HIPCHK(update(us, pfirst, reqlast, Wh, t2c, n, reqlast + 1));
                ps_open = us_open = true;
            }
            prev_split = split;
            ++hidx;
            pfirst = reqlast + 1;
            if ((int64_t)pending.size() < depth && pfirst < n) continue;
            int64_t aux = 0;
            const int64_t r = read_info(h, &aux, pq);      // after the last panel
            if (r == INT64_MIN) { rc = fail(MFLA_ERR_HIP, "status read failed"); break; }
            if (r == kGridTimeoutInfo) { rc = fail(MFLA_ERR_HIP, "ldlt! panel: a grid barrier timed out"); break; }
            if (mfk::is_panel_stop(r)) {
                const bool extended = !mfk::panel_stop_truncated(r);
                // aux: the extended panel's last column / the column where the
                // panel stopped
                const Launched *p = nullptr;
                for (const auto &pr : pending)
                    if (pr.pfirst == mfk::panel_stop_first(r)) p = &pr;
                if (!p) { rc = fail(MFLA_ERR_HIP, "ldlt! panel: the stopped panel was not launched"); break; }
                HIPCHK(gather());
                HIPCHK(mfk::status_reset(s, st));
                if (extended) {
                    HIPCHK(trailing(s, p->pfirst, aux, p->Wh));
                    pfirst = aux + 1;
                } else {
                    HIPCHK(mfk::ldlt_restore(N, s, a, n, p->pfirst, aux, p->Wh));
                    pfirst = p->pfirst;
                    // off for the rest of the call after a second stop
                    la_hold = ++la_stops >= 2 ? n : la_backoff;
                    la_backoff = std::min<int64_t>(8 * la_backoff, n);
                }
                prev_split = false;
                pending.clear();
                depth = 1;
                continue;
            }
            if (r > 0) {
                // the panels keep the lower triangle and write each row's upper
                // part when it freezes; the block from the first unfrozen row on
                // is symmetric in MFLA's matrix at the failure
                const int64_t m = std::min(aux, n);
                HIPCHK(gather());
                HIPCHK(mfk::mirror_lower(N, s, sub(a, m, m), n - m, nullptr));
            }
            if (r != 0) break;
            pending.clear();
            depth = std::min(kLdltSpecMaxDepth, 2 * depth);
        }
        HIPCHK(gather());
        if (own) { (void)hipStreamSynchronize(s); (void)hipFree(W); }
        if (rc != MFLA_OK) return rc;
    }
    const int64_t r = read_info(h);
    if (r == INT64_MIN) return fail(MFLA_ERR_HIP, "status read failed");
    *info = r;
    for (int l = 0; l < N; ++l) origmax[l] = (r == -1) ? 0.0 : om[l];
    return MFLA_OK;
}

extern "C" {

int mfla_sytrf(mfla_handle h, int N, mfla_matrix A, mfla_matrix dsub, int64_t *pivots, uint8_t *blocks, double *origmax, int64_t *info,
               const mfla_config *config, mfla_workspace ws) {
    mfla_config tmp;
    const mfla_config &c = cfg_or_default(config, tmp);
    return sytrf_run(h, N, A, dsub, pivots, blocks, origmax, info, c, ws, nullptr, 0);
}

int mfla_sytrf_noalloc(mfla_handle h, int N, mfla_matrix A, mfla_matrix dsub, int64_t *pivots, uint8_t *blocks, double *origmax,
                       int64_t *info, const mfla_config *config) {
    mfla_config tmp;
    const mfla_config &c = cfg_or_default(config, tmp);
    // size check before the handle is used: a call that does not fit is
    // rejected without touching the device
    if (check_limbs(N)) return MFLA_ERR_ARGUMENT;
    if (A.rows != A.cols) return fail(MFLA_ERR_DIMENSION, "ldlt! requires a square matrix");
    if (A.rows < 0) return fail(MFLA_ERR_DIMENSION, "A has negative dimensions");
    mfla_ldlt_plan_t plan;
    if (ldlt_plan_core(N, A.rows, c, &plan) != MFLA_OK) return MFLA_ERR_ARGUMENT;
    size_t fold_off = 0, panel_off = 0;
    const size_t need = noalloc_layout(N, A.rows, plan, &fold_off, &panel_off);
    if (need > MFLA_NOALLOC_SCRATCH_BYTES / sizeof(double))
        return fail(MFLA_ERR_WORKSPACE, "ldlt! without allocation needs %zu bytes of device scratch (limbs * n * (block + 2) "
                    "doubles plus 256 bytes), the library holds %zu", sizeof(double) * need, (size_t)MFLA_NOALLOC_SCRATCH_BYTES);
    if (!h) return fail(MFLA_ERR_ARGUMENT, "null handle");
    double *buf = nullptr;
    hipError_t e = hipErrorInvalidDevice;
    if (h->device >= 0 && h->device < kMaxDevices) {
        std::lock_guard<std::mutex> g(g_scratch_lock);
        buf = g_scratch_addr[h->device];
        e = g_scratch_status[h->device];
    }
    if (!buf) return fail(MFLA_ERR_HIP, "device scratch unavailable on device %d (%s)", h->device, hipGetErrorString(e));
    return sytrf_run(h, N, A, dsub, pivots, blocks, origmax, info, c, nullptr, buf, MFLA_NOALLOC_SCRATCH_BYTES / sizeof(double));
}

// ------------------------------------------------------------- solves
static int check_solve(int N, const mfla_matrix &F, const mfla_matrix &B) {
    if (check_limbs(N)) return MFLA_ERR_ARGUMENT;
    if (check_desc(F, "F") || check_desc(B, "B")) return MFLA_ERR_ARGUMENT;
    if (F.rows != F.cols) return fail(MFLA_ERR_DIMENSION, "solve requires a square factor");
    if (B.rows != F.rows) return fail(MFLA_ERR_DIMENSION, "right-hand side dimensions differ");
    if (extents_overlap(B, F, N)) return fail(MFLA_ERR_ALIAS, "ldiv! does not support output/input aliasing");
    return MFLA_OK;
}

int mfla_potrs(mfla_handle h, int N, mfla_matrix F, int64_t info, mfla_matrix B, const mfla_config *config) {
    mfla_config tmp;
    const mfla_config &c = cfg_or_default(config, tmp);
    if (info != 0) return fail(MFLA_ERR_FAILED_FACTOR, "PosDefException(%lld)", (long long)info);
    if (int rc = check_solve(N, F, B)) return rc;
    std::lock_guard<std::mutex> g(h->lock);
    HIPCHK(trsm_core(h, N, MFLA_LEFT, MFLA_LOWER, MFLA_NOTRANS, MFLA_NONUNIT, B, F, kOne, c, nullptr));
    HIPCHK(trsm_core(h, N, MFLA_LEFT, MFLA_LOWER, MFLA_TRANS, MFLA_NONUNIT, B, F, kOne, c, nullptr));
    return MFLA_OK;
}

int mfla_getrs(mfla_handle h, int N, mfla_matrix F, const int64_t *ipiv, int64_t info, mfla_matrix B, const mfla_config *config) {
    mfla_config tmp;
    const mfla_config &c = cfg_or_default(config, tmp);
    if (info != 0) return fail(MFLA_ERR_FAILED_FACTOR, "SingularException(%lld)", (long long)info);
    if (int rc = check_solve(N, F, B)) return rc;
    std::lock_guard<std::mutex> g(h->lock);
    HIPCHK(mfk::apply_ipiv(N, h->stream, dvn(B), F.rows, B.cols, ipiv));
    HIPCHK(trsm_core(h, N, MFLA_LEFT, MFLA_LOWER, MFLA_NOTRANS, MFLA_UNIT, B, F, kOne, c, nullptr));
    HIPCHK(trsm_core(h, N, MFLA_LEFT, MFLA_UPPER, MFLA_NOTRANS, MFLA_NONUNIT, B, F, kOne, c, nullptr));
    return MFLA_OK;
}

int mfla_sytrs(mfla_handle h, int N, mfla_matrix F, mfla_matrix dsub, const int64_t *pivots, const uint8_t *blocks, int64_t info,
               mfla_matrix B, const mfla_config *config) {
    mfla_config tmp;
    const mfla_config &c = cfg_or_default(config, tmp);
    if (info != 0) return fail(MFLA_ERR_FAILED_FACTOR, "SingularException(%lld)", (long long)info);
    if (int rc = check_solve(N, F, B)) return rc;
    const int64_t n = F.rows, nrhs = B.cols;
    std::lock_guard<std::mutex> g(h->lock);
    hipStream_t s = h->stream;
    HIPCHK(mfk::status_reset(s, h->status));
    HIPCHK(hipMemsetAsync(h->flags, 0, sizeof(int), s));
    HIPCHK(mfk::ldlt_pivots(N, s, dvn(B), n, nrhs, pivots, blocks, 1, h->flags));
    HIPCHK(trsm_core(h, N, MFLA_LEFT, MFLA_LOWER, MFLA_NOTRANS, MFLA_UNIT, B, F, kOne, c, nullptr));
    HIPCHK(mfk::ldlt_dsolve(N, s, dvn(B), n, nrhs, dvn(F), dvn(dsub), blocks, h->status));
    HIPCHK(trsm_core(h, N, MFLA_LEFT, MFLA_LOWER, MFLA_TRANS, MFLA_UNIT, B, F, kOne, c, nullptr));
    HIPCHK(mfk::ldlt_pivots(N, s, dvn(B), n, nrhs, pivots, blocks, 0, h->flags));
    int bad = 0;
    HIPCHK(hipMemcpyAsync(&bad, h->flags, sizeof(int), hipMemcpyDeviceToHost, s));
    const int64_t r = read_info(h);
    if (bad) return fail(MFLA_ERR_ARGUMENT, "invalid LDLT block structure");
    if (r != 0) return fail(MFLA_ERR_FAILED_FACTOR, "SingularException(%lld)", (long long)r);
    return MFLA_OK;
}

// --------------------------------------------------- residual / errors
static int residual_common(mfla_handle h, int N, int uplo, mfla_matrix R, mfla_matrix A, mfla_matrix X, mfla_matrix B, bool vec,
                           const mfla_config &c) {
    if (check_limbs(N)) return MFLA_ERR_ARGUMENT;
    if (uplo != MFLA_GENERAL && uplo != MFLA_LOWER && uplo != MFLA_UPPER)
        return fail(MFLA_ERR_ARGUMENT, "uplo must be :general, :lower, or :upper");
    if (check_desc(R, "R") || check_desc(A, "A") || check_desc(X, "X") || check_desc(B, "B")) return MFLA_ERR_ARGUMENT;
    const int64_t rows = A.rows, cols = A.cols;
    if (vec) {
        if (R.rows != rows || R.cols != 1) return fail(MFLA_ERR_DIMENSION, "residual output length differs");
        if (X.rows != cols || X.cols != 1) return fail(MFLA_ERR_DIMENSION, "residual solution length differs");
        if (B.rows != rows || B.cols != 1) return fail(MFLA_ERR_DIMENSION, "residual right-hand side length differs");
    } else {
        if (R.rows != B.rows || R.cols != B.cols) return fail(MFLA_ERR_DIMENSION, "residual output dimensions differ");
        if (R.rows != rows) return fail(MFLA_ERR_DIMENSION, "residual output row count differs");
        if (X.rows != cols) return fail(MFLA_ERR_DIMENSION, "residual solution row count differs");
        if (X.cols != R.cols) return fail(MFLA_ERR_DIMENSION, "residual right-hand-side count differs");
    }
    if (uplo != MFLA_GENERAL && rows != cols) return fail(MFLA_ERR_DIMENSION, "symmetric residual requires a square matrix");
    if (extents_overlap(R, X, N)) return fail(MFLA_ERR_ALIAS, "residual output must not alias x");
    if (extents_overlap(R, A, N)) return fail(MFLA_ERR_ALIAS, "residual output must not alias A");
    const bool same = same_matrix(R, B);
    if (!same && extents_overlap(R, B, N)) return fail(MFLA_ERR_ALIAS, "residual output may equal b but must not partially alias it");
    std::lock_guard<std::mutex> g(h->lock);
    hipStream_t s = h->stream;
    if (!same) HIPCHK(mfk::copy(N, s, R.rows, R.cols, dvn(R), dvn(B)));
    if (uplo == MFLA_GENERAL) {
        if (vec) HIPCHK(mfk::gemv(N, s, 0, rows, cols, dvn(A), dvn(X), dvn(R), kMinusOne, kOne, 0, nullptr));
        else HIPCHK(gemm_launch(h, N, c.gemm_strategy, c.gemm_tile, rows, R.cols, cols, dvn(A), dvn(X), dvn(R), kMinusOne, kOne, nullptr));
    } else {
        HIPCHK(mfk::symv(N, s, uplo == MFLA_UPPER, rows, R.cols, dvn(A), dvn(X), dvn(R), kMinusOne, kOne, 0));
    }
    return MFLA_OK;
}

int mfla_residual(mfla_handle h, int N, int uplo, mfla_matrix R, mfla_matrix A, mfla_matrix X, mfla_matrix B, const mfla_config *config) {
    mfla_config tmp;
    return residual_common(h, N, uplo, R, A, X, B, false, cfg_or_default(config, tmp));
}

int mfla_residual_vector(mfla_handle h, int N, int uplo, mfla_matrix r, mfla_matrix A, mfla_matrix x, mfla_matrix b,
                         const mfla_config *config) {
    mfla_config tmp;
    return residual_common(h, N, uplo, r, A, x, b, true, cfg_or_default(config, tmp));
}

int mfla_residual_mixed(mfla_handle h, int NS, int NR, int uplo, mfla_matrix R, mfla_matrix A, mfla_matrix X, mfla_matrix B,
                        const mfla_config *) {
    const bool ok = (NS == 2 && (NR == 3 || NR == 4)) || (NS == 3 && NR == 4) || (NS == 1 && NR >= 2 && NR <= 4);
    if (!ok) return fail(MFLA_ERR_ARGUMENT, "residual_mixed! supports only Float64x2 to Float64x3/x4 and Float64x3 to Float64x4");
    if (uplo != MFLA_GENERAL && uplo != MFLA_LOWER && uplo != MFLA_UPPER)
        return fail(MFLA_ERR_ARGUMENT, "uplo must be :general, :lower, or :upper");
    const int64_t rows = A.rows, cols = A.cols;
    if (R.rows != B.rows || R.cols != B.cols) return fail(MFLA_ERR_DIMENSION, "residual output dimensions differ");
    if (R.rows != rows) return fail(MFLA_ERR_DIMENSION, "residual output row count differs");
    if (X.rows != cols) return fail(MFLA_ERR_DIMENSION, "residual solution row count differs");
    if (X.cols != R.cols) return fail(MFLA_ERR_DIMENSION, "residual right-hand-side count differs");
    if (uplo != MFLA_GENERAL && rows != cols) return fail(MFLA_ERR_DIMENSION, "symmetric residual requires a square matrix");
    if (extents_overlap(R, A, NR > NS ? NS : NR) || extents_overlap(R, X, NS) || extents_overlap(R, B, NS))
        return fail(MFLA_ERR_ALIAS, "mixed residual output must not alias source operands");
    std::lock_guard<std::mutex> g(h->lock);
    HIPCHK(mfk::residual_mixed(NS, NR, h->stream, kernel_uplo(uplo), rows, cols, R.cols, dvn(A), dvn(X), dvn(B), dvn(R)));
    return MFLA_OK;
}

int mfla_add_mixed(mfla_handle h, int dst_limbs, int src_limbs, mfla_matrix X, mfla_matrix D) {
    if (check_limbs(dst_limbs) || check_limbs(src_limbs)) return MFLA_ERR_ARGUMENT;
    if (src_limbs > dst_limbs) return fail(MFLA_ERR_ARGUMENT, "correction precision exceeds the solution precision");
    if (X.rows != D.rows || X.cols != D.cols) return fail(MFLA_ERR_DIMENSION, "correction dimensions differ");
    if (check_desc(X, "X") || check_desc(D, "D")) return MFLA_ERR_ARGUMENT;
    if (extents_overlap(X, D, src_limbs)) return fail(MFLA_ERR_ALIAS, "correction must not alias the solution");
    std::lock_guard<std::mutex> g(h->lock);
    HIPCHK(mfk::add_mixed(src_limbs, dst_limbs, h->stream, X.rows, X.cols, dvn(X), dvn(D)));
    return MFLA_OK;
}

int mfla_batched_solve(mfla_handle h, int N, const mfla_batch_entry *entries, int64_t count, double *A_pool, double *B_pool,
                       double *dsub_pool, int64_t *pivot_pool, uint8_t *block_pool, int64_t *info, double *origmax,
                       unsigned int *counter, int64_t workgroups, int threads_per_system, int exact_backward,
                       const mfla_config *config) {
    mfla_config tmp;
    const mfla_config &c = cfg_or_default(config, tmp);
    if (check_limbs(N)) return MFLA_ERR_ARGUMENT;
    if (count < 0) return fail(MFLA_ERR_ARGUMENT, "negative batch count");
    const int threads = threads_per_system == 0 ? 256 : threads_per_system;
    if (threads != 64 && threads != 128 && threads != 256) return fail(MFLA_ERR_ARGUMENT, "threads_per_system must be 64, 128 or 256");
    if (count > 0 && (!entries || !A_pool || !B_pool || !info || !counter || !origmax))
        return fail(MFLA_ERR_ARGUMENT, "null batch pool");
    mfk::BatchArgs a{};
    a.entries = entries; a.count = count; a.A = A_pool; a.B = B_pool; a.dsub = dsub_pool; a.pivots = pivot_pool;
    a.blocks = block_pool; a.info = info; a.origmax = origmax; a.counter = counter;
    ldlt_alpha(N, a.alpha);
    a.chol_block = std::max<int64_t>(c.cholesky_block, 1);
    a.exact_backward = exact_backward ? 1 : 0;
    a.threads = threads;
    a.workgroups = workgroups > 0 ? workgroups : 1024 * (256 / threads);
    std::lock_guard<std::mutex> g(h->lock);
    HIPCHK(mfk::batched_solve(N, h->stream, a));
    return MFLA_OK;
}

int mfla_batched_max_order(void) { return mfk::batched_max_order(); }
int mfla_batched_max_rhs(void) { return mfk::batched_max_rhs(); }

int mfla_backward_error(mfla_handle h, int N, int uplo, mfla_matrix A, mfla_matrix X, mfla_matrix B, mfla_matrix R, double *out) {
    if (check_limbs(N)) return MFLA_ERR_ARGUMENT;
    if (uplo != MFLA_GENERAL && uplo != MFLA_LOWER && uplo != MFLA_UPPER)
        return fail(MFLA_ERR_ARGUMENT, "uplo must be :general, :lower, or :upper");
    const int64_t m = A.rows, n = A.cols, nrhs = X.cols;
    if (R.rows != B.rows || R.cols != B.cols || R.rows != m || X.rows != n || X.cols != R.cols)
        return fail(MFLA_ERR_DIMENSION, "backward error dimensions differ");
    if (uplo != MFLA_GENERAL && m != n) return fail(MFLA_ERR_DIMENSION, "symmetric residual requires a square matrix");
    std::lock_guard<std::mutex> g(h->lock);
    hipStream_t s = h->stream;
    // scratch: row sums (m), three column-max vectors (nrhs each), flags
    const size_t dbl = (size_t)N * (size_t)(m + 3 * nrhs);
    char *base = (char *)grow_scratch(h, sizeof(double) * dbl + sizeof(int) * 3 * (size_t)nrhs + 64);
    if (!base) return fail(MFLA_ERR_HIP, "scratch allocation failed");
    double *rows = (double *)base;
    double *xm = rows + (size_t)N * m, *bm = xm + (size_t)N * nrhs, *rm = bm + (size_t)N * nrhs;
    int *fin = (int *)(rm + (size_t)N * nrhs);
    HIPCHK(mfk::status_reset(s, h->status));
    HIPCHK(mfk::finite_check(N, s, m, n, dvn(A), kernel_uplo(uplo), h->status));
    HIPCHK(mfk::row_abs_sum(N, s, kernel_uplo(uplo), m, n, dvn(A), DV{rows, 1, m, m}));
    HIPCHK(mfk::col_absmax(N, s, n, nrhs, dvn(X), DV{xm, 1, nrhs, nrhs}, fin));
    HIPCHK(mfk::col_absmax(N, s, m, nrhs, dvn(B), DV{bm, 1, nrhs, nrhs}, fin + nrhs));
    HIPCHK(mfk::col_absmax(N, s, m, nrhs, dvn(R), DV{rm, 1, nrhs, nrhs}, fin + 2 * nrhs));
    std::vector<double> hb(dbl);
    std::vector<int> hf(3 * (size_t)nrhs);
    HIPCHK(hipMemcpyAsync(hb.data(), base, sizeof(double) * dbl, hipMemcpyDeviceToHost, s));
    if (nrhs > 0) HIPCHK(hipMemcpyAsync(hf.data(), fin, sizeof(int) * 3 * nrhs, hipMemcpyDeviceToHost, s));
    const int64_t st = read_info(h);
    if (st == INT64_MIN) return fail(MFLA_ERR_HIP, "status read failed");
    const bool afin = st == 0;
    double an[4];
    host_max(N, hb.data(), m, an);
    for (int64_t j = 0; j < nrhs; ++j) {
        double xn[4], bn[4], rn[4], o[4];
        for (int l = 0; l < N; ++l) {
            xn[l] = hb[(size_t)N * m + l * nrhs + j];
            bn[l] = hb[(size_t)N * (m + nrhs) + l * nrhs + j];
            rn[l] = hb[(size_t)N * (m + 2 * nrhs) + l * nrhs + j];
        }
        const bool f = afin && hf[j] && hf[nrhs + j] && hf[2 * nrhs + j];
        switch (N) {
            case 1: backward_error_host<1>(an, xn, bn, rn, f, o); break;
            case 2: backward_error_host<2>(an, xn, bn, rn, f, o); break;
            case 3: backward_error_host<3>(an, xn, bn, rn, f, o); break;
            default: backward_error_host<4>(an, xn, bn, rn, f, o); break;
        }
        for (int l = 0; l < N; ++l) out[l * nrhs + j] = o[l];
    }
    return MFLA_OK;
}

}  // extern "C"

// ======================================================================
// Binary64 factor + Float64xN iterative refinement
namespace {

DV dT(const DV &v) { return DV{v.p, v.cs, v.rs, v.ls}; }

mfk::DGemmArgs dg(int64_t m, int64_t n, int64_t k, const DV &A, const DV &B, const DV &C, double alpha, double beta) {
    mfk::DGemmArgs g{};
    g.m = m; g.n = n; g.k = k; g.A = A; g.B = B; g.C = C;
    g.alpha = alpha; g.beta = beta; g.beta_zero = beta == 0.0 ? 1 : 0;
    g.ksplit = 1;
    return g;
}

#define HIPRET(expr)                        \
    do {                                    \
        hipError_t e_ = (expr);             \
        if (e_ != hipSuccess) return e_;    \
    } while (0)

// Side stream and events of the handle for a binary64 look-ahead (none if
// the stream cannot be created: the factorizations then run in order).
struct LookAhead {
    hipStream_t side;
    hipEvent_t ev_main, ev_side;
};
LookAhead lookahead(mfla_handle h);

constexpr int64_t kDiagBlock = 64;

// Blocked right-looking binary64 Cholesky of F's lower triangle; the inverse
// of every 64 x 64 diagonal factor block lands in the same block of Inv.
// With a side stream the trailing update is split as in dgetrf_core: the
// next block's 64 columns on the main stream (ahead of its diagonal
// factorization), the rest of the lower triangle on the side stream. The
// main stream waits for the previous side update before it writes the next
// block's columns, the only region both streams touch.
hipError_t dpotrf_core(hipStream_t s, int64_t n, const DV &F, const DV &Inv, Status *st, const LookAhead *la) {
    bool side_pending = false;
    for (int64_t j0 = 0; j0 < n; j0 += kDiagBlock) {
        const int nbk = (int)std::min<int64_t>(kDiagBlock, n - j0);
        HIPRET(mfk::dpotrf_diag(s, F, Inv, j0, nbk, st));
        const int64_t jend = j0 + nbk, rest = n - jend;
        if (rest == 0) break;
        const DV A21 = sub(F, jend, j0);
        // L21 = A21 * inv(L11)^T in place: a 64-row output tile spans all nbk
        // (<= 64) columns of its rows, reads them completely before it writes,
        // and no other tile touches those rows.
        mfk::DGemmArgs g = dg(rest, nbk, nbk, A21, dT(sub(Inv, j0, j0)), A21, 1.0, 0.0);
        g.b_tri = mfk::TRI_UPPER;
        HIPRET(mfk::dgemm(s, g, 1));
        if (side_pending) {
            HIPRET(hipStreamWaitEvent(s, la->ev_side, 0));
            side_pending = false;
        }
        // A22 -= L21 * L21^T on the lower triangle: the next block's columns
        // here, the rest on the side stream
        const int64_t next = la ? std::min<int64_t>(kDiagBlock, rest) : rest;
        g = dg(rest, next, nbk, A21, dT(sub(F, jend, j0)), sub(F, jend, jend), -1.0, 1.0);
        g.c_lower = 1;
        HIPRET(mfk::dgemm(s, g, 1));
        if (next < rest) {
            HIPRET(hipEventRecord(la->ev_main, s));
            HIPRET(hipStreamWaitEvent(la->side, la->ev_main, 0));
            const DV L2 = sub(F, jend + next, j0);
            g = dg(rest - next, rest - next, nbk, L2, dT(L2), sub(F, jend + next, jend + next), -1.0, 1.0);
            g.c_lower = 1;
            HIPRET(mfk::dgemm(la->side, g, 1));
            HIPRET(hipEventRecord(la->ev_side, la->side));
            side_pending = true;
        }
    }
    if (side_pending) HIPRET(hipStreamWaitEvent(s, la->ev_side, 0));
    return hipSuccess;
}

// Blocked right-looking binary64 LU with partial pivoting of the n x n F
// (LAPACK dgetrf semantics: ipiv 1-based, info = first zero pivot column).
// Each 64-column block is factored by register-resident leaves of
// dlu_leaf_width(n - j0) columns; a leaf also interchanges and solves the
// block's other columns, and an MFMA product updates the block's later
// columns below it. The block's interchanges then reach the columns left
// and right of it (dlu_rows, which also forms U12 = L11^{-1} P A12). With a
// side stream, the trailing update A22 -= L21 U12 is split: the next
// block's 64 columns on the main stream, the rest on the side stream, where
// it overlaps the next block's leaves (the main stream waits for it before
// the next dlu_rows, the first call that touches those columns or reorders
// the rows of L21).

hipError_t dgetrf_core(hipStream_t s, int64_t n, const DV &F, int64_t *ipiv, Status *st, const LookAhead *la) {
    if (mfk::dlu_leaf_width(n) == 0) return hipErrorInvalidValue;
    bool side_pending = false;
    for (int64_t j0 = 0; j0 < n; j0 += kDiagBlock) {
        const int64_t jb = std::min<int64_t>(kDiagBlock, n - j0), jend = j0 + jb;
        const int W = mfk::dlu_leaf_width(n - j0);
        for (int64_t c0 = j0; c0 < jend; c0 += W) {
            const int w = (int)std::min<int64_t>(W, jend - c0);
            HIPRET(mfk::dlu_leaf(s, F, n, c0, w, ipiv, st, j0, (int)(c0 - j0), c0 + w, (int)(jend - c0 - w)));
            if (c0 + w < jend)
                HIPRET(mfk::dgemm(s, dg(n - c0 - w, jend - c0 - w, w, sub(F, c0 + w, c0), sub(F, c0, c0 + w), sub(F, c0 + w, c0 + w),
                                        -1.0, 1.0), 1));
        }
        if (side_pending) {
            HIPRET(hipStreamWaitEvent(s, la->ev_side, 0));
            side_pending = false;
        }
        HIPRET(mfk::dlu_rows(s, F, n, j0, (int)jb, ipiv));
        if (jend == n) break;
        const int64_t next = la ? std::min<int64_t>(kDiagBlock, n - jend) : n - jend;
        HIPRET(mfk::dgemm(s, dg(n - jend, next, jb, sub(F, jend, j0), sub(F, j0, jend), sub(F, jend, jend), -1.0, 1.0), 1));
        if (jend + next < n) {
            HIPRET(hipEventRecord(la->ev_main, s));
            HIPRET(hipStreamWaitEvent(la->side, la->ev_main, 0));
            HIPRET(mfk::dgemm(la->side, dg(n - jend, n - jend - next, jb, sub(F, jend, j0), sub(F, j0, jend + next),
                                           sub(F, jend, jend + next), -1.0, 1.0), 1));
            HIPRET(hipEventRecord(la->ev_side, la->side));
            side_pending = true;
        }
    }
    if (side_pending) HIPRET(hipStreamWaitEvent(s, la->ev_side, 0));
    return hipSuccess;
}

LookAhead lookahead(mfla_handle h) {
    if (ensure_side(h) != hipSuccess) return LookAhead{nullptr, nullptr, nullptr};
    return LookAhead{h->side, h->ev_panel, h->ev_update};
}

// Recursive doubling from inverted diagonal blocks of size nb to the full
// inverse of a triangular factor T:
//   lower: inv([T11 0; T21 T22]) = [X11 0; -X22 T21 X11  X22]
//   upper: inv([T11 T12; 0 T22]) = [X11  -X11 T12 X22; 0 X22]
// All pairs of one level run as one batched GEMM; the intermediate product
// is kept in the unused opposite triangle of Inv.
hipError_t dtrtri_core(hipStream_t s, int64_t n, int64_t nb, const DV &T, const DV &Inv, bool lower) {
    for (int64_t sz = nb; sz < n; sz *= 2) {
        const int64_t full = n / (2 * sz);
        auto level = [&](int64_t p, int64_t s2, int64_t batch) -> hipError_t {
            const int64_t istep = 2 * sz * (Inv.rs + Inv.cs), tstep = 2 * sz * (T.rs + T.cs);
            const DV X11 = sub(Inv, p, p), X22 = sub(Inv, p + sz, p + sz);
            mfk::DGemmArgs g;
            if (lower) {
                const DV W = dT(sub(Inv, p, p + sz));                       // s2 x sz, strict upper storage
                g = dg(s2, sz, sz, sub(T, p + sz, p), X11, W, 1.0, 0.0);     // W = T21 * X11
                g.b_tri = mfk::TRI_LOWER; g.za = tstep; g.zb = istep; g.zc = istep;
                HIPRET(mfk::dgemm(s, g, batch));
                g = dg(s2, sz, s2, X22, W, sub(Inv, p + sz, p), -1.0, 0.0); // X21 = -X22 * W
                g.a_tri = mfk::TRI_LOWER; g.za = istep; g.zb = istep; g.zc = istep;
                return mfk::dgemm(s, g, batch);
            }
            const DV W = dT(sub(Inv, p + sz, p));                           // sz x s2, strict lower storage
            g = dg(sz, s2, s2, sub(T, p, p + sz), X22, W, 1.0, 0.0);        // W = T12 * X22
            g.b_tri = mfk::TRI_UPPER; g.za = tstep; g.zb = istep; g.zc = istep;
            HIPRET(mfk::dgemm(s, g, batch));
            g = dg(sz, s2, sz, X11, W, sub(Inv, p, p + sz), -1.0, 0.0);     // X12 = -X11 * W
            g.a_tri = mfk::TRI_UPPER; g.za = istep; g.zb = istep; g.zc = istep;
            return mfk::dgemm(s, g, batch);
        };
        if (full > 0) HIPRET(level(0, sz, full));
        const int64_t p = full * 2 * sz;
        if (p + sz < n) HIPRET(level(p, n - p - sz, 1));
    }
    return hipSuccess;
}

struct RefineWs {
    double *F, *Inv, *InvU, *Y, *D, *R, *P, *rows, *part, *stats;
    int64_t *perm, *ipiv;
    int64_t ksplit;
    // Ozaki residual: digit planes of A (S planes of planeA bytes in the
    // fragment layout), of the correction (T x nrhs x ld int8), exponents,
    // int32 digit-pair sums
    int8_t *dA, *dD;
    int *eA, *eD;
    int32_t *planes;
    int64_t ld, planeA;
    int S, T;
};
constexpr int kOzCorrectionDigits = 6;   // = mfk::oz_narrow_digits()

// Split-k factor for the n x nrhs correction products: enough workgroups to
// cover the chip, at least 256 reduction steps per slice.
int64_t refine_ksplit(int64_t n, int64_t nrhs) {
    const int64_t tiles = mfk::cdiv(n, 64) * mfk::cdiv(nrhs, 64);
    int64_t ks = 1;
    while (tiles * ks * 2 <= 1024 && n / (ks * 2) >= 256) ks *= 2;
    return ks;
}

int64_t refine_layout(int N, int kind, int64_t n, int64_t nrhs, double *base, RefineWs *out) {
    const int64_t ks = refine_ksplit(n, nrhs);
    int64_t off = 0;
    auto take = [&](int64_t cnt) {
        double *p = base ? base + off : nullptr;
        off += (cnt + 31) / 32 * 32;   // 256-byte aligned regions
        return p;
    };
    RefineWs w{};
    w.F = take(n * n);
    w.Inv = take(n * n);
    w.InvU = (kind & 15) == MFLA_REFINE_LU ? take(n * n) : nullptr;
    w.Y = take(n * nrhs);
    w.D = take(n * nrhs);
    w.R = take((int64_t)N * n * nrhs);
    w.P = ks > 1 ? take(ks * n * nrhs) : nullptr;
    w.rows = take(n);
    w.part = take((int64_t)mfk::drow_abs_chunks() * n);
    w.stats = take(4 * nrhs + 8);
    w.perm = (int64_t *)take(n);
    w.ipiv = (int64_t *)take(n);
    w.ksplit = ks;
    if (kind & MFLA_REFINE_OZAKI) {
        w.ld = (n + mfk::oz_row_align() - 1) / mfk::oz_row_align() * mfk::oz_row_align();
        w.S = mfk::oz_digits(N);
        w.T = kOzCorrectionDigits;
        w.planeA = (n + 15) / 16 * 16 * w.ld;   // fragment layout: rows padded to 16
        w.dA = (int8_t *)take((w.S * w.planeA + 7) / 8);
        w.dD = (int8_t *)take((w.T * nrhs * w.ld + 7) / 8);
        w.eA = (int *)take((n + 1) / 2);
        w.eD = (int *)take((nrhs + 1) / 2);
        w.planes = (int32_t *)take(((int64_t)(w.S + w.T - 1) * n * nrhs + 1) / 2);
    }
    if (out) *out = w;
    return off;
}

// Events between stages; the elapsed time up to each mark is charged to the
// mark's stage.
struct StageClock {
    hipStream_t s;
    bool on;
    std::vector<std::pair<int, hipEvent_t>> marks;
    void mark(int stage) {
        if (!on) return;
        hipEvent_t e;
        if (hipEventCreate(&e) != hipSuccess) { on = false; return; }
        (void)hipEventRecord(e, s);
        marks.emplace_back(stage, e);
    }
    ~StageClock() { finish(nullptr); }
    void finish(double *out) {
        if (out) for (int i = 0; i < 5; ++i) out[i] = 0.0;
        if (!marks.empty()) (void)hipEventSynchronize(marks.back().second);
        for (size_t i = 1; i < marks.size(); ++i) {
            float ms = 0.f;
            if (out && hipEventElapsedTime(&ms, marks[i - 1].second, marks[i].second) == hipSuccess) out[marks[i].first] += ms;
        }
        for (auto &m : marks) (void)hipEventDestroy(m.second);
        marks.clear();
    }
};
enum { ST_FACTOR = 0, ST_SOLVE = 1, ST_UPDATE = 2, ST_RESIDUAL = 3, ST_CHECK = 4 };

}  // namespace

extern "C" {

int mfla_dgemm(mfla_handle h, int trans_a, int trans_b, int tri_a, int tri_b, mfla_matrix C, mfla_matrix A, mfla_matrix B,
               double alpha, double beta) {
    if (check_desc(C, "C") || check_desc(A, "A") || check_desc(B, "B")) return MFLA_ERR_ARGUMENT;
    auto tri = [](int t, int &out) {
        if (t == MFLA_GENERAL) out = mfk::TRI_NONE;
        else if (t == MFLA_LOWER) out = mfk::TRI_LOWER;
        else if (t == MFLA_UPPER) out = mfk::TRI_UPPER;
        else return false;
        return true;
    };
    int ta, tb;
    if (!tri(tri_a, ta) || !tri(tri_b, tb)) return fail(MFLA_ERR_ARGUMENT, "triangle flags must be :general, :lower, or :upper");
    if ((trans_a != MFLA_NOTRANS && trans_a != MFLA_TRANS) || (trans_b != MFLA_NOTRANS && trans_b != MFLA_TRANS))
        return fail(MFLA_ERR_ARGUMENT, "trans must be 'N' or 'T'");
    const int64_t m = trans_a == MFLA_TRANS ? A.cols : A.rows, k = trans_a == MFLA_TRANS ? A.rows : A.cols;
    const int64_t kb = trans_b == MFLA_TRANS ? B.cols : B.rows, n = trans_b == MFLA_TRANS ? B.rows : B.cols;
    if (k != kb || C.rows != m || C.cols != n) return fail(MFLA_ERR_DIMENSION, "matrix dimensions do not match");
    if (extents_overlap(C, A, 1) || extents_overlap(C, B, 1)) return fail(MFLA_ERR_ALIAS, "output must not alias an input");
    std::lock_guard<std::mutex> g(h->lock);
    mfk::DGemmArgs a = dg(m, n, k, trans_a == MFLA_TRANS ? dvt(A) : dvn(A), trans_b == MFLA_TRANS ? dvt(B) : dvn(B), dvn(C), alpha, beta);
    a.a_tri = ta; a.b_tri = tb;
    if (k == 0) { a.k = 0; }
    HIPCHK(mfk::dgemm(h->stream, a, 1));
    return MFLA_OK;
}

int mfla_dpotrf_inverse(mfla_handle h, mfla_matrix A, mfla_matrix Inv, int64_t *info) {
    if (check_desc(A, "A") || check_desc(Inv, "Inv")) return MFLA_ERR_ARGUMENT;
    if (A.rows != A.cols || Inv.rows != A.rows || Inv.cols != A.cols) return fail(MFLA_ERR_DIMENSION, "square matrices of equal order required");
    if (extents_overlap(A, Inv, 1)) return fail(MFLA_ERR_ALIAS, "the inverse must not alias the factor");
    const int64_t n = A.rows;
    std::lock_guard<std::mutex> g(h->lock);
    hipStream_t s = h->stream;
    HIPCHK(mfk::status_reset(s, h->status));
    HIPCHK(mfk::finite_check(1, s, n, n, dvn(A), 1, h->status));
    const LookAhead la = lookahead(h);
    HIPCHK(dpotrf_core(s, n, dvn(A), dvn(Inv), h->status, la.side ? &la : nullptr));
    const int64_t r = read_info(h);
    if (r == INT64_MIN) return fail(MFLA_ERR_HIP, "status read failed");
    *info = r;
    if (r == 0) {
        HIPCHK(dtrtri_core(s, n, kDiagBlock, dvn(A), dvn(Inv), true));
        HIPCHK(hipStreamSynchronize(s));
    }
    return MFLA_OK;
}

int mfla_dgetrf(mfla_handle h, mfla_matrix A, int64_t *ipiv, int64_t *info) {
    if (check_desc(A, "A")) return MFLA_ERR_ARGUMENT;
    if (A.rows != A.cols) return fail(MFLA_ERR_DIMENSION, "square matrix required");
    const int64_t n = A.rows;
    if (mfk::dlu_leaf_width(n) == 0) return fail(MFLA_ERR_ARGUMENT, "n <= 32768 required");
    std::lock_guard<std::mutex> g(h->lock);
    hipStream_t s = h->stream;
    HIPCHK(mfk::status_reset(s, h->status));
    HIPCHK(mfk::iota1(s, ipiv, n));
    HIPCHK(mfk::finite_check(1, s, n, n, dvn(A), 0, h->status));
    const LookAhead la = lookahead(h);
    HIPCHK(dgetrf_core(s, n, dvn(A), ipiv, h->status, la.side ? &la : nullptr));
    const int64_t r = read_info(h);
    if (r == INT64_MIN) return fail(MFLA_ERR_HIP, "status read failed");
    *info = r;
    return MFLA_OK;
}

int mfla_refine_workspace_size(int limbs, int kind, int64_t n, int64_t nrhs, int64_t *doubles) {
    if (limbs < 2 || limbs > 4) return fail(MFLA_ERR_ARGUMENT, "refinement needs a 2-4 limb working precision");
    if ((kind & ~MFLA_REFINE_OZAKI) != MFLA_REFINE_CHOLESKY && (kind & ~MFLA_REFINE_OZAKI) != MFLA_REFINE_LU)
        return fail(MFLA_ERR_ARGUMENT, "kind must be cholesky or lu");
    if (n < 0 || nrhs < 0) return fail(MFLA_ERR_DIMENSION, "negative dimensions");
    *doubles = refine_layout(limbs, kind, n, nrhs, nullptr, nullptr);
    return MFLA_OK;
}

int mfla_refine_solve(mfla_handle h, int N, int kind_flags, int uplo, mfla_matrix A, mfla_matrix B, mfla_matrix X, double *work,
                      int64_t work_len, double tol, int max_iter, int64_t *info, int *iterations, double *history,
                      double *stage_ms) {
    const bool ozaki = (kind_flags & MFLA_REFINE_OZAKI) != 0;
    const int kind = kind_flags & ~MFLA_REFINE_OZAKI;
    if (N < 2 || N > 4) return fail(MFLA_ERR_ARGUMENT, "refinement needs a 2-4 limb working precision");
    if (kind != MFLA_REFINE_CHOLESKY && kind != MFLA_REFINE_LU) return fail(MFLA_ERR_ARGUMENT, "kind must be cholesky or lu");
    if (ozaki && A.rows > 8192) return fail(MFLA_ERR_ARGUMENT, "Ozaki residual updates support n up to 8192");
    if (uplo != MFLA_GENERAL && uplo != MFLA_LOWER && uplo != MFLA_UPPER)
        return fail(MFLA_ERR_ARGUMENT, "uplo must be :general, :lower, or :upper");
    if (kind == MFLA_REFINE_LU && uplo != MFLA_GENERAL) return fail(MFLA_ERR_ARGUMENT, "LU refinement needs a general matrix");
    if (check_desc(A, "A") || check_desc(B, "B") || check_desc(X, "X")) return MFLA_ERR_ARGUMENT;
    const int64_t n = A.rows, nrhs = B.cols;
    if (A.cols != n || B.rows != n || X.rows != n || X.cols != nrhs) return fail(MFLA_ERR_DIMENSION, "refinement dimensions differ");
    if (extents_overlap(X, A, N) || extents_overlap(X, B, N)) return fail(MFLA_ERR_ALIAS, "the solution must not alias A or B");
    if (max_iter < 1) return fail(MFLA_ERR_ARGUMENT, "max_iter must be positive");
    if (!info || !iterations || !history) return fail(MFLA_ERR_ARGUMENT, "null output pointer");
    RefineWs w;
    const int64_t need = refine_layout(N, kind_flags, n, nrhs, work, &w);
    if (work == nullptr || work_len < need) return fail(MFLA_ERR_WORKSPACE, "refinement workspace needs %lld doubles", (long long)need);
    *iterations = 0;
    *info = 0;
    if (n == 0 || nrhs == 0) return MFLA_OK;

    const DV Ahi{A.data, 1, A.ld, A.limb_stride};
    const DV F{w.F, 1, n, 0}, Inv{w.Inv, 1, n, 0}, InvU{w.InvU, 1, n, 0};
    const DV Y{w.Y, 1, n, 0}, D{w.D, 1, n, 0};
    const DV R{w.R, 1, n, n * nrhs};
    double *bmax = w.stats, *xmax = w.stats + nrhs, *rmax = w.stats + 2 * nrhs, *dmax = w.stats + 3 * nrhs;
    hipStream_t s = h->stream;
    StageClock clk{s, stage_ms != nullptr, {}};

    // ---- binary64 factor and inverse factors
    if (kind == MFLA_REFINE_LU && mfk::dlu_leaf_width(n) == 0) return fail(MFLA_ERR_ARGUMENT, "LU refinement supports n <= 32768");
    clk.mark(ST_FACTOR);
    std::lock_guard<std::mutex> g(h->lock);
    if (kind == MFLA_REFINE_LU) {
        HIPCHK(mfk::status_reset(s, h->status));
        HIPCHK(mfk::dcopy(s, n, n, F, Ahi, 0));
        HIPCHK(mfk::finite_check(1, s, n, n, F, 0, h->status));
        const LookAhead la = lookahead(h);
        HIPCHK(dgetrf_core(s, n, F, w.ipiv, h->status, la.side ? &la : nullptr));
        const int64_t r = read_info(h);
        if (r == INT64_MIN) return fail(MFLA_ERR_HIP, "status read failed");
        *info = r;
        if (r != 0) return fail(MFLA_ERR_FAILED_FACTOR, "SingularException(%lld)", (long long)r);
    }
    if (kind == MFLA_REFINE_CHOLESKY) {
        HIPCHK(mfk::status_reset(s, h->status));
        HIPCHK(mfk::dcopy(s, n, n, F, Ahi, uplo == MFLA_UPPER ? 2 : 0));
        HIPCHK(mfk::finite_check(1, s, n, n, F, 1, h->status));
        const LookAhead la = lookahead(h);
        HIPCHK(dpotrf_core(s, n, F, Inv, h->status, la.side ? &la : nullptr));
        const int64_t r = read_info(h);
        if (r == INT64_MIN) return fail(MFLA_ERR_HIP, "status read failed");
        *info = r;
        if (r != 0) return fail(MFLA_ERR_FAILED_FACTOR, "PosDefException(%lld)", (long long)r);
        HIPCHK(dtrtri_core(s, n, kDiagBlock, F, Inv, true));
    } else {
        // P b gathers rows: perm = the ipiv row swaps applied to the identity
        std::vector<int64_t> ip(n), perm(n);
        {
            void *dst_[1] = {ip.data()};
            const void *src_[1] = {w.ipiv};
            const size_t by_[1] = {sizeof(int64_t) * (size_t)n};
            HIPCHK(read_back(h, s, 1, dst_, src_, by_));
        }
        for (int64_t i = 0; i < n; ++i) perm[i] = i;
        for (int64_t k = 0; k < n; ++k) std::swap(perm[k], perm[ip[k] - 1]);
        HIPCHK(hipMemcpyAsync(w.perm, perm.data(), sizeof(int64_t) * n, hipMemcpyHostToDevice, s));
        HIPCHK(hipStreamSynchronize(s));
        HIPCHK(mfk::trinv_blocks(1, s, n, (int)kDiagBlock, F, Inv, 1, 1));
        HIPCHK(mfk::trinv_blocks(1, s, n, (int)kDiagBlock, F, InvU, 0, 0));
        HIPCHK(dtrtri_core(s, n, kDiagBlock, F, Inv, true));
        HIPCHK(dtrtri_core(s, n, kDiagBlock, F, InvU, false));
    }
    // ---- norms for the stopping test (binary64 estimates)
    std::vector<double> hrows(n), hb(nrhs), hx(nrhs, 0.0), hr(nrhs), hd(nrhs);
    HIPCHK(mfk::drow_abs_sum(s, n, n, Ahi, kernel_uplo(uplo), w.rows, w.part));
    HIPCHK(mfk::col_absmax_hi(s, n, nrhs, DV{B.data, 1, B.ld, B.limb_stride}, bmax));
    HIPCHK(mfk::zero_fill(N, s, n, nrhs, dvn(X)));
    HIPCHK(mfk::copy(N, s, n, nrhs, R, dvn(B)));
    {
        void *dst_[2] = {hrows.data(), hb.data()};
        const void *src_[2] = {w.rows, bmax};
        const size_t by_[2] = {sizeof(double) * (size_t)n, sizeof(double) * (size_t)nrhs};
        HIPCHK(read_back(h, s, 2, dst_, src_, by_));
    }
    double anorm = 0.0;
    for (double v : hrows) anorm = std::max(anorm, v);
    if (ozaki) {   // A's digit planes, once per solve
        HIPCHK(mfk::oz_exponents(s, n, n, Ahi, kernel_uplo(uplo), w.eA, w.part));
        HIPCHK(mfk::oz_split(N, s, n, n, Ahi, kernel_uplo(uplo), w.eA, w.S, w.dA, w.ld, w.planeA, true));
    }
    clk.mark(ST_FACTOR);

    // ---- refinement loop
    const DV Rhi{w.R, 1, n, 0};
    double prev = 0.0, best = INFINITY;
    int it = 0;
    while (it < max_iter) {
        // D = F64 \ hi(R)
        mfk::DGemmArgs ga, gb;
        if (kind == MFLA_REFINE_CHOLESKY) {
            ga = dg(n, nrhs, n, Inv, Rhi, Y, 1.0, 0.0);
            ga.a_tri = mfk::TRI_LOWER;
            gb = dg(n, nrhs, n, dT(Inv), Y, D, 1.0, 0.0);
            gb.a_tri = mfk::TRI_UPPER;
        } else {
            HIPCHK(mfk::dgather_rows(s, n, nrhs, D, Rhi, w.perm));
            ga = dg(n, nrhs, n, Inv, D, Y, 1.0, 0.0);
            ga.a_tri = mfk::TRI_LOWER;
            gb = dg(n, nrhs, n, InvU, Y, D, 1.0, 0.0);
            gb.a_tri = mfk::TRI_UPPER;
        }
        ga.ksplit = gb.ksplit = (int)w.ksplit;
        ga.P = gb.P = w.P;
        HIPCHK(mfk::dgemm(s, ga, 1));
        HIPCHK(mfk::dgemm(s, gb, 1));
        clk.mark(ST_SOLVE);
        if (ozaki) {
            // D rounded to 7T-bit fixed point per column (the value added to X
            // is the value whose digits are multiplied), R -= A*D exactly up
            // to A's digit truncation, then summed in N limbs
            const DV Dt{w.D, n, 1, 0};
            HIPCHK(mfk::oz_exponents(s, nrhs, n, Dt, 0, w.eD, w.part));
            HIPCHK(mfk::oz_round_fp64(s, n, nrhs, D, w.eD, w.T, w.dD, w.ld, nrhs * w.ld));
            HIPCHK(mfk::add_mixed(1, N, s, n, nrhs, dvn(X), D));
            clk.mark(ST_UPDATE);
            const int U = w.S + w.T - 2;
            HIPCHK(mfk::oz_gemm_narrow(s, n, nrhs, w.ld, w.dA, w.planeA, w.S, w.dD, w.ld, nrhs * w.ld, w.planes));
            HIPCHK(mfk::oz_combine(N, s, n, nrhs, U, w.planes, w.eA, w.eD, R, nullptr, nullptr, 0, 1));
            clk.mark(ST_RESIDUAL);
        } else {
        // precision of A*D: |D| / |X| = 2^-e needs about (53 N - e) / 52 limbs
        HIPCHK(mfk::col_absmax_hi(s, n, nrhs, D, dmax));
        {
            void *dst[1] = {hd.data()};
            const void *src[1] = {dmax};
            const size_t by[1] = {sizeof(double) * (size_t)nrhs};
            HIPCHK(read_back(h, s, 1, dst, src, by));
        }
        clk.mark(ST_CHECK);
        double rho = 0.0;
        for (int64_t j = 0; j < nrhs; ++j) {
            const double xn = std::max(hx[j], hd[j]);
            if (xn > 0.0) rho = std::max(rho, hd[j] / xn);
        }
        int na = N;
        if (rho > 0.0 && std::isfinite(rho)) {
            const double limbs = std::ceil((52.0 * N + 8.0 + std::log2(rho)) / 52.0);
            na = (int)std::min<double>(N, std::max(1.0, limbs));
        }
        HIPCHK(mfk::add_mixed(1, N, s, n, nrhs, dvn(X), D));
        clk.mark(ST_UPDATE);
        HIPCHK(mfk::resid_update(na, N, s, kernel_uplo(uplo), n, nrhs, n, Ahi, D, R));
        clk.mark(ST_RESIDUAL);
        }
        HIPCHK(mfk::col_absmax_hi(s, n, nrhs, Rhi, rmax));
        HIPCHK(mfk::col_absmax_hi(s, n, nrhs, DV{X.data, 1, X.ld, X.limb_stride}, xmax));
        {
            void *dst[2] = {hr.data(), hx.data()};
            const void *src[2] = {rmax, xmax};
            const size_t by[2] = {sizeof(double) * (size_t)nrhs, sizeof(double) * (size_t)nrhs};
            HIPCHK(read_back(h, s, 2, dst, src, by));
        }
        clk.mark(ST_CHECK);
        double be = 0.0;
        for (int64_t j = 0; j < nrhs; ++j) {
            const double den = anorm * hx[j] + hb[j];
            const double e = den > 0.0 ? hr[j] / den : (hr[j] == 0.0 ? 0.0 : INFINITY);
            be = std::max(be, e);
            if (std::isnan(hr[j])) be = NAN;
        }
        history[it] = be;
        ++it;
        if (!std::isfinite(be)) { best = be; break; }
        if (it > 1 && be > prev) {
            // the last correction made the estimate worse: take it back
            // (X += -D, exact up to one Float64xN rounding), so X is the
            // best iterate, the one the previous estimate belongs to
            const double minus_one = -1.0;
            HIPCHK(mfk::scale(1, s, n, nrhs, D, &minus_one, nullptr));
            HIPCHK(mfk::add_mixed(1, N, s, n, nrhs, dvn(X), D));
            HIPCHK(hipStreamSynchronize(s));
            best = prev;
            break;
        }
        best = be;
        if (be <= tol || (it > 1 && be > 0.5 * prev)) break;
        prev = be;
    }
    *iterations = it;
    clk.finish(stage_ms);
    return best <= tol ? MFLA_OK : MFLA_NOT_CONVERGED;
}

}  // extern "C"

// ======================================================================
// Adjoints of solve, Cholesky and logdet (reverse-mode rules)
namespace {

template <int N> mf::V<N> h_ldexp(mf::V<N> v, int k) {
    for (int l = 0; l < N; ++l) v.x[l] = std::ldexp(v.x[l], k);
    return v;
}

// exp(x) for |x| <= 1: Taylor series of expm1 at x / 2^12, then 12 steps of
// t <- 2t + t^2 ((1 + t)^2 - 1), which keep t's relative accuracy.
template <int N> mf::V<N> h_exp(const mf::V<N> &x) {
    constexpr int kHalvings = 12;
    const mf::V<N> r = h_ldexp(x, -kHalvings);
    mf::V<N> term = r, t = r;
    for (int k = 2; k <= 64; ++k) {
        term = mf::div(mf::mul(term, r), mf::from<N>((double)k));
        t = mf::add(t, term);
        if (term.x[0] == 0.0 || std::fabs(term.x[0]) < std::ldexp(std::fabs(t.x[0]), -53 * N - 8)) break;
    }
    for (int i = 0; i < kHalvings; ++i) t = mf::add(mf::add(t, t), mf::mul(t, t));
    return mf::add(mf::from<N>(1.0), t);
}

// log(m) for m in [1/2, 2]: Newton on exp, y <- y + m exp(-y) - 1, from the
// binary64 logarithm (quadratic convergence: 53 -> 106 -> 212 bits).
template <int N> mf::V<N> h_log(const mf::V<N> &m) {
    mf::V<N> y = mf::from<N>(std::log(m.x[0]));
    for (int it = 0; it < 4; ++it) y = mf::add(y, mf::sub(mf::mul(m, h_exp(mf::neg(y))), mf::from<N>(1.0)));
    return y;
}

// logdet(S) = 2 log(prod L_ii) = 2 (e log 2 + log m)
template <int N> void logdet_host(const double *me, double *out) {
    mf::V<N> m;
    for (int l = 0; l < N; ++l) m.x[l] = me[l];
    const double e = me[N];
    mf::V<N> v;
    if (!(m.x[0] > 0.0) || !std::isfinite(m.x[0])) {
        v = mf::from<N>(m.x[0] == 0.0 ? -INFINITY : NAN);
    } else {
        const mf::V<N> ln2 = h_log(mf::from<N>(2.0));
        v = mf::add(mf::mul(mf::from<N>(e), ln2), h_log(m));
        v = mf::add(v, v);
    }
    for (int l = 0; l < N; ++l) out[l] = v.x[l];
}

int check_adjoint_factor(int N, const mfla_matrix &F, int64_t info) {
    if (check_limbs(N)) return MFLA_ERR_ARGUMENT;
    if (check_desc(F, "F")) return MFLA_ERR_ARGUMENT;
    if (F.rows != F.cols) return fail(MFLA_ERR_DIMENSION, "the Cholesky factor must be square");
    if (info != 0) return fail(MFLA_ERR_FAILED_FACTOR, "PosDefException(%lld)", (long long)info);
    return MFLA_OK;
}

// S^{-1} B in place from the Cholesky factor (the potrs path)
hipError_t potrs_core(mfla_handle h, int N, const mfla_matrix &F, const mfla_matrix &B, const mfla_config &c) {
    hipError_t e = trsm_core(h, N, MFLA_LEFT, MFLA_LOWER, MFLA_NOTRANS, MFLA_NONUNIT, B, F, kOne, c, nullptr);
    if (e != hipSuccess) return e;
    return trsm_core(h, N, MFLA_LEFT, MFLA_LOWER, MFLA_TRANS, MFLA_NONUNIT, B, F, kOne, c, nullptr);
}

}  // namespace

extern "C" {

int mfla_logdet(mfla_handle h, int N, mfla_matrix F, int64_t info, double *out) {
    if (int rc = check_adjoint_factor(N, F, info)) return rc;
    const int64_t n = F.rows;
    if (n == 0) { for (int l = 0; l < N; ++l) out[l] = 0.0; return MFLA_OK; }
    std::lock_guard<std::mutex> g(h->lock);
    double *buf = (double *)grow_scratch(h, sizeof(double) * (N + 1));
    if (!buf) return fail(MFLA_ERR_HIP, "scratch allocation failed");
    HIPCHK(mfk::diag_logprod(N, h->stream, n, dvn(F), buf));
    double me[5];
    {
        void *dst_[1] = {me};
        const void *src_[1] = {buf};
        const size_t by_[1] = {sizeof(double) * (N + 1)};
        HIPCHK(read_back(h, h->stream, 1, dst_, src_, by_));
    }
    switch (N) {
        case 1: logdet_host<1>(me, out); break;
        case 2: logdet_host<2>(me, out); break;
        case 3: logdet_host<3>(me, out); break;
        default: logdet_host<4>(me, out); break;
    }
    return MFLA_OK;
}

int mfla_solve_adjoint(mfla_handle h, int N, mfla_matrix F, int64_t info, mfla_matrix X, mfla_matrix Xbar, mfla_matrix Bbar,
                       mfla_matrix Sbar, const mfla_config *config) {
    mfla_config tmp;
    const mfla_config &c = cfg_or_default(config, tmp);
    if (int rc = check_adjoint_factor(N, F, info)) return rc;
    if (check_desc(X, "X") || check_desc(Xbar, "Xbar") || check_desc(Bbar, "Bbar") || check_desc(Sbar, "Sbar")) return MFLA_ERR_ARGUMENT;
    const int64_t n = F.rows, nrhs = X.cols;
    if (X.rows != n || Xbar.rows != n || Xbar.cols != nrhs || Bbar.rows != n || Bbar.cols != nrhs)
        return fail(MFLA_ERR_DIMENSION, "solution, cotangent and right-hand-side gradient dimensions differ");
    if (Sbar.rows != n || Sbar.cols != n) return fail(MFLA_ERR_DIMENSION, "matrix gradient dimensions differ");
    const bool inplace = same_matrix(Bbar, Xbar);
    if ((!inplace && extents_overlap(Bbar, Xbar, N)) || extents_overlap(Bbar, X, N) || extents_overlap(Bbar, F, N) ||
        extents_overlap(Sbar, Bbar, N) || extents_overlap(Sbar, X, N) || extents_overlap(Sbar, F, N) || extents_overlap(Sbar, Xbar, N))
        return fail(MFLA_ERR_ALIAS, "adjoint outputs must not alias the inputs");
    std::lock_guard<std::mutex> g(h->lock);
    hipStream_t s = h->stream;
    if (!inplace) HIPCHK(mfk::copy(N, s, n, nrhs, dvn(Bbar), dvn(Xbar)));
    HIPCHK(potrs_core(h, N, F, Bbar, c));                                    // Bbar = S^{-1} Xbar
    HIPCHK(gemm_launch(h, N, c.gemm_strategy, c.gemm_tile, n, n, nrhs, dvn(Bbar), dvt(X), dvn(Sbar), kOne, kZero, nullptr));
    const double mhalf[4] = {-0.5, 0.0, 0.0, 0.0};
    HIPCHK(mfk::symmetrize(N, s, n, dvn(Sbar), dvn(Sbar), mhalf));          // Sbar = -sym(Bbar X^T)
    return MFLA_OK;
}

int mfla_logdet_adjoint(mfla_handle h, int N, mfla_matrix F, int64_t info, const double *ybar, mfla_matrix Sbar,
                        const mfla_config *config) {
    mfla_config tmp;
    const mfla_config &c = cfg_or_default(config, tmp);
    if (int rc = check_adjoint_factor(N, F, info)) return rc;
    if (check_desc(Sbar, "Sbar")) return MFLA_ERR_ARGUMENT;
    const int64_t n = F.rows;
    if (Sbar.rows != n || Sbar.cols != n) return fail(MFLA_ERR_DIMENSION, "matrix gradient dimensions differ");
    if (extents_overlap(Sbar, F, N)) return fail(MFLA_ERR_ALIAS, "the gradient must not alias the factor");
    std::lock_guard<std::mutex> g(h->lock);
    hipStream_t s = h->stream;
    HIPCHK(mfk::identity(N, s, n, dvn(Sbar)));
    HIPCHK(potrs_core(h, N, F, Sbar, c));                                    // S^{-1}
    double half_ybar[4] = {0.0, 0.0, 0.0, 0.0};
    for (int l = 0; l < N; ++l) half_ybar[l] = 0.5 * ybar[l];
    HIPCHK(mfk::symmetrize(N, s, n, dvn(Sbar), dvn(Sbar), half_ybar));     // ybar * sym(S^{-1})
    return MFLA_OK;
}

int mfla_cholesky_adjoint(mfla_handle h, int N, mfla_matrix F, int64_t info, mfla_matrix Lbar, mfla_matrix Sbar, mfla_matrix work,
                          const mfla_config *config) {
    mfla_config tmp;
    const mfla_config &c = cfg_or_default(config, tmp);
    if (int rc = check_adjoint_factor(N, F, info)) return rc;
    if (check_desc(Lbar, "Lbar") || check_desc(Sbar, "Sbar") || check_desc(work, "work")) return MFLA_ERR_ARGUMENT;
    const int64_t n = F.rows;
    if (Lbar.rows != n || Lbar.cols != n || Sbar.rows != n || Sbar.cols != n || work.rows != n || work.cols != n)
        return fail(MFLA_ERR_DIMENSION, "adjoint operands must be n x n");
    if (extents_overlap(Sbar, F, N) || extents_overlap(Sbar, Lbar, N) || extents_overlap(work, F, N) || extents_overlap(work, Lbar, N) ||
        extents_overlap(work, Sbar, N))
        return fail(MFLA_ERR_ALIAS, "adjoint outputs must not alias the inputs");
    std::lock_guard<std::mutex> g(h->lock);
    hipStream_t s = h->stream;
    HIPCHK(mfk::tril(N, s, n, dvn(F), dvn(work), 0));                        // L (upper storage ignored)
    HIPCHK(gemm_launch(h, N, c.gemm_strategy, c.gemm_tile, n, n, n, dvt(work), dvn(Lbar), dvn(Sbar), kOne, kZero, nullptr));
    HIPCHK(mfk::tril(N, s, n, dvn(Sbar), dvn(Sbar), 1));                     // Phi(L^T Lbar)
    HIPCHK(trsm_core(h, N, MFLA_RIGHT, MFLA_LOWER, MFLA_NOTRANS, MFLA_NONUNIT, Sbar, F, kOne, c, nullptr));   // * L^{-1}
    HIPCHK(trsm_core(h, N, MFLA_LEFT, MFLA_LOWER, MFLA_TRANS, MFLA_NONUNIT, Sbar, F, kOne, c, nullptr));     // L^{-T} *
    const double half[4] = {0.5, 0.0, 0.0, 0.0};
    HIPCHK(mfk::symmetrize(N, s, n, dvn(Sbar), dvn(Sbar), half));
    return MFLA_OK;
}

}  // extern "C"

// ======================================================================
// trmm!, mfdot, syrk_packed!
extern "C" {

int mfla_trmm(mfla_handle h, int N, int side, int uplo, int trans, int diag, mfla_matrix B, mfla_matrix A, const double *alpha,
              const mfla_config *) {
    if (A.rows != A.cols) return fail(MFLA_ERR_DIMENSION, "trmm! requires a square triangular matrix");
    if (check_tr_symbols(side, uplo, trans, diag)) return MFLA_ERR_ARGUMENT;
    if (check_limbs(N)) return MFLA_ERR_ARGUMENT;
    if (check_desc(A, "A") || check_desc(B, "B")) return MFLA_ERR_ARGUMENT;
    if (side == MFLA_LEFT && B.rows != A.rows) return fail(MFLA_ERR_DIMENSION, "trmm! left dimensions differ");
    if (side == MFLA_RIGHT && B.cols != A.rows) return fail(MFLA_ERR_DIMENSION, "trmm! right dimensions differ");
    if (extents_overlap(B, A, N)) return fail(MFLA_ERR_ALIAS, "trmm! does not support output/input aliasing");
    const int64_t m = B.rows, n = B.cols, na = A.rows;
    if (m * n == 0) return MFLA_OK;
    std::lock_guard<std::mutex> g(h->lock);
    hipStream_t s = h->stream;
    // the products read B's original values: work from a copy
    double *bc = (double *)grow_scratch(h, sizeof(double) * (size_t)N * m * n);
    if (!bc) return fail(MFLA_ERR_HIP, "scratch allocation failed");
    const DV Bc{bc, 1, m, m * n};
    HIPCHK(mfk::copy(N, s, m, n, Bc, dvn(B)));
    const bool transposed = trans == MFLA_TRANS;
    const bool eff_lower = ((uplo == MFLA_LOWER) == !transposed);
    const DV opA = transposed ? dvt(A) : dvn(A);
    const int left = side == MFLA_LEFT;
    HIPCHK(mfk::trmm(N, s, left, eff_lower ? 1 : 0, diag == MFLA_UNIT, m, n, na, left ? opA : Bc, left ? Bc : opA, dvn(B),
                     alpha ? alpha : kOne));
    return MFLA_OK;
}

int mfla_dot(mfla_handle h, int N, mfla_matrix x, mfla_matrix y, double *out) {
    if (check_limbs(N)) return MFLA_ERR_ARGUMENT;
    if (check_desc(x, "x") || check_desc(y, "y")) return MFLA_ERR_ARGUMENT;
    if (x.rows != y.rows || x.cols != y.cols) return fail(MFLA_ERR_DIMENSION, "dot product shapes differ");
    std::lock_guard<std::mutex> g(h->lock);
    double *buf = (double *)grow_scratch(h, sizeof(double) * N);
    if (!buf) return fail(MFLA_ERR_HIP, "scratch allocation failed");
    HIPCHK(mfk::dot(N, h->stream, x.rows, x.cols, dvn(x), dvn(y), buf));
    {
        void *dst_[1] = {out};
        const void *src_[1] = {buf};
        const size_t by_[1] = {sizeof(double) * N};
        HIPCHK(read_back(h, h->stream, 1, dst_, src_, by_));
    }
    return MFLA_OK;
}

int mfla_syrk_packed(mfla_handle h, int N, mfla_matrix packed, mfla_matrix P, const double *alpha, const double *beta,
                     int64_t reduction_first, int64_t reduction_last, const mfla_config *config) {
    mfla_config tmp;
    const mfla_config &c = cfg_or_default(config, tmp);
    if (check_limbs(N)) return MFLA_ERR_ARGUMENT;
    if (check_desc(packed, "packed") || check_desc(P, "panel")) return MFLA_ERR_ARGUMENT;
    const int64_t n = P.cols;
    if (packed.cols != 1 || packed.rows != n * (n + 1) / 2) return fail(MFLA_ERR_DIMENSION, "syrk_packed! output length differs");
    if (!(1 <= reduction_first && reduction_first <= reduction_last && reduction_last <= P.rows))
        return fail(MFLA_ERR_BOUNDS, "reduction range %lld:%lld out of bounds for %lld rows", (long long)reduction_first,
                    (long long)reduction_last, (long long)P.rows);
    if (extents_overlap(packed, P, N)) return fail(MFLA_ERR_ALIAS, "syrk_packed! does not support output/input aliasing");
    if (n == 0) return MFLA_OK;
    std::lock_guard<std::mutex> g(h->lock);
    hipStream_t s = h->stream;
    double *dense = (double *)grow_scratch(h, sizeof(double) * (size_t)N * n * n);
    if (!dense) return fail(MFLA_ERR_HIP, "scratch allocation failed");
    const DV D{dense, 1, n, n * n};
    const double *be = beta ? beta : kZero;
    if (!scalar_is_zero(be, N)) HIPCHK(mfk::pack_lower(N, s, n, D, packed.data, packed.limb_stride, 1));
    const mfla_matrix Pr{P.data + (reduction_first - 1), reduction_last - reduction_first + 1, n, P.ld, P.limb_stride};
    HIPCHK(lower_launch(h, N, c.gemm_tile, n, Pr.rows, dvt(Pr), dvn(Pr), D, alpha ? alpha : kOne, be, nullptr));
    HIPCHK(mfk::pack_lower(N, s, n, D, packed.data, packed.limb_stride, 0));
    return MFLA_OK;
}

}  // extern "C"

// ================================================================== QR
namespace {

// Device state of one QR call, carved from the handle's scratch.
int qr_state(mfla_handle h, int N, int64_t n, mfk::QrState &q) {
    const size_t vec = (size_t)N * n, dbl = 4 * vec + (size_t)N * mfk::QR_PANEL * n + (size_t)N * mfk::QR_PANEL + N;
    const size_t bytes = sizeof(double) * dbl + sizeof(int) * 2 * (size_t)n + 64 + 64;
    char *base = (char *)grow_scratch(h, bytes);
    if (!base) return fail(MFLA_ERR_HIP, "scratch allocation failed (%zu bytes)", bytes);
    double *d = (double *)base;
    q.scale = d; q.sum = d + vec; q.tscale = d + 2 * vec; q.tsum = d + 3 * vec;
    q.F = d + 4 * vec; q.aux = q.F + (size_t)N * mfk::QR_PANEL * n; q.diag = q.aux + (size_t)N * mfk::QR_PANEL;
    char *tail = (char *)(q.diag + N);
    q.pivot = (int64_t *)tail;
    q.unreliable = (int *)(tail + 8);
    q.dirty = (int *)(tail + 64);
    q.tflag = q.dirty + n;
    return MFLA_OK;
}

int qr_check(int N, const mfla_matrix &A, const mfla_matrix &tau) {
    if (check_limbs(N)) return MFLA_ERR_ARGUMENT;
    if (check_desc(A, "A") || check_desc(tau, "tau")) return MFLA_ERR_ARGUMENT;
    if (tau.rows != std::min(A.rows, A.cols) || tau.cols != 1) return fail(MFLA_ERR_DIMENSION, "tau must hold min(m, n) reflectors");
    if (extents_overlap(A, tau, N)) return fail(MFLA_ERR_ALIAS, "tau must not alias the matrix");
    return MFLA_OK;
}

// Non-finite input (qr.jl:710-720): status -1, tau = 0, identity permutation.
int qr_finite(mfla_handle h, int N, const mfla_matrix &A, const mfla_matrix &tau, int64_t *perm, int64_t *info) {
    hipStream_t s = h->stream;
    HIPCHK(mfk::status_reset(s, h->status));
    HIPCHK(mfk::zero_fill(N, s, tau.rows, 1, dvn(tau)));
    if (perm) HIPCHK(mfk::iota1(s, perm, A.cols));
    HIPCHK(mfk::finite_check(N, s, A.rows, A.cols, dvn(A), 0, h->status));
    const int64_t r = read_info(h);
    if (r == INT64_MIN) return fail(MFLA_ERR_HIP, "status read failed");
    *info = r;
    return MFLA_OK;
}

}  // namespace

extern "C" {

int mfla_geqrf(mfla_handle h, int N, mfla_matrix A, mfla_matrix tau, int64_t *info) {
    if (int rc = qr_check(N, A, tau)) return rc;
    const int64_t m = A.rows, n = A.cols, k = std::min(m, n);
    std::lock_guard<std::mutex> g(h->lock);
    if (int rc = qr_finite(h, N, A, tau, nullptr, info)) return rc;
    if (*info != 0 || k == 0) return MFLA_OK;
    mfk::QrState q;
    if (int rc = qr_state(h, N, n, q)) return rc;
    hipStream_t s = h->stream;
    const DV a = dvn(A), t = dvn(tau);
    for (int64_t step = 0; step < k; ++step) {   // qr.jl:789-792
        HIPCHK(mfk::qr_reflector(N, s, m, step, a, t, 0, q));
        HIPCHK(mfk::qr_apply(N, s, m, step, a, step, a, step + 1, n - step - 1, t));
    }
    HIPCHK(hipStreamSynchronize(s));
    return MFLA_OK;
}

int mfla_geqp3(mfla_handle h, int N, mfla_matrix A, mfla_matrix tau, int64_t *perm, int route, int64_t *info) {
    if (int rc = qr_check(N, A, tau)) return rc;
    if (route < 0 || route > 2) return fail(MFLA_ERR_ARGUMENT, "route must be 0 (auto), 1 (unblocked) or 2 (blocked)");
    const int64_t m = A.rows, n = A.cols, k = std::min(m, n);
    std::lock_guard<std::mutex> g(h->lock);
    if (int rc = qr_finite(h, N, A, tau, perm, info)) return rc;
    if (*info != 0 || k == 0) return MFLA_OK;
    mfk::QrState q;
    if (int rc = qr_state(h, N, n, q)) return rc;
    hipStream_t s = h->stream;
    const DV a = dvn(A), t = dvn(tau);
    HIPCHK(mfk::qr_norms(N, s, m, n, a, 0, 0, n, q));
    // qr.jl:303-314: the DLAQPS panel route once there are 16 reflectors and
    // m n >= 16384
    const bool blocked = route == 2 || (route == 0 && k >= mfk::QR_PANEL && m * n >= 16384);
    if (!blocked) {
        for (int64_t step = 0; step < k; ++step) {   // qr.jl:364-398
            HIPCHK(mfk::qr_pivot_hybrid(N, s, m, n, step, a, perm, q));
            HIPCHK(mfk::qr_swap(N, s, m, n, step, 0, a, perm, q));
            HIPCHK(mfk::qr_reflector(N, s, m, step, a, t, 0, q));
            HIPCHK(mfk::qr_apply(N, s, m, step, a, step, a, step + 1, n - step - 1, t));
            HIPCHK(mfk::qr_norm_update(N, s, m, n, step, a, t, step, 0, 0, q));
        }
        HIPCHK(hipStreamSynchronize(s));
        return MFLA_OK;
    }
    const int64_t bsz = std::min<int64_t>(mfk::QR_PANEL, k);
    const DV F{q.F, 1, mfk::QR_PANEL, (int64_t)mfk::QR_PANEL * n};
    int unreliable = 0;
    for (int64_t bs = 0; bs < k;) {   // qr.jl:450-604
        const int64_t requested = std::min(bsz, k - bs);
        HIPCHK(mfk::qr_panel_reset(N, s, n, q));
        int64_t actual = 0;
        for (int64_t ls = 0; ls < requested; ++ls) {
            const int64_t step = bs + ls;
            HIPCHK(mfk::qr_pivot_downdated(N, s, n, step, perm, q));
            HIPCHK(mfk::qr_swap(N, s, m, n, step, ls, a, perm, q));
            HIPCHK(mfk::qr_delayed_column(N, s, m, n, step, bs, ls, a, q));
            HIPCHK(mfk::qr_reflector(N, s, m, step, a, t, 1, q));
            HIPCHK(mfk::qr_aux(N, s, m, n, step, bs, ls, a, t, q));
            HIPCHK(mfk::qr_frow(N, s, m, n, step, bs, ls, a, t, q));
            HIPCHK(mfk::qr_norm_update(N, s, m, n, step, a, t, bs, ls + 1, 1, q));
            actual = ls + 1;
            // a cancellation-sensitive norm ends the panel (qr.jl:588-592)
            {
                void *dst_[1] = {&unreliable};
                const void *src_[1] = {q.unreliable};
                const size_t by_[1] = {sizeof(int)};
                HIPCHK(read_back(h, s, 1, dst_, src_, by_));
            }
            if (unreliable) break;
        }
        const int64_t be = bs + actual - 1;
        if (be + 1 < m && be + 1 < n) {   // A22 -= A21 F(0:actual, be+1:n)
            const DV Fv{q.F + (be + 1) * mfk::QR_PANEL, 1, mfk::QR_PANEL, (int64_t)mfk::QR_PANEL * n};
            HIPCHK(gemm_launch(h, N, MFLA_GEMM_AUTO, -1, m - be - 1, n - be - 1, actual, sub(a, be + 1, bs), Fv, sub(a, be + 1, be + 1),
                               kMinusOne, kOne, h->status));
        }
        HIPCHK(mfk::qr_norms(N, s, m, n, a, be + 1, be + 1, n - be - 1, q));
        bs = be + 1;
    }
    (void)F;
    HIPCHK(hipStreamSynchronize(s));
    return MFLA_OK;
}

int mfla_ormqr(mfla_handle h, int N, int trans, mfla_matrix F, mfla_matrix tau, mfla_matrix B) {
    if (check_limbs(N)) return MFLA_ERR_ARGUMENT;
    if (check_desc(F, "F") || check_desc(tau, "tau") || check_desc(B, "B")) return MFLA_ERR_ARGUMENT;
    if (trans != MFLA_NOTRANS && trans != MFLA_TRANS) return fail(MFLA_ERR_ARGUMENT, "trans must be :N or :T");
    if (B.rows != F.rows) return fail(MFLA_ERR_DIMENSION, "apply_q! destination row count differs");
    if (tau.rows != std::min(F.rows, F.cols) || tau.cols != 1) return fail(MFLA_ERR_DIMENSION, "tau must hold min(m, n) reflectors");
    if (extents_overlap(B, F, N) || extents_overlap(B, tau, N))
        return fail(MFLA_ERR_ALIAS, "apply_q! destination must not alias QR storage");
    const int64_t m = F.rows, k = tau.rows, nrhs = B.cols;
    if (nrhs == 0 || k == 0) return MFLA_OK;
    std::lock_guard<std::mutex> g(h->lock);
    hipStream_t s = h->stream;
    for (int64_t i = 0; i < k; ++i) {   // qr.jl:951-954
        const int64_t step = trans == MFLA_TRANS ? i : k - 1 - i;
        HIPCHK(mfk::qr_apply(N, s, m, step, dvn(F), step, dvn(B), 0, nrhs, dvn(tau)));
    }
    return MFLA_OK;
}

}  // extern "C"
