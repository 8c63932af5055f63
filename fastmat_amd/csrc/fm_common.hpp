// Shared helpers for the fastmat_amd HIP library (gfx950 / CDNA4).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>

#include "../../include/fastmat_amd.h"

namespace fm {

// ----------------------------------------------------------------- errors
void set_error(const char *fmt, ...);

#define FM_CHECK_HIP(expr)                                                     \
    do {                                                                       \
        hipError_t _e = (expr);                                                \
        if (_e != hipSuccess) {                                                \
            ::fm::set_error("HIP error %s at %s:%d: %s", hipGetErrorName(_e),  \
                            __FILE__, __LINE__, hipGetErrorString(_e));        \
            return _e == hipErrorOutOfMemory ? FM_ERR_ALLOC : FM_ERR_HIP;      \
        }                                                                      \
    } while (0)

#define FM_REQUIRE(cond, code, ...)                                            \
    do {                                                                       \
        if (!(cond)) {                                                         \
            ::fm::set_error(__VA_ARGS__);                                      \
            return (code);                                                     \
        }                                                                      \
    } while (0)

// Launch check: reports launch-configuration errors synchronously (no sync
// with the device is done).
#define FM_CHECK_LAUNCH() FM_CHECK_HIP(hipGetLastError())

// ---------------------------------------------------------------- complex
template <typename T> struct alignas(2 * sizeof(T)) cpx {
    T x, y;
};

template <typename T>
__host__ __device__ __forceinline__ cpx<T> mk(T a, T b) {
    cpx<T> r;
    r.x = a;
    r.y = b;
    return r;
}
template <typename T>
__device__ __forceinline__ cpx<T> operator+(cpx<T> a, cpx<T> b) {
    return mk(a.x + b.x, a.y + b.y);
}
template <typename T>
__device__ __forceinline__ cpx<T> operator-(cpx<T> a, cpx<T> b) {
    return mk(a.x - b.x, a.y - b.y);
}
template <typename T>
__device__ __forceinline__ cpx<T> cmul(cpx<T> a, cpx<T> b) {
    return mk(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// a * conj(b)
template <typename T>
__device__ __forceinline__ cpx<T> cmulc(cpx<T> a, cpx<T> b) {
    return mk(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}
template <typename T> __device__ __forceinline__ cpx<T> conj(cpx<T> a) {
    return mk(a.x, -a.y);
}
template <typename T> __device__ __forceinline__ cpx<T> scale(cpx<T> a, T s) {
    return mk(a.x * s, a.y * s);
}
// multiply by -i (forward, SIGN=-1) or +i (inverse, SIGN=+1)
template <int SIGN, typename T>
__device__ __forceinline__ cpx<T> mul_i(cpx<T> a) {
    return SIGN < 0 ? mk(a.y, -a.x) : mk(-a.y, a.x);
}

// ------------------------------------------------------------ misc utils
__host__ __device__ inline int ilog2(int64_t v) {
    int r = 0;
    while ((int64_t(1) << r) < v) ++r;
    return r;
}
__host__ __device__ inline bool is_pow2(int64_t v) { return v > 0 && (v & (v - 1)) == 0; }

inline hipStream_t as_stream(void *s) { return reinterpret_cast<hipStream_t>(s); }

__host__ __device__ inline int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }

// Number of XCDs on the device (8 on MI355X).  Workgroup ids are dispatched
// round-robin over XCDs; kernels that share cache lines between neighbouring
// tiles remap blockIdx so that neighbours run on the same XCD (same L2).
constexpr int kNumXcd = 8;

// Map a hardware block id to a logical tile id such that logical tiles
// [g*per, (g+1)*per) all run on XCD g.  Exact permutation of [0, nblocks).
__device__ __forceinline__ int64_t xcd_remap(int64_t bid, int64_t nblocks) {
    if (nblocks < kNumXcd * 2) return bid;
    int64_t xcd = bid % kNumXcd;
    int64_t slot = bid / kNumXcd;
    int64_t full = nblocks / kNumXcd;     // blocks every XCD gets
    int64_t extra = nblocks % kNumXcd;    // first `extra` XCDs get one more
    int64_t start = xcd * full + (xcd < extra ? xcd : extra);
    return start + slot;
}

// ------------------------------------------------------------- profiler
struct ProfToken {
    bool active;
    hipEvent_t a, b;
};
bool prof_on();
void prof_begin(hipStream_t s, ProfToken *t);
// bytes: the launch's algorithmic bytes; pass_bytes: what the launch itself
// reads and writes (< 0: the same as bytes)
void prof_end(hipStream_t s, ProfToken *t, const char *name, double bytes, double flops,
              double pass_bytes = -1.0);
// prof_end for an elementwise kernel, labelled "<name><<element bytes>B>"
inline void prof_elem(hipStream_t s, ProfToken *t, const char *name, size_t elem_bytes,
                      double bytes) {
    if (!t->active) return;
    char label[64];
    snprintf(label, sizeof(label), "%s<%dB>", name, int(elem_bytes));
    prof_end(s, t, label, bytes, 0.0);
}

// Per-device cached tables (twiddles), keyed by (device, log2 size, precision).
const void *twiddle_table(int precision, int logn, hipStream_t stream, fm_status *st);

}  // namespace fm
