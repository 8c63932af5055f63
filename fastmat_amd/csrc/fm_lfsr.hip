// LFSRCirculant (LFSRCirculant.pyx:159-237, 316-407).
//
// The circulant matrix of a +-1 shift-register sequence is applied as
// scatter -> Walsh-Hadamard transform -> gather (the reference's _core,
// LFSRCirculant.pyx:316-396): input row t goes to the register state reached
// after t generator steps, the output reads the states reached by the
// tap-side recurrence from 1.  The state sequences are sequential
// recurrences and are produced on the host (fm_lfsr_sequence).
//
// The apply (fm_lfsr_apply) works in the row layout: the states are a
// pseudo-random permutation, so scattering single elements of each column
// would write one 4-16 byte element per cache line.  Instead the scatter
// moves whole rows (all k columns of one element, contiguous in a (2^n, k)
// row-major work array) through an LDS-transposed 64 x 64 tile; the
// Hadamard transform then runs along the strided axis of that array
// (fm_hadamard_apply_strided, inner = k), and the gather reads whole rows
// back through the same tile transposition into the caller's columns.
#include "fm_common.hpp"

namespace {

// one step of the generating register: the new top bit is the parity of the
// tapped bits (LFSRCirculant.pyx:31-40)
inline uint32_t gen_step(uint32_t state, uint32_t poly, uint32_t mask) {
    if (__builtin_parity(state & poly)) state |= mask;
    return state >> 1;
}

// one step of the tap-side register (LFSRCirculant.pyx:42-48)
inline uint32_t tap_step(uint32_t state, uint32_t poly, uint32_t mask) {
    state <<= 1;
    if (state & mask) state ^= (poly | mask);
    return state;
}

}  // namespace

extern "C" fm_status fm_lfsr_sequence(uint32_t polynomial, uint32_t start, int order,
                                      int64_t *period, int64_t *states, int64_t *taps,
                                      int64_t capacity) {
    FM_REQUIRE(order >= 1 && order <= 31, FM_ERR_VALUE,
               "Only polynomials of order 1 to 31 are supported.");
    FM_REQUIRE(period != nullptr, FM_ERR_VALUE, "NULL period");
    const uint32_t mask = uint32_t(1) << order;
    const uint32_t st = start & (mask - 1);
    FM_REQUIRE(st != 0, FM_ERR_VALUE, "Initial state must be non-zero.");
    uint32_t s = gen_step(st, polynomial, mask);
    int64_t p = 1;
    while (s != st) {
        s = gen_step(s, polynomial, mask);
        ++p;
        FM_REQUIRE(p < int64_t(mask) && s != 0, FM_ERR_VALUE,
                   "Register configuration produces invalid sequence.");
    }
    *period = p;
    if (!states && !taps) return FM_SUCCESS;
    FM_REQUIRE(capacity >= p, FM_ERR_VALUE, "sequence buffers hold %lld < %lld entries",
               (long long)capacity, (long long)p);
    if (states) {
        uint32_t a = st;
        for (int64_t t = 0; t < p; ++t) {
            states[t] = int64_t(a & (mask - 1));
            a = gen_step(a, polynomial, mask);
        }
    }
    if (taps) {
        uint32_t b = 1;
        for (int64_t t = 0; t < p; ++t) {
            taps[t] = int64_t(b);
            b = tap_step(b, polynomial, mask);
        }
    }
    return FM_SUCCESS;
}

namespace fm {

constexpr int kLT = 64;  // tile: 64 elements x kLC<S> columns, 256 threads

// columns per tile: 64, or 32 for 16-byte elements (the tile stays <= 33 KB)
template <typename S> constexpr int lfsr_cols() { return sizeof(S) > 8 ? 32 : 64; }

// Z[idx[i], c] = X[c, i]   (X: k columns of n, leading dimension ldx;
// Z: (2^order, k) row-major)
template <typename S>
__global__ __launch_bounds__(256) void lfsr_scatter_kernel(int64_t n, int64_t k,
                                                           const int64_t *__restrict__ idx,
                                                           const S *__restrict__ x, int64_t ldx,
                                                           S *__restrict__ z) {
    constexpr int TC = lfsr_cols<S>();
    __shared__ S tile[TC][kLT + 1];
    const int64_t i0 = int64_t(blockIdx.x) * kLT, c0 = int64_t(blockIdx.y) * TC;
    // load: lanes along the elements of one column (coalesced reads)
    for (int e = threadIdx.x; e < TC * kLT; e += 256) {
        const int cc = e / kLT, ii = e % kLT;
        const int64_t c = c0 + cc, i = i0 + ii;
        if (c < k && i < n) tile[cc][ii] = x[c * ldx + i];
    }
    __syncthreads();
    // store: lanes along the columns of one row (contiguous row writes)
    for (int e = threadIdx.x; e < TC * kLT; e += 256) {
        const int ii = e / TC, cc = e % TC;
        const int64_t i = i0 + ii, c = c0 + cc;
        if (i < n && c < k) z[idx[i] * k + c] = tile[cc][ii];
    }
}

// Y[c, i] = W[idx[i], c]
template <typename S>
__global__ __launch_bounds__(256) void lfsr_gather_kernel(int64_t n, int64_t k,
                                                          const int64_t *__restrict__ idx,
                                                          const S *__restrict__ w,
                                                          S *__restrict__ y, int64_t ldy) {
    constexpr int TC = lfsr_cols<S>();
    __shared__ S tile[TC][kLT + 1];
    const int64_t i0 = int64_t(blockIdx.x) * kLT, c0 = int64_t(blockIdx.y) * TC;
    for (int e = threadIdx.x; e < TC * kLT; e += 256) {
        const int ii = e / TC, cc = e % TC;
        const int64_t i = i0 + ii, c = c0 + cc;
        if (i < n && c < k) tile[cc][ii] = w[idx[i] * k + c];
    }
    __syncthreads();
    for (int e = threadIdx.x; e < TC * kLT; e += 256) {
        const int cc = e / kLT, ii = e % kLT;
        const int64_t c = c0 + cc, i = i0 + ii;
        if (c < k && i < n) y[c * ldy + i] = tile[cc][ii];
    }
}

template <typename S>
__global__ void lfsr_zero_kernel(S *z, int64_t count) {
    for (int64_t e = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; e < count;
         e += int64_t(gridDim.x) * blockDim.x)
        z[e] = S{};
}

template <typename S>
static fm_status lfsr_apply_t(int order, int dtype, int64_t n, const int64_t *sc,
                              const int64_t *ga, const void *x, int64_t ldx, void *y,
                              int64_t ldy, int64_t k, void *work, hipStream_t s) {
    const int64_t L = int64_t(1) << order;
    S *z = (S *)work;
    S *w = z + L * k;
    // rows no state reaches must be zero: only row 0 for a full period
    const int64_t nzero = (n == L - 1) ? k : L * k;
    int64_t zb = ceil_div(nzero, 256);
    if (zb > 8192) zb = 8192;
    hipLaunchKernelGGL(lfsr_zero_kernel<S>, dim3(unsigned(zb)), dim3(256), 0, s, z, nzero);
    FM_CHECK_LAUNCH();
    const dim3 grid{unsigned(ceil_div(n, kLT)), unsigned(ceil_div(k, lfsr_cols<S>())), 1u};
    ProfToken tok;
    prof_begin(s, &tok);
    hipLaunchKernelGGL(lfsr_scatter_kernel<S>, grid, dim3(256), 0, s, n, k, sc, (const S *)x, ldx,
                       z);
    FM_CHECK_LAUNCH();
    prof_elem(s, &tok, "lfsr_scatter_kernel", sizeof(S),
              2.0 * double(n) * double(k) * sizeof(S) + 8.0 * double(n));
    fm_status st = fm_hadamard_apply_strided(dtype, order, z, w, 1, k, s);
    if (st) return st;
    prof_begin(s, &tok);
    hipLaunchKernelGGL(lfsr_gather_kernel<S>, grid, dim3(256), 0, s, n, k, ga, (const S *)w, (S *)y,
                       ldy);
    FM_CHECK_LAUNCH();
    prof_elem(s, &tok, "lfsr_gather_kernel", sizeof(S),
              2.0 * double(n) * double(k) * sizeof(S) + 8.0 * double(n));
    return FM_SUCCESS;
}

}  // namespace fm

using namespace fm;

extern "C" fm_status fm_lfsr_apply(int dtype, int order, int64_t n, const int64_t *scatter_idx,
                                   const int64_t *gather_idx, const void *x, int64_t ldx,
                                   void *y, int64_t ldy, int64_t k, void *work,
                                   int64_t work_bytes, void *stream) {
    FM_REQUIRE(order >= 1 && order <= 31, FM_ERR_VALUE,
               "Only polynomials of order 1 to 31 are supported.");
    const int64_t L = int64_t(1) << order;
    FM_REQUIRE(n >= 1 && n < L, FM_ERR_VALUE, "LFSR period %lld outside [1, 2^order)",
               (long long)n);
    FM_REQUIRE(k >= 0 && k <= 65535 * 32, FM_ERR_VALUE, "bad column count");
    if (k == 0) return FM_SUCCESS;
    FM_REQUIRE(ldx >= n && ldy >= n, FM_ERR_VALUE, "leading dimension below the period");
    FM_REQUIRE(scatter_idx && gather_idx && x && y && work, FM_ERR_VALUE, "NULL argument");
    hipStream_t s = as_stream(stream);
#define FM_LFSR_CASE(ID, TYPE)                                                           \
    case ID:                                                                             \
        FM_REQUIRE(work_bytes >= 2 * L * k * int64_t(sizeof(TYPE)), FM_ERR_VALUE,        \
                   "workspace of %lld bytes is smaller than the %lld required",          \
                   (long long)work_bytes, (long long)(2 * L * k * int64_t(sizeof(TYPE)))); \
        return lfsr_apply_t<TYPE>(order, dtype, n, scatter_idx, gather_idx, x, ldx, y, ldy, k, \
                                  work, s);
    switch (dtype) {
        FM_LFSR_CASE(FM_INT8, int8_t)
        FM_LFSR_CASE(FM_INT16, int16_t)
        FM_LFSR_CASE(FM_INT32, int32_t)
        FM_LFSR_CASE(FM_INT64, int64_t)
        FM_LFSR_CASE(FM_FLOAT32, float)
        FM_LFSR_CASE(FM_FLOAT64, double)
        FM_LFSR_CASE(FM_COMPLEX64, cpx<float>)
        FM_LFSR_CASE(FM_COMPLEX128, cpx<double>)
    default:
        set_error("Unsupported type: %d", dtype);
        return FM_ERR_TYPE;
    }
#undef FM_LFSR_CASE
}
