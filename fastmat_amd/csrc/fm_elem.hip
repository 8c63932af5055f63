// Element-wise operators of the hot path: Diag multiply, row gather/scatter
// (Permutation, Partial) and scaled accumulation (Sum, Product scalars).
//
// Reference behaviour replaced:
//   Diag._forwardC / _backwardC + cmath._multiply   Diag.pyx:149-167, cmath.pyx:1294-1326
//   Permutation._forward / _backward                Permutation.pyx:121-125
//   Partial._forward / _backward (scatter/gather)   Partial.pyx:268-294
//   Sum._forwardC / _backwardC                      Sum.pyx:151-175
//   Product scalar factor                           Product.pyx:213-238
// Integer types use wrap-around (two's-complement) arithmetic like the
// reference's C loops.  All kernels are grid-stride over an n x k block of
// column-major data and stream HBM once.
#include "fm_common.hpp"

namespace fm {

template <typename S> struct ElemOps {
    using U = S;
    __device__ static S mul(S a, S b) { return a * b; }
    __device__ static S mulc(S a, S b) { return a * b; }  // real: conj(b) == b
    __device__ static S add(S a, S b) { return a + b; }
    __device__ static S conjv(S a) { return a; }
    __host__ static S from(const double *v) { return S(v[0]); }
};
template <typename I> struct IntOps {
    using U = typename std::make_unsigned<I>::type;
    __device__ static I mul(I a, I b) { return I(U(U(a) * U(b))); }
    __device__ static I mulc(I a, I b) { return mul(a, b); }
    __device__ static I add(I a, I b) { return I(U(U(a) + U(b))); }
    __device__ static I conjv(I a) { return a; }
    __host__ static I from(const double *v) { return I(int64_t(v[0])); }
};
template <> struct ElemOps<int8_t> : IntOps<int8_t> {};
template <> struct ElemOps<int16_t> : IntOps<int16_t> {};
template <> struct ElemOps<int32_t> : IntOps<int32_t> {};
template <> struct ElemOps<int64_t> : IntOps<int64_t> {};
template <typename T> struct ElemOps<cpx<T>> {
    __device__ static cpx<T> mul(cpx<T> a, cpx<T> b) { return cmul(a, b); }
    __device__ static cpx<T> mulc(cpx<T> a, cpx<T> b) { return cmulc(a, b); }
    __device__ static cpx<T> add(cpx<T> a, cpx<T> b) { return a + b; }
    __device__ static cpx<T> conjv(cpx<T> a) { return conj(a); }
    __host__ static cpx<T> from(const double *v) { return mk(T(v[0]), T(v[1])); }
};

static inline unsigned grid_for(int64_t total) {
    int64_t g = ceil_div(total, 256);
    if (g > 65536) g = 65536;
    if (g < 1) g = 1;
    return unsigned(g);
}

// y[i, j] = x[i, j] * d[i] (or conj(d[i])), column-major n x k
template <typename S>
__global__ void diag_kernel(const S *__restrict__ d, const S *__restrict__ x, int64_t ldx,
                            S *__restrict__ y, int64_t ldy, int64_t n, int64_t k, int conj_d) {
    const int64_t total = n * k;
    for (int64_t t = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; t < total;
         t += int64_t(gridDim.x) * blockDim.x) {
        const int64_t j = t / n, i = t - j * n;
        const S v = x[j * ldx + i];
        y[j * ldy + i] = conj_d ? ElemOps<S>::mulc(v, d[i]) : ElemOps<S>::mul(v, d[i]);
    }
}

template <typename S>
__global__ void gather_kernel(const int64_t *__restrict__ idx, const S *__restrict__ x,
                              int64_t ldx, S *__restrict__ y, int64_t ldy, int64_t n_out,
                              int64_t k) {
    const int64_t total = n_out * k;
    for (int64_t t = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; t < total;
         t += int64_t(gridDim.x) * blockDim.x) {
        const int64_t j = t / n_out, i = t - j * n_out;
        y[j * ldy + i] = x[j * ldx + idx[i]];
    }
}

template <typename S>
__global__ void zero_kernel(S *__restrict__ y, int64_t ldy, int64_t n, int64_t k) {
    const int64_t total = n * k;
    for (int64_t t = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; t < total;
         t += int64_t(gridDim.x) * blockDim.x) {
        const int64_t j = t / n, i = t - j * n;
        y[j * ldy + i] = S{};
    }
}

template <typename S>
__global__ void scatter_kernel(const int64_t *__restrict__ idx, const S *__restrict__ x,
                               int64_t ldx, S *__restrict__ y, int64_t ldy, int64_t n_in,
                               int64_t k) {
    const int64_t total = n_in * k;
    for (int64_t t = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; t < total;
         t += int64_t(gridDim.x) * blockDim.x) {
        const int64_t j = t / n_in, i = t - j * n_in;
        y[j * ldy + idx[i]] = x[j * ldx + i];
    }
}

template <typename S>
__global__ void axpby_kernel(S alpha, const S *__restrict__ x, int64_t ldx, int conj_x, S beta,
                             int use_beta, S *__restrict__ y, int64_t ldy, int64_t n, int64_t k) {
    const int64_t total = n * k;
    for (int64_t t = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; t < total;
         t += int64_t(gridDim.x) * blockDim.x) {
        const int64_t j = t / n, i = t - j * n;
        S v = x[j * ldx + i];
        if (conj_x) v = ElemOps<S>::conjv(v);
        v = ElemOps<S>::mul(alpha, v);
        if (use_beta) v = ElemOps<S>::add(v, ElemOps<S>::mul(beta, y[j * ldy + i]));
        y[j * ldy + i] = v;
    }
}

#define FM_DISPATCH_ALL(dtype, FN, ...)                                                 \
    switch (dtype) {                                                                  \
    case FM_INT8: return FN<int8_t>(__VA_ARGS__);                                     \
    case FM_INT16: return FN<int16_t>(__VA_ARGS__);                                   \
    case FM_INT32: return FN<int32_t>(__VA_ARGS__);                                   \
    case FM_INT64: return FN<int64_t>(__VA_ARGS__);                                   \
    case FM_FLOAT32: return FN<float>(__VA_ARGS__);                                   \
    case FM_FLOAT64: return FN<double>(__VA_ARGS__);                                  \
    case FM_COMPLEX64: return FN<cpx<float>>(__VA_ARGS__);                            \
    case FM_COMPLEX128: return FN<cpx<double>>(__VA_ARGS__);                          \
    default: set_error("Unsupported type: %d", dtype); return FM_ERR_TYPE;            \
    }

template <typename S>
static fm_status diag_t(int conj_d, int64_t n, const void *d, const void *x, int64_t ldx,
                        void *y, int64_t ldy, int64_t k, hipStream_t s) {
    ProfToken tok;
    prof_begin(s, &tok);
    hipLaunchKernelGGL(diag_kernel<S>, grid_for(n * k), 256, 0, s, (const S *)d, (const S *)x,
                       ldx, (S *)y, ldy, n, k, conj_d);
    FM_CHECK_LAUNCH();
    prof_elem(s, &tok, "diag_kernel", sizeof(S), (2.0 * double(n) * double(k) + double(n)) *
                                                     sizeof(S));
    return FM_SUCCESS;
}

template <typename S>
static fm_status gather_t(int64_t n_out, const int64_t *idx, const void *x, int64_t ldx, void *y,
                          int64_t ldy, int64_t k, hipStream_t s) {
    ProfToken tok;
    prof_begin(s, &tok);
    hipLaunchKernelGGL(gather_kernel<S>, grid_for(n_out * k), 256, 0, s, idx, (const S *)x, ldx,
                       (S *)y, ldy, n_out, k);
    FM_CHECK_LAUNCH();
    prof_elem(s, &tok, "gather_kernel", sizeof(S),
              2.0 * double(n_out) * double(k) * sizeof(S) + 8.0 * double(n_out));
    return FM_SUCCESS;
}

template <typename S>
static fm_status scatter_t(int64_t n_in, const int64_t *idx, const void *x, int64_t ldx, void *y,
                           int64_t ldy, int64_t n_out, int64_t k, hipStream_t s) {
    ProfToken tok;
    prof_begin(s, &tok);
    hipLaunchKernelGGL(zero_kernel<S>, grid_for(n_out * k), 256, 0, s, (S *)y, ldy, n_out, k);
    hipLaunchKernelGGL(scatter_kernel<S>, grid_for(n_in * k), 256, 0, s, idx, (const S *)x, ldx,
                       (S *)y, ldy, n_in, k);
    FM_CHECK_LAUNCH();
    // zero fill and scatter together: every output row written, every input
    // row and index read
    prof_elem(s, &tok, "scatter_rows", sizeof(S),
              (double(n_in) + double(n_out)) * double(k) * sizeof(S) + 8.0 * double(n_in));
    return FM_SUCCESS;
}

template <typename S>
static fm_status axpby_t(int64_t n, int64_t k, const double *alpha, const void *x, int64_t ldx,
                         int conj_x, const double *beta, void *y, int64_t ldy, hipStream_t s) {
    const int use_beta = (beta[0] != 0.0 || beta[1] != 0.0) ? 1 : 0;
    ProfToken tok;
    prof_begin(s, &tok);
    hipLaunchKernelGGL(axpby_kernel<S>, grid_for(n * k), 256, 0, s, ElemOps<S>::from(alpha),
                       (const S *)x, ldx, conj_x, ElemOps<S>::from(beta), use_beta, (S *)y, ldy,
                       n, k);
    FM_CHECK_LAUNCH();
    prof_elem(s, &tok, "axpby_kernel", sizeof(S),
              (2.0 + use_beta) * double(n) * double(k) * sizeof(S));
    return FM_SUCCESS;
}

}  // namespace fm

using namespace fm;

extern "C" fm_status fm_diag_apply(int dtype, int conj_d, int64_t n, const void *d, const void *x,
                                   int64_t ldx, void *y, int64_t ldy, int64_t k, void *stream) {
    FM_REQUIRE(n >= 0 && k >= 0, FM_ERR_VALUE, "negative size");
    if (n == 0 || k == 0) return FM_SUCCESS;
    FM_REQUIRE(d && x && y, FM_ERR_VALUE, "NULL data pointer");
    FM_REQUIRE(ldx >= n && ldy >= n, FM_ERR_VALUE, "leading dimension below n");
    FM_DISPATCH_ALL(dtype, diag_t, conj_d, n, d, x, ldx, y, ldy, k, as_stream(stream));
}

extern "C" fm_status fm_gather_rows(int dtype, int64_t n_out, const int64_t *idx, const void *x,
                                    int64_t ldx, void *y, int64_t ldy, int64_t k, void *stream) {
    FM_REQUIRE(n_out >= 0 && k >= 0, FM_ERR_VALUE, "negative size");
    if (n_out == 0 || k == 0) return FM_SUCCESS;
    FM_REQUIRE(idx && x && y, FM_ERR_VALUE, "NULL data pointer");
    FM_REQUIRE(ldy >= n_out, FM_ERR_VALUE, "leading dimension below n_out");
    FM_DISPATCH_ALL(dtype, gather_t, n_out, idx, x, ldx, y, ldy, k, as_stream(stream));
}

extern "C" fm_status fm_scatter_rows(int dtype, int64_t n_in, const int64_t *idx, const void *x,
                                     int64_t ldx, void *y, int64_t ldy, int64_t n_out, int64_t k,
                                     void *stream) {
    FM_REQUIRE(n_in >= 0 && n_out >= 0 && k >= 0, FM_ERR_VALUE, "negative size");
    if (n_out == 0 || k == 0) return FM_SUCCESS;
    FM_REQUIRE(x && y && (idx || n_in == 0), FM_ERR_VALUE, "NULL data pointer");
    FM_REQUIRE(ldy >= n_out && ldx >= n_in, FM_ERR_VALUE, "leading dimension too small");
    FM_DISPATCH_ALL(dtype, scatter_t, n_in, idx, x, ldx, y, ldy, n_out, k, as_stream(stream));
}

extern "C" fm_status fm_axpby(int dtype, int64_t n, int64_t k, const double *alpha, const void *x,
                              int64_t ldx, int conj_x, const double *beta, void *y, int64_t ldy,
                              void *stream) {
    FM_REQUIRE(n >= 0 && k >= 0, FM_ERR_VALUE, "negative size");
    if (n == 0 || k == 0) return FM_SUCCESS;
    FM_REQUIRE(alpha && beta && x && y, FM_ERR_VALUE, "NULL argument");
    FM_REQUIRE(ldx >= n && ldy >= n, FM_ERR_VALUE, "leading dimension below n");
    FM_DISPATCH_ALL(dtype, axpby_t, n, k, alpha, x, ldx, conj_x, beta, y, ldy, as_stream(stream));
}

namespace fm {
// Plain device copy in 16-byte vectors (grid-strided, four loads in flight
// per thread before their stores), the reference point for HBM bandwidth and
// for scaling the FETCH_SIZE / WRITE_SIZE counters.
__global__ __launch_bounds__(256) void copy16_kernel(const uint4 *__restrict__ src,
                                                      uint4 *__restrict__ dst, int64_t n) {
    constexpr int U = 4;
    const int64_t stride = int64_t(gridDim.x) * blockDim.x;
    int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
    for (; i + (U - 1) * stride < n; i += U * stride) {
        uint4 v[U];
#pragma unroll
        for (int u = 0; u < U; ++u) v[u] = src[i + u * stride];
#pragma unroll
        for (int u = 0; u < U; ++u) dst[i + u * stride] = v[u];
    }
    for (; i < n; i += stride) dst[i] = src[i];
}
}  // namespace fm

extern "C" fm_status fm_device_copy(void *dst, const void *src, int64_t bytes, void *stream) {
    FM_REQUIRE(bytes >= 0, FM_ERR_VALUE, "negative size");
    if (bytes == 0) return FM_SUCCESS;
    FM_REQUIRE(dst && src && (bytes % 16) == 0 && (uintptr_t(dst) % 16) == 0 &&
                   (uintptr_t(src) % 16) == 0,
               FM_ERR_VALUE, "copy needs 16-byte aligned pointers and a multiple of 16 bytes");
    const int64_t n = bytes / 16;
    int ncu = 256;
    int dev = 0;
    if (hipGetDevice(&dev) == hipSuccess)
        (void)hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
    int64_t grid = ceil_div(n, 256);
    if (grid > int64_t(ncu) * 8) grid = int64_t(ncu) * 8;
    ProfToken tok;
    prof_begin(as_stream(stream), &tok);
    hipLaunchKernelGGL(copy16_kernel, dim3(unsigned(grid)), dim3(256), 0, as_stream(stream),
                       (const uint4 *)src, (uint4 *)dst, n);
    FM_CHECK_LAUNCH();
    if (tok.active) prof_end(as_stream(stream), &tok, "copy16_kernel", 2.0 * double(bytes), 0.0);
    return FM_SUCCESS;
}
