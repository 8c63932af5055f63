// Device side of the iterative shrinkage solvers (ISTA / FISTA).
//
// Replaces the per-step numpy arithmetic of fastmat's algorithms
// (algorithms/ISTA.py:152-168, algorithms/FISTA.py:154-178): after the
// operator's forward and adjoint applies produce the gradient
// g = A^H (A y - b), one kernel forms
//     step   = y - L g
//     x_new  = step * m / (m + alpha),   m = max(|step| - alpha, 0)
//     y_next = x_new + beta (x_new - x_old)          (FISTA only)
// in the data's own precision, elementwise, with no host round trip.  The
// FISTA momentum beta of step k is read from a device table through a device
// step counter, so one captured iteration (forward, residual, adjoint,
// shrink, counter advance) can be replayed as a HIP graph for every step.
#include "fm_common.hpp"

namespace fm {

template <typename T> struct ShrinkOps {
    __device__ static T magnitude(T v) { return fabs(v); }
    __device__ static T scaled(T v, T s) { return v * s; }
    __device__ static T axpy(T a, T x, T y) { return y - a * x; }  // y - a x
    __device__ static T add(T a, T b) { return a + b; }
    __device__ static T sub(T a, T b) { return a - b; }
};
template <typename R> struct ShrinkOps<cpx<R>> {
    __device__ static R magnitude(cpx<R> v) { return hypot(v.x, v.y); }
    __device__ static cpx<R> scaled(cpx<R> v, R s) { return mk(v.x * s, v.y * s); }
    __device__ static cpx<R> axpy(R a, cpx<R> x, cpx<R> y) {
        return mk(y.x - a * x.x, y.y - a * x.y);
    }
    __device__ static cpx<R> add(cpx<R> a, cpx<R> b) { return mk(a.x + b.x, a.y + b.y); }
    __device__ static cpx<R> sub(cpx<R> a, cpx<R> b) { return mk(a.x - b.x, a.y - b.y); }
};

template <typename V, typename R>
__global__ void shrink_kernel(int64_t n, const V *y, const V *__restrict__ g, R L,
                              R alpha, V *__restrict__ step, V *x, const double *beta_tab,
                              const int64_t *counter, R beta_const, V *y_next) {
    using Op = ShrinkOps<V>;
    R beta = beta_const;
    if (beta_tab) beta = R(beta_tab[counter ? *counter : 0]);
    for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n;
         i += int64_t(gridDim.x) * blockDim.x) {
        const V s = Op::axpy(L, g[i], y[i]);
        const R m = fmax(Op::magnitude(s) - alpha, R(0));
        const V xn = Op::scaled(s, m / (m + alpha));
        if (y_next) {
            const V xo = x[i];
            y_next[i] = Op::add(xn, Op::scaled(Op::sub(xn, xo), beta));
        }
        step[i] = s;
        x[i] = xn;
    }
}

__global__ void counter_advance_kernel(int64_t *counter) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *counter += 1;
}

template <typename V, typename R>
static fm_status shrink_t(int64_t n, const void *y, const void *g, double L, double alpha,
                          void *step, void *x, const double *beta_tab, const int64_t *counter,
                          double beta, void *y_next, hipStream_t s) {
    int64_t blocks = ceil_div(n, 256);
    if (blocks > 8192) blocks = 8192;
    ProfToken tok;
    prof_begin(s, &tok);
    hipLaunchKernelGGL((shrink_kernel<V, R>), dim3(unsigned(blocks)), dim3(256), 0, s, n,
                       (const V *)y, (const V *)g, R(L), R(alpha), (V *)step, (V *)x, beta_tab,
                       counter, R(beta), (V *)y_next);
    FM_CHECK_LAUNCH();
    // reads y, g (and x_old for the momentum), writes step, x (and y_next)
    prof_elem(s, &tok, "shrink_kernel", sizeof(V),
              (y_next ? 6.0 : 4.0) * double(n) * sizeof(V));
    return FM_SUCCESS;
}

// ------------------------------------------------------------------ STELA
// Soft-thresholding with exact line search (algorithms/STELA.py:128-267),
// per column c of the k x n iterate, all in double precision:
//   stop_c  = || z - clamp(z - x, +-lambda) ||          (clamp per re / im)
//   active  = stop_c > max_error
//   bx      = soft(d x - z, lambda) / d                 (d = 1 / colnorm^2)
//   gamma_c = clamp(-(Re<res, A(bx - x)> + lambda sum(|bx| - |x|))
//                   / ||A(bx - x)||^2, 0, 1)
//   x += gamma (bx - x),  res += gamma A(bx - x),  z = A^H res   (active only)
// Column sums are two-stage and deterministic: a (parts x k) grid writes one
// partial per block, a k-thread kernel adds them in order.  The iterate x and
// the residual are real or complex double (X); z, the adjoint of the residual,
// is complex double also for a real iterate when the operator is complex
// (Z), as numpy keeps it in the reference; bx then keeps the real part, the
// reference's assignment into its real array.

__device__ __forceinline__ double np_min(double a, double b) { return (a != a || a < b) ? a : b; }
__device__ __forceinline__ double np_max(double a, double b) { return (a != a || a > b) ? a : b; }
__device__ __forceinline__ double clampl(double v, double lam) { return np_max(np_min(v, lam), -lam); }

__device__ __forceinline__ double re_(double v) { return v; }
__device__ __forceinline__ double re_(cpx<double> v) { return v.x; }
__device__ __forceinline__ double im_(double) { return 0.0; }
__device__ __forceinline__ double im_(cpx<double> v) { return v.y; }
__device__ __forceinline__ double abs2_(double v) { return v * v; }
__device__ __forceinline__ double abs2_(cpx<double> v) { return v.x * v.x + v.y * v.y; }
__device__ __forceinline__ double abs_(double v) { return fabs(v); }
__device__ __forceinline__ double abs_(cpx<double> v) { return hypot(v.x, v.y); }
template <typename V> __device__ __forceinline__ V from_ri(double r, double i);
template <> __device__ __forceinline__ double from_ri<double>(double r, double) { return r; }
template <> __device__ __forceinline__ cpx<double> from_ri<cpx<double>>(double r, double i) {
    return mk(r, i);
}

template <int NT>
__device__ __forceinline__ void block_sum2(double &a, double &b, double *sa, double *sb) {
    sa[threadIdx.x] = a;
    sb[threadIdx.x] = b;
    __syncthreads();
    for (int w = NT / 2; w > 0; w >>= 1) {
        if (int(threadIdx.x) < w) {
            sa[threadIdx.x] += sa[threadIdx.x + w];
            sb[threadIdx.x] += sb[threadIdx.x + w];
        }
        __syncthreads();
    }
    a = sa[0];
    b = sb[0];
}

constexpr int kStelaNT = 256;

// partials layout: [slot][column][part], slots 0 stop^2, 1 sum(|bx|-|x|),
// 2 Re<res, f>, 3 ||f||^2
template <typename X, typename Z>
__global__ __launch_bounds__(kStelaNT) void stela_prox_kernel(
    int64_t n, int64_t k, const X *__restrict__ x, const Z *__restrict__ z,
    const double *__restrict__ d, double lam, X *__restrict__ bx, X *__restrict__ dlt,
    double *__restrict__ part, int64_t parts) {
    __shared__ double sa[kStelaNT], sb[kStelaNT];
    const int64_t c = blockIdx.y, p = blockIdx.x;
    const int64_t chunk = ceil_div(n, parts);
    const int64_t lo = p * chunk, hi = (lo + chunk < n) ? lo + chunk : n;
    const bool cplx_iter = sizeof(X) == sizeof(cpx<double>);
    double stop2 = 0.0, l1 = 0.0;
    for (int64_t i = lo + threadIdx.x; i < hi; i += kStelaNT) {
        const int64_t e = c * n + i;
        const X xv = x[e];
        const Z zv = z[e];
        const double dv = d[i];
        // z - diff, diff clamped per component (only the real one for a
        // real iterate)
        const double dr = clampl(re_(zv) - re_(xv), lam);
        const double di = cplx_iter ? clampl(im_(zv) - im_(xv), lam) : 0.0;
        const double er = re_(zv) - dr, ei = im_(zv) - di;
        stop2 += er * er + ei * ei;
        // gradient d x - z and its soft threshold, divided by d
        const double gr = dv * re_(xv) - re_(zv), gi = dv * im_(xv) - im_(zv);
        const double mag = (sizeof(Z) == sizeof(cpx<double>)) ? hypot(gr, gi) : fabs(gr);
        const double m = np_max(mag - lam, 0.0);
        const double sc = m / (m + lam);
        const X b = from_ri<X>(sc * gr / dv, sc * gi / dv);
        bx[e] = b;
        dlt[e] = from_ri<X>(re_(b) - re_(xv), im_(b) - im_(xv));
        l1 += abs_(b) - abs_(xv);
    }
    block_sum2<kStelaNT>(stop2, l1, sa, sb);
    if (threadIdx.x == 0) {
        part[(0 * k + c) * parts + p] = stop2;
        part[(1 * k + c) * parts + p] = l1;
    }
}

template <typename X>
__global__ __launch_bounds__(kStelaNT) void stela_line_kernel(
    int64_t m, int64_t k, const X *__restrict__ res, const X *__restrict__ f,
    double *__restrict__ part, int64_t parts) {
    __shared__ double sa[kStelaNT], sb[kStelaNT];
    const int64_t c = blockIdx.y, p = blockIdx.x;
    const int64_t chunk = ceil_div(m, parts);
    const int64_t lo = p * chunk, hi = (lo + chunk < m) ? lo + chunk : m;
    double num = 0.0, den = 0.0;
    for (int64_t i = lo + threadIdx.x; i < hi; i += kStelaNT) {
        const X r = res[c * m + i], fv = f[c * m + i];
        num += re_(r) * re_(fv) + im_(r) * im_(fv);  // Re(conj(r) f)
        den += abs2_(fv);
    }
    block_sum2<kStelaNT>(num, den, sa, sb);
    if (threadIdx.x == 0) {
        part[(2 * k + c) * parts + p] = num;
        part[(3 * k + c) * parts + p] = den;
    }
}

__global__ void stela_gamma_kernel(int64_t k, const double *__restrict__ part, int64_t parts,
                                   double lam, double max_error, double *__restrict__ gamma,
                                   int32_t *__restrict__ active) {
    const int64_t c = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
    if (c >= k) return;
    double s[4];
    for (int q = 0; q < 4; ++q) {
        double acc = 0.0;
        for (int64_t p = 0; p < parts; ++p) acc += part[(q * k + c) * parts + p];
        s[q] = acc;
    }
    active[c] = sqrt(s[0]) > max_error ? 1 : 0;
    gamma[c] = np_max(np_min(-(s[2] + lam * s[1]) / s[3], 1.0), 0.0);
}

// y += gamma_c (a - b) (b may be NULL) on active columns
template <typename X>
__global__ void stela_axpy_kernel(int64_t rows, int64_t k, X *y, const X *a, const X *b,
                                  const double *__restrict__ gamma,
                                  const int32_t *__restrict__ active) {
    const int64_t total = rows * k;
    for (int64_t e = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; e < total;
         e += int64_t(gridDim.x) * blockDim.x) {
        const int64_t c = e / rows;
        if (!active[c]) continue;
        const double g = gamma[c];
        const X av = a[e];
        const double r = b ? re_(av) - re_(b[e]) : re_(av);
        const double i = b ? im_(av) - im_(b[e]) : im_(av);
        const X yv = y[e];
        y[e] = from_ri<X>(re_(yv) + g * r, im_(yv) + g * i);
    }
}

template <typename Z>
__global__ void stela_select_kernel(int64_t rows, int64_t k, Z *__restrict__ z,
                                    const Z *__restrict__ zn,
                                    const int32_t *__restrict__ active) {
    const int64_t total = rows * k;
    for (int64_t e = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; e < total;
         e += int64_t(gridDim.x) * blockDim.x)
        if (active[e / rows]) z[e] = zn[e];
}

static unsigned grid_for(int64_t total) {
    int64_t b = ceil_div(total, 256);
    if (b < 1) b = 1;
    if (b > 8192) b = 8192;
    return unsigned(b);
}

}  // namespace fm

using namespace fm;

extern "C" fm_status fm_shrink_step(int dtype, int64_t n, const void *y, const void *g, double L,
                                    double alpha, void *step, void *x, const double *beta_table,
                                    const int64_t *step_counter, double beta, void *y_next,
                                    void *stream) {
    FM_REQUIRE(n >= 0, FM_ERR_VALUE, "negative size");
    if (n == 0) return FM_SUCCESS;
    FM_REQUIRE(y && g && step && x, FM_ERR_VALUE, "NULL data pointer");
    FM_REQUIRE(alpha > 0.0, FM_ERR_VALUE, "shrinkage threshold must be positive");
    hipStream_t s = as_stream(stream);
    switch (dtype) {
    case FM_FLOAT32:
        return shrink_t<float, float>(n, y, g, L, alpha, step, x, beta_table, step_counter, beta,
                                      y_next, s);
    case FM_FLOAT64:
        return shrink_t<double, double>(n, y, g, L, alpha, step, x, beta_table, step_counter,
                                        beta, y_next, s);
    case FM_COMPLEX64:
        return shrink_t<cpx<float>, float>(n, y, g, L, alpha, step, x, beta_table, step_counter,
                                           beta, y_next, s);
    case FM_COMPLEX128:
        return shrink_t<cpx<double>, double>(n, y, g, L, alpha, step, x, beta_table,
                                             step_counter, beta, y_next, s);
    default:
        set_error("shrinkage step needs a floating-point type, got type id %d", dtype);
        return FM_ERR_TYPE;
    }
}

extern "C" fm_status fm_counter_advance(int64_t *step_counter, void *stream) {
    FM_REQUIRE(step_counter != nullptr, FM_ERR_VALUE, "NULL counter");
    hipLaunchKernelGGL(counter_advance_kernel, dim3(1), dim3(64), 0, as_stream(stream),
                       step_counter);
    FM_CHECK_LAUNCH();
    return FM_SUCCESS;
}

// ------------------------------------------------------------------ STELA ABI
static bool stela_types(int dx, int dz) {
    return (dx == FM_FLOAT64 && (dz == FM_FLOAT64 || dz == FM_COMPLEX128)) ||
           (dx == FM_COMPLEX128 && dz == FM_COMPLEX128);
}

extern "C" fm_status fm_stela_prox(int dtype_x, int dtype_z, int64_t n, int64_t k, const void *x,
                                   const void *z, const double *d, double lambda, void *bx,
                                   void *dlt, double *partials, int64_t parts, void *stream) {
    FM_REQUIRE(stela_types(dtype_x, dtype_z), FM_ERR_TYPE,
               "STELA runs in float64 / complex128 (types %d, %d)", dtype_x, dtype_z);
    FM_REQUIRE(n >= 1 && k >= 1 && k <= 65535 && parts >= 1 && parts <= 65535, FM_ERR_VALUE,
               "STELA: bad sizes n=%lld k=%lld parts=%lld", (long long)n, (long long)k,
               (long long)parts);
    FM_REQUIRE(x && z && d && bx && dlt && partials, FM_ERR_VALUE, "NULL pointer");
    hipStream_t s = as_stream(stream);
    const dim3 grid{unsigned(parts), unsigned(k), 1u};
    if (dtype_x == FM_FLOAT64 && dtype_z == FM_FLOAT64)
        hipLaunchKernelGGL((stela_prox_kernel<double, double>), grid, dim3(kStelaNT), 0, s, n, k,
                           (const double *)x, (const double *)z, d, lambda, (double *)bx,
                           (double *)dlt, partials, parts);
    else if (dtype_x == FM_FLOAT64)
        hipLaunchKernelGGL((stela_prox_kernel<double, cpx<double>>), grid, dim3(kStelaNT), 0, s,
                           n, k, (const double *)x, (const cpx<double> *)z, d, lambda,
                           (double *)bx, (double *)dlt, partials, parts);
    else
        hipLaunchKernelGGL((stela_prox_kernel<cpx<double>, cpx<double>>), grid, dim3(kStelaNT),
                           0, s, n, k, (const cpx<double> *)x, (const cpx<double> *)z, d,
                           lambda, (cpx<double> *)bx, (cpx<double> *)dlt, partials, parts);
    FM_CHECK_LAUNCH();
    return FM_SUCCESS;
}

extern "C" fm_status fm_stela_line(int dtype_x, int64_t m, int64_t k, const void *res,
                                   const void *f, double *partials, int64_t parts,
                                   double lambda, double max_error, double *gamma,
                                   int32_t *active, void *stream) {
    FM_REQUIRE(dtype_x == FM_FLOAT64 || dtype_x == FM_COMPLEX128, FM_ERR_TYPE,
               "STELA runs in float64 / complex128 (type %d)", dtype_x);
    FM_REQUIRE(m >= 1 && k >= 1 && k <= 65535 && parts >= 1 && parts <= 65535, FM_ERR_VALUE,
               "STELA: bad sizes m=%lld k=%lld parts=%lld", (long long)m, (long long)k,
               (long long)parts);
    FM_REQUIRE(res && f && partials && gamma && active, FM_ERR_VALUE, "NULL pointer");
    hipStream_t s = as_stream(stream);
    const dim3 grid{unsigned(parts), unsigned(k), 1u};
    if (dtype_x == FM_FLOAT64)
        hipLaunchKernelGGL((stela_line_kernel<double>), grid, dim3(kStelaNT), 0, s, m, k,
                           (const double *)res, (const double *)f, partials, parts);
    else
        hipLaunchKernelGGL((stela_line_kernel<cpx<double>>), grid, dim3(kStelaNT), 0, s, m, k,
                           (const cpx<double> *)res, (const cpx<double> *)f, partials, parts);
    FM_CHECK_LAUNCH();
    hipLaunchKernelGGL(stela_gamma_kernel, dim3(unsigned(ceil_div(k, 256))), dim3(256), 0, s, k,
                       (const double *)partials, parts, lambda, max_error, gamma, active);
    FM_CHECK_LAUNCH();
    return FM_SUCCESS;
}

extern "C" fm_status fm_stela_update(int dtype_x, int64_t n, int64_t m, int64_t k, const void *bx,
                                     void *x, const void *f, void *res, const double *gamma,
                                     const int32_t *active, void *stream) {
    FM_REQUIRE(dtype_x == FM_FLOAT64 || dtype_x == FM_COMPLEX128, FM_ERR_TYPE,
               "STELA runs in float64 / complex128 (type %d)", dtype_x);
    FM_REQUIRE(n >= 1 && m >= 1 && k >= 1, FM_ERR_VALUE, "STELA: bad sizes");
    FM_REQUIRE(bx && x && f && res && gamma && active, FM_ERR_VALUE, "NULL pointer");
    hipStream_t s = as_stream(stream);
    if (dtype_x == FM_FLOAT64) {
        hipLaunchKernelGGL((stela_axpy_kernel<double>), dim3(grid_for(n * k)), dim3(256), 0, s, n,
                           k, (double *)x, (const double *)bx, (const double *)x, gamma, active);
        FM_CHECK_LAUNCH();
        hipLaunchKernelGGL((stela_axpy_kernel<double>), dim3(grid_for(m * k)), dim3(256), 0, s, m,
                           k, (double *)res, (const double *)f, (const double *)nullptr, gamma,
                           active);
    } else {
        using C = cpx<double>;
        hipLaunchKernelGGL((stela_axpy_kernel<C>), dim3(grid_for(n * k)), dim3(256), 0, s, n, k,
                           (C *)x, (const C *)bx, (const C *)x, gamma, active);
        FM_CHECK_LAUNCH();
        hipLaunchKernelGGL((stela_axpy_kernel<C>), dim3(grid_for(m * k)), dim3(256), 0, s, m, k,
                           (C *)res, (const C *)f, (const C *)nullptr, gamma, active);
    }
    FM_CHECK_LAUNCH();
    return FM_SUCCESS;
}

extern "C" fm_status fm_stela_select(int dtype_z, int64_t n, int64_t k, void *z, const void *zn,
                                     const int32_t *active, void *stream) {
    FM_REQUIRE(dtype_z == FM_FLOAT64 || dtype_z == FM_COMPLEX128, FM_ERR_TYPE,
               "STELA runs in float64 / complex128 (type %d)", dtype_z);
    FM_REQUIRE(n >= 1 && k >= 1, FM_ERR_VALUE, "STELA: bad sizes");
    FM_REQUIRE(z && zn && active, FM_ERR_VALUE, "NULL pointer");
    hipStream_t s = as_stream(stream);
    if (dtype_z == FM_FLOAT64)
        hipLaunchKernelGGL((stela_select_kernel<double>), dim3(grid_for(n * k)), dim3(256), 0, s,
                           n, k, (double *)z, (const double *)zn, active);
    else
        hipLaunchKernelGGL((stela_select_kernel<cpx<double>>), dim3(grid_for(n * k)), dim3(256),
                           0, s, n, k, (cpx<double> *)z, (const cpx<double> *)zn, active);
    FM_CHECK_LAUNCH();
    return FM_SUCCESS;
}
