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

// ------------------------------------------------------------------- OMP
// Orthogonal matching pursuit (algorithms/OMP.pyx:122-240), per problem l of
// L, in double precision (V = double or complex<double>), the rank-1
// pseudo-inverse update of the reference kept step for step:
//   select:  idx = argmax_j |(A^H r)_j| / colnorm_j          (first maximum)
//   c = A e_idx (the new column, computed by the operator's forward)
//   v1 = P[:ii] c,  v2 = c - Acols[:ii]^T v1,  nrm2 = ||v2||^2
//   y = <v2, b> / nrm2,  x[:ii] -= y v1,  x[ii] += y
//   P[:ii] -= v1 (conj(v2) / nrm2)^T,  P[ii] = conj(v2) / nrm2,  Acols[ii] = c
//   r -= y v2
// P and Acols are (L, steps, N) row-major, r / b / c / v2 (L, N), x (L, steps).
namespace fm {

template <typename V> struct OmpOps;
template <> struct OmpOps<double> {
    __device__ static double mul(double a, double b) { return a * b; }
    __device__ static double conj(double a) { return a; }
    __device__ static double abs2(double a) { return a * a; }
    __device__ static double scale(double a, double s) { return a * s; }
    __device__ static double zero() { return 0.0; }
    __device__ static double one() { return 1.0; }
    __device__ static double add(double a, double b) { return a + b; }
    __device__ static double sub(double a, double b) { return a - b; }
};
template <> struct OmpOps<cpx<double>> {
    using C = cpx<double>;
    __device__ static C mul(C a, C b) { return mk(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x); }
    __device__ static C conj(C a) { return mk(a.x, -a.y); }
    __device__ static double abs2(C a) { return a.x * a.x + a.y * a.y; }
    __device__ static C scale(C a, double s) { return mk(a.x * s, a.y * s); }
    __device__ static C zero() { return mk(0.0, 0.0); }
    __device__ static C one() { return mk(1.0, 0.0); }
    __device__ static C add(C a, C b) { return mk(a.x + b.x, a.y + b.y); }
    __device__ static C sub(C a, C b) { return mk(a.x - b.x, a.y - b.y); }
};

constexpr int kOmpNT = 256;

template <typename V>
__device__ __forceinline__ V omp_block_sum(V v, V *sh) {
    using O = OmpOps<V>;
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int w = kOmpNT / 2; w > 0; w >>= 1) {
        if (int(threadIdx.x) < w) sh[threadIdx.x] = O::add(sh[threadIdx.x], sh[threadIdx.x + w]);
        __syncthreads();
    }
    V r = sh[0];
    __syncthreads();
    return r;
}

// one block per problem: first index of the largest |g_j| inv_norm_j
constexpr int kSelNT = 1024;
template <typename V>
__global__ __launch_bounds__(kSelNT) void omp_select_kernel(int64_t m, const V *__restrict__ g,
                                                            const double *__restrict__ inv_norm,
                                                            int64_t *__restrict__ idx,
                                                            int64_t *__restrict__ support,
                                                            int steps, int ii) {
    using O = OmpOps<V>;
    __shared__ double sv[kSelNT];
    __shared__ int64_t si[kSelNT];
    const int64_t l = blockIdx.x;
    double best = -1.0;
    int64_t bi = m;
    for (int64_t j = threadIdx.x; j < m; j += kSelNT) {
        const double v = sqrt(O::abs2(g[l * m + j])) * inv_norm[j];
        if (v > best || (v != v && best == best)) {  // first NaN wins, as argmax
            best = v;
            bi = j;
        }
    }
    sv[threadIdx.x] = best;
    si[threadIdx.x] = bi;
    __syncthreads();
    for (int w = kSelNT / 2; w > 0; w >>= 1) {
        if (int(threadIdx.x) < w) {
            const double a = sv[threadIdx.x], b = sv[threadIdx.x + w];
            const int64_t ia = si[threadIdx.x], ib = si[threadIdx.x + w];
            const bool a_nan = a != a, b_nan = b != b;
            bool take_b;
            if (a_nan || b_nan) take_b = b_nan && (!a_nan || ib < ia);
            else take_b = b > a || (b == a && ib < ia);
            if (take_b) {
                sv[threadIdx.x] = b;
                si[threadIdx.x] = ib;
            }
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const int64_t j = si[0] < m ? si[0] : 0;
        idx[l] = j;
        support[l * steps + ii] = j;
    }
}

template <typename V>
__global__ void omp_onehot_kernel(int64_t L, int64_t m, const int64_t *__restrict__ idx,
                                  V *__restrict__ e) {
    using O = OmpOps<V>;
    const int64_t total = L * m;
    for (int64_t q = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; q < total;
         q += int64_t(gridDim.x) * blockDim.x) {
        const int64_t l = q / m, j = q - l * m;
        e[q] = (j == idx[l]) ? O::one() : O::zero();
    }
}

// (a, ia) against (b, ib): true when b wins -- a NaN first (the lower index
// among NaNs), then the larger value, then the lower index
__device__ __forceinline__ bool omp_sel_takes(double a, int64_t ia, double b, int64_t ib) {
    const bool a_nan = a != a, b_nan = b != b;
    if (a_nan || b_nan) return b_nan && (!a_nan || ib < ia);
    return b > a || (b == a && ib < ia);
}

// tree over the block's (value, index) pairs; the winner ends in sv[0], si[0]
template <int NT>
__device__ __forceinline__ void omp_sel_reduce(double best, int64_t bi, double *sv, int64_t *si) {
    sv[threadIdx.x] = best;
    si[threadIdx.x] = bi;
    __syncthreads();
    const int t = threadIdx.x;
    for (int w = NT / 2; w > 0; w >>= 1) {
        if (t < w && omp_sel_takes(sv[t], si[t], sv[t + w], si[t + w])) {
            sv[t] = sv[t + w];
            si[t] = si[t + w];
        }
        __syncthreads();
    }
}

// selection, first stage: the winner of each chunk of columns; grid (sp, L)
template <typename V>
__global__ __launch_bounds__(kOmpNT) void omp_select_part_kernel(
    int64_t m, int64_t sp, const V *__restrict__ g, const double *__restrict__ inv_norm,
    double *__restrict__ part_v, int64_t *__restrict__ part_i) {
    using O = OmpOps<V>;
    __shared__ double sv[kOmpNT];
    __shared__ int64_t si[kOmpNT];
    const int64_t p = blockIdx.x, l = blockIdx.y;
    const int64_t chunk = ceil_div(m, sp);
    const int64_t lo = p * chunk, hi = (lo + chunk < m) ? lo + chunk : m;
    double best = -1.0;
    int64_t bi = m;
    for (int64_t j = lo + threadIdx.x; j < hi; j += kOmpNT) {
        const double v = sqrt(O::abs2(g[l * m + j])) * inv_norm[j];
        if (omp_sel_takes(best, bi, v, j)) {
            best = v;
            bi = j;
        }
    }
    omp_sel_reduce<kOmpNT>(best, bi, sv, si);
    if (threadIdx.x == 0) {
        part_v[l * sp + p] = sv[0];
        part_i[l * sp + p] = si[0];
    }
}

// selection, second stage: merge the chunk winners of each problem; grid (L)
__global__ __launch_bounds__(kOmpNT) void omp_select_merge_kernel(
    int64_t m, int64_t sp, const double *__restrict__ part_v, const int64_t *__restrict__ part_i,
    int64_t *__restrict__ idx, int64_t *__restrict__ support, int steps, int ii) {
    __shared__ double sv[kOmpNT];
    __shared__ int64_t si[kOmpNT];
    const int64_t l = blockIdx.x;
    double best = -1.0;
    int64_t bi = m;
    for (int64_t p = threadIdx.x; p < sp; p += kOmpNT) {
        const double v = part_v[l * sp + p];
        const int64_t j = part_i[l * sp + p];
        if (omp_sel_takes(best, bi, v, j)) {
            best = v;
            bi = j;
        }
    }
    omp_sel_reduce<kOmpNT>(best, bi, sv, si);
    if (threadIdx.x == 0) {
        const int64_t j = si[0] < m ? si[0] : 0;
        idx[l] = j;
        support[l * steps + ii] = j;
    }
}

// partial sums v1p[l, i, p] = sum_{j in chunk p} P[l, i, j] c[l, j];
// grid (parts, ii, L)
template <typename V>
__global__ __launch_bounds__(kOmpNT) void omp_proj_kernel(int64_t n, int steps, int64_t parts,
                                                          const V *__restrict__ P,
                                                          const V *__restrict__ c,
                                                          V *__restrict__ v1p) {
    using O = OmpOps<V>;
    __shared__ V sh[kOmpNT];
    const int64_t p = blockIdx.x, i = blockIdx.y, l = blockIdx.z;
    const int64_t chunk = ceil_div(n, parts);
    const int64_t lo = p * chunk, hi = (lo + chunk < n) ? lo + chunk : n;
    const V *row = P + (l * steps + i) * n;
    V acc = O::zero();
    for (int64_t j = lo + threadIdx.x; j < hi; j += kOmpNT)
        acc = O::add(acc, O::mul(row[j], c[l * n + j]));
    acc = omp_block_sum<V>(acc, sh);
    if (threadIdx.x == 0) v1p[(l * steps + i) * parts + p] = acc;
}

// v1[l, i] = sum_p v1p[l, i, p]; grid (ii, L)
template <typename V>
__global__ __launch_bounds__(kOmpNT) void omp_proj_sum_kernel(int steps, int64_t parts,
                                                              const V *__restrict__ v1p,
                                                              V *__restrict__ v1) {
    using O = OmpOps<V>;
    __shared__ V sh[kOmpNT];
    const int64_t i = blockIdx.x, l = blockIdx.y;
    const V *src = v1p + (l * steps + i) * parts;
    V acc = O::zero();
    for (int64_t p = threadIdx.x; p < parts; p += kOmpNT) acc = O::add(acc, src[p]);
    acc = omp_block_sum<V>(acc, sh);
    if (threadIdx.x == 0) v1[l * steps + i] = acc;
}

// v2 = c - Acols[:ii]^T v1 and partial sums of |v2|^2 and conj(v2) b;
// grid (parts, L)
template <typename V>
__global__ __launch_bounds__(kOmpNT) void omp_resid_kernel(
    int64_t n, int steps, int ii, const V *__restrict__ acols, const V *__restrict__ v1,
    const V *__restrict__ c, const V *__restrict__ b, V *__restrict__ v2,
    double *__restrict__ part_n, V *__restrict__ part_d, int64_t parts) {
    using O = OmpOps<V>;
    __shared__ V sh[kOmpNT];
    __shared__ double shn[kOmpNT];
    const int64_t p = blockIdx.x, l = blockIdx.y;
    const int64_t chunk = ceil_div(n, parts);
    const int64_t lo = p * chunk, hi = (lo + chunk < n) ? lo + chunk : n;
    double nrm = 0.0;
    V dot = O::zero();
    for (int64_t j = lo + threadIdx.x; j < hi; j += kOmpNT) {
        V v = c[l * n + j];
        for (int i = 0; i < ii; ++i)
            v = O::sub(v, O::mul(acols[(l * steps + i) * n + j], v1[l * steps + i]));
        v2[l * n + j] = v;
        nrm += O::abs2(v);
        dot = O::add(dot, O::mul(O::conj(v), b[l * n + j]));
    }
    shn[threadIdx.x] = nrm;
    __syncthreads();
    for (int w = kOmpNT / 2; w > 0; w >>= 1) {
        if (int(threadIdx.x) < w) shn[threadIdx.x] += shn[threadIdx.x + w];
        __syncthreads();
    }
    dot = omp_block_sum<V>(dot, sh);
    if (threadIdx.x == 0) {
        part_n[l * parts + p] = shn[0];
        part_d[l * parts + p] = dot;
    }
}

// per problem: nrm2, y = <v2, b> / nrm2, coefficient updates; grid (L)
template <typename V>
__global__ __launch_bounds__(kOmpNT) void omp_coef_kernel(int steps, int ii,
                                                          const double *__restrict__ part_n,
                                                          const V *__restrict__ part_d,
                                                          int64_t parts, const V *__restrict__ v1,
                                                          V *__restrict__ x,
                                                          double *__restrict__ nrm2,
                                                          V *__restrict__ y) {
    using O = OmpOps<V>;
    __shared__ V sh[kOmpNT];
    __shared__ double shn[kOmpNT];
    const int64_t l = blockIdx.x;
    double s = 0.0;
    V d = O::zero();
    for (int64_t p = threadIdx.x; p < parts; p += kOmpNT) {
        s += part_n[l * parts + p];
        d = O::add(d, part_d[l * parts + p]);
    }
    shn[threadIdx.x] = s;
    __syncthreads();
    for (int w = kOmpNT / 2; w > 0; w >>= 1) {
        if (int(threadIdx.x) < w) shn[threadIdx.x] += shn[threadIdx.x + w];
        __syncthreads();
    }
    d = omp_block_sum<V>(d, sh);
    const double tot = shn[0];
    const V yy = O::scale(d, 1.0 / tot);
    if (threadIdx.x == 0) {
        nrm2[l] = tot;
        y[l] = yy;
        x[l * steps + ii] = O::add(x[l * steps + ii], yy);
    }
    for (int i = threadIdx.x; i < ii; i += kOmpNT)
        x[l * steps + i] = O::sub(x[l * steps + i], O::mul(yy, v1[l * steps + i]));
}

// rank-1 update of P, new rows of P and Acols, residual; grid-stride over L n
template <typename V>
__global__ void omp_rank1_kernel(int64_t L, int64_t n, int steps, int ii,
                                 const V *__restrict__ v1, const V *__restrict__ v2,
                                 const V *__restrict__ c, const double *__restrict__ nrm2,
                                 const V *__restrict__ y, V *__restrict__ P,
                                 V *__restrict__ acols, V *__restrict__ r) {
    using O = OmpOps<V>;
    const int64_t total = L * n;
    for (int64_t q = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; q < total;
         q += int64_t(gridDim.x) * blockDim.x) {
        const int64_t l = q / n, j = q - l * n;
        const V vv = v2[q];
        const V w = O::conj(O::scale(vv, 1.0 / nrm2[l]));
        for (int i = 0; i < ii; ++i) {
            V &pij = P[(l * steps + i) * n + j];
            pij = O::sub(pij, O::mul(w, v1[l * steps + i]));
        }
        P[(l * steps + ii) * n + j] = w;
        acols[(l * steps + ii) * n + j] = c[q];
        r[q] = O::sub(r[q], O::mul(y[l], vv));
    }
}

// x_full[l, support[l, i]] = x[l, i], later steps overwriting earlier ones;
// x_full zero-filled first
template <typename V>
__global__ void omp_zero_kernel(V *p, int64_t count) {
    using O = OmpOps<V>;
    for (int64_t q = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; q < count;
         q += int64_t(gridDim.x) * blockDim.x)
        p[q] = O::zero();
}

template <typename V>
__global__ void omp_result_kernel(int64_t L, int64_t m, int steps,
                                  const int64_t *__restrict__ support,
                                  const V *__restrict__ x, V *__restrict__ out) {
    const int64_t l = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
    if (l >= L) return;
    for (int i = 0; i < steps; ++i) out[l * m + support[l * steps + i]] = x[l * steps + i];
}

// column chunks of the OMP reductions, about 512 elements each, so that a
// step's reductions spread over many workgroups even for a few problems
static inline int64_t omp_parts(int64_t n) {
    const int64_t p = (n + 511) / 512;
    return p < 1 ? 1 : (p > 1024 ? 1024 : p);
}

template <typename V>
static fm_status omp_step_t(int phase, int64_t L, int64_t m, int64_t n, int steps, int ii,
                            const void *g, const double *inv_norm, int64_t *idx,
                            int64_t *support, void *e, const void *c, const void *b, void *P,
                            void *acols, void *r, void *work, void *x, void *x_full,
                            hipStream_t s) {
    const int64_t parts = omp_parts(n);
    // work: v1 (L steps V) | v2 (L n V) | part_d (L parts V) | y (L V) |
    //       v1p (L steps parts V) | part_n (L parts f64) | nrm2 (L f64)
    V *v1 = (V *)work;
    V *v2 = v1 + L * steps;
    V *part_d = v2 + L * n;
    V *yv = part_d + L * parts;
    V *v1p = yv + L;
    double *part_n = (double *)(v1p + L * steps * parts);
    double *nrm2 = part_n + L * parts;
    auto grid_of = [](int64_t total) {
        int64_t bl = ceil_div(total, 256);
        if (bl < 1) bl = 1;
        if (bl > 8192) bl = 8192;
        return unsigned(bl);
    };
    if (phase == 0) {  // select the new index, build the one-hot block
        // chunk winners go to the v2 region, which phase 0 does not use
        const int64_t sp_cap = n * int64_t(sizeof(V)) / 16;
        const int64_t sp = omp_parts(m) < sp_cap ? omp_parts(m) : sp_cap;
        if (sp > 1) {
            double *sel_v = (double *)v2;
            int64_t *sel_i = (int64_t *)(sel_v + L * sp);
            hipLaunchKernelGGL((omp_select_part_kernel<V>), dim3(unsigned(sp), unsigned(L)),
                               dim3(kOmpNT), 0, s, m, sp, (const V *)g, inv_norm, sel_v, sel_i);
            FM_CHECK_LAUNCH();
            hipLaunchKernelGGL(omp_select_merge_kernel, dim3(unsigned(L)), dim3(kOmpNT), 0, s, m,
                               sp, (const double *)sel_v, (const int64_t *)sel_i, idx, support,
                               steps, ii);
            FM_CHECK_LAUNCH();
        } else {
            hipLaunchKernelGGL((omp_select_kernel<V>), dim3(unsigned(L)), dim3(kSelNT), 0, s, m,
                               (const V *)g, inv_norm, idx, support, steps, ii);
            FM_CHECK_LAUNCH();
        }
        hipLaunchKernelGGL((omp_onehot_kernel<V>), dim3(grid_of(L * m)), dim3(256), 0, s, L, m,
                           (const int64_t *)idx, (V *)e);
        FM_CHECK_LAUNCH();
        return FM_SUCCESS;
    }
    if (phase == 1) {  // pseudo-inverse update with the new column c
        if (ii > 0) {
            hipLaunchKernelGGL((omp_proj_kernel<V>),
                               dim3(unsigned(parts), unsigned(ii), unsigned(L)), dim3(kOmpNT), 0,
                               s, n, steps, parts, (const V *)P, (const V *)c, v1p);
            FM_CHECK_LAUNCH();
            hipLaunchKernelGGL((omp_proj_sum_kernel<V>), dim3(unsigned(ii), unsigned(L)),
                               dim3(kOmpNT), 0, s, steps, parts, (const V *)v1p, v1);
            FM_CHECK_LAUNCH();
        }
        hipLaunchKernelGGL((omp_resid_kernel<V>), dim3(unsigned(parts), unsigned(L)),
                           dim3(kOmpNT), 0, s, n, steps, ii, (const V *)acols,
                           (const V *)v1, (const V *)c, (const V *)b, v2, part_n, part_d,
                           parts);
        FM_CHECK_LAUNCH();
        hipLaunchKernelGGL((omp_coef_kernel<V>), dim3(unsigned(L)), dim3(kOmpNT), 0, s, steps, ii,
                           (const double *)part_n, (const V *)part_d, parts, (const V *)v1,
                           (V *)x, nrm2, yv);
        FM_CHECK_LAUNCH();
        hipLaunchKernelGGL((omp_rank1_kernel<V>), dim3(grid_of(L * n)), dim3(256), 0, s, L, n,
                           steps, ii, (const V *)v1, (const V *)v2, (const V *)c,
                           (const double *)nrm2, (const V *)yv, (V *)P, (V *)acols, (V *)r);
        FM_CHECK_LAUNCH();
        return FM_SUCCESS;
    }
    // phase 2: scatter the coefficients into the (L, m) result
    hipLaunchKernelGGL((omp_zero_kernel<V>), dim3(grid_of(L * m)), dim3(256), 0, s, (V *)x_full,
                       L * m);
    FM_CHECK_LAUNCH();
    hipLaunchKernelGGL((omp_result_kernel<V>), dim3(unsigned(ceil_div(L, 64))), dim3(64), 0, s, L,
                       m, steps, (const int64_t *)support, (const V *)x, (V *)x_full);
    FM_CHECK_LAUNCH();
    return FM_SUCCESS;
}

}  // namespace fm

extern "C" fm_status fm_omp_workspace_size(int dtype, int64_t L, int64_t n, int steps,
                                           int64_t *bytes) {
    FM_REQUIRE(bytes != nullptr, FM_ERR_VALUE, "NULL argument");
    FM_REQUIRE(dtype == FM_FLOAT64 || dtype == FM_COMPLEX128, FM_ERR_TYPE,
               "OMP runs in float64 / complex128 (type %d)", dtype);
    FM_REQUIRE(L >= 1 && n >= 1 && steps >= 1, FM_ERR_VALUE, "OMP: bad sizes");
    const int64_t vs = dtype == FM_FLOAT64 ? 8 : 16;
    const int64_t parts = fm::omp_parts(n);
    *bytes = vs * (L * steps + L * n + L * parts + L + L * steps * parts) + 8 * (L * parts + L);
    return FM_SUCCESS;
}

extern "C" fm_status fm_omp_step(int dtype, int phase, int64_t L, int64_t m, int64_t n, int steps,
                                 int ii, const void *g, const double *inv_norm, int64_t *idx,
                                 int64_t *support, void *e, const void *c, const void *b, void *P,
                                 void *acols, void *r, void *work, int64_t work_bytes, void *x,
                                 void *x_full, void *stream) {
    FM_REQUIRE(dtype == FM_FLOAT64 || dtype == FM_COMPLEX128, FM_ERR_TYPE,
               "OMP runs in float64 / complex128 (type %d)", dtype);
    FM_REQUIRE(phase >= 0 && phase <= 2, FM_ERR_VALUE, "OMP: phase %d", phase);
    FM_REQUIRE(L >= 1 && L <= 65535 && m >= 1 && n >= 1 && steps >= 1 && steps <= 65535 &&
                   ii >= 0 && ii < steps,
               FM_ERR_VALUE, "OMP: bad sizes or step");
    int64_t need = 0;
    fm_omp_workspace_size(dtype, L, n, steps, &need);
    FM_REQUIRE(work && work_bytes >= need, FM_ERR_VALUE, "OMP workspace of %lld < %lld bytes",
               (long long)work_bytes, (long long)need);
    if (phase == 0) FM_REQUIRE(g && inv_norm && idx && support && e, FM_ERR_VALUE, "NULL argument");
    if (phase == 1) FM_REQUIRE(c && b && P && acols && r && x, FM_ERR_VALUE, "NULL argument");
    if (phase == 2) FM_REQUIRE(support && x && x_full, FM_ERR_VALUE, "NULL argument");
    hipStream_t s = as_stream(stream);
    if (dtype == FM_FLOAT64)
        return omp_step_t<double>(phase, L, m, n, steps, ii, g, inv_norm, idx, support, e, c, b,
                                  P, acols, r, work, x, x_full, s);
    return omp_step_t<cpx<double>>(phase, L, m, n, steps, ii, g, inv_norm, idx, support, e, c,
                                   b, P, acols, r, work, x, x_full, s);
}
