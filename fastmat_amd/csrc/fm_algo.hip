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
__global__ void shrink_kernel(int64_t n, const V *__restrict__ y, const V *__restrict__ g, R L,
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
    hipLaunchKernelGGL((shrink_kernel<V, R>), dim3(unsigned(blocks)), dim3(256), 0, s, n,
                       (const V *)y, (const V *)g, R(L), R(alpha), (V *)step, (V *)x, beta_tab,
                       counter, R(beta), (V *)y_next);
    FM_CHECK_LAUNCH();
    return FM_SUCCESS;
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
