// Dense factors on FP32 MFMA (v_mfma_f32_16x16x4f32, FP32 accumulation).
//
// Replaces the dense Matrix transform (fastmat/Matrix.pyx:1831-1842,
// 1924-1935: ndarray.dot) for the dense factors of a Kron product
// (Kron.pyx:267-341) and for plain dense matrices.
//
// One kernel computes Y[i, c] = sum_l A_eff[i, l] * Bm[l, c] with strided
// addressing:
//   A_eff[i, l] = A[i*a_rs + l*a_cs]                  (transpose by strides)
//   Bm[l, c]    = X[s*(o*bos + l*brs + j*bjs) + p]    c -> (o, j, p)
//   Y[i, c]     = Y[s*(o*yos + i*yrs + j*yjs) + p]
// where s = 2 and p = c & 1 for complex data (real and imaginary parts become
// neighbouring real columns, so a real A multiplies complex data without
// widening) and s = 1, p = 0 otherwise.  A complex A is handled by the host
// as the real 2m x 2n matrix [[Re, -Im], [Im, Re]] interleaved to match the
// (re, im) storage of X and Y.
//
// Tiling: 256 threads (4 wave64) own a 64 x 64 output tile; each wave owns a
// 32 x 32 quarter as 2 x 2 MFMA 16x16 blocks; K advances in slabs of 16
// staged through LDS.
#include <type_traits>

#include "fm_common.hpp"

namespace fm {

int knob(const char *name, int dflt);

typedef float f32x4 __attribute__((ext_vector_type(4)));

struct GemmArgs {
    int64_t M, K, NC;
    const float *A;
    int64_t a_rs, a_cs;
    const float *X;
    int64_t J, bos, brs, bjs;
    float *Y;
    int64_t yos, yrs, yjs;
    int cplx;
};

constexpr int GT_M = 64, GT_N = 64, GT_K = 16;

static fm_status gemm_launch_tiled(const GemmArgs &g, hipStream_t stream);

__device__ __forceinline__ int64_t col_base(const GemmArgs &g, int64_t c, int64_t os, int64_t js,
                                            int *part) {
    int64_t cc = c;
    *part = 0;
    if (g.cplx) {
        *part = int(c & 1);
        cc = c >> 1;
    }
    const int64_t o = cc / g.J, j = cc - o * g.J;
    return o * os + j * js;
}

__global__ __launch_bounds__(256) void gemm_mfma_kernel(const GemmArgs g) {
    __shared__ float As[GT_K][GT_M + 4];
    __shared__ float Bs[GT_K][GT_N + 4];
    __shared__ int64_t bbase[GT_N];
    __shared__ int bpart[GT_N];

    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const int64_t tiles_n = ceil_div(g.NC, GT_N);
    const int64_t tile = xcd_remap(blockIdx.x, gridDim.x);
    const int64_t tm = tile / tiles_n, tn = tile - tm * tiles_n;
    const int64_t row0 = tm * GT_M, col0 = tn * GT_N;
    const int s = g.cplx ? 2 : 1;

    if (tid < GT_N) {
        const int64_t c = col0 + tid;
        int part = 0;
        bbase[tid] = c < g.NC ? col_base(g, c, g.bos, g.bjs, &part) : -1;
        bpart[tid] = part;
    }
    __syncthreads();

    f32x4 acc[2][2];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) acc[a][b] = f32x4{0.f, 0.f, 0.f, 0.f};

    const int wr = (wave >> 1) * 32, wc = (wave & 1) * 32;
    const bool a_kfast = g.a_cs == 1;
    const bool b_kfast = g.brs == 1;

    for (int64_t k0 = 0; k0 < g.K; k0 += GT_K) {
        // stage A (64 x 16) and B (16 x 64): 4 elements per thread each
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const int idx = tid + e * 256;
            int kk, i;
            if (a_kfast) { kk = idx & 15; i = idx >> 4; }
            else { i = idx & 63; kk = idx >> 6; }
            const int64_t gi = row0 + i, gk = k0 + kk;
            As[kk][i] = (gi < g.M && gk < g.K) ? g.A[gi * g.a_rs + gk * g.a_cs] : 0.f;
        }
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const int idx = tid + e * 256;
            int kk, c;
            if (b_kfast) { kk = idx & 15; c = idx >> 4; }
            else { c = idx & 63; kk = idx >> 6; }
            const int64_t gk = k0 + kk;
            const int64_t base = bbase[c];
            Bs[kk][c] = (base >= 0 && gk < g.K) ? g.X[s * (base + gk * g.brs) + bpart[c]] : 0.f;
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < GT_K; kk += 4) {
            const int kr = kk + (lane >> 4);
            float av[2], bv[2];
#pragma unroll
            for (int a = 0; a < 2; ++a) av[a] = As[kr][wr + a * 16 + (lane & 15)];
#pragma unroll
            for (int b = 0; b < 2; ++b) bv[b] = Bs[kr][wc + b * 16 + (lane & 15)];
#pragma unroll
            for (int a = 0; a < 2; ++a)
#pragma unroll
                for (int b = 0; b < 2; ++b)
                    acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x4f32(av[a], bv[b], acc[a][b], 0, 0, 0);
        }
        __syncthreads();
    }

    // D layout of 16x16x4 f32: lane holds column (lane & 15), rows 4*(lane>>4) + v
#pragma unroll
    for (int b = 0; b < 2; ++b) {
        const int c = wc + b * 16 + (lane & 15);
        const int64_t gc = col0 + c;
        if (gc >= g.NC) continue;
        int part = 0;
        const int64_t ybase = col_base(g, gc, g.yos, g.yjs, &part);
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int64_t gi = row0 + wr + a * 16 + 4 * (lane >> 4) + v;
                if (gi < g.M) g.Y[s * (ybase + gi * g.yrs) + part] = acc[a][b][v];
            }
    }
}

// Small factors (M, K <= 128, the dense factors of a Kron product): every
// workgroup (8 wave64) keeps A in registers in the MFMA operand layout for
// its whole life and walks 64-column tiles of the data persistently.  Wave w
// owns rows [16w, 16w + 16); per k-step of 4 it issues four
// v_mfma_f32_16x16x4f32 (one per 16-column block).  B tiles are staged
// through LDS (coalesced along the data's contiguous axis), results through
// LDS again so that stores are as contiguous as the layout allows.  The next
// tile's B is loaded into registers while the current one multiplies (its
// column table in the other half of a double-buffered table), so the loads'
// latency hides behind the MFMA work.  A in registers takes ~180 VGPRs, so
// one workgroup runs per CU (2 waves per SIMD).
constexpr int SK_N = 64;
constexpr int SK_K = 128;
constexpr int SK_T = 512;

__global__ __launch_bounds__(SK_T) void gemm_smallk_kernel(const GemmArgs g) {
    __shared__ float Bs[SK_K][SK_N + 1];  // B tile, then the output tile
    __shared__ int xcol[2][SK_N];
    __shared__ int ycol[2][SK_N];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int s = g.cplx ? 2 : 1;
    const bool b_kfast = g.brs == 1;
    const bool y_rfast = g.yrs == 1;
    const int64_t ntiles = ceil_div(g.NC, SK_N);
    // element offsets fit in 32 bits (checked by the launcher)
    const int xk = int(s * g.brs), yr = int(s * g.yrs);
    int64_t t = blockIdx.x;
    if (t >= ntiles) return;  // uniform over the workgroup

    // a[st] = A_eff[16*wave + (lane & 15)][4*st + (lane >> 4)]; zero beyond M, K
    float a[SK_K / 4];
    {
        const int64_t i = 16 * wave + (lane & 15);
#pragma unroll
        for (int st = 0; st < SK_K / 4; ++st) {
            const int64_t k = 4 * st + (lane >> 4);
            a[st] = (i < g.M && k < g.K) ? g.A[i * g.a_rs + k * g.a_cs] : 0.f;
        }
    }
    // column tables of tile tt into half h (-1: past the last column)
    auto tables = [&](int64_t tt, int h) {
        if (tid < SK_N) {
            const int64_t gc = tt * SK_N + tid;
            int xb = -1, yb = -1;
            if (gc < g.NC) {
                int64_t cc = gc;
                int part = 0;
                if (g.cplx) {
                    part = int(gc & 1);
                    cc = gc >> 1;
                }
                const int64_t o = cc / g.J, j = cc - o * g.J;
                xb = int(s * (o * g.bos + j * g.bjs) + part);
                yb = int(s * (o * g.yos + j * g.yjs) + part);
            }
            xcol[h][tid] = xb;
            ycol[h][tid] = yb;
        }
    };
    constexpr int PER = SK_K * SK_N / SK_T;
    // B elements of this thread for the tile whose table is in half h
    auto load = [&](int h, float *pre) {
#pragma unroll
        for (int e = 0; e < PER; ++e) {
            const int idx = tid + e * SK_T;
            int kk, c;
            if (b_kfast) { kk = idx & (SK_K - 1); c = idx >> 7; }
            else { c = idx & (SK_N - 1); kk = idx >> 6; }
            const int xb = xcol[h][c];
            pre[e] = (xb >= 0 && kk < g.K) ? g.X[xb + kk * xk] : 0.f;
        }
    };
    float pre[PER];
    tables(t, 0);
    __syncthreads();
    load(0, pre);
    for (int cur = 0; t < ntiles; t += gridDim.x, cur ^= 1) {
        // Bs is free: the previous tile's stores ended with a barrier
#pragma unroll
        for (int e = 0; e < PER; ++e) {
            const int idx = tid + e * SK_T;
            int kk, c;
            if (b_kfast) { kk = idx & (SK_K - 1); c = idx >> 7; }
            else { c = idx & (SK_N - 1); kk = idx >> 6; }
            Bs[kk][c] = pre[e];
        }
        // the other half was last read two barriers ago (the previous
        // tile's stores, the prefetch before them)
        const int64_t tn = t + gridDim.x;
        tables(tn, cur ^ 1);
        __syncthreads();
        if (tn < ntiles) load(cur ^ 1, pre);
        f32x4 acc[4];
#pragma unroll
        for (int cb = 0; cb < 4; ++cb) acc[cb] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int st = 0; st < SK_K / 4; ++st) {  // rows k >= K of A and B are zero
            const int kr = 4 * st + (lane >> 4);
#pragma unroll
            for (int cb = 0; cb < 4; ++cb)
                acc[cb] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[st], Bs[kr][16 * cb + (lane & 15)],
                                                               acc[cb], 0, 0, 0);
        }
        __syncthreads();  // all waves done reading Bs
        // stage the result: block cb holds rows 16*wave + 4*(lane>>4) + v, column lane&15
#pragma unroll
        for (int cb = 0; cb < 4; ++cb)
#pragma unroll
            for (int v = 0; v < 4; ++v)
                Bs[16 * wave + 4 * (lane >> 4) + v][16 * cb + (lane & 15)] = acc[cb][v];
        __syncthreads();
#pragma unroll
        for (int e = 0; e < PER; ++e) {
            const int idx = tid + e * SK_T;
            int r, c;
            if (y_rfast) { r = idx & (SK_K - 1); c = idx >> 7; }
            else { c = idx & (SK_N - 1); r = idx >> 6; }
            const int yb = ycol[cur][c];
            if (yb >= 0 && r < g.M) g.Y[yb + r * yr] = Bs[r][c];
        }
        __syncthreads();  // Bs and this half of the tables are rewritten next
    }
}

// Largest element offset a strided column of the data can reach.
static int64_t max_offset(const GemmArgs &g, int64_t os, int64_t rs, int64_t js, int64_t rows) {
    const int64_t cols = g.cplx ? g.NC / 2 : g.NC;
    const int64_t o_max = (cols - 1) / g.J, j_max = g.J - 1 < cols - 1 ? g.J - 1 : cols - 1;
    return (g.cplx ? 2 : 1) * (o_max * os + j_max * js + (rows - 1) * rs) + 1;
}

// ----------------------------------------------- other element types
// Y = A X (adjoint: A^H X) for int8/16/32/64, float64 and complex128 data of
// one type (ndarray.dot of the promoted operands, Matrix.pyx:1842, 1935).
// Integers accumulate in 64-bit unsigned arithmetic and are truncated to the
// element width at the store: the same residue modulo 2^bits as numpy's
// integer dot, which sums in a 64-bit integer and casts the sum to the
// output type.  float64 / complex128 accumulate in FP64.  Tiles of 64 rows x
// 16 columns of Y per 256-thread workgroup, K in LDS slabs of 32.
template <typename T> struct GenAcc { typedef T type; };
template <> struct GenAcc<int8_t> { typedef uint64_t type; };
template <> struct GenAcc<int16_t> { typedef uint64_t type; };
template <> struct GenAcc<int32_t> { typedef uint64_t type; };
template <> struct GenAcc<int64_t> { typedef uint64_t type; };

template <typename T> __device__ __forceinline__ typename GenAcc<T>::type gen_widen(T v) {
    if constexpr (std::is_integral<T>::value) return uint64_t(int64_t(v));
    else return v;
}
template <typename T> __device__ __forceinline__ T gen_conj(T v) { return v; }
template <> __device__ __forceinline__ cpx<double> gen_conj(cpx<double> v) { return conj(v); }
template <typename A> __device__ __forceinline__ A gen_mac(A acc, A a, A b) { return acc + a * b; }
template <>
__device__ __forceinline__ cpx<double> gen_mac(cpx<double> acc, cpx<double> a, cpx<double> b) {
    return acc + cmul(a, b);
}
template <typename T> __device__ __forceinline__ T gen_narrow(typename GenAcc<T>::type v) {
    if constexpr (std::is_integral<T>::value) return T(v);  // modulo 2^bits
    else return v;
}
template <typename T> __device__ __forceinline__ typename GenAcc<T>::type gen_zero() {
    if constexpr (std::is_same<T, cpx<double>>::value) return mk(0.0, 0.0);
    else return typename GenAcc<T>::type(0);
}

constexpr int GG_M = 64, GG_J = 16, GG_K = 32;

template <typename T>
__global__ __launch_bounds__(256) void gemm_generic_kernel(int64_t M, int64_t K, int64_t k,
                                                           const T *__restrict__ A, int64_t lda,
                                                           int adjoint, const T *__restrict__ X,
                                                           int64_t ldx, T *__restrict__ Y,
                                                           int64_t ldy) {
    typedef typename GenAcc<T>::type Acc;
    __shared__ Acc sa[GG_K][GG_M + 1];
    __shared__ Acc sx[GG_J][GG_K + 1];
    const int t = threadIdx.x;
    const int64_t i0 = int64_t(blockIdx.x) * GG_M, j0 = int64_t(blockIdx.y) * GG_J;
    const int il = t & (GG_M - 1);
    const int jl = (t >> 6) * 4;  // four columns per thread
    Acc acc[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) acc[q] = gen_zero<T>();
    for (int64_t l0 = 0; l0 < K; l0 += GG_K) {
        // A slab: rows i0.., columns l0.. of A_eff (A or A^H)
        for (int e = t; e < GG_M * GG_K; e += 256) {
            const int ii = e % GG_M, ll = e / GG_M;
            const int64_t i = i0 + ii, l = l0 + ll;
            Acc v = gen_zero<T>();
            if (i < M && l < K)
                v = adjoint ? gen_widen(gen_conj(A[l * lda + i])) : gen_widen(A[i * lda + l]);
            sa[ll][ii] = v;
        }
        for (int e = t; e < GG_J * GG_K; e += 256) {
            const int ll = e % GG_K, jj = e / GG_K;
            const int64_t l = l0 + ll, j = j0 + jj;
            sx[jj][ll] = (l < K && j < k) ? gen_widen(X[j * ldx + l]) : gen_zero<T>();
        }
        __syncthreads();
#pragma unroll 4
        for (int ll = 0; ll < GG_K; ++ll) {
            const Acc a = sa[ll][il];
#pragma unroll
            for (int q = 0; q < 4; ++q) acc[q] = gen_mac(acc[q], a, sx[jl + q][ll]);
        }
        __syncthreads();
    }
    const int64_t i = i0 + il;
    if (i >= M) return;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int64_t j = j0 + jl + q;
        if (j < k) Y[j * ldy + i] = gen_narrow<T>(acc[q]);
    }
}

template <typename T>
static fm_status gemm_generic(int adjoint, int64_t m, int64_t n, const void *A, int64_t lda,
                              const void *x, int64_t ldx, void *y, int64_t ldy, int64_t k,
                              hipStream_t s) {
    const int64_t M = adjoint ? n : m, K = adjoint ? m : n;
    FM_REQUIRE(ceil_div(M, GG_M) < (int64_t(1) << 31) && ceil_div(k, GG_J) <= 65535,
               FM_ERR_VALUE, "dense product too large for one launch");
    const dim3 grid(unsigned(ceil_div(M, GG_M)), unsigned(ceil_div(k, GG_J)));
    ProfToken tok;
    prof_begin(s, &tok);
    hipLaunchKernelGGL(gemm_generic_kernel<T>, grid, dim3(256), 0, s, M, K, k, (const T *)A, lda,
                       adjoint, (const T *)x, ldx, (T *)y, ldy);
    FM_CHECK_LAUNCH();
    if (tok.active)
        prof_end(s, &tok, "gemm_generic_kernel",
                 double(sizeof(T)) * double(M * K + K * k + M * k), 2.0 * double(M) * K * k);
    return FM_SUCCESS;
}

static fm_status dense_generic(int dtype, int adjoint, int64_t m, int64_t n, const void *A,
                               int64_t lda, const void *x, int64_t ldx, void *y, int64_t ldy,
                               int64_t k, hipStream_t s) {
    switch (dtype) {
    case FM_INT8: return gemm_generic<int8_t>(adjoint, m, n, A, lda, x, ldx, y, ldy, k, s);
    case FM_INT16: return gemm_generic<int16_t>(adjoint, m, n, A, lda, x, ldx, y, ldy, k, s);
    case FM_INT32: return gemm_generic<int32_t>(adjoint, m, n, A, lda, x, ldx, y, ldy, k, s);
    case FM_INT64: return gemm_generic<int64_t>(adjoint, m, n, A, lda, x, ldx, y, ldy, k, s);
    case FM_FLOAT64: return gemm_generic<double>(adjoint, m, n, A, lda, x, ldx, y, ldy, k, s);
    case FM_COMPLEX128:
        return gemm_generic<cpx<double>>(adjoint, m, n, A, lda, x, ldx, y, ldy, k, s);
    default: break;
    }
    set_error("dense apply: unsupported element type %d", dtype);
    return FM_ERR_TYPE;
}

static fm_status gemm_launch(const GemmArgs &g, hipStream_t stream) {
    const bool small = g.M <= SK_K && g.K <= SK_K && g.NC > 0 &&
                       max_offset(g, g.bos, g.brs, g.bjs, g.K) < (int64_t(1) << 31) &&
                       max_offset(g, g.yos, g.yrs, g.yjs, g.M) < (int64_t(1) << 31);
    if (small) {
        const int64_t tiles = ceil_div(g.NC, SK_N);
        if (tiles == 0) return FM_SUCCESS;
        int dev = 0, ncu = 256;
        if (hipGetDevice(&dev) == hipSuccess)
            (void)hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
        // one workgroup fits a CU (A in registers: ~180 VGPRs, 2 waves per
        // SIMD), so the grid is one workgroup per CU walking its tiles
        static const int per_cu = knob("FASTMAT_AMD_SMALLK_WG_PER_CU", 1);
        int64_t grid = int64_t(ncu) * (per_cu >= 1 && per_cu <= 4 ? per_cu : 1);
        if (grid > tiles) grid = tiles;
        ProfToken tok;
        prof_begin(stream, &tok);
        hipLaunchKernelGGL(gemm_smallk_kernel, dim3(unsigned(grid)), dim3(SK_T), 0, stream, g);
        FM_CHECK_LAUNCH();
        if (tok.active) {
            const double bytes = 4.0 * (double(g.M) * g.K + double(g.K) * g.NC + double(g.M) * g.NC);
            prof_end(stream, &tok, "gemm_smallk_kernel", bytes, 2.0 * double(g.M) * g.K * g.NC);
        }
        return FM_SUCCESS;
    }
    return gemm_launch_tiled(g, stream);
}

static fm_status gemm_launch_tiled(const GemmArgs &g, hipStream_t stream) {
    const int64_t tiles = ceil_div(g.M, GT_M) * ceil_div(g.NC, GT_N);
    FM_REQUIRE(tiles < (int64_t(1) << 31), FM_ERR_VALUE, "GEMM too large");
    if (tiles == 0) return FM_SUCCESS;
    ProfToken tok;
    prof_begin(stream, &tok);
    hipLaunchKernelGGL(gemm_mfma_kernel, dim3(unsigned(tiles)), dim3(256), 0, stream, g);
    FM_CHECK_LAUNCH();
    if (tok.active) {
        const double bytes = 4.0 * (double(g.M) * g.K + double(g.K) * g.NC + double(g.M) * g.NC);
        prof_end(stream, &tok, "gemm_mfma_kernel", bytes, 2.0 * double(g.M) * g.K * g.NC);
    }
    return FM_SUCCESS;
}

}  // namespace fm

using namespace fm;

// a_dtype FLOAT32 with x FLOAT32 or COMPLEX64 run directly; COMPLEX64 A is
// passed by the host as its real 2m x 2n interleaved form with a_dtype
// FLOAT32 and x COMPLEX64 viewed as real (x_dtype FLOAT32, n and m doubled).
extern "C" fm_status fm_dense_apply(int a_dtype, int x_dtype, int adjoint, int64_t m, int64_t n,
                                    const void *A, int64_t lda, const void *x, int64_t ldx,
                                    void *y, int64_t ldy, int64_t k, void *stream) {
    const bool generic = a_dtype == x_dtype && a_dtype != FM_FLOAT32 && a_dtype != FM_COMPLEX64 &&
                         a_dtype >= FM_INT8 && a_dtype <= FM_COMPLEX128;
    FM_REQUIRE(generic || a_dtype == FM_FLOAT32, FM_ERR_TYPE,
               "dense apply supports float32 matrices (complex ones via their real form) and "
               "int8/16/32/64, float64, complex128 matrices with data of the same type");
    FM_REQUIRE(generic || x_dtype == FM_FLOAT32 || x_dtype == FM_COMPLEX64, FM_ERR_TYPE,
               "dense apply of a float32 matrix supports float32 or complex64 data");
    FM_REQUIRE(m >= 0 && n >= 0 && k >= 0, FM_ERR_VALUE, "negative size");
    const int64_t out_rows = adjoint ? n : m, in_rows = adjoint ? m : n;
    if (out_rows == 0 || k == 0) return FM_SUCCESS;
    FM_REQUIRE(A && x && y, FM_ERR_VALUE, "NULL data pointer");
    FM_REQUIRE(ldx >= in_rows && ldy >= out_rows && lda >= n, FM_ERR_VALUE,
               "leading dimension too small");
    if (generic)
        return dense_generic(a_dtype, adjoint, m, n, A, lda, x, ldx, y, ldy, k, as_stream(stream));
    GemmArgs g;
    memset(&g, 0, sizeof(g));
    g.M = out_rows;
    g.K = in_rows;
    g.A = (const float *)A;
    g.a_rs = adjoint ? 1 : lda;
    g.a_cs = adjoint ? lda : 1;
    g.X = (const float *)x;
    g.cplx = x_dtype == FM_COMPLEX64;
    g.NC = g.cplx ? 2 * k : k;
    g.J = k;
    g.bos = 0; g.brs = 1; g.bjs = ldx;
    g.Y = (float *)y;
    g.yos = 0; g.yrs = 1; g.yjs = ldy;
    return gemm_launch(g, as_stream(stream));
}

extern "C" fm_status fm_dense_apply_strided(int a_dtype, int x_dtype, int adjoint, int64_t n,
                                            const void *A, int64_t lda, const void *x, void *y,
                                            int64_t outer, int64_t inner, void *stream) {
    FM_REQUIRE(a_dtype == FM_FLOAT32, FM_ERR_TYPE, "dense apply supports float32 matrices");
    FM_REQUIRE(x_dtype == FM_FLOAT32 || x_dtype == FM_COMPLEX64, FM_ERR_TYPE,
               "dense apply supports float32 or complex64 data");
    FM_REQUIRE(n >= 1 && outer >= 0 && inner >= 1, FM_ERR_VALUE, "bad strided geometry");
    if (outer == 0) return FM_SUCCESS;
    FM_REQUIRE(A && x && y && x != y, FM_ERR_VALUE, "bad data pointers (in-place not allowed)");
    GemmArgs g;
    memset(&g, 0, sizeof(g));
    g.M = n;
    g.K = n;
    g.A = (const float *)A;
    g.a_rs = adjoint ? 1 : lda;
    g.a_cs = adjoint ? lda : 1;
    g.X = (const float *)x;
    g.cplx = x_dtype == FM_COMPLEX64;
    g.J = inner;
    g.NC = (g.cplx ? 2 : 1) * outer * inner;
    g.bos = n * inner; g.brs = inner; g.bjs = 1;
    g.Y = (float *)y;
    g.yos = n * inner; g.yrs = inner; g.yjs = 1;
    return gemm_launch(g, as_stream(stream));
}
