// Fast Walsh-Hadamard transform (natural / Sylvester order) for gfx950.
//
// Replaces Hadamard._forwardC / _backwardC and the typed butterfly loop
// _hadamardCore (fastmat/Hadamard.pyx:36-61, 164-239).  The reference runs
// `order` radix-2 stages with butterfly distance d = 1, 2, 4, ... and stores
// (a + b, a - b).  This kernel runs the same stages in the same order, so
//   * integer types (computed in unsigned arithmetic of the storage width or
//     wider, i.e. two's-complement wrap) are bit-exact, and
//   * float types see every output formed by the same sequence of IEEE adds,
//     hence equal the reference bit for bit as well.
//
// Structure: the index bits are split into passes of consecutive bit ranges
// [lo, hi), lowest bits first.  A pass transforms vectors of length 2^(hi-lo)
// with element stride 2^lo, staged through LDS in tiles of B vectors; inside a
// tile the stages run in groups of up to three bits, eight elements per
// thread in registers, in place (the WHT does not reorder).  Order <= 14
// (4-byte types) is one pass; 2^20 is two passes of 10 bits.
#include "fm_common.hpp"

namespace fm {

// Compute type per storage type: wrap-around integer types, or the float itself.
template <typename S> struct WhtTraits;
template <> struct WhtTraits<int8_t> { using C = uint32_t; };
template <> struct WhtTraits<int16_t> { using C = uint32_t; };
template <> struct WhtTraits<int32_t> { using C = uint32_t; };
template <> struct WhtTraits<int64_t> { using C = uint64_t; };
template <> struct WhtTraits<float> { using C = float; };
template <> struct WhtTraits<double> { using C = double; };
template <> struct WhtTraits<cpx<float>> { using C = cpx<float>; };
template <> struct WhtTraits<cpx<double>> { using C = cpx<double>; };

template <typename C> __device__ __forceinline__ C wadd(C a, C b) { return a + b; }
template <typename C> __device__ __forceinline__ C wsub(C a, C b) { return a - b; }

template <typename S, typename C> __device__ __forceinline__ C to_c(S v) {
    if constexpr (std::is_integral<S>::value) return C(typename std::make_unsigned<S>::type(v));
    else return v;
}
template <typename S, typename C> __device__ __forceinline__ S to_s(C v) {
    if constexpr (std::is_integral<S>::value) return S(typename std::make_unsigned<S>::type(v));
    else return v;
}

struct WhtArgs {
    int logN;          // vector length 2^logN for this pass
    int B;             // vectors per tile
    int64_t O, J;      // vector space: v = o*J + j; element n at o*os + j + n*ns
    int64_t ntiles;
    int64_t in_os, in_ns, out_os, out_ns;
    int seqfast;
    const void *x;
    void *y;
};

// In-place WHT of B vectors of length N in LDS (vector b at lds + b*S),
// stages d = 1, 2, 4, ... in groups of up to three bits.
template <typename C, int NT>
__device__ void lds_wht(C *lds, int S, int logN, int B) {
    int q = 0;
    while (q < logN) {
        const int g = (logN - q) >= 3 ? 3 : (logN - q);
        const int G = 1 << g;
        const int ngroups = (B << logN) >> g;   // butterfly groups in the tile
        const int per_seq_log = logN - g;
        for (int u = threadIdx.x; u < ngroups; u += NT) {
            const int b = u >> per_seq_log;
            const int w = u & ((1 << per_seq_log) - 1);
            const int low = w & ((1 << q) - 1);
            const int high = w >> q;
            C *base = lds + b * S + ((high << (q + g)) | low);
            C v[8];
#pragma unroll
            for (int t = 0; t < 8; ++t)
                if (t < G) v[t] = base[t << q];
            // stages in increasing distance, pairs (t, t ^ h)
#pragma unroll
            for (int h = 1; h < 8; h <<= 1) {
                if (h < G) {
#pragma unroll
                    for (int t = 0; t < 8; ++t)
                        if (t < G && !(t & h)) {
                            C a = v[t], c = v[t | h];
                            v[t] = wadd(a, c);
                            v[t | h] = wsub(a, c);
                        }
                }
            }
#pragma unroll
            for (int t = 0; t < 8; ++t)
                if (t < G) base[t << q] = v[t];
        }
        __syncthreads();
        q += g;
    }
}

template <typename S, int NT>
__global__ __launch_bounds__(NT) void wht_tile_kernel(const WhtArgs a) {
    using C = typename WhtTraits<S>::C;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    C *lds = reinterpret_cast<C *>(smem_raw);
    const int N = 1 << a.logN;
    const int S_ = N + (sizeof(C) >= 8 ? 1 : 1);
    const int B = a.B;
    int logB = 0;
    while ((1 << logB) < B) ++logB;
    const int64_t tile = xcd_remap(blockIdx.x, gridDim.x);
    const int64_t s0 = tile * B;
    const int64_t nseq = a.O * a.J;
    const S *x = reinterpret_cast<const S *>(a.x);
    S *y = reinterpret_cast<S *>(a.y);
    const int tot = B << a.logN;
    for (int idx = threadIdx.x; idx < tot; idx += NT) {
        int b, n;
        if (a.seqfast) { b = idx & (B - 1); n = idx >> logB; }
        else { n = idx & (N - 1); b = idx >> a.logN; }
        const int64_t s = s0 + b;
        if (s < nseq) {
            const int64_t o = s / a.J, j = s - o * a.J;
            lds[b * S_ + n] = to_c<S, C>(x[o * a.in_os + j + n * a.in_ns]);
        }
    }
    __syncthreads();
    lds_wht<C, NT>(lds, S_, a.logN, B);
    for (int idx = threadIdx.x; idx < tot; idx += NT) {
        int b, n;
        if (a.seqfast) { b = idx & (B - 1); n = idx >> logB; }
        else { n = idx & (N - 1); b = idx >> a.logN; }
        const int64_t s = s0 + b;
        if (s < nseq) {
            const int64_t o = s / a.J, j = s - o * a.J;
            y[o * a.out_os + j + n * a.out_ns] = to_s<S, C>(lds[b * S_ + n]);
        }
    }
}

template <typename S>
static fm_status wht_launch(const WhtArgs &a, hipStream_t stream) {
    using C = typename WhtTraits<S>::C;
    const int N = 1 << a.logN;
    const size_t lds = size_t(a.B) * (N + 1) * sizeof(C);
    FM_REQUIRE(lds <= 160 * 1024, FM_ERR_VALUE, "Hadamard tile needs %zu bytes of LDS", lds);
    static bool attr = false;
    if (!attr) {
        (void)hipFuncSetAttribute((const void *)wht_tile_kernel<S, 256>,
                            hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
        attr = true;
    }
    FM_REQUIRE(a.ntiles > 0 && a.ntiles < (int64_t(1) << 31), FM_ERR_VALUE,
               "tile count out of range");
    ProfToken tok;
    prof_begin(stream, &tok);
    hipLaunchKernelGGL((wht_tile_kernel<S, 256>), dim3(unsigned(a.ntiles)), dim3(256), lds,
                       stream, a);
    FM_CHECK_LAUNCH();
    if (tok.active) {
        const double elems = double(a.O) * double(a.J) * double(N);
        char name[64];
        snprintf(name, sizeof(name), "wht_tile_kernel<%dB>", int(sizeof(S)));
        prof_end(stream, &tok, name, 2.0 * elems * sizeof(S), elems * a.logN);
    }
    return FM_SUCCESS;
}

// Transform along bits [0, order) of vectors v[i] = x[(o*2^order + i)*inner + j].
// Passes: bit ranges from the lowest, each at most `maxbits` wide.
template <typename S>
static fm_status wht_run(int order, const void *x, int64_t ldx, void *y, int64_t ldy,
                         int64_t k, int64_t inner, hipStream_t stream) {
    using C = typename WhtTraits<S>::C;
    // 64 KB tiles of the compute type
    const int tile_elems = int(65536 / sizeof(C));
    int maxbits = 0;
    while ((2 << maxbits) <= tile_elems) ++maxbits;
    const int npass = order == 0 ? 1 : (order + maxbits - 1) / maxbits;
    int lo = 0;
    for (int p = 0; p < npass; ++p) {
        const int bits = (order - lo) / (npass - p);
        const int hi = lo + bits;
        WhtArgs a;
        memset(&a, 0, sizeof(a));
        a.logN = bits;
        // vector space: (col, high bits) x (low bits [0, lo) and inner)
        const int64_t Jlow = (int64_t(1) << lo) * inner;
        a.O = k << (order - hi);
        a.J = Jlow;
        // element n of vector (o, j): col = o >> (order-hi), hb = o & mask
        // address = col*ld + ((hb << hi) + (n << lo)) * inner + j
        // expressed with a single os requires ld == 2^order * inner for k > 1
        a.in_os = (int64_t(1) << hi) * inner;
        a.out_os = a.in_os;
        a.in_ns = Jlow;
        a.out_ns = Jlow;
        int B = 1;
        const int want = Jlow > 1 ? 16 : 8;
        while (B < want && (int64_t(B * 2) << bits) <= tile_elems && B * 2 <= a.O * a.J) B *= 2;
        a.B = B;
        a.seqfast = (Jlow > 1 && B > 1) ? 1 : 0;
        a.ntiles = ceil_div(a.O * a.J, B);
        a.x = p == 0 ? x : y;
        a.y = y;
        (void)ldx;
        (void)ldy;
        fm_status st = wht_launch<S>(a, stream);
        if (st) return st;
        lo = hi;
    }
    return FM_SUCCESS;
}

template <typename S>
static fm_status wht_columns(int order, const void *x, int64_t ldx, void *y, int64_t ldy,
                             int64_t k, hipStream_t stream) {
    const int64_t n = int64_t(1) << order;
    if (ldx == n && ldy == n) return wht_run<S>(order, x, n, y, n, k, 1, stream);
    // general leading dimensions: one launch sequence per column
    for (int64_t c = 0; c < k; ++c) {
        fm_status st = wht_run<S>(order, (const S *)x + c * ldx, n, (S *)y + c * ldy, n, 1, 1,
                                  stream);
        if (st) return st;
    }
    return FM_SUCCESS;
}

static fm_status wht_dispatch(int dtype, int order, const void *x, int64_t ldx, void *y,
                              int64_t ldy, int64_t k, int64_t inner, bool strided,
                              hipStream_t s) {
#define FM_WHT_CASE(ID, TYPE)                                                          \
    case ID:                                                                           \
        return strided ? wht_run<TYPE>(order, x, 0, y, 0, k, inner, s)                 \
                       : wht_columns<TYPE>(order, x, ldx, y, ldy, k, s);
    switch (dtype) {
        FM_WHT_CASE(FM_INT8, int8_t)
        FM_WHT_CASE(FM_INT16, int16_t)
        FM_WHT_CASE(FM_INT32, int32_t)
        FM_WHT_CASE(FM_INT64, int64_t)
        FM_WHT_CASE(FM_FLOAT32, float)
        FM_WHT_CASE(FM_FLOAT64, double)
        FM_WHT_CASE(FM_COMPLEX64, cpx<float>)
        FM_WHT_CASE(FM_COMPLEX128, cpx<double>)
    default:
        set_error("Unsupported type: %d", dtype);
        return FM_ERR_TYPE;
    }
#undef FM_WHT_CASE
}

}  // namespace fm

using namespace fm;

extern "C" fm_status fm_hadamard_apply(int dtype, int order, const void *x, int64_t ldx, void *y,
                                       int64_t ldy, int64_t k, void *stream) {
    FM_REQUIRE(order >= 1, FM_ERR_VALUE, "Hadamard: Order must be larger than 0.");
    FM_REQUIRE(order <= 40, FM_ERR_VALUE, "Hadamard: Order %d exceeds the supported maximum",
               order);
    FM_REQUIRE(k >= 0, FM_ERR_VALUE, "negative column count");
    if (k == 0) return FM_SUCCESS;
    FM_REQUIRE(x && y, FM_ERR_VALUE, "NULL data pointer");
    const int64_t n = int64_t(1) << order;
    FM_REQUIRE(ldx >= n && ldy >= n, FM_ERR_VALUE, "leading dimension below 2^order");
    FM_REQUIRE(!(x == y && ldx != ldy), FM_ERR_VALUE,
               "in-place transform needs equal leading dimensions");
    return wht_dispatch(dtype, order, x, ldx, y, ldy, k, 1, false, as_stream(stream));
}

extern "C" fm_status fm_hadamard_apply_strided(int dtype, int order, const void *x, void *y,
                                               int64_t outer, int64_t inner, void *stream) {
    FM_REQUIRE(order >= 1 && order <= 40, FM_ERR_VALUE, "Hadamard: bad order %d", order);
    FM_REQUIRE(outer >= 0 && inner >= 1, FM_ERR_VALUE, "bad strided geometry");
    if (outer == 0) return FM_SUCCESS;
    FM_REQUIRE(x && y, FM_ERR_VALUE, "NULL data pointer");
    return wht_dispatch(dtype, order, x, 0, y, 0, outer, inner, true, as_stream(stream));
}
