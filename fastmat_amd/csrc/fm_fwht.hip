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
// [lo, hi), lowest bits first (one pass up to 2^14 for 4-byte types, two
// passes of 10 bits for 2^20).  A pass kernel is compiled per pass length; a
// thread holds E = 32 / 16 / 8 elements (4 / 8 / 16-byte compute type) and
// runs the pass as stages of at most log2(E) bits in registers, in place
// (the WHT does not reorder), exchanging through a padded LDS tile between
// stages.  Strided passes read their first stage straight from HBM with
// neighbouring lanes on neighbouring vectors; contiguous passes load the
// tile coalesced into LDS first.  The last stage stores straight to HBM.
#include <algorithm>
#include <vector>

#include "fm_fft_core.hpp"  // lpad, shared helpers

namespace fm {

int knob(const char *name, int dflt);

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

// Per-thread element count (power of two) by compute-type size.
template <typename C> constexpr int wht_loge() {
    return sizeof(C) <= 4 ? 5 : (sizeof(C) == 8 ? 4 : 3);
}

struct WhtArgs {
    int logN;          // vector length 2^logN for this pass
    int64_t O, J;      // vector space: v = o*J + j; element n at o*os + j + n*ns
    int logJ;          // log2(J) if J is a power of two, else -1
    int64_t ntiles;
    int64_t os, ns;
    int seqfast;       // neighbouring vectors are adjacent in memory (J > 1)
    const void *x;
    void *y;
    int npass;         // passes of the whole transform (profiler attribution)
};

// Stage plan: S stages of at most LOGE bits, lowest bits first (the
// reference's order, distance 1, 2, 4, ...).
template <int LOGN, int LOGE> struct WhtPlan {
    static constexpr int S = LOGN == 0 ? 0 : (LOGN + LOGE - 1) / LOGE;
    static constexpr int bits(int s) { return LOGN / S + (s < LOGN % S ? 1 : 0); }
    static constexpr int q(int s) {
        int t = 0;
        for (int k = 0; k < s; ++k) t += bits(k);
        return t;
    }
};

// Tile shape of a length-2^LOGN pass with NT threads holding 2^LOGE elements
// each: B = NT * 2^LOGE / 2^LOGN vectors per tile.
template <int LOGN, int NT, int LOGE> struct WhtShape {
    static_assert(NT == 256 || NT == 512 || NT == 1024, "tile of 256, 512 or 1024 threads");
    static constexpr int LOGNT = NT == 256 ? 8 : (NT == 512 ? 9 : 10);
    static constexpr int LOGB = LOGNT + LOGE - LOGN;
    static constexpr int B = 1 << (LOGB > 0 ? LOGB : 0);
    static constexpr int N = 1 << LOGN;
    static constexpr int SB = N + (N >> 5) + 1;  // padded LDS row
};

// Position of element t of butterfly group u for a stage on bits [Q, Q+R).
template <int Q, int R> __host__ __device__ __forceinline__ int wht_pos(int u, int t) {
    const int low = u & ((1 << Q) - 1);
    const int high = u >> Q;
    return (high << (Q + R)) | (t << Q) | low;
}

// In-register WHT over the R = 2^LR elements of a group, stages h = 1, 2, ...
template <typename C, int LR> __device__ __forceinline__ void wht_regs(C *v) {
#pragma unroll
    for (int h = 1; h < (1 << LR); h <<= 1) {
#pragma unroll
        for (int t = 0; t < (1 << LR); ++t) {
            if (!(t & h)) {
                const C a = v[t], c = v[t | h];
                v[t] = wadd(a, c);
                v[t | h] = wsub(a, c);
            }
        }
    }
}

__device__ __forceinline__ void wht_seq(const WhtArgs &a, int64_t s, bool &ok, int64_t &base) {
    ok = s < a.O * a.J;
    int64_t o, j;
    if (a.logJ >= 0) {
        o = s >> a.logJ;
        j = s & ((int64_t(1) << a.logJ) - 1);
    } else {
        o = s / a.J;
        j = s - o * a.J;
    }
    base = o * a.os + j;
}

template <int LOGB, int LOGU>
__host__ __device__ __forceinline__ void wht_map(int g, bool seqfast, int &b, int &u) {
    if (seqfast) {
        b = g & ((1 << LOGB) - 1);
        u = g >> LOGB;
    } else {
        u = g & ((1 << LOGU) - 1);
        b = g >> LOGU;
    }
}

// Strided last-stage re-stage: thread `tid` stores element
// p = (tid >> LOGB) + e * (NT >> LOGB) of vector b = tid mod B, e < E.
template <int LOGB, int NT>
__host__ __device__ __forceinline__ void wht_restage_pos(int tid, int e, int &b, int &p) {
    b = tid & ((1 << LOGB) - 1);
    p = (tid >> LOGB) + e * (NT >> LOGB);
}

// One stage s of the plan on registers v.  SRC: 0 = v holds the inputs
// (loaded in mapping sf), 1 = read from LDS.  DST: 0 = write back to LDS,
// 1 = keep in v.
template <typename C, int LOGN, int LOGE, int NT, int LOGB, int s, int SRC, int DST>
__device__ __forceinline__ void wht_stage(C *v, C *lds, bool sf) {
    using P = WhtPlan<LOGN, LOGE>;
    constexpr int LR = P::bits(s), R = 1 << LR, Q = P::q(s);
    constexpr int E = 1 << LOGE, NB = E / R, LOGU = LOGN - LR;
    constexpr int SB = (1 << LOGN) + ((1 << LOGN) >> 5) + 1;
    if constexpr (SRC == 1) sf = false;
    if constexpr (SRC == 1) {
#pragma unroll
        for (int e = 0; e < NB; ++e) {
            int b, u;
            wht_map<LOGB, LOGU>(threadIdx.x + e * NT, sf, b, u);
#pragma unroll
            for (int t = 0; t < R; ++t) v[e * R + t] = lds[b * SB + lpad(wht_pos<Q, LR>(u, t))];
        }
    }
#pragma unroll
    for (int e = 0; e < NB; ++e) wht_regs<C, LR>(v + e * R);
    if constexpr (DST == 0) {
        __syncthreads();
#pragma unroll
        for (int e = 0; e < NB; ++e) {
            int b, u;
            wht_map<LOGB, LOGU>(threadIdx.x + e * NT, sf, b, u);
#pragma unroll
            for (int t = 0; t < R; ++t) lds[b * SB + lpad(wht_pos<Q, LR>(u, t))] = v[e * R + t];
        }
        __syncthreads();
    }
}

template <typename C, int LOGN, int LOGE, int NT, int LOGB, int FIRST_SRC, int s = 0>
__device__ __forceinline__ void wht_all(C *v, C *lds, bool sf0) {
    using P = WhtPlan<LOGN, LOGE>;
    if constexpr (s < P::S) {
        constexpr int src = s == 0 ? FIRST_SRC : 1;
        constexpr int dst = s == P::S - 1 ? 1 : 0;
        wht_stage<C, LOGN, LOGE, NT, LOGB, s, src, dst>(v, lds, sf0);
        wht_all<C, LOGN, LOGE, NT, LOGB, FIRST_SRC, s + 1>(v, lds, false);
    }
}

// Element load / store of a pass; STREAM bit 1 makes the loads and bit 0 the
// stores non-temporal (the data is touched once per pass, so it need not
// displace other lines from the caches).
template <int STREAM, typename S> __device__ __forceinline__ S wht_ld(const S *x, int64_t i) {
    const FM_GLOBAL S *q = (const FM_GLOBAL S *)x + i;
    if constexpr (STREAM & 2) return __builtin_nontemporal_load(q);
    else return *q;
}
template <int STREAM, typename S> __device__ __forceinline__ void wht_st(S *y, int64_t i, S v) {
    FM_GLOBAL S *q = (FM_GLOBAL S *)y + i;
    if constexpr (STREAM & 1) __builtin_nontemporal_store(v, q);
    else *q = v;
}

template <typename S, int LOGN, int NT, int STREAM = 0>
__global__ __launch_bounds__(NT) void wht_pass_kernel(const WhtArgs a) {
    using C = typename WhtTraits<S>::C;
    constexpr int LOGE = wht_loge<C>(), E = 1 << LOGE;
    constexpr int N = 1 << LOGN;
    using P = WhtPlan<LOGN, LOGE>;
    using Sh = WhtShape<LOGN, NT, LOGE>;
    constexpr int LOGB = Sh::LOGB;
    constexpr int B = Sh::B;
    constexpr int SB = Sh::SB;
    static_assert(LOGB >= 0, "tile must hold a vector");
    extern __shared__ __align__(16) unsigned char smem_raw[];
    C *lds = reinterpret_cast<C *>(smem_raw);
    const S *__restrict__ x = reinterpret_cast<const S *>(a.x);
    S *__restrict__ y = reinterpret_cast<S *>(a.y);
    const int64_t s0 = xcd_remap(blockIdx.x, gridDim.x) << LOGB;
    const bool seqfast = a.seqfast && B > 1;
    C v[E];
    if (seqfast) {
        // strided vectors: stage 0 straight from HBM, neighbouring lanes on
        // neighbouring vectors (contiguous in memory)
        constexpr int LR = P::bits(0), R = 1 << LR, NB = E / R, LOGU = LOGN - LR;
#pragma unroll
        for (int e = 0; e < NB; ++e) {
            int b, u;
            wht_map<LOGB, LOGU>(threadIdx.x + e * NT, true, b, u);
            bool ok;
            int64_t base;
            wht_seq(a, s0 + b, ok, base);
#pragma unroll
            for (int t = 0; t < R; ++t)
                v[e * R + t] = ok ? to_c<S, C>(wht_ld<STREAM>(x, base + int64_t(wht_pos<0, LR>(u, t)) * a.ns))
                                  : C{};
        }
        wht_all<C, LOGN, LOGE, NT, LOGB, 0>(v, lds, true);
    } else {
        // contiguous vectors: coalesced natural-order load into LDS first
#pragma unroll
        for (int e = 0; e < E; ++e) {
            const int idx = threadIdx.x + e * NT;
            const int b = idx >> LOGN, n = idx & (N - 1);
            bool ok;
            int64_t base;
            wht_seq(a, s0 + b, ok, base);
            v[e] = ok ? to_c<S, C>(wht_ld<STREAM>(x, base + int64_t(n) * a.ns)) : C{};
        }
#pragma unroll
        for (int e = 0; e < E; ++e) {
            const int idx = threadIdx.x + e * NT;
            lds[(idx >> LOGN) * SB + lpad(idx & (N - 1))] = v[e];
        }
        __syncthreads();
        wht_all<C, LOGN, LOGE, NT, LOGB, 1>(v, lds, false);
    }
    // last stage (bits [q, LOGN)): element t of group u sits at (t << q) | u,
    // so u-fastest lanes store contiguous runs; strided vectors store
    // sequence-fastest when the last stage is also the first
    constexpr int LRL = P::bits(P::S - 1), RL = 1 << LRL, QL = P::q(P::S - 1);
    constexpr int NBL = E / RL, LOGUL = LOGN - LRL;
    if (seqfast && P::S > 1) {
        // strided vectors: re-stage the last stage through LDS so that
        // neighbouring lanes store neighbouring vectors (B-element runs per
        // row) instead of one element per row each
        __syncthreads();
#pragma unroll
        for (int e = 0; e < NBL; ++e) {
            int b, u;
            wht_map<LOGB, LOGUL>(threadIdx.x + e * NT, false, b, u);
#pragma unroll
            for (int t = 0; t < RL; ++t) lds[b * SB + lpad(wht_pos<QL, LRL>(u, t))] = v[e * RL + t];
        }
        __syncthreads();
        // idx = tid + e NT: vector b = idx mod B is the same for all e
        int b, p;
        wht_restage_pos<LOGB, NT>(threadIdx.x, 0, b, p);
        bool ok;
        int64_t base;
        wht_seq(a, s0 + b, ok, base);
        if (ok) {
#pragma unroll
            for (int e = 0; e < E; ++e) {
                wht_restage_pos<LOGB, NT>(threadIdx.x, e, b, p);
                wht_st<STREAM>(y, base + int64_t(p) * a.ns, to_s<S, C>(lds[b * SB + lpad(p)]));
            }
        }
        return;
    }
    const bool sfl = P::S == 1 ? seqfast : false;
#pragma unroll
    for (int e = 0; e < NBL; ++e) {
        int b, u;
        wht_map<LOGB, LOGUL>(threadIdx.x + e * NT, sfl, b, u);
        bool ok;
        int64_t base;
        wht_seq(a, s0 + b, ok, base);
        if (!ok) continue;
#pragma unroll
        for (int t = 0; t < RL; ++t)
            wht_st<STREAM>(y, base + int64_t(wht_pos<QL, LRL>(u, t)) * a.ns, to_s<S, C>(v[e * RL + t]));
    }
}

template <typename S, int LOGN, int NT, int STREAM = 0>
static fm_status wht_launch_nt(WhtArgs a, hipStream_t stream) {
    using C = typename WhtTraits<S>::C;
    constexpr int LOGE = wht_loge<C>();
    using Sh = WhtShape<LOGN, NT, LOGE>;
    constexpr int B = Sh::B;
    constexpr int N = Sh::N;
    const size_t lds = size_t(B) * Sh::SB * sizeof(C);
    static bool attr = false;
    if (!attr) {
        (void)hipFuncSetAttribute((const void *)wht_pass_kernel<S, LOGN, NT, STREAM>,
                                  hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
        attr = true;
    }
    a.ntiles = ceil_div(a.O * a.J, B);
    a.logJ = is_pow2(a.J) ? ilog2(a.J) : -1;
    FM_REQUIRE(a.ntiles > 0 && a.ntiles < (int64_t(1) << 31), FM_ERR_VALUE,
               "tile count out of range");
    ProfToken tok;
    prof_begin(stream, &tok);
    hipLaunchKernelGGL((wht_pass_kernel<S, LOGN, NT, STREAM>), dim3(unsigned(a.ntiles)), dim3(NT), lds,
                       stream, a);
    FM_CHECK_LAUNCH();
    if (tok.active) {
        const double elems = double(a.O) * double(a.J) * double(N);
        char name[64];
        snprintf(name, sizeof(name), "wht_pass_kernel<%dB,%d>", int(sizeof(S)), LOGN);
        // algorithmic: the transform's input and output, split over its passes
        const double own = 2.0 * elems * sizeof(S);
        prof_end(stream, &tok, name, own / (a.npass > 0 ? a.npass : 1), elems * LOGN, own);
    }
    return FM_SUCCESS;
}

// Launch with plain accesses, or (4-byte numbers, nt) non-temporal ones.
template <typename S, int LOGN, int NT>
static fm_status wht_launch_sel(WhtArgs a, hipStream_t stream, bool nt) {
    if constexpr (sizeof(S) == 4 && std::is_arithmetic<S>::value) {
        if (nt) return wht_launch_nt<S, LOGN, NT, 3>(a, stream);
    }
    return wht_launch_nt<S, LOGN, NT>(a, stream);
}

template <typename S, int LOGN>
static fm_status wht_launch_t(WhtArgs a, hipStream_t stream) {
    using C = typename WhtTraits<S>::C;
    constexpr int LOGE = wht_loge<C>();
    constexpr int NT = LOGN > LOGE + 8 ? 512 : 256;
    // strided vectors: wider tiles (more vectors side by side, so every row
    // segment a wave moves is longer) when they fit the LDS: 1024 threads
    // (one tile per CU; for 4-byte types 32 vectors, 128-byte segments) by
    // default, else 512 (two tiles per CU).  Order 20, 64 columns, int32 and
    // float32 forward + adjoint: 1.177 ms per step with 512, 1.016 with 1024.
    // FASTMAT_AMD_WHT_STRIDED_NT=512 selects the narrower tile.
    // 4-byte numbers: non-temporal loads and stores in the wide strided
    // pass (mode 2) or in every pass (mode 3); mode 1 plain accesses.  Order
    // 20, 64 columns, int32 + float32 forward + adjoint: 1.02 ms per step
    // plain, 0.91 with mode 2, 0.95 with mode 3 (the contiguous pass is
    // slower with non-temporal accesses).  FASTMAT_AMD_WHT_STREAM selects it.
    const int stream_mode = knob("FASTMAT_AMD_WHT_STREAM", 2);
    if constexpr (NT == 256) {
        const int want = knob("FASTMAT_AMD_WHT_STRIDED_NT", 1024);
        if (a.seqfast && want >= 1024) {
            using S4 = WhtShape<LOGN, 1024, LOGE>;
            if constexpr (S4::LOGB >= 0) {
                constexpr size_t lds4 = size_t(S4::B) * S4::SB * sizeof(C);
                if (a.J >= S4::B && lds4 <= 160 * 1024)
                    return wht_launch_sel<S, LOGN, 1024>(a, stream, stream_mode >= 2);
            }
        }
        using S2 = WhtShape<LOGN, 512, LOGE>;
        constexpr size_t lds2 = size_t(S2::B) * S2::SB * sizeof(C);
        if (a.seqfast && a.J >= S2::B && lds2 <= 80 * 1024)
            return wht_launch_sel<S, LOGN, 512>(a, stream, stream_mode >= 3);
    }
    return wht_launch_sel<S, LOGN, NT>(a, stream, stream_mode >= 3);
}

// Host-side check of a tile's index mapping (the formulas the kernel runs):
// every (vector, element) of the B x N tile must be loaded by the strided
// stage-0 mapping, written by every stage's LDS exchange and stored by the
// strided re-stage exactly once.  The tile runs NT threads; its shape
// (vectors per tile B, the re-stage strides) is derived for SHAPE_NT threads,
// which the kernel takes equal to NT (a shape derived for another thread
// count is the error a kernel with a wrong log2(NT) makes).  Returns the
// number of slots that are not hit exactly once (0 = a bijection).
template <int LOGN, int NT, int LOGE, int SHAPE_NT = NT> static int64_t wht_tile_check() {
    using Sh = WhtShape<LOGN, SHAPE_NT, LOGE>;
    using P = WhtPlan<LOGN, LOGE>;
    constexpr int LOGB = Sh::LOGB, B = Sh::B, N = Sh::N, E = 1 << LOGE;
    if constexpr (LOGB < 0) {
        return -1;
    } else {
        std::vector<int> hits(size_t(B) * N);
        int64_t bad = 0;
        auto flush = [&]() {
            for (int h : hits) bad += h != 1;
            std::fill(hits.begin(), hits.end(), 0);
        };
        auto mark = [&](int b, int pos) {
            if (b < 0 || b >= B || pos < 0 || pos >= N) ++bad;
            else ++hits[size_t(b) * N + pos];
        };
        // stage 0, strided vectors (sequence-fastest)
        {
            constexpr int LR = P::bits(0), R = 1 << LR, NB = E / R, LOGU = LOGN - LR;
            for (int tid = 0; tid < NT; ++tid)
                for (int e = 0; e < NB; ++e) {
                    int b, u;
                    wht_map<LOGB, LOGU>(tid + e * NT, B > 1, b, u);
                    for (int t = 0; t < R; ++t) mark(b, wht_pos<0, LR>(u, t));
                }
            flush();
        }
        // every stage's exchange (butterfly-fastest after stage 0)
        auto stage = [&](auto sc) {
            constexpr int st = decltype(sc)::value;
            constexpr int LR = P::bits(st), R = 1 << LR, Q = P::q(st), NB = E / R;
            constexpr int LOGU = LOGN - LR;
            for (int tid = 0; tid < NT; ++tid)
                for (int e = 0; e < NB; ++e) {
                    int b, u;
                    wht_map<LOGB, LOGU>(tid + e * NT, false, b, u);
                    for (int t = 0; t < R; ++t) mark(b, wht_pos<Q, LR>(u, t));
                }
            flush();
        };
        if constexpr (P::S > 0) stage(std::integral_constant<int, 0>());
        if constexpr (P::S > 1) stage(std::integral_constant<int, 1>());
        if constexpr (P::S > 2) stage(std::integral_constant<int, 2>());
        if constexpr (P::S > 3) stage(std::integral_constant<int, 3>());
        if constexpr (P::S > 4) stage(std::integral_constant<int, 4>());
        static_assert(P::S <= 5, "more stages than the check covers");
        // strided re-stage of the last stage (the kernel takes it when the
        // pass has more than one stage, i.e. N > E, so NT >= B)
        if constexpr (P::S > 1) {
            for (int tid = 0; tid < NT; ++tid)
                for (int e = 0; e < E; ++e) {
                    int b, pp;
                    wht_restage_pos<LOGB, SHAPE_NT>(tid, e, b, pp);
                    mark(b, pp);
                }
            flush();
        }
        return bad;
    }
}

template <int NT, int LOGE, int SHAPE_NT = NT> static int64_t wht_tile_check_n(int logn) {
    switch (logn) {
#define FM_WC(L) case L: return wht_tile_check<L, NT, LOGE, SHAPE_NT>();
        FM_WC(1) FM_WC(2) FM_WC(3) FM_WC(4) FM_WC(5) FM_WC(6) FM_WC(7) FM_WC(8) FM_WC(9)
        FM_WC(10) FM_WC(11) FM_WC(12) FM_WC(13) FM_WC(14)
#undef FM_WC
    default:
        return -2;
    }
}

template <typename S> constexpr int wht_maxlog() {
    return wht_loge<typename WhtTraits<S>::C>() + 9;
}

template <typename S> static fm_status wht_launch(WhtArgs a, hipStream_t stream) {
    switch (a.logN) {
#define FM_W(L)                                                                  \
    case L:                                                                      \
        if constexpr (L <= wht_maxlog<S>()) return wht_launch_t<S, L>(a, stream); \
        break;
        FM_W(1) FM_W(2) FM_W(3) FM_W(4) FM_W(5) FM_W(6) FM_W(7) FM_W(8) FM_W(9) FM_W(10)
        FM_W(11) FM_W(12) FM_W(13) FM_W(14)
#undef FM_W
    default:
        break;
    }
    set_error("Hadamard pass of 2^%d elements unsupported", a.logN);
    return FM_ERR_VALUE;
}

// Transform along bits [0, order) of vectors v[i] = x[(o*2^order + i)*inner + j]:
// passes over consecutive bit ranges from the lowest, each at most
// wht_maxlog bits, split as evenly as possible.
template <typename S>
static fm_status wht_run(int order, const void *x, int64_t ldx, void *y, int64_t ldy,
                         int64_t k, int64_t inner, hipStream_t stream) {
    (void)ldx;
    (void)ldy;
    const int maxbits = wht_maxlog<S>();
    const int npass = (order + maxbits - 1) / maxbits;
    int lo = 0;
    for (int p = 0; p < npass; ++p) {
        const int bits = (order - lo) / (npass - p);
        const int hi = lo + bits;
        WhtArgs a;
        memset(&a, 0, sizeof(a));
        a.logN = bits;
        const int64_t Jlow = (int64_t(1) << lo) * inner;
        // vector (col, high bits) x (low bits, inner); element n at stride Jlow
        a.O = k << (order - hi);
        a.J = Jlow;
        a.os = (int64_t(1) << hi) * inner;
        a.ns = Jlow;
        a.seqfast = Jlow > 1 ? 1 : 0;
        a.x = p == 0 ? x : y;
        a.y = y;
        a.npass = npass;
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

extern "C" fm_status fm_wht_tile_check(int logn, int threads, int shape_threads, int loge,
                                       int64_t *bad) {
    FM_REQUIRE(bad != nullptr, FM_ERR_VALUE, "NULL argument");
    int64_t r = -2;
#define FM_WCN(NT, LE)                                                              \
    if (threads == NT && loge == LE) {                                              \
        if (shape_threads == NT) r = wht_tile_check_n<NT, LE>(logn);                \
        else if (shape_threads == 512 && NT == 1024) r = wht_tile_check_n<1024, LE, 512>(logn); \
    }
    FM_WCN(256, 5) FM_WCN(512, 5) FM_WCN(1024, 5)
    FM_WCN(256, 4) FM_WCN(512, 4) FM_WCN(1024, 4)
    FM_WCN(256, 3) FM_WCN(512, 3) FM_WCN(1024, 3)
#undef FM_WCN
    FM_REQUIRE(r != -2, FM_ERR_VALUE,
               "no tile of %d threads (shape of %d), 2^%d elements, 2^%d per thread", threads,
               shape_threads, logn, loge);
    *bad = r;
    return FM_SUCCESS;
}
