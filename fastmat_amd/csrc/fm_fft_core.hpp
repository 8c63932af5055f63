// FFT engine: the tile body used by every Fourier / Circulant / Toeplitz /
// Bluestein pass (per-operator kernels and the batched persistent kernels).
//
// Method (written from the published Stockham autosort formulation): a
// length-N transform (N = 2^LOGN, compile-time) is split into S stages of
// radix 2^bits(s) <= E, E = 32 (complex64) or 16 (complex128) elements per
// thread.  Stage with radix R and sub-length p: butterfly i in [0, N/R) takes
// X[i + r N/R], multiplies input r by W_{pR}^{r k} (k = i mod p), runs an
// R-point DFT in registers and produces Y[(i - k) R + k + r p]; outputs come
// out in natural order.  Stage 0 loads its inputs straight from HBM, the last
// stage stores straight to HBM, and the S-1 exchanges in between go through a
// padded LDS tile (one spare slot per 32 elements).  1024 = 32 x 32 needs one
// exchange.
//
// Tile: B = 256 * E / N consecutive sequences s = o*J + j of an (O x J)
// sequence space (512 threads for the largest N).  Element n of sequence
// (o, j) is read at  x[o*in_os + j*in_js + n*in_ns]  and written at
// y[o*out_os + j*out_js + n*out_ns]; this one addressing form covers
// contiguous columns, the strided and the transposed sides of the four-step
// passes, and Kron mode axes.  Global accesses map neighbouring lanes to
// neighbouring sequences where those are adjacent in memory (strided sides)
// and to neighbouring elements otherwise.  Work fused around the transform:
//     load:  zero-pad (global index n*in_gn + j*in_gj >= n_valid), conj,
//            pre-multiply by a chirp, then FFT(sign1)
//     conv:  multiply by the spectrum D[n*d_gn + j*d_gj] (or conj) in
//            registers, then FFT(-sign1) with the radices reversed so it
//            starts from the forward transform's registers
//     store: twiddle W_L^{+-j n} of an enclosing four-step, crop
//            (n*out_gn + j*out_gj >= m_valid), post-multiply, conj, scale.
#pragma once

#include <type_traits>

#include "fm_common.hpp"

namespace fm {

// ------------------------------------------------------------ small DFTs
template <int SIGN, typename T> __device__ __forceinline__ void dft2(cpx<T> *v) {
    cpx<T> a = v[0];
    v[0] = a + v[1];
    v[1] = a - v[1];
}

template <int SIGN, typename T> __device__ __forceinline__ void dft4(cpx<T> *v) {
    cpx<T> t0 = v[0] + v[2];
    cpx<T> t1 = v[0] - v[2];
    cpx<T> t2 = v[1] + v[3];
    cpx<T> t3 = mul_i<SIGN>(v[1] - v[3]);
    v[0] = t0 + t2;
    v[2] = t0 - t2;
    v[1] = t1 + t3;
    v[3] = t1 - t3;
}

template <int SIGN, typename T> __device__ __forceinline__ void dft8(cpx<T> *v) {
    cpx<T> e[4] = {v[0], v[2], v[4], v[6]};
    cpx<T> o[4] = {v[1], v[3], v[5], v[7]};
    dft4<SIGN>(e);
    dft4<SIGN>(o);
    const T h = T(0.70710678118654752440084436210485);
    cpx<T> a = o[1], c = o[3];
    if (SIGN < 0) {
        o[1] = mk(h * (a.x + a.y), h * (a.y - a.x));        // * h(1 - i)
        o[3] = mk(h * (c.y - c.x), -h * (c.x + c.y));       // * h(-1 - i)
    } else {
        o[1] = mk(h * (a.x - a.y), h * (a.x + a.y));        // * h(1 + i)
        o[3] = mk(-h * (c.x + c.y), h * (c.x - c.y));       // * h(-1 + i)
    }
    o[2] = mul_i<SIGN>(o[2]);
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        v[q] = e[q] + o[q];
        v[q + 4] = e[q] - o[q];
    }
}

// W_32^k = exp(-2 pi i k / 32), k < 16 (cos, sin)
__device__ __forceinline__ constexpr double w32c(int k) {
    constexpr double c[16] = {1.0, 0.98078528040323043, 0.92387953251128674, 0.83146961230254524,
                              0.70710678118654757, 0.55557023301960229, 0.38268343236508984,
                              0.19509032201612833, 0.0, -0.19509032201612819, -0.38268343236508973,
                              -0.55557023301960196, -0.70710678118654746, -0.83146961230254535,
                              -0.92387953251128674, -0.98078528040323043};
    return c[k];
}
__device__ __forceinline__ constexpr double w32s(int k) {
    constexpr double s[16] = {0.0, 0.19509032201612825, 0.38268343236508978, 0.55557023301960218,
                              0.70710678118654746, 0.83146961230254524, 0.92387953251128674,
                              0.98078528040323043, 1.0, 0.98078528040323043, 0.92387953251128674,
                              0.83146961230254546, 0.70710678118654757, 0.55557023301960218,
                              0.38268343236508989, 0.19509032201612861};
    return s[k];
}

// x * W_R^k (forward, SIGN < 0) or x * conj(W_R^k) (inverse), k < R/2, R <= 32;
// k is a compile-time constant after unrolling, so the branches fold away.
template <int R, int SIGN, typename T>
__device__ __forceinline__ cpx<T> wmul(cpx<T> x, int k) {
    if (k == 0) return x;
    if (4 * k == R) return mul_i<SIGN>(x);
    const int q = k * (32 / R);
    const T c = T(w32c(q)), s = T(w32s(q));
    // forward W = c - i s ; inverse conj(W) = c + i s
    return SIGN < 0 ? mk(x.x * c + x.y * s, x.y * c - x.x * s)
                    : mk(x.x * c - x.y * s, x.y * c + x.x * s);
}

// In-register DFT of length 2^LOGR (LOGR <= 5), natural order in and out.
template <int LOGR, int SIGN, typename T> __device__ __forceinline__ void dftr(cpx<T> *v) {
    if constexpr (LOGR == 1) dft2<SIGN>(v);
    else if constexpr (LOGR == 2) dft4<SIGN>(v);
    else if constexpr (LOGR == 3) dft8<SIGN>(v);
    else if constexpr (LOGR >= 4) {
        constexpr int R = 1 << LOGR, H = R / 2;
        cpx<T> e[H], o[H];
#pragma unroll
        for (int i = 0; i < H; ++i) {
            e[i] = v[2 * i];
            o[i] = v[2 * i + 1];
        }
        dftr<LOGR - 1, SIGN>(e);
        dftr<LOGR - 1, SIGN>(o);
#pragma unroll
        for (int k = 0; k < H; ++k) {
            const cpx<T> t = wmul<R, SIGN>(o[k], k);
            v[k] = e[k] + t;
            v[k + H] = e[k] - t;
        }
    }
}

template <int SIGN, typename T>
__device__ __forceinline__ cpx<T> twiddle_mul(cpx<T> v, cpx<T> w) {
    // w is the forward twiddle exp(-2 pi i e / N); the inverse uses conj(w)
    return SIGN < 0 ? cmul(v, w) : cmulc(v, w);
}

// Stage plan of a length-2^LOGN transform for threads holding 2^LOGE
// elements: S stages with radices 2^bits(s) <= 2^LOGE, bits as even as
// possible, larger radices first.
template <int LOGN, int LOGE, bool REV = false> struct FftPlan {
    static constexpr int S = LOGN == 0 ? 0 : (LOGN + LOGE - 1) / LOGE;
    static constexpr int SD = S > 0 ? S : 1;  // (LOGN == 0 has no stages)
    static constexpr int fbits(int s) { return LOGN / SD + (s < LOGN % SD ? 1 : 0); }
    // REV runs the same radices last-to-first (any order is a valid Stockham
    // factorisation); a transform whose first radix equals the previous
    // transform's last radix can start from that transform's registers.
    static constexpr int bits(int s) { return REV ? fbits(S - 1 - s) : fbits(s); }
    static constexpr int logp(int s) {
        int t = 0;
        for (int q = 0; q < s; ++q) t += bits(q);
        return t;
    }
};

// Padded LDS index: one spare slot every 32 elements keeps the strided
// exchanges of the stages free of most bank conflicts.
__device__ __forceinline__ int lpad(int n) { return n + (n >> 5); }

// ------------------------------------------------------------ tile kernel
template <typename T> struct TileArgs {
    int logN;          // transform length N = 2^logN
    int B;             // sequences per tile (power of two)
    int64_t O, J;      // sequence space, s = o*J + j
    int64_t ntiles;
    // input
    const cpx<T> *x;
    int64_t in_os, in_js, in_ns;
    int64_t in_gn, in_gj, n_valid;  // zero-pad predicate
    const cpx<T> *pre;              // optional chirp, indexed by global input index
    int conj_in;
    int seqfast_in;
    // transform
    int sign1;                      // -1 forward, +1 inverse
    const cpx<double> *tw;          // W_N^e table (FP64)
    // spectral multiply + second transform (conv mode)
    const cpx<T> *D;
    int64_t d_gn, d_gj;
    int conj_d;
    // enclosing four-step twiddle applied before the store: W_L^{tw_sign * j * n}
    const cpx<double> *twb_hi, *twb_lo;  // W_L^{a * 2^twb_shift}, W_L^{b} (FP64)
    int twb_shift;
    int twb_sign;
    // output
    cpx<T> *y;
    int64_t out_os, out_js, out_ns;
    int64_t out_gn, out_gj, m_valid;
    const cpx<T> *post;
    int conj_out;
    T out_scale;
    int seqfast_out;
    // batch engine: this pass writes the caller's output of its direction;
    // a redirected apply moves y by a byte shift and multiplies out_scale by
    // a device scalar for such passes only
    int redirect;
    // non-temporal (streaming) accesses (the batch engine's passes): bit 0
    // the output stores, bit 1 the input loads
    int stream_out;
    int logJ;                       // log2(J) when J is a power of two, else -1
    int64_t o_group;                // > 1: tiles walk groups of o_group o's (see tile_first_seq)
    // optional residual epilogue (target != NULL): the store writes
    // out_scale * (u - t) with t read from `target` at the output's position
    // (same layout as y), and the tile's sum of |u - t|^2 over the stored
    // elements goes to res_part[res_base + tile] (FP64)
    const cpx<T> *target;
    double *res_part;
    int64_t res_base;
    // optional residual between the two transforms of a conv-mode pass, in
    // place of the spectrum (mid_t != NULL): element n of sequence (o, j)
    // of the first transform's output stands at index g = n*mid_gn + j*mid_gj
    // of column o; for g < mid_valid it becomes mid_scale * (u - t) with
    // t = mid_t[o*mid_os + g] (|u - t|^2 adds to the tile's residual sum
    // res_part[res_base + tile]), beyond mid_valid it becomes 0
    const cpx<T> *mid_t;
    int64_t mid_os, mid_gn, mid_gj, mid_valid;
    T mid_scale;
    // host only: this pass's share of its transform's algorithmic bytes (the
    // input and output blocks and the generator, split over the passes),
    // for the library's profiler; 0 = count the pass's own reads and writes
    double alg_bytes;
};

// Index arithmetic of a pass: int64_t in general, uint32_t when every
// element offset and logical index of the launch is below 2^32 / sizeof
// (32-bit offsets from a uniform base become saddr + voffset addressing).
template <bool I32> using IdxT = typename std::conditional<I32, uint32_t, int64_t>::type;

// Load / store element `idx` of a uniform base pointer, as global memory:
// every data pointer of the library is device memory, and kernels that read
// their pointers from descriptors in memory (the batch engine) would
// otherwise get flat loads and stores, which count against the LDS wait
// counter too and so serialise the tile's LDS exchanges behind its
// outstanding HBM traffic.
// (the host pass of the same source sees plain pointers)
#if defined(__HIP_DEVICE_COMPILE__)
#define FM_GLOBAL __attribute__((address_space(1)))
#else
#define FM_GLOBAL
#endif
template <typename Ix, typename V>
__device__ __forceinline__ V ldg(const V *base, Ix idx) {
    const FM_GLOBAL char *b = (const FM_GLOBAL char *)base;
    if constexpr (sizeof(Ix) == 4)
        return *(const FM_GLOBAL V *)(b + uint32_t(idx * uint32_t(sizeof(V))));
    else
        return ((const FM_GLOBAL V *)b)[idx];
}
template <typename Ix, typename V>
__device__ __forceinline__ void stg(V *base, Ix idx, V val) {
    FM_GLOBAL char *b = (FM_GLOBAL char *)base;
    if constexpr (sizeof(Ix) == 4)
        *(FM_GLOBAL V *)(b + uint32_t(idx * uint32_t(sizeof(V)))) = val;
    else
        ((FM_GLOBAL V *)b)[idx] = val;
}

// Non-temporal store of one complex element (see stg).
template <typename Ix, typename T>
__device__ __forceinline__ void stg_stream(cpx<T> *base, Ix idx, cpx<T> val) {
    typedef T v2 __attribute__((ext_vector_type(2)));
    FM_GLOBAL char *b = (FM_GLOBAL char *)base;
    FM_GLOBAL v2 *q;
    if constexpr (sizeof(Ix) == 4)
        q = (FM_GLOBAL v2 *)(b + uint32_t(idx * uint32_t(sizeof(cpx<T>))));
    else
        q = (FM_GLOBAL v2 *)(b + idx * int64_t(sizeof(cpx<T>)));
    v2 v = {val.x, val.y};
    __builtin_nontemporal_store(v, q);
}

// Non-temporal load of one complex element (see ldg).
template <typename Ix, typename T>
__device__ __forceinline__ cpx<T> ldg_stream(const cpx<T> *base, Ix idx) {
    typedef T v2 __attribute__((ext_vector_type(2)));
    const FM_GLOBAL char *b = (const FM_GLOBAL char *)base;
    const FM_GLOBAL v2 *q;
    if constexpr (sizeof(Ix) == 4)
        q = (const FM_GLOBAL v2 *)(b + uint32_t(idx * uint32_t(sizeof(cpx<T>))));
    else
        q = (const FM_GLOBAL v2 *)(b + idx * int64_t(sizeof(cpx<T>)));
    const v2 v = __builtin_nontemporal_load(q);
    return mk(v.x, v.y);
}

// Bound of a logical index in the launch's index type.
template <typename Ix> __device__ __forceinline__ Ix idx_bound(int64_t v) {
    if constexpr (sizeof(Ix) == 4) return v > int64_t(0xffffffffu) ? Ix(0xffffffffu) : Ix(v);
    else return v;
}

// Sequence (o, j) of tile-local sequence b.
template <typename Ix> struct SeqPos {
    Ix in_base, out_base, j;
    bool ok;
};

template <typename Ix, typename T>
__device__ __forceinline__ SeqPos<Ix> seq_pos(const TileArgs<T> &a, int64_t s) {
    SeqPos<Ix> p;
    p.ok = s < a.O * a.J;
    const Ix si = Ix(s);
    Ix o, j;
    if (a.logJ >= 0) {
        o = si >> a.logJ;
        j = si & ((Ix(1) << a.logJ) - 1);
    } else {
        o = si / Ix(a.J);
        j = si - o * Ix(a.J);
    }
    p.in_base = o * Ix(a.in_os) + j * Ix(a.in_js);
    p.out_base = o * Ix(a.out_os) + j * Ix(a.out_js);
    p.j = j;
    return p;
}

// Map butterfly g of a stage to (sequence b, butterfly index i): sequence-
// fastest for strided global sides (neighbouring sequences are adjacent in
// memory), butterfly-fastest otherwise.
template <int LOGB, int LOGNR>
__device__ __forceinline__ void bfly_pos(int g, bool seqfast, int &b, int &i) {
    if (seqfast) {
        b = g & ((1 << LOGB) - 1);
        i = g >> LOGB;
    } else {
        i = g & ((1 << LOGNR) - 1);
        b = g >> LOGNR;
    }
}

// One radix-2^LOGR stage with sub-length 2^LOGP on registers v (E per thread,
// NB = E / R butterflies).  SRC: 0 = v already holds this stage's inputs,
// 1 = read them from LDS.  DST: 0 = write outputs to LDS, 1 = leave them in v.
template <typename T, int LOGN, int LOGR, int LOGP, int SIGN, int NT, int E, int LOGB, int SRC,
          int DST>
__device__ __forceinline__ void fft_stage2(cpx<T> *v, cpx<T> *lds,
                                           const cpx<double> *__restrict__ tw, bool sf) {
    // sf: mapping of the butterflies held in v when SRC == 0 (see bfly_pos);
    // inputs read from LDS always use the butterfly-fastest mapping.
    if constexpr (SRC == 1) sf = false;
    constexpr int R = 1 << LOGR, NB = E / R;
    constexpr int LOGNR = LOGN - LOGR, NR = 1 << LOGNR;
    constexpr int P = 1 << LOGP;
    constexpr int SB = (1 << LOGN) + ((1 << LOGN) >> 5) + 1;
    if constexpr (SRC == 1) {
#pragma unroll
        for (int e = 0; e < NB; ++e) {
            int b, i;
            bfly_pos<LOGB, LOGNR>(threadIdx.x + e * NT, false, b, i);
#pragma unroll
            for (int r = 0; r < R; ++r) v[e * R + r] = lds[b * SB + lpad(i + r * NR)];
        }
    }
#pragma unroll
    for (int e = 0; e < NB; ++e) {
        int b, i;
        bfly_pos<LOGB, LOGNR>(threadIdx.x + e * NT, sf, b, i);
        if constexpr (LOGP > 0) {
            const int k = i & (P - 1);
            // W_{pR}^{r k} = W_N^{r k N / (p R)}: one FP64 lookup of
            // W_{pR}^k, its powers by FP64 products, rounded once to T
            const int step = k << (LOGN - LOGP - LOGR);
            const cpx<double> w1 = ldg(tw, uint32_t(step));
            cpx<double> w = w1;
#pragma unroll
            for (int r = 1; r < R; ++r) {
                v[e * R + r] = twiddle_mul<SIGN>(v[e * R + r], mk(T(w.x), T(w.y)));
                if (r + 1 < R) w = cmul(w, w1);
            }
        }
        dftr<LOGR, SIGN>(v + e * R);
    }
    if constexpr (DST == 0) {
        __syncthreads();  // every thread has read its inputs of this stage
#pragma unroll
        for (int e = 0; e < NB; ++e) {
            int b, i;
            bfly_pos<LOGB, LOGNR>(threadIdx.x + e * NT, sf, b, i);
            const int k = i & (P - 1);
            const int base = ((i - k) << LOGR) + k;
#pragma unroll
            for (int r = 0; r < R; ++r) lds[b * SB + lpad(base + r * P)] = v[e * R + r];
        }
        __syncthreads();
    }
}

// All stages of one transform.  FIRST_SRC: 0 = v holds stage-0 inputs
// (loaded from global), 1 = they are in LDS in natural order.  LAST_DST:
// 0 = leave natural-order outputs in LDS, 1 = leave the last stage's outputs
// in v (butterfly-fastest mapping).
template <typename T, int LOGN, int LOGE, int SIGN, int NT, int LOGB, int FIRST_SRC, int LAST_DST,
          bool REV = false, int s = 0>
__device__ __forceinline__ void fft_all(cpx<T> *v, cpx<T> *lds,
                                        const cpx<double> *__restrict__ tw, bool sf0) {
    using P = FftPlan<LOGN, LOGE, REV>;
    constexpr int E = 1 << LOGE;
    if constexpr (s < P::S) {
        constexpr int src = s == 0 ? FIRST_SRC : 1;
        constexpr int dst = (s == P::S - 1) ? (LAST_DST == 1 ? 1 : 0) : 0;
        fft_stage2<T, LOGN, P::bits(s), P::logp(s), SIGN, NT, E, LOGB, src, dst>(v, lds, tw, sf0);
        fft_all<T, LOGN, LOGE, SIGN, NT, LOGB, FIRST_SRC, LAST_DST, REV, s + 1>(v, lds, tw, false);
    }
}

// Compile-time shape of one tile of a length-2^LOGN pass with NT threads.
template <typename T, int LOGN, int NT, int LE = -1> struct TileShape {
    // elements per thread: 2^LE, by default 32 (complex64) / 16 (complex128)
    static constexpr int LOGE = LE >= 0 ? LE : (sizeof(T) == 4 ? 5 : 4);
    static constexpr int E = 1 << LOGE;
    static constexpr int N = 1 << LOGN;
    using P = FftPlan<LOGN, LOGE>;
    static constexpr int LOGR0 = P::S > 0 ? P::bits(0) : 0;
    static constexpr int R0 = 1 << LOGR0;
    static constexpr int LOGNT = NT == 64 ? 6 : NT == 128 ? 7 : NT == 256 ? 8 : NT == 512 ? 9 : 10;
    // sequences per tile (NT * E / N, at least one)
    static constexpr int LOGB = LOGNT + LOGE - LOGN;
    static constexpr int B = 1 << LOGB;
    static_assert(LOGB >= 0, "tile must hold at least one sequence");
};

// First sequence of a tile.  Plain order: consecutive tiles step through the
// j-blocks of one o.  With o_group = G > 1 (G divides O, B divides J) the o's
// form groups of G and, inside a group, consecutive tiles step through the G
// o's of one j-block before the next j-block: tiles that share per-j data (a
// spectrum row, chirp or twiddle values) run close together and reuse it
// from cache, while each o still streams through neighbouring j-blocks.
template <typename T, int LOGB>
__device__ __forceinline__ int64_t tile_first_seq(const TileArgs<T> &a, int64_t tile) {
    if (a.o_group <= 1) return tile << LOGB;
    const int64_t G = a.o_group;
    const int64_t per = G * (a.J >> LOGB);  // tiles of one group
    const int64_t grp = tile / per, rem = tile - grp * per;
    const int64_t jb = rem / G, o = grp * G + (rem - jb * G);
    return o * a.J + (jb << LOGB);
}

// Stage-0 inputs of one tile straight from global memory (zero outside the
// valid range); only loads, so a caller can issue them ahead of compute.
template <typename T, int LOGN, int NT, bool I32 = false, int LE = -1>
__device__ __forceinline__ void fft_tile_load(const TileArgs<T> &a, int64_t tile, cpx<T> *v) {
    using S = TileShape<T, LOGN, NT, LE>;
    using Ix = IdxT<I32>;
    constexpr int NB = S::E / S::R0;
    constexpr int LOGNR = LOGN - S::LOGR0, NR = 1 << LOGNR;
    const int64_t s0 = tile_first_seq<T, S::LOGB>(a, tile);
    const Ix nv = idx_bound<Ix>(a.n_valid);
    const Ix in_gn = Ix(a.in_gn), in_gj = Ix(a.in_gj), in_ns = Ix(a.in_ns);
    // one loop per load kind (uniform), so that all the tile's loads issue
    // before the first use
    auto body = [&](auto load) {
#pragma unroll
        for (int e = 0; e < NB; ++e) {
            int b, i;
            bfly_pos<S::LOGB, LOGNR>(threadIdx.x + e * NT, a.seqfast_in, b, i);
            const SeqPos<Ix> sp = seq_pos<Ix>(a, s0 + b);
#pragma unroll
            for (int r = 0; r < S::R0; ++r) {
                const Ix n = Ix(i + r * NR);
                const Ix gidx = n * in_gn + sp.j * in_gj;
                const bool ok = sp.ok && gidx < nv;
                v[e * S::R0 + r] = ok ? load(Ix(sp.in_base + n * in_ns)) : mk(T(0), T(0));
            }
        }
    };
    if (a.stream_out & 2)
        body([&](Ix k) { return ldg_stream(a.x, k); });
    else
        body([&](Ix k) { return ldg(a.x, k); });
}

// The tile's sum of the threads' res2 (sums of |u - t|^2) into
// a.res_part[a.res_base + tile]: shuffle within each wave, one value per
// wave through LDS, summed in wave order by thread 0 (deterministic).  Every
// thread of the tile calls it; it waits for the tile's earlier LDS reads.
template <typename T, int NT>
__device__ __forceinline__ void tile_residual_sum(const TileArgs<T> &a, int64_t tile, double res2,
                                                  cpx<T> *lds) {
    double sum = res2;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) sum += __shfl_down(sum, off, 64);
    __syncthreads();  // the tile's LDS data is no longer read
    double *red = reinterpret_cast<double *>(lds);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = sum;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int q = 0; q < NT / 64; ++q) t += red[q];
        stg(a.res_part, int64_t(a.res_base + tile), t);
    }
    __syncthreads();  // red is read before the LDS is written again
}

// Store the last stage's outputs of one tile: v[e*RL + r] holds element
// i + r*N/RL (RL = 2^LOGRL, the radix of the stage that produced them) of
// sequence b in mapping cur_sf; re-stage through LDS when the store wants the
// other mapping, then apply the enclosing four-step twiddle / post-multiply /
// conj / scale and write, cropping at m_valid.
template <typename T, int LOGN, int NT, int LOGRL, bool I32 = false, int LE = -1>
__device__ __forceinline__ void fft_tile_store(const TileArgs<T> &a, int64_t tile, cpx<T> *v,
                                               cpx<T> *lds, bool cur_sf) {
    using S = TileShape<T, LOGN, NT, LE>;
    using Ix = IdxT<I32>;
    constexpr int E = S::E;
    constexpr int N = S::N;
    constexpr int RL = 1 << LOGRL;
    constexpr int LOGB = S::LOGB;
    constexpr int B = S::B;
    constexpr int SB = N + (N >> 5) + 1;
    const int64_t s0 = tile_first_seq<T, LOGB>(a, tile);
    constexpr int NBL = E / RL;
    constexpr int LOGNRL = LOGN - LOGRL, NRL = 1 << LOGNRL;
    const bool want_sf = a.seqfast_out && B > 1;
    if (want_sf != cur_sf) {
        __syncthreads();
#pragma unroll
        for (int e = 0; e < NBL; ++e) {
            int b, i;
            bfly_pos<LOGB, LOGNRL>(threadIdx.x + e * NT, cur_sf, b, i);
#pragma unroll
            for (int r = 0; r < RL; ++r) lds[b * SB + lpad(i + r * NRL)] = v[e * RL + r];
        }
        __syncthreads();
#pragma unroll
        for (int e = 0; e < NBL; ++e) {
            int b, i;
            bfly_pos<LOGB, LOGNRL>(threadIdx.x + e * NT, want_sf, b, i);
#pragma unroll
            for (int r = 0; r < RL; ++r) v[e * RL + r] = lds[b * SB + lpad(i + r * NRL)];
        }
        cur_sf = want_sf;
    }
    const Ix lo_mask = (Ix(1) << a.twb_shift) - 1;
    const Ix mv = idx_bound<Ix>(a.m_valid);
    const Ix out_gn = Ix(a.out_gn), out_gj = Ix(a.out_gj), out_ns = Ix(a.out_ns);
    double res2 = 0.0;  // residual epilogue: sum of |u - t|^2 of this thread
#pragma unroll
    for (int e = 0; e < NBL; ++e) {
        int b, i;
        bfly_pos<LOGB, LOGNRL>(threadIdx.x + e * NT, cur_sf, b, i);
        const SeqPos<Ix> sp = seq_pos<Ix>(a, s0 + b);
        if (!sp.ok) continue;
        cpx<T> w[RL];
        if (a.twb_hi || a.post) {
            // four-step twiddle W_L^{j n} for n = i + r NRL: the FP64 base
            // W_L^{j i} times successive FP64 powers of W_L^{j NRL} (two
            // table lookups each, j n < L), rounded once to T
            cpx<double> tw = mk(1.0, 0.0), step = mk(1.0, 0.0);
            if (a.twb_hi) {
                const Ix e0 = sp.j * Ix(i), e1 = sp.j * Ix(NRL);
                tw = cmul(ldg(a.twb_hi, Ix(e0 >> a.twb_shift)), ldg(a.twb_lo, Ix(e0 & lo_mask)));
                step = cmul(ldg(a.twb_hi, Ix(e1 >> a.twb_shift)), ldg(a.twb_lo, Ix(e1 & lo_mask)));
            }
#pragma unroll
            for (int r = 0; r < RL; ++r) {
                const Ix n = Ix(i + r * NRL);
                cpx<T> t = mk(T(1), T(0));
                if (a.twb_hi) {
                    t = mk(T(tw.x), T(a.twb_sign > 0 ? -tw.y : tw.y));
                    if (r + 1 < RL) tw = cmul(tw, step);
                }
                const Ix gidx = n * out_gn + sp.j * out_gj;
                if (a.post && gidx < mv) t = cmul(t, ldg(a.post, gidx));
                w[r] = t;
            }
        }
#pragma unroll
        for (int r = 0; r < RL; ++r) {
            const Ix n = Ix(i + r * NRL);
            const Ix gidx = n * out_gn + sp.j * out_gj;
            if (gidx >= mv) continue;
            cpx<T> u = v[e * RL + r];
            if (a.twb_hi || a.post) u = cmul(u, w[r]);
            if (a.conj_out) u = conj(u);
            if (a.target) {
                const cpx<T> d = u - ldg(a.target, Ix(sp.out_base + n * out_ns));
                res2 += double(d.x) * double(d.x) + double(d.y) * double(d.y);
                u = d;
            }
            if (a.stream_out & 1)
                stg_stream(a.y, Ix(sp.out_base + n * out_ns), scale(u, a.out_scale));
            else
                stg(a.y, Ix(sp.out_base + n * out_ns), scale(u, a.out_scale));
        }
    }
    if (a.target) tile_residual_sum<T, NT>(a, tile, res2, lds);
}

// Everything after the stage-0 loads of one tile: conj / pre-multiply, the
// transform (conv: forward, spectrum, inverse), twiddle / post-multiply and
// the stores.  `lds` holds B * (N + N/32 + 1) elements.
template <typename T, int LOGN, int NT, bool CONV, bool I32 = false, int LE = -1>
__device__ __forceinline__ void fft_tile_compute(const TileArgs<T> &a, int64_t tile, cpx<T> *v,
                                                 cpx<T> *lds) {
    using S = TileShape<T, LOGN, NT, LE>;
    using Ix = IdxT<I32>;
    constexpr int LOGE = S::LOGE;
    constexpr int E = S::E;
    constexpr int N = S::N;
    using P = typename S::P;
    constexpr int LOGR0 = S::LOGR0;
    // radix of the stage that produces the stored outputs (the inverse of a
    // conv runs the radices reversed, so it ends with bits(0))
    constexpr int LOGRL = P::S > 0 ? (CONV ? P::bits(0) : P::bits(P::S - 1)) : 0;
    constexpr int R0 = 1 << LOGR0, RL = 1 << LOGRL;
    constexpr int LOGB = S::LOGB;
    constexpr int B = S::B;
    constexpr int SB = N + (N >> 5) + 1;
    const int64_t s0 = tile_first_seq<T, LOGB>(a, tile);

    // ---- conj, pre-multiply (zero-padded positions hold zeros already)
    if (a.conj_in) {
#pragma unroll
        for (int q = 0; q < E; ++q) v[q] = conj(v[q]);
    }
    if (a.pre) {
        constexpr int NB = E / R0;
        constexpr int LOGNR = LOGN - LOGR0, NR = 1 << LOGNR;
#pragma unroll
        for (int e = 0; e < NB; ++e) {
            int b, i;
            bfly_pos<LOGB, LOGNR>(threadIdx.x + e * NT, a.seqfast_in, b, i);
            const SeqPos<Ix> sp = seq_pos<Ix>(a, s0 + b);
#pragma unroll
            for (int r = 0; r < R0; ++r) {
                const Ix n = Ix(i + r * NR);
                const Ix gidx = n * Ix(a.in_gn) + sp.j * Ix(a.in_gj);
                const bool ok = sp.ok && gidx < idx_bound<Ix>(a.n_valid);
                v[e * R0 + r] = cmul(v[e * R0 + r], ok ? ldg(a.pre, gidx) : mk(T(0), T(0)));
            }
        }
    }
    // mapping of the butterflies currently held in v (see bfly_pos)
    bool cur_sf = a.seqfast_in && B > 1;
    if constexpr (P::S == 0) {
        // N == 1: identity transform (conv: multiply by the single spectrum value)
        if constexpr (CONV) {
#pragma unroll
            for (int e = 0; e < E; ++e) {
                int b, i;
                bfly_pos<LOGB, 0>(threadIdx.x + e * NT, cur_sf, b, i);
                const SeqPos<Ix> sp = seq_pos<Ix>(a, s0 + b);
                if (sp.ok) {
                    const cpx<T> d = ldg(a.D, Ix(sp.j * Ix(a.d_gj)));
                    v[e] = a.conj_d ? cmulc(v[e], d) : cmul(v[e], d);
                }
            }
        }
    } else if constexpr (!CONV) {
        if (a.sign1 < 0) fft_all<T, LOGN, LOGE, -1, NT, LOGB, 0, 1>(v, lds, a.tw, cur_sf);
        else fft_all<T, LOGN, LOGE, +1, NT, LOGB, 0, 1>(v, lds, a.tw, cur_sf);
        if (P::S > 1) cur_sf = false;
    } else {
        // forward transform ends in registers (last radix RL); multiply by the
        // spectrum there and run the inverse with the radices reversed, so its
        // first stage (radix RL, p = 1) consumes exactly those registers
        if (a.sign1 < 0) fft_all<T, LOGN, LOGE, -1, NT, LOGB, 0, 1>(v, lds, a.tw, cur_sf);
        else fft_all<T, LOGN, LOGE, +1, NT, LOGB, 0, 1>(v, lds, a.tw, cur_sf);
        if (P::S > 1) cur_sf = false;
        {
            constexpr int RF = 1 << P::bits(P::S - 1);
            constexpr int NBF = E / RF;
            constexpr int LOGNRF = LOGN - P::bits(P::S - 1), NRF = 1 << LOGNRF;
            if (a.mid_t) {
                // residual in place of the spectrum (see TileArgs::mid_t)
                const Ix mv = idx_bound<Ix>(a.mid_valid);
                double res2 = 0.0;
#pragma unroll
                for (int e = 0; e < NBF; ++e) {
                    int b, i;
                    bfly_pos<LOGB, LOGNRF>(threadIdx.x + e * NT, cur_sf, b, i);
                    const SeqPos<Ix> sp = seq_pos<Ix>(a, s0 + b);
                    if (!sp.ok) continue;
                    const Ix si = Ix(s0 + b);
                    const Ix o = a.logJ >= 0 ? Ix(si >> a.logJ) : Ix(si / Ix(a.J));
                    const Ix tb = o * Ix(a.mid_os);
#pragma unroll
                    for (int r = 0; r < RF; ++r) {
                        const Ix n = Ix(i + r * NRF);
                        const Ix g = n * Ix(a.mid_gn) + sp.j * Ix(a.mid_gj);
                        cpx<T> u = mk(T(0), T(0));
                        if (g < mv) {
                            const cpx<T> d = v[e * RF + r] - ldg(a.mid_t, Ix(tb + g));
                            res2 += double(d.x) * double(d.x) + double(d.y) * double(d.y);
                            u = scale(d, a.mid_scale);
                        }
                        v[e * RF + r] = u;
                    }
                }
                tile_residual_sum<T, NT>(a, tile, res2, lds);
            } else {
#pragma unroll
                for (int e = 0; e < NBF; ++e) {
                    int b, i;
                    bfly_pos<LOGB, LOGNRF>(threadIdx.x + e * NT, cur_sf, b, i);
                    const SeqPos<Ix> sp = seq_pos<Ix>(a, s0 + b);
                    if (!sp.ok) continue;
#pragma unroll
                    for (int r = 0; r < RF; ++r) {
                        const Ix n = Ix(i + r * NRF);
                        const cpx<T> d = ldg(a.D, Ix(n * Ix(a.d_gn) + sp.j * Ix(a.d_gj)));
                        v[e * RF + r] =
                            a.conj_d ? cmulc(v[e * RF + r], d) : cmul(v[e * RF + r], d);
                    }
                }
            }
        }
        if (a.sign1 < 0) fft_all<T, LOGN, LOGE, +1, NT, LOGB, 0, 1, true>(v, lds, a.tw, cur_sf);
        else fft_all<T, LOGN, LOGE, -1, NT, LOGB, 0, 1, true>(v, lds, a.tw, cur_sf);
        if (P::S > 1) cur_sf = false;
    }

    fft_tile_store<T, LOGN, NT, LOGRL, I32, LE>(a, tile, v, lds, cur_sf);
}

template <typename T, int LOGN, int NT, bool CONV, bool I32 = false, int LE = -1>
__device__ __forceinline__ void fft_pass_body(const TileArgs<T> &a, int64_t tile, cpx<T> *lds) {
    cpx<T> v[TileShape<T, LOGN, NT, LE>::E];
    fft_tile_load<T, LOGN, NT, I32, LE>(a, tile, v);
    fft_tile_compute<T, LOGN, NT, CONV, I32, LE>(a, tile, v, lds);
}

template <typename T, int LOGN, int NT, bool CONV, bool I32, int LE = -1>
__global__ __launch_bounds__(NT) void fft_pass_kernel(const TileArgs<T> a) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    fft_pass_body<T, LOGN, NT, CONV, I32, LE>(a, xcd_remap(blockIdx.x, gridDim.x),
                                              reinterpret_cast<cpx<T> *>(smem_raw));
}

// Host-side launcher for one tile pass (defined in fm_fft.hip).
template <typename T>
fm_status launch_tile(const TileArgs<T> &a, bool conv, hipStream_t stream);

// Largest tile (B*N elements) the engine supports for one precision.
template <typename T> constexpr int max_tile_elems() {
    return sizeof(T) == 4 ? 16384 : 8192;
}

// Launch shape of a length-2^LOGN pass.  A thread holds 32 complex64 (16
// complex128) elements by default.  Complex64 passes of 2-256 and of
// 2048-8192 points hold 16 with twice the threads instead: the tile (and its
// LDS) is unchanged, registers drop to ~90-125 per thread and a CU runs twice
// the waves.  Up to 256 points the stage count is the same either way; at
// 2048-8192 the extra waves pay for the extra stage; at 512 and 1024 points
// the 32-element shape's one LDS exchange wins (measured on the uneven batch,
// the Kron modes and the 2048-point Bluestein spectrum pass).
template <typename T, int LOGN> struct PassCfg {
    // (16384 points would need 1024 threads at <= 128 registers: it spills)
    static constexpr bool e16 =
        sizeof(T) == 4 && ((LOGN >= 11 && LOGN <= 13) || (LOGN >= 1 && LOGN <= 8));
    static constexpr int LE = e16 ? 4 : -1;
    static constexpr int NT =
        e16 ? 512 : ((sizeof(T) == 4 ? LOGN >= 14 : LOGN >= 13) ? 512 : 256);
};

// Sequences per tile of the pass kernel for a length-2^logN transform.
template <typename T> inline int tile_batch(int logN) {
    const int loge = sizeof(T) == 4 ? 5 : 4;
    const int lognt = (sizeof(T) == 4 ? logN >= 14 : logN >= 13) ? 9 : 8;
    const int logb = lognt + loge - logN;
    return logb > 0 ? (1 << logb) : 1;
}

}  // namespace fm
