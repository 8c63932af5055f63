// Plans and pipelines of the FFT engine, shared by the single-operator
// entry points (fm_fft.hip), the uneven batch (fm_batch.hip) and the
// generator-to-spectrum programs: the geometry of a length-L transform
// (one tile pass or a four-step), the passes a transform or convolution
// runs as, and the convolution plan itself.
#pragma once

#include <vector>

#include "fm_fft_core.hpp"

namespace fm {

// out[q] = s * in[perm(q)]; with logN2 > 0 the output is in the order the
// four-step's row pass reads it: q = k1 * N2 + k2 holds bin k1 + N1 * k2.
template <typename T>
__global__ void scale_cast_kernel(cpx<T> *out, const cpx<double> *in, int64_t n, double s,
                                  int logN1, int logN2) {
    const int64_t q = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
    if (q >= n) return;
    int64_t src = q;
    if (logN2 > 0) {
        const int64_t k1 = q >> logN2, k2 = q & ((int64_t(1) << logN2) - 1);
        src = k1 + (k2 << logN1);
    }
    out[q] = mk(T(in[src].x * s), T(in[src].y * s));
}

inline unsigned grid1d(int64_t n, int bs = 256) { return unsigned(ceil_div(n, bs)); }

inline size_t esize(int prec) { return prec == FM_COMPLEX64 ? 8 : 16; }

// Tile-shape knob: environment override (used by the tuning sweeps), else
// the built-in default.
int knob(const char *name, int dflt);

// ------------------------------------------------------------ tile launch
// Whether every element offset (x, y) and logical index (zero-pad / crop
// bounds, pre / post / spectrum tables, twiddle exponents) of a pass stays
// below 2^32 / sizeof(element), so the kernel may run 32-bit index math.
template <typename T> inline bool fits_idx32(const TileArgs<T> &a, int64_t N) {
    const int64_t lim = int64_t((uint64_t(1) << 32) / sizeof(cpx<T>));
    const int64_t O1 = a.O - 1, J1 = a.J - 1, N1 = N - 1;
    const int64_t st[] = {a.in_os, a.in_js, a.in_ns, a.out_os, a.out_js, a.out_ns, a.in_gn,
                          a.in_gj, a.out_gn, a.out_gj, a.d_gn, a.d_gj};
    for (int64_t v : st)
        if (v < 0 || v >= lim) return false;
    auto below = [lim](double v) { return v < double(lim); };
    return below(double(a.O) * double(a.J) + 512.0) &&
           below(double(O1) * a.in_os + double(J1) * a.in_js + double(N1) * a.in_ns) &&
           below(double(O1) * a.out_os + double(J1) * a.out_js + double(N1) * a.out_ns) &&
           below(double(N1) * a.in_gn + double(J1) * a.in_gj) &&
           below(double(N1) * a.out_gn + double(J1) * a.out_gj) &&
           below(double(N1) * a.d_gn + double(J1) * a.d_gj) && below(double(a.J) * double(N)) &&
           (a.mid_t == nullptr ||
            (a.mid_os >= 0 && a.mid_gn >= 0 && a.mid_gj >= 0 &&
             below(double(O1) * a.mid_os + double(N1) * a.mid_gn + double(J1) * a.mid_gj)));
}

// Fill in the tile shape (batch, tile count, log2 J) of a pass.
template <typename T> inline fm_status normalize_pass(TileArgs<T> &a) {
    const int maxlog = sizeof(T) == 4 ? 14 : 13;
    FM_REQUIRE(a.logN >= 0 && a.logN <= maxlog, FM_ERR_VALUE,
               "sub-transform length 2^%d exceeds the tile limit", a.logN);
    a.B = tile_batch<T>(a.logN);
    a.ntiles = ceil_div(a.O * a.J, a.B);
    a.logJ = is_pow2(a.J) ? ilog2(a.J) : -1;
    if (a.o_group > 1 && ((a.J % a.B) != 0 || (a.O % a.o_group) != 0)) a.o_group = 1;
    return FM_SUCCESS;
}

template <typename T> inline TileArgs<T> blank_args() {
    TileArgs<T> a;
    memset(&a, 0, sizeof(a));
    a.out_scale = T(1);
    a.sign1 = -1;
    a.twb_shift = 0;
    a.n_valid = INT64_MAX;
    a.m_valid = INT64_MAX;
    return a;
}

inline int pick_batch(int logN, int64_t nseq, int max_tile, int want) {
    int B = 1;
    while (B < want && (int64_t(B) * 2 << logN) <= max_tile && B * 2 <= nseq) B *= 2;
    while ((int64_t(B) << logN) > max_tile && B > 1) B /= 2;
    return B;
}

// ------------------------------------------------------------- transform
// One length-L transform (L = 2^logL) over k columns, as a single tile pass
// or as a four-step; optional Bluestein chirps and spectrum.
struct Geometry {
    int prec;
    int logL;
    int passes;  // 1 (single tile) or multi (2 plain / 3 conv)
    int logN1, logN2;
    const void *tw_full, *tw1, *tw2, *twb_hi, *twb_lo;
    int twb_shift;
};

fm_status make_geometry(int prec, int logL, bool conv, hipStream_t stream, Geometry *g);

struct PipeIO {
    const void *x;
    int64_t ldx;
    void *y;
    int64_t ldy;
    int64_t k;
    int64_t n_in;   // valid input length (zero-pad beyond)
    int64_t m_out;  // stored output length (crop)
    const void *pre, *post;  // chirps (Bluestein) or null
    const void *D;           // spectrum (conv) or null
    int conj_d;
    int conj_io;             // conj on load and on store (adjoint DFT)
    void *work;              // L*k elements for multi-pass
};

template <typename T> struct PassSpec {
    TileArgs<T> a;
    bool conv;
};

// The passes of one transform (at most three), in a fixed-capacity list: the
// apply paths build their pipelines without touching the heap.
template <typename T> struct PassList {
    static constexpr int kMax = 4;
    PassSpec<T> v[kMax];
    int n = 0;
    void push_back(const PassSpec<T> &p) { v[n++] = p; }
    int size() const { return n; }
    PassSpec<T> &operator[](int i) { return v[i]; }
    const PassSpec<T> &operator[](int i) const { return v[i]; }
    PassSpec<T> &back() { return v[n - 1]; }
    void pop_front() {
        for (int i = 1; i < n; ++i) v[i - 1] = v[i];
        --n;
    }
    PassSpec<T> *begin() { return v; }
    PassSpec<T> *end() { return v + n; }
    const PassSpec<T> *begin() const { return v; }
    const PassSpec<T> *end() const { return v + n; }
};

template <typename T>
inline fm_status build_pipeline_raw(const Geometry &g, const PipeIO &io,
                                PassList<T> *out) {
    const bool conv = io.D != nullptr;
    const int64_t L = int64_t(1) << g.logL;
    const int max_tile = max_tile_elems<T>();
    if (g.passes == 1) {
        TileArgs<T> a = blank_args<T>();
        a.logN = g.logL;
        a.O = io.k;
        a.J = 1;
        a.B = pick_batch(g.logL, io.k, 4096, 64);
        a.ntiles = ceil_div(io.k, a.B);
        a.x = (const cpx<T> *)io.x;
        a.in_os = io.ldx; a.in_js = 0; a.in_ns = 1;
        a.in_gn = 1; a.in_gj = 0; a.n_valid = io.n_in;
        a.pre = (const cpx<T> *)io.pre;
        a.conj_in = io.conj_io;
        a.tw = (const cpx<double> *)g.tw_full;
        a.D = (const cpx<T> *)io.D;
        a.d_gn = 1; a.d_gj = 0; a.conj_d = io.conj_d;
        a.y = (cpx<T> *)io.y;
        a.out_os = io.ldy; a.out_js = 0; a.out_ns = 1;
        a.out_gn = 1; a.out_gj = 0; a.m_valid = io.m_out;
        a.post = (const cpx<T> *)io.post;
        a.conj_out = io.conj_io;
        out->push_back(PassSpec<T>{a, conv});
        return FM_SUCCESS;
    }
    FM_REQUIRE(io.work != nullptr, FM_ERR_VALUE, "multi-pass transform needs a workspace");
    // a.B below only sets the vector-fast addressing flags; the sequences
    // per tile are fixed by the pass kernel's tile shape (normalize_pass)
    const int64_t N1 = int64_t(1) << g.logN1, N2 = int64_t(1) << g.logN2;
    cpx<T> *tmp = (cpx<T> *)io.work;
    // pass A: strided N1-point FFTs over n1 for every n2, twiddle W_L^{n2 k1}
    {
        TileArgs<T> a = blank_args<T>();
        a.logN = g.logN1;
        a.O = io.k; a.J = N2;
        a.B = pick_batch(g.logN1, N2, max_tile, 16);
        a.ntiles = ceil_div(io.k * N2, a.B);
        a.x = (const cpx<T> *)io.x;
        a.in_os = io.ldx; a.in_js = 1; a.in_ns = N2;
        a.in_gn = N2; a.in_gj = 1; a.n_valid = io.n_in;
        a.pre = (const cpx<T> *)io.pre;
        a.conj_in = io.conj_io;
        a.seqfast_in = a.B > 1;
        a.tw = (const cpx<double> *)g.tw1;
        a.twb_hi = (const cpx<double> *)g.twb_hi; a.twb_lo = (const cpx<double> *)g.twb_lo;
        a.twb_shift = g.twb_shift; a.twb_sign = -1;
        a.y = tmp;
        a.out_os = L; a.out_js = 1; a.out_ns = N2;
        a.seqfast_out = a.B > 1;
        out->push_back(PassSpec<T>{a, false});
    }
    if (!conv) {
        // pass B: N2-point FFTs along rows k1, transposed store y[k1 + N1 k2]
        TileArgs<T> a = blank_args<T>();
        a.logN = g.logN2;
        a.O = io.k; a.J = N1;
        a.B = pick_batch(g.logN2, N1, max_tile, 16);
        a.ntiles = ceil_div(io.k * N1, a.B);
        a.x = tmp;
        a.in_os = L; a.in_js = N2; a.in_ns = 1;
        a.tw = (const cpx<double> *)g.tw2;
        a.y = (cpx<T> *)io.y;
        a.out_os = io.ldy; a.out_js = 1; a.out_ns = N1;
        a.out_gn = N1; a.out_gj = 1; a.m_valid = io.m_out;
        a.post = (const cpx<T> *)io.post;
        a.conj_out = io.conj_io;
        a.seqfast_out = a.B > 1;
        out->push_back(PassSpec<T>{a, false});
        return FM_SUCCESS;
    }
    // pass B (conv): rows k1: FFT, spectrum D[k1 + N1 k2], inverse FFT,
    // twiddle W_L^{-n2 k1}; in place in the workspace
    {
        TileArgs<T> a = blank_args<T>();
        a.logN = g.logN2;
        a.O = io.k; a.J = N1;
        a.B = pick_batch(g.logN2, N1, max_tile, 2);
        a.ntiles = ceil_div(io.k * N1, a.B);
        a.x = tmp;
        a.in_os = L; a.in_js = N2; a.in_ns = 1;
        a.tw = (const cpx<double> *)g.tw2;
        a.D = (const cpx<T> *)io.D;
        a.d_gn = 1; a.d_gj = N2; a.conj_d = io.conj_d;  // spectrum in row-pass order
        a.o_group = io.k;  // all columns of a row block reuse its spectrum rows
        a.twb_hi = (const cpx<double> *)g.twb_hi; a.twb_lo = (const cpx<double> *)g.twb_lo;
        a.twb_shift = g.twb_shift; a.twb_sign = +1;
        a.y = tmp;
        a.out_os = L; a.out_js = N2; a.out_ns = 1;
        out->push_back(PassSpec<T>{a, true});
    }
    // pass C: strided inverse N1-point FFTs over k1 -> y[n2 + N2 n1], crop
    {
        TileArgs<T> a = blank_args<T>();
        a.logN = g.logN1;
        a.O = io.k; a.J = N2;
        a.B = pick_batch(g.logN1, N2, max_tile, 16);
        a.ntiles = ceil_div(io.k * N2, a.B);
        a.x = tmp;
        a.in_os = L; a.in_js = 1; a.in_ns = N2;
        a.seqfast_in = a.B > 1;
        a.sign1 = +1;
        a.tw = (const cpx<double> *)g.tw1;
        a.y = (cpx<T> *)io.y;
        a.out_os = io.ldy; a.out_js = 1; a.out_ns = N2;
        a.out_gn = N2; a.out_gj = 1; a.m_valid = io.m_out;
        a.post = (const cpx<T> *)io.post;
        a.conj_out = io.conj_io;
        a.seqfast_out = a.B > 1;
        out->push_back(PassSpec<T>{a, false});
        return FM_SUCCESS;
    }
}


// The passes of one transform (see build_pipeline_raw), each tagged with its
// share of the transform's algorithmic bytes: input and output blocks
// (n_in + m_out) x k elements, split evenly over the passes.
template <typename T>
inline fm_status build_pipeline(const Geometry &g, const PipeIO &io, PassList<T> *out) {
    const int first = out->size();
    fm_status st = build_pipeline_raw<T>(g, io, out);
    if (st) return st;
    const int np = out->size() - first;
    const double alg = double(io.n_in + io.m_out) * double(io.k) * double(sizeof(cpx<T>));
    for (int q = first; q < out->size(); ++q) (*out)[q].a.alg_bytes = alg / np;
    return FM_SUCCESS;
}


// A second default-priority stream on which a plan runs every other column
// chunk of a multi-pass transform, forked from and joined back into the
// caller's stream with events (so the apply stays stream-ordered and
// capturable).  With chunks sized for the memory-side cache, one chunk's
// passes overlap the next chunk's while each chunk's intermediate is still
// cached.  Used when FASTMAT_AMD_CHUNK_LANES is 2 and the columns are chunked.
struct Lanes {
    hipStream_t side = nullptr;
    hipEvent_t fork = nullptr, join = nullptr;
    bool ok = false;
};
void lanes_create(Lanes *l);
void lanes_destroy(Lanes *l);
// Lanes a chunked transform uses (1 or 2).
int chunk_lanes(const Lanes *l);
// Columns per chunk of a multi-pass transform (lanes: the plan's second
// stream, or NULL).
int64_t chunk_columns(const Geometry &g, int64_t k, size_t esz, const Lanes *l);

template <typename T>
inline fm_status run_pipeline(const Geometry &g, const PipeIO &io, hipStream_t stream,
                              const Lanes *lanes = nullptr) {
    const int64_t cc = chunk_columns(g, io.k, sizeof(cpx<T>), lanes);
    const int nl = cc < io.k ? chunk_lanes(lanes) : 1;
    const int64_t lane_bytes = (int64_t(1) << g.logL) * cc * int64_t(sizeof(cpx<T>));
    if (nl > 1) {
        FM_CHECK_HIP(hipEventRecord(lanes->fork, stream));
        FM_CHECK_HIP(hipStreamWaitEvent(lanes->side, lanes->fork, 0));
    }
    int64_t q = 0;
    for (int64_t c0 = 0; c0 < io.k; c0 += cc, ++q) {
        const int lane = int(q % nl);
        hipStream_t s = lane ? lanes->side : stream;
        PipeIO sub = io;
        sub.k = io.k - c0 < cc ? io.k - c0 : cc;
        sub.x = (const cpx<T> *)io.x + c0 * io.ldx;
        sub.y = (cpx<T> *)io.y + c0 * io.ldy;
        if (io.work) sub.work = (char *)io.work + lane * lane_bytes;
        PassList<T> passes;
        fm_status st = build_pipeline<T>(g, sub, &passes);
        for (int p = 0; !st && p < passes.size(); ++p)
            st = launch_tile<T>(passes[p].a, passes[p].conv, s);
        if (st) return st;
    }
    if (nl > 1) {
        FM_CHECK_HIP(hipEventRecord(lanes->join, lanes->side));
        FM_CHECK_HIP(hipStreamWaitEvent(stream, lanes->join, 0));
    }
    return FM_SUCCESS;
}

inline fm_status run_pipeline_any(const Geometry &g, const PipeIO &io, hipStream_t stream,
                                  const Lanes *lanes = nullptr) {
    return g.prec == FM_COMPLEX64 ? run_pipeline<float>(g, io, stream, lanes)
                                  : run_pipeline<double>(g, io, stream, lanes);
}

// Spectrum D = fft(t')/L of a complex128 generator vector, stored in `prec`
// (natural order for one-pass geometries, the row pass's order otherwise).
fm_status spectrum_from(cpx<double> *t, const Geometry &g, int prec, void *D,
                        hipStream_t stream);

}  // namespace fm

struct fm_conv_plan {
    int device;
    int prec;
    int kind;          // 0 circulant, 1 toeplitz, 2 two-level circulant
    int64_t rows, cols, gen_len, L;
    // embedding of the generator into the length-L circulant:
    // t' = [gen[0:head], 0, ..., 0, gen[tail_src:tail_src+tail]]
    // (Circulant: head = n, tail = n-1 or 0, tail_src = 1;
    //  Toeplitz: head = m, tail = n-1, tail_src = m)
    int64_t head, tail, tail_src;
    fm::Geometry geo;      // conv pipeline (1 or 3 passes)
    fm::Geometry geo_fft;  // plain length-L FFT (1 or 2 passes): generator
                           // gradient and generator -> spectrum
    void *D;           // fft(t')/L in prec, L entries (row-pass order for 3 passes)
};

namespace fm {

// Generator element as complex128 (real generators have zero imaginary part).
template <typename G> __device__ __forceinline__ cpx<double> gen_elem(const G *g, int64_t i);
template <> __device__ __forceinline__ cpx<double> gen_elem<float>(const float *g, int64_t i) {
    return mk(double(ldg(g, i)), 0.0);
}
template <> __device__ __forceinline__ cpx<double> gen_elem<double>(const double *g, int64_t i) {
    return mk(ldg(g, i), 0.0);
}
template <>
__device__ __forceinline__ cpx<double> gen_elem<cpx<float>>(const cpx<float> *g, int64_t i) {
    const cpx<float> v = ldg(g, i);
    return mk(double(v.x), double(v.y));
}
template <>
__device__ __forceinline__ cpx<double> gen_elem<cpx<double>>(const cpx<double> *g, int64_t i) {
    return ldg(g, i);
}

// Embedded generator t'[i] (see fm_conv_plan) for i < L.
template <typename G>
__device__ __forceinline__ cpx<double> embedded_at(const G *gen, int64_t i, int64_t L, int64_t head,
                                                   int64_t tail, int64_t tail_src) {
    if (i < head) return gen_elem(gen, i);
    if (i >= L - tail) return gen_elem(gen, tail_src + (i - (L - tail)));
    return mk(0.0, 0.0);
}

// t'[i] for i < L, written in precision T.
template <typename G, typename T>
__global__ void embed_gen_kernel(cpx<T> *t, const G *gen, int64_t L, int64_t head, int64_t tail,
                                 int64_t tail_src) {
    const int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
    if (i >= L) return;
    const cpx<double> v = embedded_at(gen, i, L, head, tail, tail_src);
    t[i] = mk(T(v.x), T(v.y));
}

// The passes that turn an embedded generator t (L elements of T, read
// only) into the plan's spectrum D = fft(t)/L in the order the convolution
// pipeline reads it: natural order for a one-pass geometry; for a four-step
// the row pass stores untransposed (q = k1 N2 + k2 holds bin k1 + N1 k2),
// which is the convolution row pass's order.  `tmp` holds L elements for a
// four-step.  The 1/L scale is applied by the last store.
template <typename T>
inline fm_status build_spectrum_pipeline(const Geometry &g, const cpx<T> *t, cpx<T> *tmp,
                                         cpx<T> *D, PassList<T> *out) {
    const int64_t L = int64_t(1) << g.logL;
    PipeIO io;
    memset(&io, 0, sizeof(io));
    io.x = t; io.ldx = L; io.y = D; io.ldy = L; io.k = 1; io.n_in = L; io.m_out = L;
    io.work = tmp;
    fm_status st = build_pipeline<T>(g, io, out);
    if (st) return st;
    PassSpec<T> &last = out->back();
    last.a.out_scale = T(1.0 / double(L));
    if (g.passes > 1) {
        const int64_t N2 = int64_t(1) << g.logN2;
        last.a.out_js = N2;   // row k1 ...
        last.a.out_ns = 1;    // ... stored contiguously: q = k1 N2 + k2
        last.a.out_gn = 1;
        last.a.out_gj = N2;
        last.a.seqfast_out = 0;
    }
    return FM_SUCCESS;
}

}  // namespace fm
