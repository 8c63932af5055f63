// Fourier, Circulant and Toeplitz operators on the FFT tile engine.
//
// Reference behaviour replaced (fastmat 0.2.2):
//   Fourier.__init__/_forward/_backward   Fourier.pyx:63-238
//   Circulant 1-D construction             Circulant.pyx:104-137
//   Toeplitz 1-D construction              Toeplitz.pyx:146-316
//   Partial(Product(F.H, Diag, F)) apply   Partial.pyx:268-294, Product.pyx:205-240
//
// Design: see fm_fft_core.hpp.  Twiddles, Bluestein chirps and generator
// spectra are produced on the device in FP64 (sincospi on exactly reduced
// arguments) and rounded once to the plan precision.
#include <cstdarg>
#include <map>
#include <tuple>
#include <algorithm>
#include <vector>

#include "fm_fft_core.hpp"

namespace fm {

// ------------------------------------------------------------ error state
static thread_local std::string g_last_error;

void set_error(const char *fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_last_error = buf;
}

// ------------------------------------------------------- table generation
// W_M^e = exp(-2 pi i e / M) for e < count, M = 2^logm; FP64 sincospi of the
// exact argument -2e/M.
template <typename T>
__global__ void gen_twiddle_kernel(cpx<T> *out, int logm, int64_t count) {
    const int64_t e = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
    if (e >= count) return;
    double s, c;
    sincospi(-2.0 * double(e) / double(int64_t(1) << logm), &s, &c);
    out[e] = mk(T(c), T(s));
}

// Bluestein chirp p[n] = exp(-i pi n^2 / N) with n^2 reduced modulo 2N in
// int64 before forming the angle (the angle stays in [0, 2 pi)).
template <typename T>
__global__ void gen_chirp_kernel(cpx<T> *out, int64_t N, double sign) {
    const int64_t n = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
    if (n >= N) return;
    const int64_t r = (n * n) % (2 * N);  // exact: N < 2^31
    double s, c;
    sincospi(sign * double(r) / double(N), &s, &c);
    out[n] = mk(T(c), T(s));
}

// Wrapped Bluestein kernel g[n] = exp(+i pi n^2 / N), g[L-n] = g[n]
// (Fourier.pyx:137-153), in complex128.
__global__ void gen_bluestein_kernel(cpx<double> *g, int64_t N, int64_t L) {
    const int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
    if (i >= L) return;
    int64_t n = -1;
    if (i < N) n = i;
    else if (i > L - N) n = L - i;
    cpx<double> v = mk(0.0, 0.0);
    if (n >= 0) {
        const int64_t r = (n * n) % (2 * N);  // exact: N < 2^31
        double s, c;
        sincospi(double(r) / double(N), &s, &c);
        v = mk(c, s);
    }
    g[i] = v;
}

// Circulant / Toeplitz generator embedding into a length-L circulant:
// t'[i] = gen[i] for i < head, t'[L - tail + q] = gen[head + q] for q < tail.
__global__ void embed_kernel(cpx<double> *t, const cpx<double> *gen, int64_t L,
                             int64_t head, int64_t tail) {
    const int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
    if (i >= L) return;
    cpx<double> v = mk(0.0, 0.0);
    if (i < head) v = gen[i];
    else if (i >= L - tail) v = gen[head + (i - (L - tail))];
    t[i] = v;
}

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

static inline unsigned grid1d(int64_t n, int bs = 256) { return unsigned(ceil_div(n, bs)); }

static inline size_t esize(int prec) { return prec == FM_COMPLEX64 ? 8 : 16; }

// ---------------------------------------------------------- twiddle cache
namespace {
std::mutex g_tw_mutex;
std::map<std::tuple<int, int, int, int64_t>, void *> g_tw_tables;
}  // namespace

// Table of W_{2^logm}^e for e < count, cached per device for the process.
static const void *table(int prec, int logm, int64_t count, hipStream_t stream,
                         fm_status *st) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) {
        set_error("hipGetDevice failed");
        *st = FM_ERR_HIP;
        return nullptr;
    }
    std::lock_guard<std::mutex> lock(g_tw_mutex);
    auto key = std::make_tuple(dev, prec, logm, count);
    auto it = g_tw_tables.find(key);
    if (it != g_tw_tables.end()) return it->second;
    const size_t esz = prec == FM_COMPLEX64 ? 8 : 16;
    void *p = nullptr;
    if (hipMalloc(&p, esz * size_t(count)) != hipSuccess) {
        set_error("hipMalloc of a %lld-entry twiddle table failed", (long long)count);
        *st = FM_ERR_ALLOC;
        return nullptr;
    }
    if (prec == FM_COMPLEX64)
        hipLaunchKernelGGL(gen_twiddle_kernel<float>, grid1d(count), 256, 0, stream,
                           (cpx<float> *)p, logm, count);
    else
        hipLaunchKernelGGL(gen_twiddle_kernel<double>, grid1d(count), 256, 0, stream,
                           (cpx<double> *)p, logm, count);
    if (hipGetLastError() != hipSuccess || hipStreamSynchronize(stream) != hipSuccess) {
        set_error("twiddle generation failed");
        *st = FM_ERR_HIP;
        hipFree(p);
        return nullptr;
    }
    g_tw_tables[key] = p;
    return p;
}

const void *twiddle_table(int precision, int logn, hipStream_t stream, fm_status *st) {
    return table(precision, logn, int64_t(1) << logn, stream, st);
}

// ------------------------------------------------------------ tile launch
// Whether every element offset (x, y) and logical index (zero-pad / crop
// bounds, pre / post / spectrum tables, twiddle exponents) of a pass stays
// below 2^32 / sizeof(element), so the kernel may run 32-bit index math.
template <typename T> static bool fits_idx32(const TileArgs<T> &a, int64_t N) {
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
           below(double(N1) * a.d_gn + double(J1) * a.d_gj) && below(double(a.J) * double(N));
}

template <typename T, int LOGN, bool CONV>
static fm_status launch_pass(const TileArgs<T> &a, hipStream_t stream) {
    constexpr int NT = PassCfg<T, LOGN>::NT;
    constexpr int LE = PassCfg<T, LOGN>::LE;
    constexpr int N = 1 << LOGN;
    const size_t lds = size_t(a.B) * (N + (N >> 5) + 1) * sizeof(cpx<T>);
    static bool attr_set = false;
    if (!attr_set) {
        (void)hipFuncSetAttribute((const void *)fft_pass_kernel<T, LOGN, NT, CONV, false, LE>,
                                  hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
        (void)hipFuncSetAttribute((const void *)fft_pass_kernel<T, LOGN, NT, CONV, true, LE>,
                                  hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
        attr_set = true;
    }
    FM_REQUIRE(lds <= 160 * 1024, FM_ERR_VALUE, "tile of %d x %d needs %zu bytes of LDS", a.B, N,
               lds);
    FM_REQUIRE(a.ntiles > 0 && a.ntiles < (int64_t(1) << 31), FM_ERR_VALUE,
               "tile count %lld out of range", (long long)a.ntiles);
    const bool i32 = fits_idx32(a, N);
    // 1024-point complex64 passes that also multiply by a chirp (Bluestein
    // pass A / C) run the 16-element, 512-thread shape: with the chirp loads
    // they gain more from the extra waves than they lose to the extra stage
    // (plain 1024-point passes are faster with 32 elements per thread)
    bool e16 = false;
    if constexpr (sizeof(T) == 4 && LOGN == 10) e16 = i32 && (a.pre || a.post);
    ProfToken tok;
    prof_begin(stream, &tok);
    if constexpr (sizeof(T) == 4 && LOGN == 10) {
        if (e16) {
            static bool attr16 = false;
            if (!attr16) {
                (void)hipFuncSetAttribute((const void *)fft_pass_kernel<T, LOGN, 512, CONV, true, 4>,
                                          hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
                attr16 = true;
            }
            hipLaunchKernelGGL((fft_pass_kernel<T, LOGN, 512, CONV, true, 4>),
                               dim3(unsigned(a.ntiles)), dim3(512), lds, stream, a);
        }
    }
    if (!e16) {
        if (i32)
            hipLaunchKernelGGL((fft_pass_kernel<T, LOGN, NT, CONV, true, LE>),
                               dim3(unsigned(a.ntiles)), dim3(NT), lds, stream, a);
        else
            hipLaunchKernelGGL((fft_pass_kernel<T, LOGN, NT, CONV, false, LE>),
                               dim3(unsigned(a.ntiles)), dim3(NT), lds, stream, a);
    }
    FM_CHECK_LAUNCH();
    if (tok.active) {
        const double span = double(a.J) * N;
        const double esz = double(sizeof(cpx<T>));
        const double in = double(a.O) * (double(a.n_valid) < span ? double(a.n_valid) : span);
        const double out = double(a.O) * (double(a.m_valid) < span ? double(a.m_valid) : span);
        const double spec = CONV ? (a.d_gj ? span : double(N)) : 0.0;
        char name[96];
        snprintf(name, sizeof(name), "fft_pass_kernel<%s,%d,%d,%s>",
                 sizeof(T) == 4 ? "float" : "double", LOGN, e16 ? 512 : NT,
                 CONV ? "conv" : "fft");
        const double nseq = double(a.O) * double(a.J);
        const double flops = nseq * 5.0 * N * LOGN * (CONV ? 2.0 : 1.0);
        prof_end(stream, &tok, name, (in + out + spec) * esz, flops);
    }
    return FM_SUCCESS;
}

// Fill in the tile shape (batch, tile count, log2 J) of a pass.
template <typename T> static fm_status normalize_pass(TileArgs<T> &a) {
    const int maxlog = sizeof(T) == 4 ? 14 : 13;
    FM_REQUIRE(a.logN >= 0 && a.logN <= maxlog, FM_ERR_VALUE,
               "sub-transform length 2^%d exceeds the tile limit", a.logN);
    a.B = tile_batch<T>(a.logN);
    a.ntiles = ceil_div(a.O * a.J, a.B);
    a.logJ = is_pow2(a.J) ? ilog2(a.J) : -1;
    if (a.o_group > 1 && ((a.J % a.B) != 0 || (a.O % a.o_group) != 0)) a.o_group = 1;
    return FM_SUCCESS;
}

template <typename T>
fm_status launch_tile(const TileArgs<T> &a_in, bool conv, hipStream_t stream) {
    TileArgs<T> a = a_in;
    fm_status nst = normalize_pass(a);
    if (nst) return nst;
#define FM_PASS_CASE(L)                                                                  \
    case L:                                                                              \
        return conv ? launch_pass<T, L, true>(a, stream) : launch_pass<T, L, false>(a, stream);
    switch (a.logN) {
        FM_PASS_CASE(0) FM_PASS_CASE(1) FM_PASS_CASE(2) FM_PASS_CASE(3) FM_PASS_CASE(4)
        FM_PASS_CASE(5) FM_PASS_CASE(6) FM_PASS_CASE(7) FM_PASS_CASE(8) FM_PASS_CASE(9)
        FM_PASS_CASE(10) FM_PASS_CASE(11) FM_PASS_CASE(12) FM_PASS_CASE(13)
    case 14:
        if constexpr (sizeof(T) == 4)
            return conv ? launch_pass<T, 14, true>(a, stream) : launch_pass<T, 14, false>(a, stream);
        break;
    default:
        break;
    }
#undef FM_PASS_CASE
    set_error("unsupported sub-transform length 2^%d", a.logN);
    return FM_ERR_VALUE;
}

template fm_status launch_tile<float>(const TileArgs<float> &, bool, hipStream_t);
template fm_status launch_tile<double>(const TileArgs<double> &, bool, hipStream_t);

template <typename T> static TileArgs<T> blank_args() {
    TileArgs<T> a;
    memset(&a, 0, sizeof(a));
    a.out_scale = T(1);
    a.sign1 = -1;
    a.twb_shift = 0;
    a.n_valid = INT64_MAX;
    a.m_valid = INT64_MAX;
    return a;
}

// Tile-shape knob: environment override (used by the tuning sweeps), else
// the built-in default.
static int knob(const char *name, int dflt) {
    const char *v = getenv(name);
    if (!v || !*v) return dflt;
    int x = atoi(v);
    return x > 0 ? x : dflt;
}

static int pick_batch(int logN, int64_t nseq, int max_tile, int want) {
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

static fm_status make_geometry(int prec, int logL, bool conv, hipStream_t stream,
                               Geometry *g) {
    memset(g, 0, sizeof(*g));
    g->prec = prec;
    g->logL = logL;
    const int max_tile = prec == FM_COMPLEX64 ? max_tile_elems<float>() : max_tile_elems<double>();
    // single pass while one column fills at most 64 KB of LDS
    const int single_max = prec == FM_COMPLEX64 ? 8192 : 4096;
    fm_status st = FM_SUCCESS;
    if ((int64_t(1) << logL) <= single_max) {
        g->passes = 1;
        g->tw_full = twiddle_table(FM_COMPLEX128, logL, stream, &st);
        return st;
    }
    g->logN1 = logL / 2;
    g->logN2 = logL - g->logN1;
    FM_REQUIRE((int64_t(1) << g->logN1) <= max_tile && (int64_t(1) << g->logN2) <= max_tile,
               FM_ERR_NOT_IMPLEMENTED, "FFT length 2^%d exceeds the two-level limit", logL);
    g->passes = conv ? 3 : 2;
    g->tw1 = twiddle_table(FM_COMPLEX128, g->logN1, stream, &st);
    if (st) return st;
    g->tw2 = twiddle_table(FM_COMPLEX128, g->logN2, stream, &st);
    if (st) return st;
    g->twb_shift = (logL + 1) / 2;
    // hi: W_L^{a 2^s} = W_{L/2^s}^a (full table), lo: W_L^b for b < 2^s; in
    // FP64 for both precisions (the pass kernels form the twiddles in FP64)
    g->twb_hi = table(FM_COMPLEX128, logL - g->twb_shift, int64_t(1) << (logL - g->twb_shift),
                      stream, &st);
    if (st) return st;
    g->twb_lo = table(FM_COMPLEX128, logL, int64_t(1) << g->twb_shift, stream, &st);
    return st;
}

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

template <typename T>
static fm_status build_pipeline(const Geometry &g, const PipeIO &io,
                                std::vector<PassSpec<T>> *out) {
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

// Columns per chunk of a multi-pass transform.  All columns go in one chunk
// unless FASTMAT_AMD_CHUNK_MB sets a budget for one chunk's intermediate
// (L x columns), e.g. to keep it in the memory-side cache (MALL).  Measured
// on 2^20 x 64 and Bluestein 1,000,003 x 64 (complex64), a 96 MB budget was
// 14% slower than one chunk: the smaller grids lose more to tails than the
// cache saves.
static int64_t chunk_columns(const Geometry &g, int64_t k, size_t esz) {
    if (g.passes == 1) return k;
    const int64_t mb = knob("FASTMAT_AMD_CHUNK_MB", 0);
    if (mb <= 0) return k;
    const int64_t budget = mb << 20;
    const int64_t per_col = (int64_t(1) << g.logL) * int64_t(esz);
    int64_t c = budget / per_col;
    if (c < 1) c = 1;
    return c < k ? c : k;
}

template <typename T>
static fm_status run_pipeline(const Geometry &g, const PipeIO &io, hipStream_t stream) {
    const int64_t cc = chunk_columns(g, io.k, sizeof(cpx<T>));
    for (int64_t c0 = 0; c0 < io.k; c0 += cc) {
        PipeIO sub = io;
        sub.k = io.k - c0 < cc ? io.k - c0 : cc;
        sub.x = (const cpx<T> *)io.x + c0 * io.ldx;
        sub.y = (cpx<T> *)io.y + c0 * io.ldy;
        std::vector<PassSpec<T>> passes;
        fm_status st = build_pipeline<T>(g, sub, &passes);
        for (size_t q = 0; !st && q < passes.size(); ++q)
            st = launch_tile<T>(passes[q].a, passes[q].conv, stream);
        if (st) return st;
    }
    return FM_SUCCESS;
}

static fm_status run_pipeline_any(const Geometry &g, const PipeIO &io, hipStream_t stream) {
    return g.prec == FM_COMPLEX64 ? run_pipeline<float>(g, io, stream)
                                  : run_pipeline<double>(g, io, stream);
}

// Strided single-pass application along a mode axis (Kron): vectors
// v[i] = x[(o*N + i)*inner + j].
template <typename T>
static fm_status run_strided(const Geometry &g, const void *x, void *y, int64_t outer,
                             int64_t inner, int conj_io, const void *D, int conj_d,
                             hipStream_t stream) {
    FM_REQUIRE(g.passes == 1, FM_ERR_NOT_IMPLEMENTED,
               "strided mode application needs a single-pass transform");
    const int64_t N = int64_t(1) << g.logL;
    TileArgs<T> a = blank_args<T>();
    a.logN = g.logL;
    a.O = outer; a.J = inner;
    int want = inner >= 16 ? 16 : 1;
    a.B = pick_batch(g.logL, outer * inner, max_tile_elems<T>() / 2, inner > 1 ? want : 64);
    a.ntiles = ceil_div(outer * inner, a.B);
    a.x = (const cpx<T> *)x;
    a.in_os = N * inner; a.in_js = 1; a.in_ns = inner;
    a.conj_in = conj_io;
    a.seqfast_in = (inner > 1 && a.B > 1);
    a.tw = (const cpx<double> *)g.tw_full;
    a.D = (const cpx<T> *)D;
    a.d_gn = 1; a.d_gj = 0; a.conj_d = conj_d;
    a.y = (cpx<T> *)y;
    a.out_os = N * inner; a.out_js = 1; a.out_ns = inner;
    a.conj_out = conj_io;
    a.seqfast_out = a.seqfast_in;
    return launch_tile<T>(a, D != nullptr, stream);
}

// FP64 in-place forward FFT of one length-L complex128 vector (plan set-up
// only; synchronises the stream before freeing its workspace).
static fm_status fft64_inplace(cpx<double> *v, int logL, hipStream_t stream) {
    Geometry g;
    fm_status st = make_geometry(FM_COMPLEX128, logL, false, stream, &g);
    if (st) return st;
    const int64_t L = int64_t(1) << logL;
    void *work = nullptr;
    if (g.passes > 1) FM_CHECK_HIP(hipMalloc(&work, size_t(L) * 16));
    void *out = nullptr;
    if (g.passes > 1) {
        // pass A reads v and pass B writes y: keep them distinct
        FM_CHECK_HIP(hipMalloc(&out, size_t(L) * 16));
    }
    PipeIO io;
    memset(&io, 0, sizeof(io));
    io.x = v; io.ldx = L; io.y = out ? out : v; io.ldy = L; io.k = 1;
    io.n_in = L; io.m_out = L; io.work = work;
    st = run_pipeline<double>(g, io, stream);
    if (!st && out)
        if (hipMemcpyAsync(v, out, size_t(L) * 16, hipMemcpyDeviceToDevice, stream) != hipSuccess)
            st = FM_ERR_HIP;
    hipError_t e = hipStreamSynchronize(stream);
    if (work) hipFree(work);
    if (out) hipFree(out);
    if (!st && e != hipSuccess) {
        set_error("FP64 set-up FFT failed: %s", hipGetErrorString(e));
        st = FM_ERR_HIP;
    }
    return st;
}

// Spectrum D = fft(t')/L of a complex128 generator vector, stored in `prec`.
// The spectrum is stored in natural order for single-pass geometries and in
// the row pass's order (see scale_cast_kernel) for four-step geometries, so
// that the row pass reads it contiguously.
static fm_status spectrum_from(cpx<double> *t, const Geometry &g, int prec, void *D,
                               hipStream_t stream) {
    const int logL = g.logL;
    const int64_t L = int64_t(1) << logL;
    fm_status st = fft64_inplace(t, logL, stream);
    if (st) return st;
    const int l1 = g.passes > 1 ? g.logN1 : 0, l2 = g.passes > 1 ? g.logN2 : 0;
    if (prec == FM_COMPLEX64)
        hipLaunchKernelGGL(scale_cast_kernel<float>, grid1d(L), 256, 0, stream,
                           (cpx<float> *)D, t, L, 1.0 / double(L), l1, l2);
    else
        hipLaunchKernelGGL(scale_cast_kernel<double>, grid1d(L), 256, 0, stream,
                           (cpx<double> *)D, t, L, 1.0 / double(L), l1, l2);
    FM_CHECK_LAUNCH();
    FM_CHECK_HIP(hipStreamSynchronize(stream));
    return FM_SUCCESS;
}



}  // namespace fm

using namespace fm;

// ======================================================================
//                                 Fourier
// ======================================================================
struct fm_fourier_plan {
    int device;
    int prec;
    int64_t N;
    int bluestein;
    int64_t L;
    Geometry geo;
    void *chirp;  // p[n] = exp(-i pi n^2 / N), N entries (Bluestein)
    void *D;      // fft(g)/L, L entries (Bluestein)
};

extern "C" int fm_abi_version(void) { return FM_ABI_VERSION; }
extern "C" const char *fm_get_last_error(void) { return g_last_error.c_str(); }

extern "C" fm_status fm_fourier_plan_create(int64_t order, int precision, void *stream,
                                            fm_fourier_plan **plan) {
    FM_REQUIRE(plan != nullptr, FM_ERR_VALUE, "plan pointer is NULL");
    *plan = nullptr;
    FM_REQUIRE(order >= 1, FM_ERR_VALUE, "Fourier order cannot be smaller than 1.");
    FM_REQUIRE(order < (int64_t(1) << 30), FM_ERR_NOT_IMPLEMENTED,
               "Fourier order %lld exceeds the supported maximum 2^30", (long long)order);
    FM_REQUIRE(precision == FM_COMPLEX64 || precision == FM_COMPLEX128, FM_ERR_TYPE,
               "Fourier precision must be complex64 or complex128");
    hipStream_t s = as_stream(stream);
    auto *p = new fm_fourier_plan();
    memset(p, 0, sizeof(*p));
    hipGetDevice(&p->device);
    p->prec = precision;
    p->N = order;
    fm_status st = FM_SUCCESS;
    if (is_pow2(order)) {
        p->L = order;
        st = make_geometry(precision, ilog2(order), false, s, &p->geo);
    } else {
        p->bluestein = 1;
        int logL = ilog2(2 * order - 1);
        p->L = int64_t(1) << logL;
        st = make_geometry(precision, logL, true, s, &p->geo);
        if (!st && hipMalloc(&p->chirp, esize(precision) * size_t(order)) != hipSuccess) st = FM_ERR_ALLOC;
        if (!st && hipMalloc(&p->D, esize(precision) * size_t(p->L)) != hipSuccess) st = FM_ERR_ALLOC;
        cpx<double> *g = nullptr;
        if (!st && hipMalloc(&g, 16 * size_t(p->L)) != hipSuccess) st = FM_ERR_ALLOC;
        if (st == FM_ERR_ALLOC) set_error("device allocation for Bluestein plan failed");
        if (!st) {
            if (precision == FM_COMPLEX64)
                hipLaunchKernelGGL(gen_chirp_kernel<float>, grid1d(order), 256, 0, s,
                                   (cpx<float> *)p->chirp, order, -1.0);
            else
                hipLaunchKernelGGL(gen_chirp_kernel<double>, grid1d(order), 256, 0, s,
                                   (cpx<double> *)p->chirp, order, -1.0);
            hipLaunchKernelGGL(gen_bluestein_kernel, grid1d(p->L), 256, 0, s, g, order, p->L);
            if (hipGetLastError() != hipSuccess) st = FM_ERR_HIP;
        }
        if (!st) st = spectrum_from(g, p->geo, precision, p->D, s);
        if (g) hipFree(g);
    }
    if (st) {
        fm_fourier_plan_destroy(p);
        return st;
    }
    *plan = p;
    return FM_SUCCESS;
}

extern "C" fm_status fm_fourier_plan_info(const fm_fourier_plan *p, int64_t *fft_length,
                                          int *bluestein, int *passes) {
    FM_REQUIRE(p != nullptr, FM_ERR_VALUE, "plan is NULL");
    if (fft_length) *fft_length = p->L;
    if (bluestein) *bluestein = p->bluestein;
    if (passes) *passes = p->geo.passes;
    return FM_SUCCESS;
}

extern "C" fm_status fm_fourier_workspace_size(const fm_fourier_plan *p, int64_t k,
                                               int64_t *bytes) {
    FM_REQUIRE(p != nullptr && bytes != nullptr, FM_ERR_VALUE, "NULL argument");
    FM_REQUIRE(k >= 0, FM_ERR_VALUE, "negative column count");
    *bytes = p->geo.passes > 1 ? int64_t(esize(p->prec)) * p->L *
                                     chunk_columns(p->geo, k, esize(p->prec))
                               : 0;
    return FM_SUCCESS;
}

extern "C" fm_status fm_fourier_apply(const fm_fourier_plan *p, int adjoint, const void *x,
                                      int64_t ldx, void *y, int64_t ldy, int64_t k,
                                      void *workspace, int64_t workspace_bytes, void *stream) {
    FM_REQUIRE(p != nullptr, FM_ERR_VALUE, "plan is NULL");
    FM_REQUIRE(k >= 0, FM_ERR_VALUE, "negative column count");
    if (k == 0) return FM_SUCCESS;
    FM_REQUIRE(x && y, FM_ERR_VALUE, "NULL data pointer");
    FM_REQUIRE(ldx >= p->N && ldy >= p->N, FM_ERR_VALUE,
               "leading dimensions (%lld, %lld) below the order %lld", (long long)ldx,
               (long long)ldy, (long long)p->N);
    int64_t need = 0;
    fm_fourier_workspace_size(p, k, &need);
    FM_REQUIRE(workspace_bytes >= need, FM_ERR_VALUE,
               "workspace of %lld bytes is smaller than the %lld required",
               (long long)workspace_bytes, (long long)need);
    PipeIO io;
    memset(&io, 0, sizeof(io));
    io.x = x; io.ldx = ldx; io.y = y; io.ldy = ldy; io.k = k;
    io.n_in = p->N; io.m_out = p->N;
    io.conj_io = adjoint ? 1 : 0;
    io.work = workspace;
    if (p->bluestein) {
        io.pre = p->chirp;
        io.post = p->chirp;
        io.D = p->D;
    }
    return run_pipeline_any(p->geo, io, as_stream(stream));
}

extern "C" fm_status fm_fourier_apply_strided(const fm_fourier_plan *p, int adjoint,
                                              const void *x, void *y, int64_t outer,
                                              int64_t inner, void *stream) {
    FM_REQUIRE(p != nullptr, FM_ERR_VALUE, "plan is NULL");
    FM_REQUIRE(outer >= 0 && inner >= 1, FM_ERR_VALUE, "bad strided geometry");
    FM_REQUIRE(!p->bluestein, FM_ERR_NOT_IMPLEMENTED,
               "strided application needs a power-of-two order");
    if (outer == 0) return FM_SUCCESS;
    if (p->prec == FM_COMPLEX64)
        return run_strided<float>(p->geo, x, y, outer, inner, adjoint, nullptr, 0, as_stream(stream));
    return run_strided<double>(p->geo, x, y, outer, inner, adjoint, nullptr, 0, as_stream(stream));
}

extern "C" fm_status fm_fourier_plan_destroy(fm_fourier_plan *p) {
    if (!p) return FM_SUCCESS;
    if (p->chirp) hipFree(p->chirp);
    if (p->D) hipFree(p->D);
    delete p;
    return FM_SUCCESS;
}

// ======================================================================
//                          Circulant / Toeplitz
// ======================================================================
struct fm_conv_plan {
    int device;
    int prec;
    int kind;          // 0 circulant, 1 toeplitz
    int64_t rows, cols, gen_len, L;
    int64_t head, tail;  // t' = [gen[:head], 0..., gen[head:head+tail]]
    Geometry geo;      // conv pipeline (1 or 3 passes)
    Geometry geo_fft;  // plain length-L FFT (1 or 2 passes), generator gradient
    void *D;           // fft(t')/L in prec, L entries
};

static fm_status conv_plan_create(int prec, int kind, int64_t m, int64_t n, int64_t L,
                                  int64_t head, int64_t tail, const void *gen,
                                  int64_t gen_len, hipStream_t s, fm_conv_plan **out) {
    auto *p = new fm_conv_plan();
    memset(p, 0, sizeof(*p));
    hipGetDevice(&p->device);
    p->prec = prec; p->kind = kind; p->rows = m; p->cols = n; p->L = L;
    p->gen_len = gen_len;
    p->head = head;
    p->tail = tail;
    const int logL = ilog2(L);
    fm_status st = make_geometry(prec, logL, true, s, &p->geo);
    if (!st) st = make_geometry(prec, logL, false, s, &p->geo_fft);
    cpx<double> *t = nullptr;
    if (!st && hipMalloc(&p->D, esize(prec) * size_t(L)) != hipSuccess) st = FM_ERR_ALLOC;
    if (!st && hipMalloc(&t, 16 * size_t(L)) != hipSuccess) st = FM_ERR_ALLOC;
    if (st == FM_ERR_ALLOC) set_error("device allocation for convolution plan failed");
    if (!st) {
        hipLaunchKernelGGL(embed_kernel, grid1d(L), 256, 0, s, t, (const cpx<double> *)gen, L,
                           head, tail);
        if (hipGetLastError() != hipSuccess) st = FM_ERR_HIP;
    }
    if (!st) st = spectrum_from(t, p->geo, prec, p->D, s);
    if (t) hipFree(t);
    if (st) {
        fm_conv_plan_destroy(p);
        return st;
    }
    *out = p;
    return FM_SUCCESS;
}

extern "C" fm_status fm_circulant_plan_create(int precision, int64_t n, const void *generator,
                                              void *stream, fm_conv_plan **plan) {
    FM_REQUIRE(plan != nullptr, FM_ERR_VALUE, "plan pointer is NULL");
    *plan = nullptr;
    FM_REQUIRE(n >= 1, FM_ERR_VALUE, "Column-definition tensor must be at least 1D.");
    FM_REQUIRE(generator != nullptr, FM_ERR_VALUE, "generator is NULL");
    FM_REQUIRE(precision == FM_COMPLEX64 || precision == FM_COMPLEX128, FM_ERR_TYPE,
               "Circulant precision must be complex64 or complex128");
    int64_t L = is_pow2(n) ? n : (int64_t(1) << ilog2(2 * n - 1));
    // t' = c (L == n) or [c, 0, ..., 0, c[1:]] (L >= 2n - 1)
    int64_t head = n, tail = L == n ? 0 : n - 1;
    // the tail copies c[1:], i.e. generator entries head-... : use an offset view
    // by embedding from gen = c with head = n and tail entries c[1..n-1]
    fm_status st;
    if (tail == 0) {
        st = conv_plan_create(precision, 0, n, n, L, head, 0, generator, n, as_stream(stream), plan);
    } else {
        // build [c, c[1:]] so the generic embed (head, tail) applies
        cpx<double> *gen2 = nullptr;
        FM_CHECK_HIP(hipMalloc(&gen2, 16 * size_t(2 * n - 1)));
        hipStream_t s = as_stream(stream);
        hipMemcpyAsync(gen2, generator, 16 * size_t(n), hipMemcpyDeviceToDevice, s);
        hipMemcpyAsync(gen2 + n, (const cpx<double> *)generator + 1, 16 * size_t(n - 1),
                       hipMemcpyDeviceToDevice, s);
        st = conv_plan_create(precision, 0, n, n, L, head, tail, gen2, n, s, plan);
        hipStreamSynchronize(s);
        hipFree(gen2);
    }
    return st;
}

extern "C" fm_status fm_toeplitz_plan_create(int precision, int64_t m, int64_t n,
                                             const void *generator, void *stream,
                                             fm_conv_plan **plan) {
    FM_REQUIRE(plan != nullptr, FM_ERR_VALUE, "plan pointer is NULL");
    *plan = nullptr;
    FM_REQUIRE(m >= 1 && n >= 1, FM_ERR_VALUE, "Toeplitz dimensions must be positive");
    FM_REQUIRE(generator != nullptr, FM_ERR_VALUE, "generator is NULL");
    FM_REQUIRE(precision == FM_COMPLEX64 || precision == FM_COMPLEX128, FM_ERR_TYPE,
               "Toeplitz precision must be complex64 or complex128");
    const int64_t d = m + n - 1;
    const int64_t L = int64_t(1) << ilog2(d);
    return conv_plan_create(precision, 1, m, n, L, m, n - 1, generator, d, as_stream(stream), plan);
}

// Two-level circulant of n0 x n1 (powers of two) as the three-pass
// pipeline of a 1-D convolution with the four-step twiddles left out: pass A
// transforms level 0 (strided), pass B transforms level 1, multiplies the
// spectrum and transforms back, pass C transforms level 0 back.  Without
// the twiddle W_L^{n2 k1} the two passes are exactly the 2-D DFT, and the
// spectrum, fftn(t)/(n0 n1) in C order, is already in the row pass's order.
extern "C" fm_status fm_circulant2_plan_create(int precision, int64_t n0, int64_t n1,
                                               const void *spectrum, void *stream,
                                               fm_conv_plan **plan) {
    FM_REQUIRE(plan != nullptr, FM_ERR_VALUE, "plan pointer is NULL");
    *plan = nullptr;
    FM_REQUIRE(precision == FM_COMPLEX64 || precision == FM_COMPLEX128, FM_ERR_TYPE,
               "two-level convolution computes in complex64 or complex128");
    FM_REQUIRE(spectrum != nullptr, FM_ERR_VALUE, "spectrum is NULL");
    FM_REQUIRE(n0 >= 2 && n1 >= 2 && is_pow2(n0) && is_pow2(n1), FM_ERR_VALUE,
               "two-level lengths must be powers of two >= 2 (got %lld x %lld)",
               (long long)n0, (long long)n1);
    const int max_log = precision == FM_COMPLEX64 ? 13 : 12;
    FM_REQUIRE(ilog2(n0) <= max_log && ilog2(n1) <= max_log, FM_ERR_NOT_IMPLEMENTED,
               "two-level lengths above 2^%d", max_log);
    hipStream_t s = as_stream(stream);
    auto *p = new fm_conv_plan();
    memset(p, 0, sizeof(*p));
    hipGetDevice(&p->device);
    const int64_t L = n0 * n1;
    p->prec = precision; p->kind = 2; p->rows = L; p->cols = L; p->L = L; p->gen_len = L;
    Geometry &g = p->geo;
    g.prec = precision;
    g.logL = ilog2(L);
    g.passes = 3;
    g.logN1 = ilog2(n0);
    g.logN2 = ilog2(n1);
    fm_status st = FM_SUCCESS;
    g.tw1 = twiddle_table(FM_COMPLEX128, g.logN1, s, &st);
    if (!st) g.tw2 = twiddle_table(FM_COMPLEX128, g.logN2, s, &st);
    p->geo_fft = g;
    p->geo_fft.passes = 0;  // no generator gradient for this kind
    if (!st && hipMalloc(&p->D, esize(precision) * size_t(L)) != hipSuccess) {
        set_error("device allocation for convolution plan failed");
        st = FM_ERR_ALLOC;
    }
    if (!st) {
        if (precision == FM_COMPLEX64)
            hipLaunchKernelGGL(scale_cast_kernel<float>, grid1d(L), 256, 0, s, (cpx<float> *)p->D,
                               (const cpx<double> *)spectrum, L, 1.0, 0, 0);
        else
            hipLaunchKernelGGL(scale_cast_kernel<double>, grid1d(L), 256, 0, s,
                               (cpx<double> *)p->D, (const cpx<double> *)spectrum, L, 1.0, 0, 0);
        if (hipGetLastError() != hipSuccess) {
            set_error("spectrum cast launch failed");
            st = FM_ERR_HIP;
        }
    }
    if (!st && hipStreamSynchronize(s) != hipSuccess) {
        set_error("stream synchronisation failed");
        st = FM_ERR_HIP;
    }
    if (st) {
        fm_conv_plan_destroy(p);
        return st;
    }
    *plan = p;
    return FM_SUCCESS;
}

extern "C" fm_status fm_conv_plan_info(const fm_conv_plan *p, int64_t *rows, int64_t *cols,
                                       int64_t *fft_length, int *passes) {
    FM_REQUIRE(p != nullptr, FM_ERR_VALUE, "plan is NULL");
    if (rows) *rows = p->rows;
    if (cols) *cols = p->cols;
    if (fft_length) *fft_length = p->L;
    if (passes) *passes = p->geo.passes;
    return FM_SUCCESS;
}

extern "C" fm_status fm_conv_workspace_size(const fm_conv_plan *p, int64_t k, int64_t *bytes) {
    FM_REQUIRE(p != nullptr && bytes != nullptr, FM_ERR_VALUE, "NULL argument");
    FM_REQUIRE(k >= 0, FM_ERR_VALUE, "negative column count");
    *bytes = p->geo.passes > 1 ? int64_t(esize(p->prec)) * p->L *
                                     chunk_columns(p->geo, k, esize(p->prec))
                               : 0;
    return FM_SUCCESS;
}

extern "C" fm_status fm_conv_apply(const fm_conv_plan *p, int adjoint, const void *x,
                                   int64_t ldx, void *y, int64_t ldy, int64_t k, void *workspace,
                                   int64_t workspace_bytes, void *stream) {
    FM_REQUIRE(p != nullptr, FM_ERR_VALUE, "plan is NULL");
    FM_REQUIRE(k >= 0, FM_ERR_VALUE, "negative column count");
    if (k == 0) return FM_SUCCESS;
    FM_REQUIRE(x && y, FM_ERR_VALUE, "NULL data pointer");
    const int64_t nin = adjoint ? p->rows : p->cols;
    const int64_t mout = adjoint ? p->cols : p->rows;
    FM_REQUIRE(ldx >= nin && ldy >= mout, FM_ERR_VALUE, "leading dimension too small");
    int64_t need = 0;
    fm_conv_workspace_size(p, k, &need);
    FM_REQUIRE(workspace_bytes >= need, FM_ERR_VALUE,
               "workspace of %lld bytes is smaller than the %lld required",
               (long long)workspace_bytes, (long long)need);
    PipeIO io;
    memset(&io, 0, sizeof(io));
    io.x = x; io.ldx = ldx; io.y = y; io.ldy = ldy; io.k = k;
    io.n_in = nin; io.m_out = mout;
    io.D = p->D; io.conj_d = adjoint ? 1 : 0;
    io.work = workspace;
    return run_pipeline_any(p->geo, io, as_stream(stream));
}

extern "C" fm_status fm_conv_apply_strided(const fm_conv_plan *p, int adjoint, const void *x,
                                           void *y, int64_t outer, int64_t inner, void *stream) {
    FM_REQUIRE(p == nullptr || p->kind != 2, FM_ERR_NOT_IMPLEMENTED,
               "strided apply of a two-level convolution plan");
    FM_REQUIRE(p != nullptr, FM_ERR_VALUE, "plan is NULL");
    FM_REQUIRE(p->rows == p->cols && p->L == p->rows, FM_ERR_NOT_IMPLEMENTED,
               "strided application needs a square plan without padding");
    FM_REQUIRE(outer >= 0 && inner >= 1, FM_ERR_VALUE, "bad strided geometry");
    if (outer == 0) return FM_SUCCESS;
    if (p->prec == FM_COMPLEX64)
        return run_strided<float>(p->geo, x, y, outer, inner, 0, p->D, adjoint, as_stream(stream));
    return run_strided<double>(p->geo, x, y, outer, inner, 0, p->D, adjoint, as_stream(stream));
}

namespace fm {
// acc[q] = sum_c G^[c, q] * conj(X^[c, q])
template <typename T>
__global__ void corr_reduce_kernel(const cpx<T> *__restrict__ gh, const cpx<T> *__restrict__ xh,
                                   cpx<T> *__restrict__ acc, int64_t L, int64_t k) {
    const int64_t q = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
    if (q >= L) return;
    cpx<T> s = mk(T(0), T(0));
    for (int64_t c = 0; c < k; ++c) s = s + cmulc(gh[c * L + q], xh[c * L + q]);
    acc[q] = s;
}

// Fold the gradient over t' (length L, unnormalised inverse transform r)
// back onto the generator: Circulant c[j] <- r[j] (+ r[L - tail + j - 1] for
// j >= 1 when padded); Toeplitz tenT[i] <- r[i] (i < head) or
// r[L - tail + i - head].
template <typename T>
__global__ void fold_grad_kernel(const cpx<T> *__restrict__ r, cpx<T> *__restrict__ out,
                                 int kind, int64_t L, int64_t head, int64_t tail,
                                 int64_t gen_len, T s) {
    const int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
    if (i >= gen_len) return;
    cpx<T> v;
    if (kind == 0) {
        v = r[i];
        if (tail > 0 && i >= 1) v = v + r[L - tail + i - 1];
    } else {
        v = i < head ? r[i] : r[L - tail + (i - head)];
    }
    out[i] = scale(v, s);
}

template <typename T>
static fm_status conv_grad(const fm_conv_plan *p, const void *x, int64_t ldx, const void *g,
                           int64_t ldg, int64_t k, void *grad, void *ws, hipStream_t s) {
    const int64_t L = p->L;
    cpx<T> *xh = (cpx<T> *)ws;
    cpx<T> *gh = xh + L * k;
    cpx<T> *tmp = gh + L * k;
    cpx<T> *acc = tmp + L * k;
    cpx<T> *r = acc + L;
    PipeIO io;
    memset(&io, 0, sizeof(io));
    io.x = x; io.ldx = ldx; io.y = xh; io.ldy = L; io.k = k;
    io.n_in = p->cols; io.m_out = L; io.work = tmp;
    fm_status st = run_pipeline<T>(p->geo_fft, io, s);
    if (st) return st;
    io.x = g; io.ldx = ldg; io.y = gh; io.n_in = p->rows;
    st = run_pipeline<T>(p->geo_fft, io, s);
    if (st) return st;
    hipLaunchKernelGGL(corr_reduce_kernel<T>, grid1d(L), 256, 0, s, gh, xh, acc, L, k);
    FM_CHECK_LAUNCH();
    // unnormalised inverse transform: conj(FFT(conj(acc)))
    io.x = acc; io.ldx = L; io.y = r; io.ldy = L; io.k = 1; io.n_in = L; io.m_out = L;
    io.conj_io = 1; io.work = tmp;
    st = run_pipeline<T>(p->geo_fft, io, s);
    if (st) return st;
    hipLaunchKernelGGL(fold_grad_kernel<T>, grid1d(p->gen_len), 256, 0, s, r, (cpx<T> *)grad,
                       p->kind, L, p->head, p->tail, p->gen_len, T(1.0 / double(L)));
    FM_CHECK_LAUNCH();
    return FM_SUCCESS;
}
}  // namespace fm

extern "C" fm_status fm_conv_generator_grad_workspace_size(const fm_conv_plan *p, int64_t k,
                                                           int64_t *bytes) {
    FM_REQUIRE(p != nullptr && bytes != nullptr, FM_ERR_VALUE, "NULL argument");
    FM_REQUIRE(k >= 0, FM_ERR_VALUE, "negative column count");
    const int64_t kk = k > 1 ? k : 1;
    *bytes = int64_t(esize(p->prec)) * p->L * (3 * kk + 2);
    return FM_SUCCESS;
}

extern "C" fm_status fm_conv_generator_grad(const fm_conv_plan *p, const void *x, int64_t ldx,
                                            const void *g, int64_t ldg, int64_t k, void *grad,
                                            void *workspace, int64_t workspace_bytes,
                                            void *stream) {
    FM_REQUIRE(p != nullptr, FM_ERR_VALUE, "plan is NULL");
    FM_REQUIRE(p->kind != 2, FM_ERR_NOT_IMPLEMENTED,
               "generator gradient of a two-level convolution plan");
    FM_REQUIRE(k >= 1, FM_ERR_VALUE, "generator gradient needs at least one column");
    FM_REQUIRE(x && g && grad && workspace, FM_ERR_VALUE, "NULL data pointer");
    FM_REQUIRE(ldx >= p->cols && ldg >= p->rows, FM_ERR_VALUE, "leading dimension too small");
    int64_t need = 0;
    fm_conv_generator_grad_workspace_size(p, k, &need);
    FM_REQUIRE(workspace_bytes >= need, FM_ERR_VALUE,
               "workspace of %lld bytes is smaller than the %lld required",
               (long long)workspace_bytes, (long long)need);
    if (p->prec == FM_COMPLEX64)
        return conv_grad<float>(p, x, ldx, g, ldg, k, grad, workspace, as_stream(stream));
    return conv_grad<double>(p, x, ldx, g, ldg, k, grad, workspace, as_stream(stream));
}

extern "C" fm_status fm_conv_plan_destroy(fm_conv_plan *p) {
    if (!p) return FM_SUCCESS;
    if (p->D) hipFree(p->D);
    delete p;
    return FM_SUCCESS;
}

// ======================================================================
//                    Uneven batch of Circulant / Toeplitz
// ======================================================================
// Every operator's transform is a list of passes (one fused pass, or the
// three passes A, B, C of the four-step); pass q of all operators forms
// phase q.  One persistent kernel per phase walks a device table of
// (descriptor, tile) work items in decreasing-cost order, grid-strided, and
// runs the compile-time FFT body the descriptor's length selects.  Items of
// one phase are independent, so no workgroup ever waits for another.
namespace fm {

template <typename T> struct BatchDesc {
    TileArgs<T> a;
    int conv;
};

// One persistent kernel per (length, conv) class: grid-strided over the
// class's (descriptor, tile) items, most expensive first.
template <typename T, int LOGN, bool CONV, bool I32>
__global__ __launch_bounds__((PassCfg<T, LOGN>::NT)) void batch_pass_kernel(
    const BatchDesc<T> *__restrict__ descs, const int2 *__restrict__ items, int64_t nitems) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    cpx<T> *lds = reinterpret_cast<cpx<T> *>(smem_raw);
    for (int64_t it = blockIdx.x; it < nitems; it += gridDim.x) {
        const int2 w = items[it];
        const TileArgs<T> a = descs[w.x].a;
        fft_pass_body<T, LOGN, PassCfg<T, LOGN>::NT, CONV, I32, PassCfg<T, LOGN>::LE>(a, w.y, lds);
        __syncthreads();  // the next item reuses the LDS tile
    }
}

template <typename T> constexpr int batch_max_log() { return sizeof(T) == 4 ? 13 : 12; }

// Work items of one (phase, length, conv) class, in launch order.
struct BatchGroup {
    int phase, logN, conv;
    int idx32;  // every descriptor of the group fits 32-bit index arithmetic
    void *items;
    int64_t nitems;
    double bytes;
};

struct BatchDir {
    void *descs = nullptr;
    std::vector<BatchGroup> groups;
    bool bound = false;
};

}  // namespace fm

struct fm_batch_plan {
    int device;
    int prec;
    int count;
    std::vector<const fm_conv_plan *> plans;
    std::vector<int64_t> ks;
    std::vector<int64_t> ws_off;
    int64_t ws_bytes;
    int grid_per_cu;
    int ncu;
    BatchDir dir[2];
    // generator-gradient program: forward FFTs of the inputs and residuals,
    // the per-operator correlation sum, inverse FFTs, the fold
    BatchDir gdir[2];
    void *gtab;
    int64_t g_maxL, g_maxgen;
    // side streams (default priority) for the independent length classes of
    // one phase, forked from and joined back into the caller's stream
    static constexpr int kSide = 3;
    hipStream_t side[kSide];
    hipEvent_t fork_ev, join_ev[kSide];
    bool has_side;
};

static void batch_free_dir(BatchDir &d) {
    if (d.descs) (void)hipFree(d.descs);
    for (auto &g : d.groups)
        if (g.items) (void)hipFree(g.items);
    d = BatchDir();
}

extern "C" fm_status fm_batch_plan_create(int count, const fm_conv_plan *const *plans,
                                          const int64_t *ks, fm_batch_plan **out) {
    FM_REQUIRE(out != nullptr, FM_ERR_VALUE, "plan pointer is NULL");
    *out = nullptr;
    FM_REQUIRE(count >= 1 && plans && ks, FM_ERR_VALUE, "empty batch");
    for (int i = 0; i < count; ++i)
        FM_REQUIRE(plans[i] && plans[i]->kind != 2, FM_ERR_NOT_IMPLEMENTED,
                   "batch entry %d: only one-level convolution plans batch", i);
    const int prec = plans[0]->prec;
    auto *p = new fm_batch_plan();
    p->prec = prec;
    p->count = count;
    p->ws_bytes = 0;
    p->grid_per_cu = 2;
    (void)hipGetDevice(&p->device);
    int ncu = 0;
    if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, p->device) != hipSuccess ||
        ncu <= 0)
        ncu = 256;
    p->ncu = ncu;
    {
        bool ok = true;
        for (int q = 0; q < fm_batch_plan::kSide && ok; ++q) {
            p->side[q] = nullptr;
            p->join_ev[q] = nullptr;
            ok = hipStreamCreateWithFlags(&p->side[q], hipStreamNonBlocking) == hipSuccess &&
                 hipEventCreateWithFlags(&p->join_ev[q], hipEventDisableTiming) == hipSuccess;
        }
        p->fork_ev = nullptr;
        ok = ok && hipEventCreateWithFlags(&p->fork_ev, hipEventDisableTiming) == hipSuccess;
        p->has_side = ok;
    }
    for (int i = 0; i < count; ++i) {
        if (!plans[i] || plans[i]->prec != prec || ks[i] < 1) {
            delete p;
            set_error("batch entry %d: missing plan, mixed precision or no columns", i);
            return FM_ERR_VALUE;
        }
        const int maxlog = prec == FM_COMPLEX64 ? batch_max_log<float>() : batch_max_log<double>();
        const Geometry &g = plans[i]->geo;
        const int worst = g.passes == 1 ? g.logL : (g.logN1 > g.logN2 ? g.logN1 : g.logN2);
        if (worst > maxlog) {
            delete p;
            set_error("batch entry %d: transform length exceeds the batched kernel limit", i);
            return FM_ERR_NOT_IMPLEMENTED;
        }
        p->plans.push_back(plans[i]);
        p->ks.push_back(ks[i]);
        p->ws_off.push_back(p->ws_bytes);
        if (g.passes > 1) {
            int64_t b = int64_t(esize(prec)) * plans[i]->L * ks[i];
            p->ws_bytes += (b + 255) / 256 * 256;
        }
    }
    *out = p;
    return FM_SUCCESS;
}

extern "C" fm_status fm_batch_workspace_size(const fm_batch_plan *p, int64_t *bytes) {
    FM_REQUIRE(p && bytes, FM_ERR_VALUE, "NULL argument");
    *bytes = p->ws_bytes;
    return FM_SUCCESS;
}

extern "C" fm_status fm_batch_set_grid(fm_batch_plan *p, int workgroups_per_cu) {
    FM_REQUIRE(p != nullptr, FM_ERR_VALUE, "plan is NULL");
    FM_REQUIRE(workgroups_per_cu >= 1 && workgroups_per_cu <= 16, FM_ERR_VALUE,
               "workgroups per CU must be in [1, 16]");
    p->grid_per_cu = workgroups_per_cu;
    return FM_SUCCESS;
}

// Descriptors and work-item groups of one batched program: every pipeline's
// passes, pass q of a multi-pass pipeline in phase q and the one-pass
// pipelines in the last phase (they depend on nothing; measured on the
// uneven batch: 2.75 ms per step in the first phase, 2.68 in the last).
template <typename T>
static fm_status batch_build_dir(BatchDir &d, std::vector<std::vector<PassSpec<T>>> &pipes,
                                 hipStream_t s) {
    batch_free_dir(d);
    std::vector<BatchDesc<T>> descs;
    std::vector<int> phase_of;
    int last_phase = 0;
    for (auto &pl : pipes)
        if (int(pl.size()) - 1 > last_phase) last_phase = int(pl.size()) - 1;
    for (auto &pl : pipes) {
        for (size_t q = 0; q < pl.size(); ++q) {
            TileArgs<T> a = pl[q].a;
            fm_status st = normalize_pass(a);
            if (st) return st;
            descs.push_back(BatchDesc<T>{a, pl[q].conv ? 1 : 0});
            phase_of.push_back(pl.size() == 1 ? last_phase : int(q));
        }
    }
    // work items grouped by (phase, length, conv); groups of a phase run
    // largest first, items inside a group in descriptor order
    std::map<std::tuple<int, int, int>, std::vector<int2>> groups;
    std::map<std::tuple<int, int, int>, double> gbytes;
    std::map<std::tuple<int, int, int>, bool> gfit;
    for (size_t di = 0; di < descs.size(); ++di) {
        const TileArgs<T> &a = descs[di].a;
        auto key = std::make_tuple(phase_of[di], -a.logN, descs[di].conv ? 0 : 1);
        const bool fit = fits_idx32(a, int64_t(1) << a.logN);
        auto fit_it = gfit.find(key);
        if (fit_it == gfit.end()) gfit[key] = fit;
        else fit_it->second = fit_it->second && fit;
        auto &v = groups[key];
        for (int64_t t = 0; t < a.ntiles; ++t) {
            int2 w;
            w.x = int(di);
            w.y = int(t);
            v.push_back(w);
        }
        const double span = double(a.J) * double(int64_t(1) << a.logN);
        const double in = double(a.O) * (double(a.n_valid) < span ? double(a.n_valid) : span);
        const double outb = double(a.O) * (double(a.m_valid) < span ? double(a.m_valid) : span);
        gbytes[key] += (in + outb) * sizeof(cpx<T>);
    }
    FM_CHECK_HIP(hipMalloc(&d.descs, sizeof(BatchDesc<T>) * descs.size()));
    FM_CHECK_HIP(hipMemcpyAsync(d.descs, descs.data(), sizeof(BatchDesc<T>) * descs.size(),
                                hipMemcpyHostToDevice, s));
    for (auto &kv : groups) {
        BatchGroup g;
        g.phase = std::get<0>(kv.first);
        g.logN = -std::get<1>(kv.first);
        g.conv = std::get<2>(kv.first) == 0 ? 1 : 0;
        g.nitems = int64_t(kv.second.size());
        g.bytes = gbytes[kv.first];
        g.idx32 = gfit[kv.first] ? 1 : 0;
        g.items = nullptr;
        FM_CHECK_HIP(hipMalloc(&g.items, sizeof(int2) * kv.second.size()));
        FM_CHECK_HIP(hipMemcpyAsync(g.items, kv.second.data(), sizeof(int2) * kv.second.size(),
                                    hipMemcpyHostToDevice, s));
        d.groups.push_back(g);
    }
    FM_CHECK_HIP(hipStreamSynchronize(s));
    d.bound = true;
    return FM_SUCCESS;
}

template <typename T>
static fm_status batch_bind_t(fm_batch_plan *p, int adjoint, const void *const *ins,
                              const int64_t *ld_in, void *const *outs, const int64_t *ld_out,
                              void *workspace, hipStream_t s) {
    BatchDir &d = p->dir[adjoint ? 1 : 0];
    std::vector<std::vector<PassSpec<T>>> pipes(p->count);
    for (int i = 0; i < p->count; ++i) {
        const fm_conv_plan *c = p->plans[i];
        const int64_t nin = adjoint ? c->rows : c->cols;
        const int64_t mout = adjoint ? c->cols : c->rows;
        FM_REQUIRE(ins[i] && outs[i], FM_ERR_VALUE, "batch entry %d: NULL buffer", i);
        FM_REQUIRE(ld_in[i] >= nin && ld_out[i] >= mout, FM_ERR_VALUE,
                   "batch entry %d: leading dimension too small", i);
        PipeIO io;
        memset(&io, 0, sizeof(io));
        io.x = ins[i]; io.ldx = ld_in[i]; io.y = outs[i]; io.ldy = ld_out[i]; io.k = p->ks[i];
        io.n_in = nin; io.m_out = mout; io.D = c->D; io.conj_d = adjoint ? 1 : 0;
        io.work = c->geo.passes > 1 ? (char *)workspace + p->ws_off[i] : nullptr;
        FM_REQUIRE(c->geo.passes == 1 || workspace, FM_ERR_VALUE, "batch needs a workspace");
        fm_status st = build_pipeline<T>(c->geo, io, &pipes[i]);
        if (st) return st;
    }
    return batch_build_dir<T>(d, pipes, s);
}

extern "C" fm_status fm_batch_bind(fm_batch_plan *p, int adjoint, const void *const *ins,
                                   const int64_t *ld_in, void *const *outs,
                                   const int64_t *ld_out, void *workspace,
                                   int64_t workspace_bytes, void *stream) {
    FM_REQUIRE(p && ins && ld_in && outs && ld_out, FM_ERR_VALUE, "NULL argument");
    FM_REQUIRE(workspace_bytes >= p->ws_bytes, FM_ERR_VALUE,
               "workspace of %lld bytes is smaller than the %lld required",
               (long long)workspace_bytes, (long long)p->ws_bytes);
    if (p->prec == FM_COMPLEX64)
        return batch_bind_t<float>(p, adjoint, ins, ld_in, outs, ld_out, workspace,
                                   as_stream(stream));
    return batch_bind_t<double>(p, adjoint, ins, ld_in, outs, ld_out, workspace,
                                as_stream(stream));
}

template <typename T, int LOGN, bool CONV>
static fm_status batch_launch(const fm_batch_plan *p, const BatchDir &d, const BatchGroup &g,
                              hipStream_t s) {
    constexpr int N = 1 << LOGN;
    const size_t lds = size_t(tile_batch<T>(LOGN)) * (N + (N >> 5) + 1) * sizeof(cpx<T>);
    static bool attr = false;
    if (!attr) {
        (void)hipFuncSetAttribute((const void *)batch_pass_kernel<T, LOGN, CONV, false>,
                                  hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
        (void)hipFuncSetAttribute((const void *)batch_pass_kernel<T, LOGN, CONV, true>,
                                  hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
        attr = true;
    }
    int64_t grid = int64_t(p->ncu) * p->grid_per_cu;
    if (grid > g.nitems) grid = g.nitems;
    ProfToken tok;
    prof_begin(s, &tok);
    constexpr int NT = PassCfg<T, LOGN>::NT;
    if (g.idx32)
        hipLaunchKernelGGL((batch_pass_kernel<T, LOGN, CONV, true>), dim3(unsigned(grid)), dim3(NT),
                           lds, s, (const BatchDesc<T> *)d.descs, (const int2 *)g.items, g.nitems);
    else
        hipLaunchKernelGGL((batch_pass_kernel<T, LOGN, CONV, false>), dim3(unsigned(grid)),
                           dim3(NT), lds, s, (const BatchDesc<T> *)d.descs,
                           (const int2 *)g.items, g.nitems);
    FM_CHECK_LAUNCH();
    if (tok.active) {
        char name[80];
        snprintf(name, sizeof(name), "batch_pass_kernel<%s,%d,%s>",
                 sizeof(T) == 4 ? "float" : "double", LOGN, CONV ? "conv" : "fft");
        prof_end(s, &tok, name, g.bytes, 0.0);
    }
    return FM_SUCCESS;
}

template <typename T>
static fm_status batch_launch_group(const fm_batch_plan *p, const BatchDir &d,
                                    const BatchGroup &g, hipStream_t s);

// Groups of one phase touch different operators' data, so they may run at
// once: the largest stays on the caller's stream, the others go round-robin
// to the side streams, which fork from and join back into the caller's
// stream (event record / wait only, so the sequence is graph-capturable).
// Measured on the 256-operator uneven batch: 3.10 -> 2.81 ms per step.
template <typename T>
static fm_status batch_apply_t(const fm_batch_plan *p, const BatchDir &d, hipStream_t s) {
    // FASTMAT_AMD_BATCH_STREAMS=1 keeps every group on the caller's stream
    const bool concurrent = p->has_side && knob("FASTMAT_AMD_BATCH_STREAMS", 2) > 1;
    size_t q0 = 0;
    while (q0 < d.groups.size()) {
        size_t q1 = q0;
        while (q1 < d.groups.size() && d.groups[q1].phase == d.groups[q0].phase) ++q1;
        if (!concurrent || q1 - q0 == 1) {
            for (size_t q = q0; q < q1; ++q) {
                fm_status st = batch_launch_group<T>(p, d, d.groups[q], s);
                if (st) return st;
            }
        } else {
            size_t big = q0;
            for (size_t q = q0; q < q1; ++q)
                if (d.groups[q].bytes > d.groups[big].bytes) big = q;
            FM_CHECK_HIP(hipEventRecord(p->fork_ev, s));
            int used = 0;
            for (int k = 0; k < fm_batch_plan::kSide && k < int(q1 - q0) - 1; ++k) {
                FM_CHECK_HIP(hipStreamWaitEvent(p->side[k], p->fork_ev, 0));
                ++used;
            }
            int rr = 0;
            for (size_t q = q0; q < q1; ++q) {
                hipStream_t sq = s;
                if (q != big) {
                    sq = p->side[rr % used];
                    ++rr;
                }
                fm_status st = batch_launch_group<T>(p, d, d.groups[q], sq);
                if (st) return st;
            }
            for (int k = 0; k < used; ++k) {
                FM_CHECK_HIP(hipEventRecord(p->join_ev[k], p->side[k]));
                FM_CHECK_HIP(hipStreamWaitEvent(s, p->join_ev[k], 0));
            }
        }
        q0 = q1;
    }
    return FM_SUCCESS;
}

template <typename T>
static fm_status batch_launch_group(const fm_batch_plan *p, const BatchDir &d,
                                    const BatchGroup &g, hipStream_t s) {
    {
        fm_status st = FM_ERR_VALUE;
#define FM_BATCH_CASE(L)                                                                 \
    case L:                                                                              \
        st = g.conv ? batch_launch<T, L, true>(p, d, g, s) : batch_launch<T, L, false>(p, d, g, s); \
        break;
        switch (g.logN) {
            FM_BATCH_CASE(0) FM_BATCH_CASE(1) FM_BATCH_CASE(2) FM_BATCH_CASE(3)
            FM_BATCH_CASE(4) FM_BATCH_CASE(5) FM_BATCH_CASE(6) FM_BATCH_CASE(7)
            FM_BATCH_CASE(8) FM_BATCH_CASE(9) FM_BATCH_CASE(10) FM_BATCH_CASE(11)
            FM_BATCH_CASE(12)
        case 13:
            if constexpr (sizeof(T) == 4)
                st = g.conv ? batch_launch<T, 13, true>(p, d, g, s)
                            : batch_launch<T, 13, false>(p, d, g, s);
            break;
        default:
            break;
        }
#undef FM_BATCH_CASE
        if (st) return st;
    }
    return FM_SUCCESS;
}

extern "C" fm_status fm_batch_apply(const fm_batch_plan *p, int adjoint, void *stream) {
    FM_REQUIRE(p != nullptr, FM_ERR_VALUE, "plan is NULL");
    const BatchDir &d = p->dir[adjoint ? 1 : 0];
    FM_REQUIRE(d.bound, FM_ERR_VALUE, "batch direction %d is not bound to buffers", adjoint);
    if (p->prec == FM_COMPLEX64) return batch_apply_t<float>(p, d, as_stream(stream));
    return batch_apply_t<double>(p, d, as_stream(stream));
}

// ------------------------------------------------- batched generator gradient
// For every operator i of the batch (the per-operator fm_conv_generator_grad
// at once): xh = FFT(x'), gh = FFT(g') over its k_i columns, acc = sum_c
// gh conj(xh), r = unnormalised inverse FFT of acc, folded onto the generator
// and scaled by 1/L.  The FFTs run as two batched programs on the batch
// engine (descriptor table, persistent kernels per length class); the sum
// and the fold are one launch each over all operators.
namespace fm {
template <typename T> struct GradEntry {
    const cpx<T> *gh, *xh, *r;
    cpx<T> *acc, *out;
    int64_t L, k, head, tail, gen_len;
    int kind;
};

template <typename T>
__global__ void batch_corr_reduce_kernel(const GradEntry<T> *__restrict__ tab) {
    const GradEntry<T> e = tab[blockIdx.y];
    const int64_t q = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
    if (q >= e.L) return;
    cpx<T> acc = mk(T(0), T(0));
    for (int64_t c = 0; c < e.k; ++c) acc = acc + cmulc(e.gh[c * e.L + q], e.xh[c * e.L + q]);
    e.acc[q] = acc;
}

template <typename T>
__global__ void batch_fold_kernel(const GradEntry<T> *__restrict__ tab) {
    const GradEntry<T> e = tab[blockIdx.y];
    const int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
    if (i >= e.gen_len) return;
    cpx<T> v;
    if (e.kind == 0) {
        v = e.r[i];
        if (e.tail > 0 && i >= 1) v = v + e.r[e.L - e.tail + i - 1];
    } else {
        v = i < e.head ? e.r[i] : e.r[e.L - e.tail + (i - e.head)];
    }
    e.out[i] = scale(v, T(1.0 / double(e.L)));
}

static int64_t grad_ws_entry(const fm_conv_plan *c, int64_t k, int64_t *parts) {
    const int64_t e = int64_t(esize(c->prec)), L = c->L;
    const bool multi = c->geo_fft.passes > 1;
    auto up = [](int64_t b) { return (b + 255) / 256 * 256; };
    parts[0] = up(e * L * k);                 // xh
    parts[1] = up(e * L * k);                 // gh
    parts[2] = multi ? up(e * L * k) : 0;     // four-step intermediate of FFT(x')
    parts[3] = multi ? up(e * L * k) : 0;     // ... of FFT(g')
    parts[4] = up(e * L);                     // acc
    parts[5] = up(e * L);                     // r
    parts[6] = multi ? up(e * L) : 0;         // intermediate of the inverse
    int64_t t = 0;
    for (int q = 0; q < 7; ++q) t += parts[q];
    return t;
}

template <typename T>
static fm_status batch_grad_bind_t(fm_batch_plan *p, const void *const *xs, const int64_t *ldx,
                                   const void *const *gs, const int64_t *ldg,
                                   void *const *grads, char *ws, hipStream_t s) {
    std::vector<std::vector<PassSpec<T>>> fwd(2 * size_t(p->count)), inv(p->count);
    std::vector<GradEntry<T>> tab(p->count);
    int64_t off = 0, maxL = 1, maxgen = 1;
    for (int i = 0; i < p->count; ++i) {
        const fm_conv_plan *c = p->plans[i];
        const int64_t k = p->ks[i], L = c->L;
        FM_REQUIRE(xs[i] && gs[i] && grads[i], FM_ERR_VALUE, "batch entry %d: NULL buffer", i);
        FM_REQUIRE(ldx[i] >= c->cols && ldg[i] >= c->rows, FM_ERR_VALUE,
                   "batch entry %d: leading dimension too small", i);
        int64_t part[7];
        const int64_t tot = grad_ws_entry(c, k, part);
        char *b = ws + off;
        cpx<T> *xh = (cpx<T> *)b, *gh = (cpx<T> *)(b + part[0]);
        char *tx = b + part[0] + part[1], *tg = tx + part[2];
        cpx<T> *acc = (cpx<T> *)(tg + part[3]), *r = (cpx<T> *)((char *)acc + part[4]);
        char *tr = (char *)r + part[5];
        off += tot;
        PipeIO io;
        memset(&io, 0, sizeof(io));
        io.x = xs[i]; io.ldx = ldx[i]; io.y = xh; io.ldy = L; io.k = k;
        io.n_in = c->cols; io.m_out = L; io.work = part[2] ? tx : nullptr;
        fm_status st = build_pipeline<T>(c->geo_fft, io, &fwd[2 * size_t(i)]);
        if (st) return st;
        io.x = gs[i]; io.ldx = ldg[i]; io.y = gh; io.n_in = c->rows; io.work = part[3] ? tg : nullptr;
        st = build_pipeline<T>(c->geo_fft, io, &fwd[2 * size_t(i) + 1]);
        if (st) return st;
        PipeIO iv;
        memset(&iv, 0, sizeof(iv));
        iv.x = acc; iv.ldx = L; iv.y = r; iv.ldy = L; iv.k = 1; iv.n_in = L; iv.m_out = L;
        iv.conj_io = 1; iv.work = part[6] ? tr : nullptr;
        st = build_pipeline<T>(c->geo_fft, iv, &inv[i]);
        if (st) return st;
        tab[i] = GradEntry<T>{gh, xh, r, acc, (cpx<T> *)grads[i], L, k, c->head, c->tail,
                              c->gen_len, c->kind};
        if (L > maxL) maxL = L;
        if (c->gen_len > maxgen) maxgen = c->gen_len;
    }
    fm_status st = batch_build_dir<T>(p->gdir[0], fwd, s);
    if (st) return st;
    st = batch_build_dir<T>(p->gdir[1], inv, s);
    if (st) return st;
    if (p->gtab) (void)hipFree(p->gtab);
    p->gtab = nullptr;
    FM_CHECK_HIP(hipMalloc(&p->gtab, sizeof(GradEntry<T>) * tab.size()));
    FM_CHECK_HIP(hipMemcpyAsync(p->gtab, tab.data(), sizeof(GradEntry<T>) * tab.size(),
                                hipMemcpyHostToDevice, s));
    FM_CHECK_HIP(hipStreamSynchronize(s));
    p->g_maxL = maxL;
    p->g_maxgen = maxgen;
    return FM_SUCCESS;
}

template <typename T>
static fm_status batch_grad_apply_t(const fm_batch_plan *p, hipStream_t s) {
    fm_status st = batch_apply_t<T>(p, p->gdir[0], s);
    if (st) return st;
    hipLaunchKernelGGL(batch_corr_reduce_kernel<T>,
                       dim3(unsigned(ceil_div(p->g_maxL, 256)), unsigned(p->count)), dim3(256), 0,
                       s, (const GradEntry<T> *)p->gtab);
    FM_CHECK_LAUNCH();
    st = batch_apply_t<T>(p, p->gdir[1], s);
    if (st) return st;
    hipLaunchKernelGGL(batch_fold_kernel<T>,
                       dim3(unsigned(ceil_div(p->g_maxgen, 256)), unsigned(p->count)), dim3(256),
                       0, s, (const GradEntry<T> *)p->gtab);
    FM_CHECK_LAUNCH();
    return FM_SUCCESS;
}
}  // namespace fm

extern "C" fm_status fm_batch_grad_workspace_size(const fm_batch_plan *p, int64_t *bytes) {
    FM_REQUIRE(p && bytes, FM_ERR_VALUE, "NULL argument");
    int64_t t = 0, part[7];
    for (int i = 0; i < p->count; ++i) t += grad_ws_entry(p->plans[i], p->ks[i], part);
    *bytes = t;
    return FM_SUCCESS;
}

extern "C" fm_status fm_batch_grad_bind(fm_batch_plan *p, const void *const *xs,
                                        const int64_t *ldx, const void *const *gs,
                                        const int64_t *ldg, void *const *grads, void *workspace,
                                        int64_t workspace_bytes, void *stream) {
    FM_REQUIRE(p && xs && ldx && gs && ldg && grads && workspace, FM_ERR_VALUE, "NULL argument");
    FM_REQUIRE(p->count <= 65535, FM_ERR_VALUE, "batch gradient: more than 65535 operators");
    int64_t need = 0;
    fm_batch_grad_workspace_size(p, &need);
    FM_REQUIRE(workspace_bytes >= need, FM_ERR_VALUE,
               "workspace of %lld bytes is smaller than the %lld required",
               (long long)workspace_bytes, (long long)need);
    if (p->prec == FM_COMPLEX64)
        return batch_grad_bind_t<float>(p, xs, ldx, gs, ldg, grads, (char *)workspace,
                                        as_stream(stream));
    return batch_grad_bind_t<double>(p, xs, ldx, gs, ldg, grads, (char *)workspace,
                                     as_stream(stream));
}

extern "C" fm_status fm_batch_grad_apply(const fm_batch_plan *p, void *stream) {
    FM_REQUIRE(p != nullptr, FM_ERR_VALUE, "plan is NULL");
    FM_REQUIRE(p->gdir[0].bound && p->gdir[1].bound && p->gtab, FM_ERR_VALUE,
               "batch gradient used before fm_batch_grad_bind");
    if (p->prec == FM_COMPLEX64) return batch_grad_apply_t<float>(p, as_stream(stream));
    return batch_grad_apply_t<double>(p, as_stream(stream));
}

extern "C" fm_status fm_batch_plan_destroy(fm_batch_plan *p) {
    if (!p) return FM_SUCCESS;
    batch_free_dir(p->dir[0]);
    batch_free_dir(p->dir[1]);
    batch_free_dir(p->gdir[0]);
    batch_free_dir(p->gdir[1]);
    if (p->gtab) (void)hipFree(p->gtab);
    for (int q = 0; q < fm_batch_plan::kSide; ++q) {
        if (p->side[q]) (void)hipStreamDestroy(p->side[q]);
        if (p->join_ev[q]) (void)hipEventDestroy(p->join_ev[q]);
    }
    if (p->fork_ev) (void)hipEventDestroy(p->fork_ev);
    delete p;
    return FM_SUCCESS;
}
