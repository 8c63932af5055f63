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

#include "fm_fft_plan.hpp"

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
        const double own = (in + out + spec) * esz;
        prof_end(stream, &tok, name, a.alg_bytes > 0 ? a.alg_bytes : own, flops, own);
    }
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

// Tile-shape knob: environment override (used by the tuning sweeps), else
// the built-in default.
int knob(const char *name, int dflt) {
    const char *v = getenv(name);
    if (!v || !*v) return dflt;
    int x = atoi(v);
    return x > 0 ? x : dflt;
}

fm_status make_geometry(int prec, int logL, bool conv, hipStream_t stream, Geometry *g) {
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

// A second stream for chunked multi-pass applies (see Lanes).
void lanes_create(Lanes *l) {
    l->ok = hipStreamCreateWithFlags(&l->side, hipStreamNonBlocking) == hipSuccess &&
            hipEventCreateWithFlags(&l->fork, hipEventDisableTiming) == hipSuccess &&
            hipEventCreateWithFlags(&l->join, hipEventDisableTiming) == hipSuccess;
}

void lanes_destroy(Lanes *l) {
    if (l->side) (void)hipStreamDestroy(l->side);
    if (l->fork) (void)hipEventDestroy(l->fork);
    if (l->join) (void)hipEventDestroy(l->join);
    *l = Lanes();
}

int chunk_lanes(const Lanes *l) {
    if (!l || !l->ok) return 1;
    return knob("FASTMAT_AMD_CHUNK_LANES", 2) >= 2 ? 2 : 1;
}

// Columns per chunk of a multi-pass transform.  With two lanes (a plan that
// owns a second stream) the default budget is 64 MB of intermediate per
// chunk: chunks alternate between the lanes, so one chunk's passes overlap
// the next one's while its intermediate is still in the memory-side cache
// (measured on Fourier 2^20 and 1,000,003 x 64 columns, complex64: 4.56 ->
// 4.36 ms per forward + adjoint step; 32 MB: 5.07, 128 MB: 4.76, 192 MB:
// 4.59).  Without lanes all columns go in one chunk (serial 96 MB chunks:
// 5.30 ms).  FASTMAT_AMD_CHUNK_MB overrides the budget (0: one chunk).
int64_t chunk_columns(const Geometry &g, int64_t k, size_t esz, const Lanes *l) {
    if (g.passes == 1) return k;
    const char *env = getenv("FASTMAT_AMD_CHUNK_MB");
    int64_t mb = chunk_lanes(l) > 1 ? 64 : 0;
    if (env && *env) mb = atoll(env);
    if (mb <= 0) return k;
    const int64_t budget = mb << 20;
    const int64_t per_col = (int64_t(1) << g.logL) * int64_t(esz);
    int64_t c = budget / per_col;
    if (c < 1) c = 1;
    return c < k ? c : k;
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
fm_status spectrum_from(cpx<double> *t, const Geometry &g, int prec, void *D,
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
    Lanes lanes;  // second stream for chunked multi-pass applies
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
    if (p->geo.passes > 1) lanes_create(&p->lanes);
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
    const int64_t cc = chunk_columns(p->geo, k, esize(p->prec), &p->lanes);
    const int nl = cc < k ? chunk_lanes(&p->lanes) : 1;
    *bytes = p->geo.passes > 1 ? int64_t(esize(p->prec)) * p->L * cc * nl : 0;
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
    return run_pipeline_any(p->geo, io, as_stream(stream), &p->lanes);
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
    lanes_destroy(&p->lanes);
    delete p;
    return FM_SUCCESS;
}

// ======================================================================
//                          Circulant / Toeplitz
// ======================================================================

static fm_status conv_plan_create(int prec, int kind, int64_t m, int64_t n, int64_t L,
                                  int64_t head, int64_t tail, int64_t tail_src, const void *gen,
                                  int64_t gen_len, hipStream_t s, fm_conv_plan **out) {
    auto *p = new fm_conv_plan();
    memset(p, 0, sizeof(*p));
    hipGetDevice(&p->device);
    p->prec = prec; p->kind = kind; p->rows = m; p->cols = n; p->L = L;
    p->gen_len = gen_len;
    p->head = head;
    p->tail = tail;
    p->tail_src = tail_src;
    const int logL = ilog2(L);
    fm_status st = make_geometry(prec, logL, true, s, &p->geo);
    if (!st) st = make_geometry(prec, logL, false, s, &p->geo_fft);
    cpx<double> *t = nullptr;
    if (!st && hipMalloc(&p->D, esize(prec) * size_t(L)) != hipSuccess) st = FM_ERR_ALLOC;
    if (!st && hipMalloc(&t, 16 * size_t(L)) != hipSuccess) st = FM_ERR_ALLOC;
    if (st == FM_ERR_ALLOC) set_error("device allocation for convolution plan failed");
    if (!st) {
        // construction: the spectrum of the host-given generator in FP64
        hipLaunchKernelGGL((embed_gen_kernel<cpx<double>, double>), grid1d(L), 256, 0, s, t,
                           (const cpx<double> *)gen, L, head, tail, tail_src);
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
    const int64_t L = is_pow2(n) ? n : (int64_t(1) << ilog2(2 * n - 1));
    // t' = c (L == n) or [c, 0, ..., 0, c[1:]] (L >= 2n - 1)
    const int64_t tail = L == n ? 0 : n - 1;
    return conv_plan_create(precision, 0, n, n, L, n, tail, 1, generator, n, as_stream(stream),
                            plan);
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
    return conv_plan_create(precision, 1, m, n, L, m, n - 1, m, generator, d, as_stream(stream),
                            plan);
}

// A private copy of a plan (same geometry, its own spectrum buffer), e.g. for
// a caller that rewrites the spectrum with fm_conv_set_generator while the
// original keeps its own.
extern "C" fm_status fm_conv_plan_clone(const fm_conv_plan *p, void *stream, fm_conv_plan **out) {
    FM_REQUIRE(p != nullptr && out != nullptr, FM_ERR_VALUE, "NULL argument");
    *out = nullptr;
    auto *q = new fm_conv_plan(*p);
    q->D = nullptr;
    if (hipMalloc(&q->D, esize(p->prec) * size_t(p->L)) != hipSuccess) {
        delete q;
        set_error("device allocation for a plan copy failed");
        return FM_ERR_ALLOC;
    }
    hipStream_t s = as_stream(stream);
    if (hipMemcpyAsync(q->D, p->D, esize(p->prec) * size_t(p->L), hipMemcpyDeviceToDevice, s) !=
            hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess) {
        (void)hipFree(q->D);
        delete q;
        set_error("copying the spectrum of a plan failed");
        return FM_ERR_HIP;
    }
    *out = q;
    return FM_SUCCESS;
}

namespace fm {
static int64_t align256(int64_t b) { return (b + 255) / 256 * 256; }

// Workspace of fm_conv_set_generator: the embedded generator (L elements)
// and, for a four-step, the intermediate between its two passes.
static int64_t gen_ws_bytes(const fm_conv_plan *p) {
    const int64_t e = int64_t(esize(p->prec));
    return align256(e * p->L) + (p->geo_fft.passes > 1 ? align256(e * p->L) : 0);
}

template <typename T, typename G>
static fm_status set_generator_t(fm_conv_plan *p, const G *gen, char *ws, hipStream_t s) {
    const int64_t L = p->L;
    cpx<T> *t = (cpx<T> *)ws;
    cpx<T> *tmp = p->geo_fft.passes > 1 ? (cpx<T> *)(ws + align256(int64_t(sizeof(cpx<T>)) * L))
                                        : nullptr;
    hipLaunchKernelGGL((embed_gen_kernel<G, T>), grid1d(L), 256, 0, s, t, gen, L, p->head, p->tail,
                       p->tail_src);
    FM_CHECK_LAUNCH();
    PassList<T> passes;
    fm_status st = build_spectrum_pipeline<T>(p->geo_fft, t, tmp, (cpx<T> *)p->D, &passes);
    if (st) return st;
    for (int q = 0; q < passes.size(); ++q) {
        fm_status st = launch_tile<T>(passes[q].a, passes[q].conv, s);
        if (st) return st;
    }
    return FM_SUCCESS;
}

template <typename T>
static fm_status set_generator_any(fm_conv_plan *p, int gen_dtype, const void *gen, char *ws,
                                   hipStream_t s) {
    switch (gen_dtype) {
    case FM_FLOAT32: return set_generator_t<T>(p, (const float *)gen, ws, s);
    case FM_FLOAT64: return set_generator_t<T>(p, (const double *)gen, ws, s);
    case FM_COMPLEX64: return set_generator_t<T>(p, (const cpx<float> *)gen, ws, s);
    case FM_COMPLEX128: return set_generator_t<T>(p, (const cpx<double> *)gen, ws, s);
    default: break;
    }
    set_error("generator element type %d: expected float32, float64, complex64 or complex128",
              gen_dtype);
    return FM_ERR_TYPE;
}
}  // namespace fm

extern "C" fm_status fm_conv_set_generator_workspace_size(const fm_conv_plan *p, int64_t *bytes) {
    FM_REQUIRE(p != nullptr && bytes != nullptr, FM_ERR_VALUE, "NULL argument");
    *bytes = p->kind == 2 ? 0 : gen_ws_bytes(p);
    return FM_SUCCESS;
}

extern "C" fm_status fm_conv_set_generator(fm_conv_plan *p, int gen_dtype, const void *generator,
                                           void *workspace, int64_t workspace_bytes,
                                           void *stream) {
    FM_REQUIRE(p != nullptr, FM_ERR_VALUE, "plan is NULL");
    FM_REQUIRE(p->kind != 2, FM_ERR_NOT_IMPLEMENTED,
               "the spectrum of a two-level convolution plan is set at creation");
    FM_REQUIRE(generator != nullptr && workspace != nullptr, FM_ERR_VALUE, "NULL data pointer");
    FM_REQUIRE(workspace_bytes >= gen_ws_bytes(p), FM_ERR_VALUE,
               "workspace of %lld bytes is smaller than the %lld required",
               (long long)workspace_bytes, (long long)gen_ws_bytes(p));
    if (p->prec == FM_COMPLEX64)
        return set_generator_any<float>(p, gen_dtype, generator, (char *)workspace,
                                        as_stream(stream));
    return set_generator_any<double>(p, gen_dtype, generator, (char *)workspace, as_stream(stream));
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
                                     chunk_columns(p->geo, k, esize(p->prec), nullptr)
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