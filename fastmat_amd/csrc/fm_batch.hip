// Uneven batch of Circulant / Toeplitz operators on the FFT tile engine
// (pattern (b)), its batched generator gradient (pattern (a)) and its
// batched generator-to-spectrum program.
//
// No single reference call does this; it replaces the per-operator loop of
// forward()/backward() calls a fastmat user writes (Matrix.pyx:1844-2007 once
// per operator, each one Partial(Product(F.H, Diag, F)), Partial.pyx:268-294,
// Product.pyx:205-240).
//
// Programs.  Every operator's transform is a list of passes (one fused pass,
// or the passes A, B, C of the four-step).  A program is a list of launches;
// each launch is one persistent kernel of one (kind, length) class that walks
// a device table of (descriptor, tile) work items grid-strided, largest
// first.  Launches go to lanes (the caller's stream and side streams forked
// from it and joined back with events, all default priority); launches of
// one lane run in order, lanes run concurrently.  Nothing waits inside a
// kernel, and an apply only launches, so it is graph-capturable.
//
// Schedule of a direction (memory-side-cache grouping).  The four-step
// intermediates of all operators together (1 GB per direction for the
// 256-operator uneven batch) exceed the 256 MB memory-side cache (MALL), so
// passes A, B and C of all operators in turn would stream them through HBM
// four times.  Instead the multi-pass operators, longest first, are cut into
// groups whose intermediates fit a budget (group_mb, default 96 MB); a group
// runs A, B, C back to back, so pass B and C read what the pass before wrote
// while it is still in the cache, and consecutive groups alternate between
// `lanes` lanes so one group's passes overlap the next group's.  The one-pass
// operators run on a lane of their own beside them.  The schedule (grid,
// group budget, lanes) is set by fm_batch_set_schedule; the autotuner in
// fastmat_amd/batch.py times candidate schedules on the real batch.
//
// Workspace.  Each direction owns two intermediates per multi-pass operator:
// T1 = pass A's output (the column transforms with the four-step twiddle,
// i.e. the first half of a plain FFT of the input) and T2 = pass B's output.
// Keeping T1 lets the generator gradient reuse the forward's column
// transforms of x and the adjoint's of g.
#include <algorithm>
#include <deque>
#include <map>
#include <tuple>
#include <vector>

#include "fm_fft_plan.hpp"

using namespace fm;

namespace fm {

template <typename T> struct BatchDesc {
    TileArgs<T> a;
    int conv;
};

// Generator-gradient correlation of one operator (see corr_tile).
enum CorrMode {
    CORR_FULL = 0,     // sum over the columns, inverse transform, store
    CORR_PARTIAL = 1,  // sum over a chunk of columns, store the spectrum
    CORR_SUM = 2       // sum stored partial spectra, inverse transform, store
};
template <typename T> struct CorrDesc {
    TileArgs<T> a;       // x side load (column 0) and the store
    const cpx<T> *g;     // g side, column 0
    int64_t x_col, g_col;
    int64_t g_valid;     // zero-pad bound of the g side
    int64_t k;           // columns (or partial spectra) summed over
    int mode;            // CorrMode
};

// One persistent kernel per (length, conv) class: grid-strided over the
// class's (descriptor, tile) items.
template <typename T, int LOGN, bool CONV, bool I32>
__global__ __launch_bounds__((PassCfg<T, LOGN>::NT)) void batch_pass_kernel(
    const BatchDesc<T> *__restrict__ descs, const int2 *__restrict__ items, int64_t nitems,
    int64_t out_shift, const T *__restrict__ out_mul) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    cpx<T> *lds = reinterpret_cast<cpx<T> *>(smem_raw);
    for (int64_t it = blockIdx.x; it < nitems; it += gridDim.x) {
        const int2 w = items[it];
        TileArgs<T> a = descs[w.x].a;
        if (a.redirect && (out_shift != 0 || out_mul != nullptr)) {
            // redirected apply: the output block moved by out_shift bytes,
            // scaled by the device scalar *out_mul
            a.y = reinterpret_cast<cpx<T> *>(reinterpret_cast<char *>(a.y) + out_shift);
            if (out_mul) a.out_scale *= *out_mul;
        }
        fft_pass_body<T, LOGN, PassCfg<T, LOGN>::NT, CONV, I32, PassCfg<T, LOGN>::LE>(a, w.y, lds);
        __syncthreads();  // the next item reuses the LDS tile
    }
}

// Elements per thread of the correlation tiles: three transforms' worth of
// registers are live (x, g and the running sum), so 8 (a quarter of the
// complex64 pass kernels'), which leaves room for two or more waves per
// SIMD; complex64 transforms of 8192 points hold 16, as a tile has at most
// 512 threads.
template <typename T> constexpr int corr_loge(int logN) {
    return sizeof(T) == 4 && logN > 12 ? 4 : 3;
}

// One tile of the generator-gradient correlation: for the tile's B
// sequences (rows of a four-step, or the one sequence of a one-pass
// transform), over the k columns c: X = FFT(x_c rows), G = FFT(g_c rows),
// acc += G conj(X); then (CORR_FULL) the inverse transform of acc (radices
// reversed, from the registers) and the store of `a` (four-step twiddle
// W_L^{-n2 k1} for rows), or (CORR_PARTIAL) the store of acc itself, a
// partial spectrum in natural order.  CORR_SUM loads k partial spectra,
// sums them and stores their inverse transform.  In four-step terms: x_c rows
// are the forward's pass-A output, so FFT(x_c rows) are bins k1 + N1 k2 of
// x_c's length-L transform, and the stored rows followed by the column
// inverse (pass C) give the unnormalised inverse transform of
// sum_c G_c conj(X_c), i.e. L times the correlation
// sum_c sum_i g_c[i] conj(x_c[(i - d) mod L]).
template <typename T, int LOGN, int NT, bool I32>
__device__ __forceinline__ void corr_tile(const CorrDesc<T> &d, int64_t tile, cpx<T> *lds) {
    constexpr int LE = corr_loge<T>(LOGN);
    using S = TileShape<T, LOGN, NT, LE>;
    using P = typename S::P;
    constexpr int E = S::E;
    constexpr int LOGE = S::LOGE;
    constexpr int LOGB = S::LOGB;
    // radix of the last stage of a forward (natural-radix) transform, and of
    // the radix-reversed inverse
    constexpr int LOGR_FWD = P::S > 0 ? P::bits(P::S - 1) : 0;
    constexpr int LOGR_REV = P::S > 0 ? P::bits(0) : 0;
    cpx<T> acc[E];
#pragma unroll
    for (int q = 0; q < E; ++q) acc[q] = mk(T(0), T(0));
    TileArgs<T> ax = d.a;
    if (d.mode == CORR_SUM) {
        for (int64_t c = 0; c < d.k; ++c) {
            ax.x = d.a.x + c * d.x_col;
            cpx<T> v[E];
            fft_tile_load<T, LOGN, NT, I32, LE>(ax, tile, v);
#pragma unroll
            for (int q = 0; q < E; ++q) acc[q] = acc[q] + v[q];
        }
        if constexpr (P::S > 0) fft_all<T, LOGN, LOGE, +1, NT, LOGB, 0, 1>(acc, lds, d.a.tw, false);
        fft_tile_store<T, LOGN, NT, LOGR_FWD, I32, LE>(d.a, tile, acc, lds, false);
        return;
    }
    TileArgs<T> ag = d.a;
    ag.n_valid = d.g_valid;
    for (int64_t c = 0; c < d.k; ++c) {
        ax.x = d.a.x + c * d.x_col;
        ag.x = d.g + c * d.g_col;
        cpx<T> vx[E], vg[E];
        fft_tile_load<T, LOGN, NT, I32, LE>(ax, tile, vx);
        fft_tile_load<T, LOGN, NT, I32, LE>(ag, tile, vg);
        if constexpr (P::S > 0) {
            fft_all<T, LOGN, LOGE, -1, NT, LOGB, 0, 1>(vx, lds, d.a.tw, false);
            fft_all<T, LOGN, LOGE, -1, NT, LOGB, 0, 1>(vg, lds, d.a.tw, false);
        }
#pragma unroll
        for (int q = 0; q < E; ++q) acc[q] = acc[q] + cmulc(vg[q], vx[q]);
    }
    if (d.mode == CORR_PARTIAL) {
        fft_tile_store<T, LOGN, NT, LOGR_FWD, I32, LE>(d.a, tile, acc, lds, false);
        return;
    }
    if constexpr (P::S > 0)
        fft_all<T, LOGN, LOGE, +1, NT, LOGB, 0, 1, true>(acc, lds, d.a.tw, false);
    fft_tile_store<T, LOGN, NT, LOGR_REV, I32, LE>(d.a, tile, acc, lds, false);
}

template <typename T, int LOGN, int NT, bool I32>
__global__ __launch_bounds__(NT) void batch_corr_kernel(const CorrDesc<T> *__restrict__ descs,
                                                        const int2 *__restrict__ items,
                                                        int64_t nitems) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    cpx<T> *lds = reinterpret_cast<cpx<T> *>(smem_raw);
    for (int64_t it = blockIdx.x; it < nitems; it += gridDim.x) {
        const int2 w = items[it];
        const CorrDesc<T> dsc = descs[w.x];
        corr_tile<T, LOGN, NT, I32>(dsc, w.y, lds);
        __syncthreads();
    }
}

template <typename T> constexpr int batch_max_log() { return sizeof(T) == 4 ? 13 : 12; }

// Threads of a correlation tile: rows of a four-step get 8 sequences per
// tile (at least 64 threads), one-pass transforms one sequence per tile.
constexpr int corr_threads(int logN, int loge, bool rows) {
    const int per_seq = logN > loge ? (1 << (logN - loge)) : 1;
    const int nt = rows ? (per_seq * 8 < 256 ? 256 : per_seq * 8) : per_seq;
    return nt < 64 ? 64 : (nt > 512 ? 512 : nt);
}

// Columns per correlation tile of a one-pass transform: an operator's
// columns are split into chunks that run as separate tiles (partial spectra),
// then summed by one more tile.
constexpr int kCorrChunk = 4;

enum LaunchKind { LK_FFT = 0, LK_CONV = 1, LK_CORR = 2 };

// One launch of a program: the work items of one kernel class on one lane.
struct Launch {
    int kind, logN, nt, lane;
    int early;  // from a pass group that needs no spectrum
    int idx32;  // every descriptor of the launch fits 32-bit index arithmetic
    void *items;
    int64_t nitems;
    double bytes;       // algorithmic bytes, for the library's profiler
    double pass_bytes;  // what the launch reads and writes
};

struct Program {
    void *descs = nullptr;   // BatchDesc<T>[]
    void *cdescs = nullptr;  // CorrDesc<T>[]
    std::vector<Launch> launches;
    int lanes = 1;
    bool bound = false;
};

// Builder of a program: descriptors, and "pass groups" (a lane and the
// descriptors that form one step of that lane).  Inside a pass group the
// descriptors are split by kernel class, and the classes are issued largest
// first.
template <typename T> struct ProgBuilder {
    std::vector<BatchDesc<T>> descs;
    std::vector<CorrDesc<T>> cdescs;
    struct Ref {
        int corr;
        int index;
    };
    struct PassGroup {
        int lane;
        std::vector<Ref> refs;
        bool early;  // needs no spectrum: may run while the spectra are formed
    };
    std::deque<PassGroup> groups;  // references stay valid as groups are added
    int lanes = 1;

    int add(const PassSpec<T> &ps) {
        TileArgs<T> a = ps.a;
        descs.push_back(BatchDesc<T>{a, ps.conv ? 1 : 0});
        return int(descs.size()) - 1;
    }
    int add_corr(const CorrDesc<T> &cd) {
        cdescs.push_back(cd);
        return int(cdescs.size()) - 1;
    }
    PassGroup &group(int lane, bool early = false) {
        groups.push_back(PassGroup{lane, {}, early});
        if (lane + 1 > lanes) lanes = lane + 1;
        return groups.back();
    }
};

}  // namespace fm

struct fm_batch_plan {
    int device;
    int prec;
    int count;
    std::vector<const fm_conv_plan *> plans;
    std::vector<int64_t> ks;
    // per direction and operator: offsets of T1 and T2 in the workspace
    std::vector<int64_t> t1_off[2], t2_off[2];
    int64_t ws_bytes;
    // schedule
    int grid_per_cu;
    int group_mb;
    int lanes;
    int ncu;
    // bound programs: 0 forward, 1 adjoint
    Program dir[2];
    void *ws_bound[2];                       // workspace of the bound direction
    std::vector<const void *> ins_bound[2];  // its input blocks
    std::vector<int64_t> ld_bound[2];
    // generator gradient: pass A of x and of g (into the directions' T1),
    // the correlations, pass C of the four-step rows, the fold
    Program gpassA, gcorr, gpassC;
    void *gtab;
    int64_t g_maxgen;
    void *g_ws;                               // the workspace the gradient reads T1 from
    std::vector<const void *> g_xs, g_gs;     // its x and g blocks
    std::vector<int64_t> g_ldx, g_ldg;
    bool g_bound;
    // generator-to-spectrum program: one embedding launch over all
    // operators, then every operator's spectrum FFT
    Program sprog;
    void *stab;
    int s_gen_dtype;
    int64_t s_maxL;
    // residual mode of the forward direction (fm_batch_bind_residual): tile
    // partial sums of |A x - t|^2
    double *res_part;
    int64_t n_res;
    bool res_bound;
    // both directions bound by fm_batch_bind_lsq (the adjoint starts at its
    // row pass; its column transforms of the residual come from the forward)
    bool lsq_bound;
    // side streams (default priority) for the lanes, forked from and joined
    // back into the caller's stream
    static constexpr int kSide = 3;
    hipStream_t side[kSide];
    hipEvent_t fork_ev, join_ev[kSide];
    bool has_side;
};

static void free_program(Program &d) {
    if (d.descs) (void)hipFree(d.descs);
    if (d.cdescs) (void)hipFree(d.cdescs);
    for (auto &l : d.launches)
        if (l.items) (void)hipFree(l.items);
    d = Program();
}

namespace fm {

static double pass_bytes(const TileArgs<float> &a, size_t esz) {
    const double span = double(a.J) * double(int64_t(1) << a.logN);
    const double in = double(a.O) * (double(a.n_valid) < span ? double(a.n_valid) : span);
    const double out = double(a.O) * (double(a.m_valid) < span ? double(a.m_valid) : span);
    return (in + out) * double(esz);
}
static double pass_bytes(const TileArgs<double> &a, size_t esz) {
    const double span = double(a.J) * double(int64_t(1) << a.logN);
    const double in = double(a.O) * (double(a.n_valid) < span ? double(a.n_valid) : span);
    const double out = double(a.O) * (double(a.m_valid) < span ? double(a.m_valid) : span);
    return (in + out) * double(esz);
}

// Normalise the descriptors, group work items into launches and upload.
template <typename T>
static fm_status finalize(ProgBuilder<T> &b, Program &prog, int ncu_lanes_cap, hipStream_t s) {
    free_program(prog);
    for (auto &d : b.descs) {
        fm_status st = normalize_pass(d.a);
        if (st) return st;
        // the batch's passes write gigabytes per launch that the next pass
        // reads back long after: non-temporal stores and loads (the
        // 256-point conv pass 126 -> 121 -> 119 us alone; the
        // single-operator Fourier passes got slower with such stores)
        d.a.stream_out = 3;
    }
    for (auto &c : b.cdescs) {
        // the correlations read each column transform once: non-temporal
        // loads (the 256-point row correlation 304 -> 289 us alone)
        c.a.stream_out = 2;
        // the correlation tiles use corr_loge elements per thread: set the
        // tile shape (B sequences) for that shape
        const int loge = corr_loge<T>(c.a.logN);
        const bool rows = c.a.J > 1;
        const int nt = corr_threads(c.a.logN, loge, rows);
        int logb = 0;
        while ((1 << logb) * (int64_t(1) << c.a.logN) < int64_t(nt) << loge) ++logb;
        c.a.B = 1 << logb;
        c.a.ntiles = ceil_div(c.a.O * c.a.J, c.a.B);
        c.a.logJ = is_pow2(c.a.J) ? ilog2(c.a.J) : -1;
        c.a.o_group = 1;
    }
    prog.lanes = b.lanes < ncu_lanes_cap ? b.lanes : ncu_lanes_cap;
    if (prog.lanes < 1) prog.lanes = 1;
    if (!b.descs.empty()) {
        FM_CHECK_HIP(hipMalloc(&prog.descs, sizeof(BatchDesc<T>) * b.descs.size()));
        FM_CHECK_HIP(hipMemcpyAsync(prog.descs, b.descs.data(),
                                    sizeof(BatchDesc<T>) * b.descs.size(),
                                    hipMemcpyHostToDevice, s));
    }
    if (!b.cdescs.empty()) {
        FM_CHECK_HIP(hipMalloc(&prog.cdescs, sizeof(CorrDesc<T>) * b.cdescs.size()));
        FM_CHECK_HIP(hipMemcpyAsync(prog.cdescs, b.cdescs.data(),
                                    sizeof(CorrDesc<T>) * b.cdescs.size(),
                                    hipMemcpyHostToDevice, s));
    }
    const size_t esz = sizeof(cpx<T>);
    std::vector<std::vector<int2>> staged;  // host item lists, kept until the sync below
    for (auto &pg : b.groups) {
        // classes of this pass group: (kind, logN, nt) -> items
        std::map<std::tuple<int, int, int>, std::vector<int2>> cls;
        std::map<std::tuple<int, int, int>, double> cbytes, calg;
        std::map<std::tuple<int, int, int>, bool> cfit;
        for (const auto &r : pg.refs) {
            const TileArgs<T> &a = r.corr ? b.cdescs[r.index].a : b.descs[r.index].a;
            int kind = r.corr ? LK_CORR : (b.descs[r.index].conv ? LK_CONV : LK_FFT);
            int nt = r.corr ? corr_threads(a.logN, corr_loge<T>(a.logN), a.J > 1) : 0;
            auto key = std::make_tuple(kind, a.logN, nt);
            bool fit = fits_idx32(a, int64_t(1) << a.logN);
            if (r.corr) {
                const CorrDesc<T> &c = b.cdescs[r.index];
                // column offsets of both sides stay in 32 bits as well
                const double lim = double((uint64_t(1) << 32) / sizeof(cpx<T>));
                fit = fit && double(c.k) * double(c.x_col) < lim &&
                      double(c.k) * double(c.g_col) < lim;
            }
            auto fi = cfit.find(key);
            if (fi == cfit.end()) cfit[key] = fit;
            else fi->second = fi->second && fit;
            auto &v = cls[key];
            for (int64_t t = 0; t < a.ntiles; ++t) {
                int2 w;
                w.x = r.index;
                w.y = int(t);
                v.push_back(w);
            }
            double by = pass_bytes(a, esz);
            if (r.corr) {
                const CorrDesc<T> &c = b.cdescs[r.index];
                by = (double(c.k) * 2.0 * double(int64_t(1) << a.logN) * double(a.J) +
                      double(a.J) * double(int64_t(1) << a.logN)) * double(esz);
            }
            cbytes[key] += by;
            calg[key] += a.alg_bytes;
        }
        std::vector<std::tuple<int, int, int>> order;
        for (auto &kv : cls) order.push_back(kv.first);
        std::sort(order.begin(), order.end(),
                  [&](const std::tuple<int, int, int> &x, const std::tuple<int, int, int> &y) {
                      return cbytes[x] > cbytes[y];
                  });
        for (auto &key : order) {
            Launch l;
            l.kind = std::get<0>(key);
            l.logN = std::get<1>(key);
            l.nt = std::get<2>(key);
            l.lane = pg.lane < prog.lanes ? pg.lane : pg.lane % prog.lanes;
            l.early = pg.early ? 1 : 0;
            l.idx32 = cfit[key] ? 1 : 0;
            l.nitems = int64_t(cls[key].size());
            l.bytes = calg[key];
            l.pass_bytes = cbytes[key];
            l.items = nullptr;
            FM_CHECK_HIP(hipMalloc(&l.items, sizeof(int2) * cls[key].size()));
            staged.push_back(std::move(cls[key]));
            FM_CHECK_HIP(hipMemcpyAsync(l.items, staged.back().data(),
                                        sizeof(int2) * staged.back().size(),
                                        hipMemcpyHostToDevice, s));
            prog.launches.push_back(l);
        }
    }
    FM_CHECK_HIP(hipStreamSynchronize(s));
    prog.bound = true;
    return FM_SUCCESS;
}

// ------------------------------------------------------------- launching
template <typename T, int LOGN, bool CONV>
static fm_status launch_pass_class(const fm_batch_plan *p, const Program &d, const Launch &l,
                                   hipStream_t s, int64_t out_shift, const T *out_mul) {
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
    if (grid > l.nitems) grid = l.nitems;
    ProfToken tok;
    prof_begin(s, &tok);
    constexpr int NT = PassCfg<T, LOGN>::NT;
    if (l.idx32)
        hipLaunchKernelGGL((batch_pass_kernel<T, LOGN, CONV, true>), dim3(unsigned(grid)), dim3(NT),
                           lds, s, (const BatchDesc<T> *)d.descs, (const int2 *)l.items, l.nitems,
                           out_shift, out_mul);
    else
        hipLaunchKernelGGL((batch_pass_kernel<T, LOGN, CONV, false>), dim3(unsigned(grid)),
                           dim3(NT), lds, s, (const BatchDesc<T> *)d.descs, (const int2 *)l.items,
                           l.nitems, out_shift, out_mul);
    FM_CHECK_LAUNCH();
    if (tok.active) {
        char name[80];
        snprintf(name, sizeof(name), "batch_pass_kernel<%s,%d,%s>",
                 sizeof(T) == 4 ? "float" : "double", LOGN, CONV ? "conv" : "fft");
        prof_end(s, &tok, name, l.bytes, 0.0, l.pass_bytes);
    }
    return FM_SUCCESS;
}

template <typename T, int LOGN, int NT>
static fm_status launch_corr_nt(const fm_batch_plan *p, const Program &d, const Launch &l,
                                hipStream_t s) {
    constexpr int N = 1 << LOGN;
    constexpr int LE = corr_loge<T>(LOGN);
    using S = TileShape<T, LOGN, NT, LE>;
    const size_t lds = size_t(S::B) * (N + (N >> 5) + 1) * sizeof(cpx<T>);
    static bool attr = false;
    if (!attr) {
        (void)hipFuncSetAttribute((const void *)batch_corr_kernel<T, LOGN, NT, false>,
                                  hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
        (void)hipFuncSetAttribute((const void *)batch_corr_kernel<T, LOGN, NT, true>,
                                  hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
        attr = true;
    }
    FM_REQUIRE(lds <= 160 * 1024, FM_ERR_VALUE, "correlation tile needs %zu bytes of LDS", lds);
    int64_t grid = int64_t(p->ncu) * p->grid_per_cu;
    if (grid > l.nitems) grid = l.nitems;
    ProfToken tok;
    prof_begin(s, &tok);
    if (l.idx32)
        hipLaunchKernelGGL((batch_corr_kernel<T, LOGN, NT, true>), dim3(unsigned(grid)), dim3(NT),
                           lds, s, (const CorrDesc<T> *)d.cdescs, (const int2 *)l.items,
                           l.nitems);
    else
        hipLaunchKernelGGL((batch_corr_kernel<T, LOGN, NT, false>), dim3(unsigned(grid)),
                           dim3(NT), lds, s, (const CorrDesc<T> *)d.cdescs,
                           (const int2 *)l.items, l.nitems);
    FM_CHECK_LAUNCH();
    if (tok.active) {
        char name[80];
        snprintf(name, sizeof(name), "batch_corr_kernel<%s,%d,%d>",
                 sizeof(T) == 4 ? "float" : "double", LOGN, NT);
        prof_end(s, &tok, name, l.bytes, 0.0, l.pass_bytes);
    }
    return FM_SUCCESS;
}

template <typename T, int LOGN>
static fm_status launch_corr(const fm_batch_plan *p, const Program &d, const Launch &l,
                             hipStream_t s) {
    // the two tile shapes corr_threads chooses for this length
    constexpr int NT_ROWS = corr_threads(LOGN, corr_loge<T>(LOGN), true);
    constexpr int NT_ONE = corr_threads(LOGN, corr_loge<T>(LOGN), false);
    if (l.nt == NT_ROWS) return launch_corr_nt<T, LOGN, NT_ROWS>(p, d, l, s);
    if constexpr (NT_ONE != NT_ROWS)
        if (l.nt == NT_ONE) return launch_corr_nt<T, LOGN, NT_ONE>(p, d, l, s);
    set_error("correlation launch with %d threads", l.nt);
    return FM_ERR_VALUE;
}

template <typename T>
static fm_status launch_one(const fm_batch_plan *p, const Program &d, const Launch &l,
                            hipStream_t s, int64_t out_shift = 0, const T *out_mul = nullptr) {
    fm_status st = FM_ERR_VALUE;
#define FM_BATCH_CASE(L)                                                                     \
    case L:                                                                                  \
        st = l.kind == LK_CORR   ? launch_corr<T, L>(p, d, l, s)                             \
             : l.kind == LK_CONV ? launch_pass_class<T, L, true>(p, d, l, s, out_shift, out_mul) \
                                 : launch_pass_class<T, L, false>(p, d, l, s, out_shift, out_mul); \
        break;
    switch (l.logN) {
        FM_BATCH_CASE(0) FM_BATCH_CASE(1) FM_BATCH_CASE(2) FM_BATCH_CASE(3)
        FM_BATCH_CASE(4) FM_BATCH_CASE(5) FM_BATCH_CASE(6) FM_BATCH_CASE(7)
        FM_BATCH_CASE(8) FM_BATCH_CASE(9) FM_BATCH_CASE(10) FM_BATCH_CASE(11)
        FM_BATCH_CASE(12)
    case 13:
        if constexpr (sizeof(T) == 4)
            st = l.kind == LK_CORR   ? launch_corr<T, 13>(p, d, l, s)
                 : l.kind == LK_CONV ? launch_pass_class<T, 13, true>(p, d, l, s, out_shift, out_mul)
                                     : launch_pass_class<T, 13, false>(p, d, l, s, out_shift,
                                                                       out_mul);
        break;
    default:
        break;
    }
#undef FM_BATCH_CASE
    if (st == FM_ERR_VALUE && !(l.logN >= 0 && l.logN <= batch_max_log<T>()))
        set_error("batched transform length 2^%d out of range", l.logN);
    return st;
}

template <typename T>
static int program_lanes(const fm_batch_plan *p, const Program &d) {
    return p->has_side ? (d.lanes < fm_batch_plan::kSide + 1 ? d.lanes : fm_batch_plan::kSide + 1)
                       : 1;
}

// Issue the launches of a program on `lanes` lanes (lane 0 is the caller's
// stream) without forking or joining.
template <typename T>
static fm_status issue_launches(const fm_batch_plan *p, const Program &d, int lanes, hipStream_t s,
                                bool skip_early, int64_t out_shift, const T *out_mul) {
    for (const Launch &l : d.launches) {
        if (skip_early && l.early) continue;
        const int lane = lanes > 1 ? l.lane % lanes : 0;
        hipStream_t sq = lane == 0 ? s : p->side[lane - 1];
        fm_status st = launch_one<T>(p, d, l, sq, out_shift, out_mul);
        if (st) return st;
    }
    return FM_SUCCESS;
}

static fm_status fork_lanes(const fm_batch_plan *p, int lanes, hipStream_t s) {
    if (lanes > 1) {
        FM_CHECK_HIP(hipEventRecord(p->fork_ev, s));
        for (int q = 0; q < lanes - 1; ++q) FM_CHECK_HIP(hipStreamWaitEvent(p->side[q], p->fork_ev, 0));
    }
    return FM_SUCCESS;
}

static fm_status join_lanes(const fm_batch_plan *p, int lanes, hipStream_t s) {
    if (lanes > 1) {
        for (int q = 0; q < lanes - 1; ++q) {
            FM_CHECK_HIP(hipEventRecord(p->join_ev[q], p->side[q]));
            FM_CHECK_HIP(hipStreamWaitEvent(s, p->join_ev[q], 0));
        }
    }
    return FM_SUCCESS;
}

// Issue a program: lanes beyond the first fork from the caller's stream and
// join back into it (event record / wait only, so it is graph-capturable).
// skip_early leaves out the launches of the groups that need no spectrum.
template <typename T>
static fm_status run_program(const fm_batch_plan *p, const Program &d, hipStream_t s,
                             bool skip_early = false, int64_t out_shift = 0,
                             const T *out_mul = nullptr) {
    const int lanes = program_lanes<T>(p, d);
    fm_status st = fork_lanes(p, lanes, s);
    if (!st) st = issue_launches<T>(p, d, lanes, s, skip_early, out_shift, out_mul);
    if (!st) st = join_lanes(p, lanes, s);
    return st;
}

static int64_t up256(int64_t b) { return (b + 255) / 256 * 256; }

}  // namespace fm

// ======================================================================
//                               plan
// ======================================================================
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
    p->group_mb = 96;
    p->lanes = 2;
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
    const int maxlog = prec == FM_COMPLEX64 ? batch_max_log<float>() : batch_max_log<double>();
    int64_t half = 0;
    std::vector<int64_t> t1, t2;
    for (int i = 0; i < count; ++i) {
        if (!plans[i] || plans[i]->prec != prec || ks[i] < 1) {
            fm_batch_plan_destroy(p);
            set_error("batch entry %d: missing plan, mixed precision or no columns", i);
            return FM_ERR_VALUE;
        }
        const Geometry &g = plans[i]->geo;
        const int worst = g.passes == 1 ? g.logL : (g.logN1 > g.logN2 ? g.logN1 : g.logN2);
        if (worst > maxlog) {
            fm_batch_plan_destroy(p);
            set_error("batch entry %d: transform length exceeds the batched kernel limit", i);
            return FM_ERR_NOT_IMPLEMENTED;
        }
        p->plans.push_back(plans[i]);
        p->ks.push_back(ks[i]);
        int64_t o1 = -1, o2 = -1;
        if (g.passes > 1) {
            const int64_t b = up256(int64_t(esize(prec)) * plans[i]->L * ks[i]);
            o1 = half;
            o2 = half + b;
            half += 2 * b;
        }
        t1.push_back(o1);
        t2.push_back(o2);
    }
    for (int d = 0; d < 2; ++d) {
        p->t1_off[d] = t1;
        p->t2_off[d] = t2;
        for (int i = 0; i < count; ++i)
            if (t1[i] >= 0) {
                p->t1_off[d][i] += d * half;
                p->t2_off[d][i] += d * half;
            }
    }
    p->ws_bytes = 2 * half;
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

extern "C" fm_status fm_batch_set_schedule(fm_batch_plan *p, int workgroups_per_cu, int group_mb,
                                           int lanes) {
    FM_REQUIRE(p != nullptr, FM_ERR_VALUE, "plan is NULL");
    FM_REQUIRE(workgroups_per_cu >= 1 && workgroups_per_cu <= 16, FM_ERR_VALUE,
               "workgroups per CU must be in [1, 16]");
    FM_REQUIRE(group_mb >= 0, FM_ERR_VALUE, "negative group budget");
    FM_REQUIRE(lanes >= 1 && lanes <= fm_batch_plan::kSide + 1, FM_ERR_VALUE,
               "lanes must be in [1, %d]", fm_batch_plan::kSide + 1);
    p->grid_per_cu = workgroups_per_cu;
    const bool regroup = group_mb != p->group_mb || lanes != p->lanes;
    p->group_mb = group_mb;
    p->lanes = lanes;
    if (regroup) {
        // the directions' programs encode the grouping: bind again
        p->dir[0].bound = false;
        p->dir[1].bound = false;
    }
    return FM_SUCCESS;
}

// ======================================================================
//                       forward / adjoint programs
// ======================================================================
// Pass groups of a direction: the one-pass operators on a lane of their own
// (when there is one to spare); the multi-pass operators, longest transform
// first, in groups whose intermediates (T1 + T2) stay within the budget, each
// group's phases (pass q of every operator of the group) in order on one
// lane, the groups alternating between the lanes.
template <typename T>
static void group_pipes(const fm_batch_plan *p, std::vector<PassList<T>> &pipes,
                        ProgBuilder<T> &b) {
    std::vector<int> multi, single;
    for (int i = 0; i < p->count; ++i) (pipes[i].size() > 1 ? multi : single).push_back(i);
    std::stable_sort(multi.begin(), multi.end(), [&](int a, int c) {
        return p->plans[a]->L > p->plans[c]->L;
    });
    const int64_t budget = int64_t(p->group_mb) << 20;
    const int lanes = p->lanes;
    if (!single.empty()) {
        auto &pg = b.group(lanes > 1 ? lanes : 0);
        for (int i : single) pg.refs.push_back({0, b.add(pipes[i][0])});
    }
    size_t q = 0;
    int gi = 0;
    while (q < multi.size()) {
        size_t r = q;
        int64_t used = 0;
        while (r < multi.size()) {
            const fm_conv_plan *c = p->plans[multi[r]];
            const int64_t need = 2 * int64_t(sizeof(cpx<T>)) * c->L * p->ks[multi[r]];
            if (r > q && budget > 0 && used + need > budget) break;
            used += need;
            ++r;
        }
        const int lane = gi % lanes;
        const int phases = int(pipes[multi[q]].size());
        for (int ph = 0; ph < phases; ++ph) {
            // the column transforms of the input (a first pass that is no
            // convolution) need no spectrum
            auto &pg = b.group(lane, ph == 0 && !pipes[multi[q]][0].conv);
            for (size_t t = q; t < r; ++t) pg.refs.push_back({0, b.add(pipes[multi[t]][ph])});
        }
        q = r;
        ++gi;
    }
}

template <typename T>
static fm_status batch_bind_t(fm_batch_plan *p, int adjoint, const void *const *ins,
                              const int64_t *ld_in, void *const *outs, const int64_t *ld_out,
                              void *workspace, hipStream_t s, const void *const *targets = nullptr,
                              const int64_t *ld_t = nullptr, double res_scale = 1.0) {
    const int dd = adjoint ? 1 : 0;
    std::vector<PassList<T>> pipes(p->count);
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
        if (c->geo.passes > 1) {
            FM_REQUIRE(workspace != nullptr, FM_ERR_VALUE, "batch needs a workspace");
            io.work = (char *)workspace + p->t1_off[dd][i];
        }
        fm_status st = build_pipeline<T>(c->geo, io, &pipes[i]);
        if (st) return st;
        // algorithmic bytes: input and output blocks and the generator,
        // split over the passes
        for (auto &ps : pipes[i])
            ps.a.alg_bytes = double((nin + mout) * p->ks[i] + c->gen_len) * double(sizeof(cpx<T>)) /
                             double(pipes[i].size());
        if (c->geo.passes > 1) {
            // pass B writes T2 (T1 keeps pass A's column transforms), pass C
            // reads T2
            cpx<T> *t2 = (cpx<T> *)((char *)workspace + p->t2_off[dd][i]);
            pipes[i][1].a.y = t2;
            pipes[i][2].a.x = t2;
        }
        pipes[i].back().a.redirect = 1;  // writes outs[i]
    }
    if (targets) {
        // residual mode: the last pass of every operator stores
        // res_scale * (A x - t) and its tiles' sums of |A x - t|^2
        int64_t slots = 0;
        for (int i = 0; i < p->count; ++i) {
            FM_REQUIRE(targets[i] != nullptr && ld_t[i] == ld_out[i], FM_ERR_VALUE,
                       "batch entry %d: target missing or with another leading dimension", i);
            TileArgs<T> a = pipes[i].back().a;
            fm_status st = normalize_pass(a);
            if (st) return st;
            slots += a.ntiles;
        }
        if (p->res_part) (void)hipFree(p->res_part);
        p->res_part = nullptr;
        FM_CHECK_HIP(hipMalloc(&p->res_part, sizeof(double) * size_t(slots > 0 ? slots : 1)));
        int64_t base = 0;
        for (int i = 0; i < p->count; ++i) {
            TileArgs<T> &a = pipes[i].back().a;
            a.target = (const cpx<T> *)targets[i];
            a.res_part = p->res_part;
            a.res_base = base;
            a.out_scale = T(res_scale);
            TileArgs<T> n = a;
            (void)normalize_pass(n);
            base += n.ntiles;
        }
        p->n_res = slots;
    }
    ProgBuilder<T> b;
    group_pipes<T>(p, pipes, b);
    fm_status st = finalize<T>(b, p->dir[dd], fm_batch_plan::kSide + 1, s);
    if (st) return st;
    if (dd == 0) p->res_bound = targets != nullptr;
    p->lsq_bound = false;
    p->ws_bound[dd] = workspace;
    p->ins_bound[dd].assign(ins, ins + p->count);
    p->ld_bound[dd].assign(ld_in, ld_in + p->count);
    return FM_SUCCESS;
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

extern "C" fm_status fm_batch_bind_residual(fm_batch_plan *p, const void *const *ins,
                                            const int64_t *ld_in, const void *const *targets,
                                            const int64_t *ld_t, void *const *outs,
                                            const int64_t *ld_out, double scale, void *workspace,
                                            int64_t workspace_bytes, void *stream) {
    FM_REQUIRE(p && ins && ld_in && targets && ld_t && outs && ld_out, FM_ERR_VALUE,
               "NULL argument");
    FM_REQUIRE(workspace_bytes >= p->ws_bytes, FM_ERR_VALUE,
               "workspace of %lld bytes is smaller than the %lld required",
               (long long)workspace_bytes, (long long)p->ws_bytes);
    if (p->prec == FM_COMPLEX64)
        return batch_bind_t<float>(p, 0, ins, ld_in, outs, ld_out, workspace, as_stream(stream),
                                   targets, ld_t, scale);
    return batch_bind_t<double>(p, 0, ins, ld_in, outs, ld_out, workspace, as_stream(stream),
                                targets, ld_t, scale);
}

namespace fm {
// Both directions of a least-squares step (fm_batch_bind_lsq).  Forward:
// the one-pass operators store scale * (A x - t) (residual epilogue); the
// multi-pass operators run passes A and B, then one fused pass that is the
// forward's column inverse (pass C) and, on the same columns, the adjoint's
// column transform (pass A) of the residual: between the two transforms the
// tile forms scale * (A x - t) (0 beyond the m rows) and its |A x - t|^2,
// and it stores the adjoint's T1 directly, so the residual of a multi-pass
// operator is never written.  Adjoint: the one-pass operators read the
// residual blocks; the multi-pass operators start at pass B.
template <typename T>
static fm_status batch_bind_lsq_t(fm_batch_plan *p, const void *const *ins, const int64_t *ld_in,
                                  const void *const *targets, const int64_t *ld_t,
                                  void *const *routs, const int64_t *ld_r, void *const *gouts,
                                  const int64_t *ld_g, double scale, void *ws, hipStream_t s) {
    std::vector<PassList<T>> fw(p->count), ad(p->count);
    const double esz = double(sizeof(cpx<T>));
    for (int i = 0; i < p->count; ++i) {
        const fm_conv_plan *c = p->plans[i];
        const int64_t k = p->ks[i], m = c->rows, n = c->cols;
        FM_REQUIRE(ins[i] && targets[i] && routs[i] && gouts[i], FM_ERR_VALUE,
                   "batch entry %d: NULL buffer", i);
        FM_REQUIRE(ld_in[i] >= n && ld_r[i] >= m && ld_g[i] >= n && ld_t[i] == ld_r[i],
                   FM_ERR_VALUE,
                   "batch entry %d: leading dimension too small, or targets laid out unlike "
                   "the residual",
                   i);
        const bool multi = c->geo.passes > 1;
        FM_REQUIRE(!multi || ws != nullptr, FM_ERR_VALUE, "batch needs a workspace");
        for (int dir = 0; dir < 2; ++dir) {
            PipeIO io;
            memset(&io, 0, sizeof(io));
            io.x = dir ? (const void *)routs[i] : ins[i];
            io.ldx = dir ? ld_r[i] : ld_in[i];
            io.y = dir ? gouts[i] : routs[i];
            io.ldy = dir ? ld_g[i] : ld_r[i];
            io.k = k;
            io.n_in = dir ? m : n;
            io.m_out = dir ? n : m;
            io.D = c->D;
            io.conj_d = dir;
            if (multi) io.work = (char *)ws + p->t1_off[dir][i];
            PassList<T> &pl = dir ? ad[i] : fw[i];
            fm_status st = build_pipeline<T>(c->geo, io, &pl);
            if (st) return st;
            for (auto &ps : pl)
                ps.a.alg_bytes = double((n + m) * k + c->gen_len) * esz / double(pl.size());
            if (multi) {
                cpx<T> *t2 = (cpx<T> *)((char *)ws + p->t2_off[dir][i]);
                pl[1].a.y = t2;
                pl[2].a.x = t2;
            }
        }
        if (multi) {
            // the fused pass: pass C's input side and inverse transform, the
            // residual, pass A's forward transform, twiddle and output side
            TileArgs<T> f = fw[i][2].a;
            const TileArgs<T> &pa = ad[i][0].a;
            f.y = pa.y;
            f.out_os = pa.out_os; f.out_js = pa.out_js; f.out_ns = pa.out_ns;
            f.out_gn = pa.out_gn; f.out_gj = pa.out_gj; f.m_valid = pa.m_valid;
            f.post = pa.post; f.conj_out = pa.conj_out; f.seqfast_out = pa.seqfast_out;
            f.twb_hi = pa.twb_hi; f.twb_lo = pa.twb_lo;
            f.twb_shift = pa.twb_shift; f.twb_sign = pa.twb_sign;
            f.out_scale = T(1);
            f.mid_t = (const cpx<T> *)targets[i];
            f.mid_os = ld_t[i];
            f.mid_gn = fw[i][2].a.out_gn;
            f.mid_gj = fw[i][2].a.out_gj;
            f.mid_valid = m;
            f.mid_scale = T(scale);
            f.alg_bytes = fw[i][2].a.alg_bytes + pa.alg_bytes;
            fw[i][2] = PassSpec<T>{f, true};
            ad[i].pop_front();
        } else {
            TileArgs<T> &a = fw[i][0].a;
            a.target = (const cpx<T> *)targets[i];
            a.out_scale = T(scale);
            a.redirect = 1;  // writes routs[i]
        }
        ad[i].back().a.redirect = 1;  // writes gouts[i]
    }
    // residual slots: one per tile of every operator's residual pass
    int64_t slots = 0;
    std::vector<int64_t> base(p->count);
    for (int i = 0; i < p->count; ++i) {
        TileArgs<T> a = fw[i].back().a;
        fm_status st = normalize_pass(a);
        if (st) return st;
        base[i] = slots;
        slots += a.ntiles;
    }
    if (p->res_part) (void)hipFree(p->res_part);
    p->res_part = nullptr;
    FM_CHECK_HIP(hipMalloc(&p->res_part, sizeof(double) * size_t(slots > 0 ? slots : 1)));
    for (int i = 0; i < p->count; ++i) {
        fw[i].back().a.res_part = p->res_part;
        fw[i].back().a.res_base = base[i];
    }
    p->n_res = slots;
    ProgBuilder<T> bf, ba;
    group_pipes<T>(p, fw, bf);
    group_pipes<T>(p, ad, ba);
    fm_status st = finalize<T>(bf, p->dir[0], fm_batch_plan::kSide + 1, s);
    if (!st) st = finalize<T>(ba, p->dir[1], fm_batch_plan::kSide + 1, s);
    if (st) return st;
    p->res_bound = true;
    p->lsq_bound = true;
    for (int d = 0; d < 2; ++d) p->ws_bound[d] = ws;
    p->ins_bound[0].assign(ins, ins + p->count);
    p->ld_bound[0].assign(ld_in, ld_in + p->count);
    p->ins_bound[1].assign(routs, routs + p->count);
    p->ld_bound[1].assign(ld_r, ld_r + p->count);
    return FM_SUCCESS;
}

// loss = sum of the tile partial sums, in slot order (deterministic).
__global__ __launch_bounds__(256) void residual_sum_kernel(const double *__restrict__ part,
                                                           int64_t n, double *__restrict__ out) {
    __shared__ double red[256];
    // eight independent running sums per thread keep eight loads in flight
    // (one workgroup: the sum's order is fixed, so the loss is deterministic)
    double s[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) s[u] = 0.0;
    int64_t i = threadIdx.x;
    for (; i + 7 * 256 < n; i += 8 * 256) {
#pragma unroll
        for (int u = 0; u < 8; ++u) s[u] += part[i + u * 256];
    }
    for (; i < n; i += 256) s[0] += part[i];
    red[threadIdx.x] = ((s[0] + s[1]) + (s[2] + s[3])) + ((s[4] + s[5]) + (s[6] + s[7]));
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (int(threadIdx.x) < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = red[0];
}
}  // namespace fm

extern "C" fm_status fm_batch_bind_lsq(fm_batch_plan *p, const void *const *ins,
                                       const int64_t *ld_in, const void *const *targets,
                                       const int64_t *ld_t, void *const *residuals,
                                       const int64_t *ld_r, void *const *grads_x,
                                       const int64_t *ld_gx, double scale, void *workspace,
                                       int64_t workspace_bytes, void *stream) {
    FM_REQUIRE(p && ins && ld_in && targets && ld_t && residuals && ld_r && grads_x && ld_gx,
               FM_ERR_VALUE, "NULL argument");
    FM_REQUIRE(workspace_bytes >= p->ws_bytes, FM_ERR_VALUE,
               "workspace of %lld bytes is smaller than the %lld required",
               (long long)workspace_bytes, (long long)p->ws_bytes);
    if (p->prec == FM_COMPLEX64)
        return batch_bind_lsq_t<float>(p, ins, ld_in, targets, ld_t, residuals, ld_r, grads_x,
                                       ld_gx, scale, workspace, as_stream(stream));
    return batch_bind_lsq_t<double>(p, ins, ld_in, targets, ld_t, residuals, ld_r, grads_x, ld_gx,
                                    scale, workspace, as_stream(stream));
}

extern "C" fm_status fm_batch_residual_loss(const fm_batch_plan *p, double *loss, void *stream) {
    FM_REQUIRE(p != nullptr && loss != nullptr, FM_ERR_VALUE, "NULL argument");
    FM_REQUIRE(p->res_bound && p->res_part, FM_ERR_VALUE,
               "the forward direction is not bound in residual mode");
    hipLaunchKernelGGL(residual_sum_kernel, dim3(1), dim3(256), 0, as_stream(stream),
                       (const double *)p->res_part, p->n_res, loss);
    FM_CHECK_LAUNCH();
    return FM_SUCCESS;
}

extern "C" fm_status fm_batch_apply(const fm_batch_plan *p, int adjoint, void *stream) {
    FM_REQUIRE(p != nullptr, FM_ERR_VALUE, "plan is NULL");
    const Program &d = p->dir[adjoint ? 1 : 0];
    FM_REQUIRE(d.bound, FM_ERR_VALUE, "batch direction %d is not bound to buffers", adjoint);
    if (p->prec == FM_COMPLEX64) return run_program<float>(p, d, as_stream(stream));
    return run_program<double>(p, d, as_stream(stream));
}

extern "C" fm_status fm_batch_apply_redirected(const fm_batch_plan *p, int adjoint,
                                               int64_t shift_bytes, const void *scale,
                                               void *stream) {
    FM_REQUIRE(p != nullptr, FM_ERR_VALUE, "plan is NULL");
    const Program &d = p->dir[adjoint ? 1 : 0];
    FM_REQUIRE(d.bound, FM_ERR_VALUE, "batch direction %d is not bound to buffers", adjoint);
    const int64_t esz = p->prec == FM_COMPLEX64 ? 8 : 16;
    FM_REQUIRE(shift_bytes % esz == 0, FM_ERR_VALUE,
               "output shift of %lld bytes is not a whole number of elements",
               (long long)shift_bytes);
    if (p->prec == FM_COMPLEX64)
        return run_program<float>(p, d, as_stream(stream), false, shift_bytes,
                                  (const float *)scale);
    return run_program<double>(p, d, as_stream(stream), false, shift_bytes,
                               (const double *)scale);
}

// ======================================================================
//                     batched generator gradient
// ======================================================================
// For every operator i (what fm_conv_generator_grad returns for it alone):
//   multi-pass: T1x = pass A of x (the forward direction's T1), T1g = pass A
//     of g (the adjoint direction's T1); one correlation tile per block of
//     rows k1 sums FFT(T1g rows) conj(FFT(T1x rows)) over the columns,
//     transforms back and applies W_L^{-n2 k1}; pass C (column inverse)
//     gives r = the unnormalised inverse transform of
//     sum_c FFT_L(g_c') conj(FFT_L(x_c')).
//   one-pass: one correlation tile over the operator's columns straight
//     from x and g (zero-padded to L), stored as r.
//   then r is folded onto the generator and scaled by 1/L.
// fm_batch_grad_apply_reuse skips both pass-A launches: valid when the
// forward direction was last applied with inputs xs and the adjoint with
// inputs gs (their T1 hold exactly those column transforms).
namespace fm {
template <typename T> struct GradEntry {
    const cpx<T> *r;
    cpx<T> *out;
    int64_t L, head, tail, gen_len;
    int kind;
};

template <typename T>
__global__ void batch_fold_kernel(const GradEntry<T> *__restrict__ tab, int64_t out_shift,
                                  const T *__restrict__ out_mul) {
    GradEntry<T> e = tab[blockIdx.y];
    T s = T(1.0 / double(e.L));
    if (out_shift != 0 || out_mul != nullptr) {
        // least-squares backward: into another buffer of the same layout,
        // scaled by the device scalar *out_mul
        e.out = reinterpret_cast<cpx<T> *>(reinterpret_cast<char *>(e.out) + out_shift);
        if (out_mul) s *= *out_mul;
    }
    const int64_t stride = int64_t(gridDim.x) * blockDim.x;
    for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < e.gen_len; i += stride) {
        cpx<T> v;
        if (e.kind == 0) {
            v = ldg(e.r, i);
            if (e.tail > 0 && i >= 1) v = v + ldg(e.r, int64_t(e.L - e.tail + i - 1));
        } else {
            v = i < e.head ? ldg(e.r, i) : ldg(e.r, int64_t(e.L - e.tail + (i - e.head)));
        }
        stg(e.out, i, scale(v, s));
    }
}

// Gradient workspace of one operator: rows of the correlation (four-step
// only) and r.
static int64_t corr_chunks(const fm_conv_plan *c, int64_t k) {
    return c->geo.passes > 1 ? 1 : ceil_div(k, kCorrChunk);
}

static int64_t grad_ws_entry(const fm_conv_plan *c, int64_t k) {
    const int64_t e = int64_t(esize(c->prec)), L = c->L;
    const int64_t nch = corr_chunks(c, k);
    // rows (four-step) or partial spectra (one-pass, several chunks), and r
    const int64_t first = c->geo.passes > 1 ? up256(e * L) : (nch > 1 ? up256(e * L * nch) : 0);
    return first + up256(e * L);
}

template <typename T>
static fm_status batch_grad_bind_t(fm_batch_plan *p, const void *const *xs, const int64_t *ldx,
                                   const void *const *gs, const int64_t *ldg,
                                   void *const *grads, char *ws, hipStream_t s) {
    FM_REQUIRE(p->ws_bound[0] != nullptr || p->ws_bound[1] != nullptr || p->ws_bytes == 0,
               FM_ERR_VALUE, "bind the batch's directions (fm_batch_bind) before its gradient");
    char *bws = (char *)(p->ws_bound[0] ? p->ws_bound[0] : p->ws_bound[1]);
    ProgBuilder<T> ba, bc, bp;
    auto &pa_x = ba.group(0);
    auto &pa_g = ba.group(p->lanes > 1 ? 1 : 0);
    // the four-step operators' correlations: the longest rows on lane 0, the
    // shorter ones on lane 2; the one-pass operators' on lane 1 (they share
    // no data, and each alone leaves the chip partly idle)
    auto &pc = bc.group(0);
    auto &pc1 = bc.group(p->lanes > 1 ? 1 : 0);
    auto &pc2 = bc.group(p->lanes > 1 ? 2 : 0);
    int max_log2 = 0;
    for (int i = 0; i < p->count; ++i)
        if (p->plans[i]->geo.passes > 1 && p->plans[i]->geo.logN2 > max_log2)
            max_log2 = p->plans[i]->geo.logN2;
    auto &pp = bp.group(0);
    auto &pp1 = bp.group(p->lanes > 1 ? 1 : 0);
    std::vector<GradEntry<T>> tab(p->count);
    int64_t off = 0, maxgen = 1;
    for (int i = 0; i < p->count; ++i) {
        const fm_conv_plan *c = p->plans[i];
        const int64_t k = p->ks[i], L = c->L;
        FM_REQUIRE(xs[i] && gs[i] && grads[i], FM_ERR_VALUE, "batch entry %d: NULL buffer", i);
        FM_REQUIRE(ldx[i] >= c->cols && ldg[i] >= c->rows, FM_ERR_VALUE,
                   "batch entry %d: leading dimension too small", i);
        char *e = ws + off;
        off += grad_ws_entry(c, k);
        const bool multi = c->geo.passes > 1;
        const int64_t nch = corr_chunks(c, k);
        cpx<T> *rows = multi ? (cpx<T> *)e : nullptr;
        cpx<T> *parts = (!multi && nch > 1) ? (cpx<T> *)e : nullptr;
        const int64_t first = multi ? up256(int64_t(sizeof(cpx<T>)) * L)
                                    : (nch > 1 ? up256(int64_t(sizeof(cpx<T>)) * L * nch) : 0);
        cpx<T> *r = (cpx<T> *)(e + first);
        CorrDesc<T> cd;
        memset(&cd, 0, sizeof(cd));
        TileArgs<T> a = blank_args<T>();
        if (multi) {
            const Geometry &g = c->geo;
            const int64_t N1 = int64_t(1) << g.logN1, N2 = int64_t(1) << g.logN2;
            cpx<T> *t1x = (cpx<T> *)(bws + p->t1_off[0][i]);
            cpx<T> *t1g = (cpx<T> *)(bws + p->t1_off[1][i]);
            // pass A of x and of g: the conv pipelines' pass A into T1
            for (int side = 0; side < 2; ++side) {
                PipeIO io;
                memset(&io, 0, sizeof(io));
                io.x = side ? gs[i] : xs[i];
                io.ldx = side ? ldg[i] : ldx[i];
                io.y = r;  // unused by pass A
                io.ldy = L;
                io.k = k;
                io.n_in = side ? c->rows : c->cols;
                io.m_out = L;
                io.D = c->D;
                io.work = side ? t1g : t1x;
                PassList<T> pl;
                fm_status st = build_pipeline<T>(g, io, &pl);
                if (st) return st;
                pl[0].a.alg_bytes = 0;  // the inputs were counted by the applies
                (side ? pa_g : pa_x).refs.push_back({0, ba.add(pl[0])});
            }
            // correlation over rows k1 (contiguous N2-element rows of T1)
            a.logN = g.logN2;
            a.O = 1;
            a.J = N1;
            a.x = t1x;
            a.in_os = 0; a.in_js = N2; a.in_ns = 1;
            a.tw = (const cpx<double> *)g.tw2;
            a.twb_hi = (const cpx<double> *)g.twb_hi;
            a.twb_lo = (const cpx<double> *)g.twb_lo;
            a.twb_shift = g.twb_shift;
            a.twb_sign = +1;
            a.y = rows;
            a.out_os = 0; a.out_js = N2; a.out_ns = 1;
            a.alg_bytes = double(c->gen_len) * double(sizeof(cpx<T>));
            cd.a = a;
            cd.g = t1g;
            cd.x_col = L;
            cd.g_col = L;
            cd.g_valid = INT64_MAX;
            cd.k = k;
            cd.mode = CORR_FULL;
            (g.logN2 >= max_log2 ? pc : pc2).refs.push_back({1, bc.add_corr(cd)});
            // pass C of the rows (column inverse), natural order into r
            PipeIO io;
            memset(&io, 0, sizeof(io));
            io.x = rows; io.ldx = L; io.y = r; io.ldy = L; io.k = 1;
            io.n_in = L; io.m_out = L; io.D = c->D; io.work = rows;
            PassList<T> pl;
            fm_status st = build_pipeline<T>(g, io, &pl);
            if (st) return st;
            pl[2].a.alg_bytes = 0;
            pp.refs.push_back({0, bp.add(pl[2])});
        } else {
            const Geometry &g = c->geo;
            a.logN = g.logL;
            a.O = 1;
            a.J = 1;
            a.x = (const cpx<T> *)xs[i];
            a.in_os = 0; a.in_js = 0; a.in_ns = 1;
            a.in_gn = 1; a.in_gj = 0; a.n_valid = c->cols;
            a.tw = (const cpx<double> *)g.tw_full;
            a.y = r;
            a.out_os = 0; a.out_js = 0; a.out_ns = 1;
            a.out_gn = 1; a.out_gj = 0; a.m_valid = L;
            a.alg_bytes = double(c->gen_len) * double(sizeof(cpx<T>));
            cd.g_valid = c->rows;
            cd.x_col = ldx[i];
            cd.g_col = ldg[i];
            if (nch == 1) {
                cd.a = a;
                cd.g = (const cpx<T> *)gs[i];
                cd.k = k;
                cd.mode = CORR_FULL;
                pc1.refs.push_back({1, bc.add_corr(cd)});
            } else {
                // chunks of columns -> partial spectra, then their sum
                for (int64_t q = 0; q < nch; ++q) {
                    const int64_t c0 = q * kCorrChunk;
                    CorrDesc<T> pd = cd;
                    pd.a = a;
                    pd.a.x = (const cpx<T> *)xs[i] + c0 * ldx[i];
                    pd.a.y = parts + q * L;
                    pd.a.alg_bytes = q == 0 ? a.alg_bytes : 0.0;
                    pd.g = (const cpx<T> *)gs[i] + c0 * ldg[i];
                    pd.k = k - c0 < kCorrChunk ? k - c0 : kCorrChunk;
                    pd.mode = CORR_PARTIAL;
                    pc1.refs.push_back({1, bc.add_corr(pd)});
                }
                CorrDesc<T> sd;
                memset(&sd, 0, sizeof(sd));
                sd.a = a;
                sd.a.x = parts;
                sd.a.n_valid = INT64_MAX;
                sd.a.alg_bytes = 0;
                sd.x_col = L;
                sd.k = nch;
                sd.g_valid = INT64_MAX;
                sd.mode = CORR_SUM;
                pp1.refs.push_back({1, bp.add_corr(sd)});
            }
        }
        tab[i] = GradEntry<T>{r, (cpx<T> *)grads[i], L, c->head, c->tail, c->gen_len, c->kind};
        if (c->gen_len > maxgen) maxgen = c->gen_len;
    }
    fm_status st = finalize<T>(ba, p->gpassA, fm_batch_plan::kSide + 1, s);
    if (!st) st = finalize<T>(bc, p->gcorr, fm_batch_plan::kSide + 1, s);
    if (!st) st = finalize<T>(bp, p->gpassC, fm_batch_plan::kSide + 1, s);
    if (st) return st;
    if (p->gtab) (void)hipFree(p->gtab);
    p->gtab = nullptr;
    FM_CHECK_HIP(hipMalloc(&p->gtab, sizeof(GradEntry<T>) * tab.size()));
    FM_CHECK_HIP(hipMemcpyAsync(p->gtab, tab.data(), sizeof(GradEntry<T>) * tab.size(),
                                hipMemcpyHostToDevice, s));
    FM_CHECK_HIP(hipStreamSynchronize(s));
    p->g_maxgen = maxgen;
    p->g_ws = bws;
    p->g_xs.assign(xs, xs + p->count);
    p->g_gs.assign(gs, gs + p->count);
    p->g_ldx.assign(ldx, ldx + p->count);
    p->g_ldg.assign(ldg, ldg + p->count);
    p->g_bound = true;
    return FM_SUCCESS;
}

template <typename T>
static fm_status batch_grad_apply_t(const fm_batch_plan *p, bool run_pass_a, hipStream_t s,
                                    bool with_adjoint = false, int64_t out_shift = 0,
                                    const T *out_mul = nullptr, int64_t gen_shift = 0) {
    fm_status st = FM_SUCCESS;
    if (run_pass_a) st = run_program<T>(p, p->gpassA, s);
    if (!st && with_adjoint) {
        // least-squares backward: the adjoint (from its row pass on) and the
        // correlations read the same column transforms and write disjoint
        // buffers, so they share one fork of the lanes; a lane that is done
        // with its share of the adjoint goes on with its correlations while
        // the others finish
        const Program &ad = p->dir[1], &gc = p->gcorr;
        const int la = program_lanes<T>(p, ad), lc = program_lanes<T>(p, gc);
        const int lanes = la > lc ? la : lc;
        constexpr int NL = fm_batch_plan::kSide + 1;
        // each correlation lane (its launches depend on one another in
        // stream order, so they move together) goes to the group lane with
        // the least adjoint work so far, largest first (bytes moved as the
        // measure of work); not to the one-pass operators' lane, whose few
        // bytes take long while the persistent group launches hold the CUs
        double load[NL] = {}, cl[NL] = {};
        for (const Launch &l : ad.launches) load[lanes > 1 ? l.lane % lanes : 0] += l.pass_bytes;
        for (const Launch &l : gc.launches) cl[lanes > 1 ? l.lane % lanes : 0] += l.pass_bytes;
        int lane_of[NL];
        bool placed[NL] = {};
        for (int u = 0; u < NL; ++u) lane_of[u] = u;
        for (int n = 0; n < lanes; ++n) {
            int u = -1;
            for (int v = 0; v < lanes; ++v)
                if (!placed[v] && cl[v] > 0 && (u < 0 || cl[v] > cl[u])) u = v;
            if (u < 0) break;
            const int skip = p->lanes < lanes ? p->lanes : -1;  // the one-pass lane
            int q = -1;
            for (int v = 0; v < lanes; ++v)
                if (v != skip && (q < 0 || load[v] < load[q])) q = v;
            placed[u] = true;
            lane_of[u] = q;
            load[q] += cl[u];
        }
        st = fork_lanes(p, lanes, s);
        if (!st) st = issue_launches<T>(p, ad, lanes, s, false, out_shift, out_mul);
        for (const Launch &l : gc.launches) {
            if (st) break;
            const int q = lane_of[lanes > 1 ? l.lane % lanes : 0];
            st = launch_one<T>(p, gc, l, q == 0 ? s : p->side[q - 1]);
        }
        if (!st) st = join_lanes(p, lanes, s);
    } else if (!st) {
        st = run_program<T>(p, p->gcorr, s);
    }
    if (!st) st = run_program<T>(p, p->gpassC, s);
    if (st) return st;
    int64_t gx = ceil_div(p->g_maxgen, 256);
    if (gx > 256) gx = 256;
    hipLaunchKernelGGL(batch_fold_kernel<T>, dim3(unsigned(gx), unsigned(p->count)), dim3(256), 0,
                       s, (const GradEntry<T> *)p->gtab, gen_shift,
                       with_adjoint ? out_mul : (const T *)nullptr);
    FM_CHECK_LAUNCH();
    return FM_SUCCESS;
}
}  // namespace fm

extern "C" fm_status fm_batch_grad_workspace_size(const fm_batch_plan *p, int64_t *bytes) {
    FM_REQUIRE(p && bytes, FM_ERR_VALUE, "NULL argument");
    int64_t t = 0;
    for (int i = 0; i < p->count; ++i) t += grad_ws_entry(p->plans[i], p->ks[i]);
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
    FM_REQUIRE(p->g_bound && p->gtab, FM_ERR_VALUE,
               "batch gradient used before fm_batch_grad_bind");
    if (p->prec == FM_COMPLEX64) return batch_grad_apply_t<float>(p, true, as_stream(stream));
    return batch_grad_apply_t<double>(p, true, as_stream(stream));
}

extern "C" fm_status fm_batch_grad_apply_reuse(const fm_batch_plan *p, void *stream) {
    FM_REQUIRE(p != nullptr, FM_ERR_VALUE, "plan is NULL");
    FM_REQUIRE(p->g_bound && p->gtab, FM_ERR_VALUE,
               "batch gradient used before fm_batch_grad_bind");
    const bool ok = p->dir[0].bound && p->dir[1].bound && p->ws_bound[0] == p->g_ws &&
                    p->ws_bound[1] == p->g_ws && p->ins_bound[0] == p->g_xs &&
                    p->ld_bound[0] == p->g_ldx && p->ins_bound[1] == p->g_gs &&
                    p->ld_bound[1] == p->g_ldg;
    FM_REQUIRE(ok, FM_ERR_VALUE,
               "batch gradient reuse needs the forward direction bound to the gradient's x "
               "blocks and the adjoint direction to its g blocks");
    if (p->prec == FM_COMPLEX64) return batch_grad_apply_t<float>(p, false, as_stream(stream));
    return batch_grad_apply_t<double>(p, false, as_stream(stream));
}

extern "C" fm_status fm_batch_lsq_backward(const fm_batch_plan *p, int64_t shift_bytes,
                                           const void *scale, int64_t gen_shift_bytes,
                                           void *stream) {
    FM_REQUIRE(p != nullptr, FM_ERR_VALUE, "plan is NULL");
    FM_REQUIRE(p->lsq_bound, FM_ERR_VALUE,
               "the least-squares backward needs both directions bound by fm_batch_bind_lsq");
    FM_REQUIRE(p->g_bound && p->gtab, FM_ERR_VALUE,
               "batch gradient used before fm_batch_grad_bind");
    const bool ok = p->ws_bound[0] == p->g_ws && p->ws_bound[1] == p->g_ws &&
                    p->ins_bound[0] == p->g_xs && p->ld_bound[0] == p->g_ldx &&
                    p->ins_bound[1] == p->g_gs && p->ld_bound[1] == p->g_ldg;
    FM_REQUIRE(ok, FM_ERR_VALUE,
               "the gradient must be bound to the least-squares step's inputs and residuals");
    const int64_t esz = p->prec == FM_COMPLEX64 ? 8 : 16;
    FM_REQUIRE(shift_bytes % esz == 0 && gen_shift_bytes % esz == 0, FM_ERR_VALUE,
               "output shift of %lld / %lld bytes is not a whole number of elements",
               (long long)shift_bytes, (long long)gen_shift_bytes);
    if (p->prec == FM_COMPLEX64)
        return batch_grad_apply_t<float>(p, false, as_stream(stream), true, shift_bytes,
                                         (const float *)scale, gen_shift_bytes);
    return batch_grad_apply_t<double>(p, false, as_stream(stream), true, shift_bytes,
                                      (const double *)scale, gen_shift_bytes);
}

// ======================================================================
//                  batched generator -> spectrum
// ======================================================================
// For every operator i: t_i = the embedded generator (fm_conv_plan) in the
// batch precision, then D_i = fft(t_i)/L_i in the order its convolution
// pipeline reads it (build_spectrum_pipeline), written into the plan's
// spectrum buffer.  One embedding launch over all operators, then the
// spectrum FFTs as one batched program.
namespace fm {
struct GenEntry {
    const void *gen;
    void *t;
    int64_t L, head, tail, tail_src;
};

template <typename G, typename T>
__global__ void batch_embed_kernel(const GenEntry *__restrict__ tab) {
    const GenEntry e = tab[blockIdx.y];
    const int64_t stride = int64_t(gridDim.x) * blockDim.x;
    for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < e.L; i += stride) {
        const cpx<double> v = embedded_at((const G *)e.gen, i, e.L, e.head, e.tail, e.tail_src);
        stg((cpx<T> *)e.t, i, mk(T(v.x), T(v.y)));
    }
}

static int64_t spec_ws_entry(const fm_conv_plan *c) {
    const int64_t e = int64_t(esize(c->prec));
    return up256(e * c->L) * (c->geo_fft.passes > 1 ? 2 : 1);
}

template <typename T>
static fm_status batch_spec_bind_t(fm_batch_plan *p, const void *const *gens, char *ws,
                                   hipStream_t s) {
    std::vector<GenEntry> tab(p->count);
    ProgBuilder<T> b;
    auto &pa = b.group(0);
    auto &pb = b.group(0);
    auto &p1 = b.group(p->lanes > 1 ? 1 : 0);  // one-pass spectra, independent of pa/pb
    int64_t off = 0, maxL = 1;
    for (int i = 0; i < p->count; ++i) {
        const fm_conv_plan *c = p->plans[i];
        FM_REQUIRE(gens[i] != nullptr, FM_ERR_VALUE, "batch entry %d: NULL generator", i);
        const int64_t L = c->L;
        cpx<T> *t = (cpx<T> *)(ws + off);
        cpx<T> *tmp = c->geo_fft.passes > 1
                          ? (cpx<T> *)(ws + off + up256(int64_t(sizeof(cpx<T>)) * L))
                          : nullptr;
        off += spec_ws_entry(c);
        PassList<T> pl;
        fm_status st = build_spectrum_pipeline<T>(c->geo_fft, t, tmp, (cpx<T> *)c->D, &pl);
        if (st) return st;
        for (auto &ps : pl) ps.a.alg_bytes = 0;  // the generator read is the applies'
        if (pl.size() == 1) {
            p1.refs.push_back({0, b.add(pl[0])});
        } else {
            pa.refs.push_back({0, b.add(pl[0])});
            pb.refs.push_back({0, b.add(pl[1])});
        }
        tab[i] = GenEntry{gens[i], t, L, c->head, c->tail, c->tail_src};
        if (L > maxL) maxL = L;
    }
    fm_status st = finalize<T>(b, p->sprog, fm_batch_plan::kSide + 1, s);
    if (st) return st;
    if (p->stab) (void)hipFree(p->stab);
    p->stab = nullptr;
    FM_CHECK_HIP(hipMalloc(&p->stab, sizeof(GenEntry) * tab.size()));
    FM_CHECK_HIP(hipMemcpyAsync(p->stab, tab.data(), sizeof(GenEntry) * tab.size(),
                                hipMemcpyHostToDevice, s));
    FM_CHECK_HIP(hipStreamSynchronize(s));
    p->s_maxL = maxL;
    return FM_SUCCESS;
}

template <typename T>
static fm_status batch_spec_apply_t(const fm_batch_plan *p, hipStream_t s) {
    int64_t gx = ceil_div(p->s_maxL, 256);
    if (gx > 64) gx = 64;
    const dim3 grid(unsigned(gx), unsigned(p->count));
    const GenEntry *tab = (const GenEntry *)p->stab;
    switch (p->s_gen_dtype) {
    case FM_FLOAT32:
        hipLaunchKernelGGL((batch_embed_kernel<float, T>), grid, dim3(256), 0, s, tab);
        break;
    case FM_FLOAT64:
        hipLaunchKernelGGL((batch_embed_kernel<double, T>), grid, dim3(256), 0, s, tab);
        break;
    case FM_COMPLEX64:
        hipLaunchKernelGGL((batch_embed_kernel<cpx<float>, T>), grid, dim3(256), 0, s, tab);
        break;
    default:
        hipLaunchKernelGGL((batch_embed_kernel<cpx<double>, T>), grid, dim3(256), 0, s, tab);
        break;
    }
    FM_CHECK_LAUNCH();
    return run_program<T>(p, p->sprog, s);
}
}  // namespace fm

extern "C" fm_status fm_batch_gen_workspace_size(const fm_batch_plan *p, int64_t *bytes) {
    FM_REQUIRE(p && bytes, FM_ERR_VALUE, "NULL argument");
    int64_t t = 0;
    for (int i = 0; i < p->count; ++i) t += spec_ws_entry(p->plans[i]);
    *bytes = t;
    return FM_SUCCESS;
}

extern "C" fm_status fm_batch_gen_bind(fm_batch_plan *p, int gen_dtype, const void *const *gens,
                                       void *workspace, int64_t workspace_bytes, void *stream) {
    FM_REQUIRE(p && gens && workspace, FM_ERR_VALUE, "NULL argument");
    FM_REQUIRE(p->count <= 65535, FM_ERR_VALUE, "batch generators: more than 65535 operators");
    FM_REQUIRE(gen_dtype == FM_FLOAT32 || gen_dtype == FM_FLOAT64 || gen_dtype == FM_COMPLEX64 ||
                   gen_dtype == FM_COMPLEX128,
               FM_ERR_TYPE,
               "generator element type %d: expected float32, float64, complex64 or complex128",
               gen_dtype);
    int64_t need = 0;
    fm_batch_gen_workspace_size(p, &need);
    FM_REQUIRE(workspace_bytes >= need, FM_ERR_VALUE,
               "workspace of %lld bytes is smaller than the %lld required",
               (long long)workspace_bytes, (long long)need);
    p->s_gen_dtype = gen_dtype;
    if (p->prec == FM_COMPLEX64)
        return batch_spec_bind_t<float>(p, gens, (char *)workspace, as_stream(stream));
    return batch_spec_bind_t<double>(p, gens, (char *)workspace, as_stream(stream));
}

extern "C" fm_status fm_batch_gen_apply(const fm_batch_plan *p, void *stream) {
    FM_REQUIRE(p != nullptr, FM_ERR_VALUE, "plan is NULL");
    FM_REQUIRE(p->sprog.bound && p->stab, FM_ERR_VALUE,
               "batch generators used before fm_batch_gen_bind");
    if (p->prec == FM_COMPLEX64) return batch_spec_apply_t<float>(p, as_stream(stream));
    return batch_spec_apply_t<double>(p, as_stream(stream));
}

namespace fm {
// The spectra from the generators and the forward apply, with the forward's
// column transforms of the inputs (which need no spectrum) running on a
// side stream while the spectra are formed: fork, early launches on the
// last side stream, embedding + spectrum program on the caller's stream,
// join, the rest of the forward program.  Without a free side stream it is
// the two calls in sequence.
template <typename T>
static fm_status batch_gen_forward_t(const fm_batch_plan *p, hipStream_t s) {
    const Program &d = p->dir[0];
    constexpr int K = fm_batch_plan::kSide;
    bool any_early = false;
    for (const Launch &l : d.launches) any_early = any_early || l.early;
    // the spectrum program runs on the caller's stream and its first
    // sprog.lanes - 1 side streams; the early launches go to the others
    const int first = p->sprog.lanes > 1 ? p->sprog.lanes - 1 : 0;
    const bool overlap = p->has_side && any_early && first < K;
    if (!overlap) {
        fm_status st = batch_spec_apply_t<T>(p, s);
        return st ? st : run_program<T>(p, d, s);
    }
    const int nfree = K - first;
    FM_CHECK_HIP(hipEventRecord(p->fork_ev, s));
    for (int q = first; q < K; ++q) FM_CHECK_HIP(hipStreamWaitEvent(p->side[q], p->fork_ev, 0));
    for (const Launch &l : d.launches) {
        if (!l.early) continue;
        fm_status st = launch_one<T>(p, d, l, p->side[first + l.lane % nfree]);
        if (st) return st;
    }
    for (int q = first; q < K; ++q) FM_CHECK_HIP(hipEventRecord(p->join_ev[q], p->side[q]));
    fm_status st = batch_spec_apply_t<T>(p, s);
    if (st) return st;
    for (int q = first; q < K; ++q) FM_CHECK_HIP(hipStreamWaitEvent(s, p->join_ev[q], 0));
    return run_program<T>(p, d, s, true);
}
}  // namespace fm

extern "C" fm_status fm_batch_gen_forward_apply(const fm_batch_plan *p, void *stream) {
    FM_REQUIRE(p != nullptr, FM_ERR_VALUE, "plan is NULL");
    FM_REQUIRE(p->sprog.bound && p->stab, FM_ERR_VALUE,
               "batch generators used before fm_batch_gen_bind");
    FM_REQUIRE(p->dir[0].bound, FM_ERR_VALUE, "batch forward direction is not bound to buffers");
    if (p->prec == FM_COMPLEX64) return batch_gen_forward_t<float>(p, as_stream(stream));
    return batch_gen_forward_t<double>(p, as_stream(stream));
}

extern "C" fm_status fm_batch_plan_destroy(fm_batch_plan *p) {
    if (!p) return FM_SUCCESS;
    free_program(p->dir[0]);
    free_program(p->dir[1]);
    free_program(p->gpassA);
    free_program(p->gcorr);
    free_program(p->gpassC);
    free_program(p->sprog);
    if (p->gtab) (void)hipFree(p->gtab);
    if (p->stab) (void)hipFree(p->stab);
    if (p->res_part) (void)hipFree(p->res_part);
    for (int q = 0; q < fm_batch_plan::kSide; ++q) {
        if (p->side[q]) (void)hipStreamDestroy(p->side[q]);
        if (p->join_ev[q]) (void)hipEventDestroy(p->join_ev[q]);
    }
    if (p->fork_ev) (void)hipEventDestroy(p->fork_ev);
    delete p;
    return FM_SUCCESS;
}
