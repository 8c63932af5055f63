// The last two factors of a Kron product, a 128-point Circulant and a dense
// 128 x 128 float32 matrix, applied to complex64 data in one pass over HBM.
//
// Replaces two mode passes of Kron._forward / _backward (Kron.pyx:267-303:
// one factor per mode, each a full read and write of the data) for
// x viewed as (outer, n = 128, m = 128) with the dense axis contiguous: for
// every outer index the 128 x 128 slab (128 KB, contiguous in HBM) is read
// once, transformed by both factors on the chip and written once.
//
// Per slab, one 1024-thread workgroup:
//   1. loads the slab into registers in the FFT tile layout (sequences along
//      the contiguous axis, 128 of them, elements along the circulant axis);
//   2. runs the circulant as forward FFT, spectrum, inverse FFT on the
//      registers and LDS (the tile engine of fm_fft_core.hpp);
//   3. writes the slab to LDS row-major (rows = circulant index, padded to
//      132 complex so that the MFMA operand reads below hit 64 distinct
//      banks) and multiplies it by the dense factor on FP32 MFMA: the 256
//      real columns (circulant index, real / imaginary part) in two halves,
//      16 waves as 8 row blocks of A (loaded into registers in the MFMA
//      operand layout) x 2 halves, each half eight 16-column blocks, K = 128
//      in 32 steps of v_mfma_f32_16x16x4f32;
//   4. stages the result through LDS and stores it as contiguous 16-byte
//      vectors.
// The factors act on different axes, so their order does not matter; the
// adjoint of (Circulant x A) is the conjugate spectrum with A^T, which the
// caller passes as the matrix (A is real).
#include "fm_fft_plan.hpp"

namespace fm {

typedef float kf_f32x4 __attribute__((ext_vector_type(4)));

constexpr int KF_LOGN = 7;             // circulant length 128 (middle axis)
constexpr int KF_N = 1 << KF_LOGN;
constexpr int KF_M = 128;              // dense size (contiguous axis)
constexpr int KF_NT = 1024;            // 16 wave64
constexpr int KF_LE = 4;               // 16 elements per thread in the transform
constexpr int KF_ROW = KF_M + 4;       // padded LDS row of the slab (complex)
constexpr int KF_SLAB = KF_N * KF_M;   // elements per slab

struct KronPairArgs {
    TileArgs<float> t;  // the slab as FFT sequences (load side, twiddles, spectrum)
    cpx<float> *y;
    const float *A;     // the dense factor as applied, 128 x 128 row-major
};

template <bool I32>
__global__ __launch_bounds__(KF_NT) void kron_conv_dense_kernel(const KronPairArgs k) {
    using S = TileShape<float, KF_LOGN, KF_NT, KF_LE>;
    using P = typename S::P;
    using Ix = IdxT<I32>;
    constexpr int E = S::E, LOGE = S::LOGE, LOGB = S::LOGB;
    static_assert(S::B == KF_M, "one tile is one slab");
    // radices of the forward transform's last stage and of the reversed inverse's
    constexpr int LOGRF = P::bits(P::S - 1), RF = 1 << LOGRF, NBF = E / RF;
    constexpr int LOGNRF = KF_LOGN - LOGRF, NRF = 1 << LOGNRF;
    constexpr int LOGRL = P::bits(0), RL = 1 << LOGRL, NBL = E / RL;
    constexpr int LOGNRL = KF_LOGN - LOGRL, NRL = 1 << LOGNRL;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    cpx<float> *lds = reinterpret_cast<cpx<float> *>(smem_raw);
    const float *ldsf = reinterpret_cast<const float *>(smem_raw);
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const TileArgs<float> &t = k.t;
    int64_t tile = blockIdx.x;
    if (tile >= t.ntiles) return;  // uniform over the workgroup

    for (; tile < t.ntiles; tile += gridDim.x) {
        cpx<float> v[E];
        fft_tile_load<float, KF_LOGN, KF_NT, I32, KF_LE>(t, tile, v);
        // ---- circulant along the middle axis: forward, spectrum, inverse
        bool sf = true;  // the load mapping: neighbouring lanes on neighbouring sequences
        fft_all<float, KF_LOGN, LOGE, -1, KF_NT, LOGB, 0, 1>(v, lds, t.tw, sf);
        if (P::S > 1) sf = false;
#pragma unroll
        for (int e = 0; e < NBF; ++e) {
            int b, i;
            bfly_pos<LOGB, LOGNRF>(tid + e * KF_NT, sf, b, i);
#pragma unroll
            for (int r = 0; r < RF; ++r) {
                const cpx<float> d = t.D[i + r * NRF];
                v[e * RF + r] = t.conj_d ? cmulc(v[e * RF + r], d) : cmul(v[e * RF + r], d);
            }
        }
        fft_all<float, KF_LOGN, LOGE, +1, KF_NT, LOGB, 0, 1, true>(v, lds, t.tw, sf);
        if (P::S > 1) sf = false;
        // ---- the slab to LDS, row n (circulant index), column b (dense input index)
        __syncthreads();  // the transform's last LDS reads are done
#pragma unroll
        for (int e = 0; e < NBL; ++e) {
            int b, i;
            bfly_pos<LOGB, LOGNRL>(tid + e * KF_NT, sf, b, i);
#pragma unroll
            for (int r = 0; r < RL; ++r) lds[(i + r * NRL) * KF_ROW + b] = v[e * RL + r];
        }
        // ---- dense factor along the contiguous axis on FP32 MFMA:
        // D[i][c] = sum_j A[i][j] * slab[c >> 1][j].part(c & 1); wave w owns
        // output rows 16 (w & 7) .. + 15 and the columns of half w >> 3
        // (128 real columns: slab rows 64 (w >> 3) .. + 63), as eight 16 x 16
        // blocks, K = 128 in 32 steps
        const int rb = wave & 7, h = wave >> 3, q = lane & 15;
        // this wave's rows of A in the MFMA operand layout, a[st] =
        // A[16 rb + (lane & 15)][4 st + (lane >> 4)] (from L2: 64 KB shared by
        // every workgroup)
        float a[KF_M / 4];
        {
            // one base offset per lane, the k-steps at constant distances
            const uint32_t base = (16 * rb + q) * KF_M + (lane >> 4);
#pragma unroll
            for (int st = 0; st < KF_M / 4; ++st) a[st] = ldg(k.A, uint32_t(base + 4 * st));
        }
        // every wave's slab rows are in LDS before any wave reads them (the
        // A loads above are in flight meanwhile)
        __syncthreads();
        // column c = 128 h + 16 cb + q is row c >> 1 = 64 h + 8 cb + (q >> 1),
        // part q & 1; with j = 4 st + (lane >> 4) every operand is this lane's
        // base plus a constant
        const float *bp =
            ldsf + (((64 * h + (q >> 1)) * KF_ROW + (lane >> 4)) * 2 + (q & 1));
        kf_f32x4 acc[8];
#pragma unroll
        for (int cb = 0; cb < 8; ++cb) acc[cb] = kf_f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int st = 0; st < KF_M / 4; ++st) {
#pragma unroll
            for (int cb = 0; cb < 8; ++cb) {
                const float bv = bp[(8 * cb * KF_ROW + 4 * st) * 2];
                acc[cb] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[st], bv, acc[cb], 0, 0, 0);
            }
            // keep the scheduler from hoisting every k-step's operand reads
            // (the unrolled loop would hold them all in registers; grouping
            // 2 or 4 k-steps per barrier spilled and measured 9% slower)
            __builtin_amdgcn_sched_barrier(0);
        }
        __syncthreads();  // every wave is done reading the slab
        // ---- stage the result: D[i][c], i = 16 rb + 4 (lane >> 4) + u, goes
        // to row c >> 1, column i
        {
            float *wp = reinterpret_cast<float *>(smem_raw) +
                        (((64 * h + (q >> 1)) * KF_ROW + 16 * rb + 4 * (lane >> 4)) * 2 + (q & 1));
#pragma unroll
            for (int cb = 0; cb < 8; ++cb)
#pragma unroll
                for (int u = 0; u < 4; ++u) wp[(8 * cb * KF_ROW + u) * 2] = acc[cb][u];
        }
        __syncthreads();
        {
            const Ix base = Ix(tile) * Ix(KF_SLAB);
            constexpr int PER = KF_SLAB / 2 / KF_NT;  // pairs of elements per thread
#pragma unroll
            for (int e = 0; e < PER; ++e) {
                const int p = tid + e * KF_NT;        // pair index in the slab
                const int n = (2 * p) / KF_M, i = (2 * p) % KF_M;
                const float4 w = *reinterpret_cast<const float4 *>(&lds[n * KF_ROW + i]);
                if (t.stream_out & 1) {
                    FM_GLOBAL kf_f32x4 *q = (FM_GLOBAL kf_f32x4 *)(FM_GLOBAL char *)k.y +
                                            (base / 2 + Ix(p));
                    __builtin_nontemporal_store(kf_f32x4{w.x, w.y, w.z, w.w}, q);
                } else {
                    stg(reinterpret_cast<float4 *>(k.y), Ix(base / 2 + Ix(p)), w);
                }
            }
        }
        __syncthreads();  // the LDS is rewritten by the next slab's transform
    }
}

}  // namespace fm

using namespace fm;

extern "C" fm_status fm_kron_conv_dense_apply(const fm_conv_plan *plan, int conj_circulant,
                                              const void *M, int64_t ldm, const void *x, void *y,
                                              int64_t outer, void *stream) {
    FM_REQUIRE(plan != nullptr && M != nullptr, FM_ERR_VALUE, "NULL plan or matrix");
    FM_REQUIRE(plan->kind != 2 && plan->prec == FM_COMPLEX64 && plan->rows == KF_N &&
                   plan->cols == KF_N && plan->L == KF_N,
               FM_ERR_NOT_IMPLEMENTED,
               "the fused Kron pair needs a complex64 128-point Circulant plan");
    FM_REQUIRE(ldm == KF_M, FM_ERR_VALUE, "the fused Kron pair needs a contiguous 128 x 128 matrix");
    FM_REQUIRE(outer >= 0, FM_ERR_VALUE, "negative slab count");
    if (outer == 0) return FM_SUCCESS;
    FM_REQUIRE(x != nullptr && y != nullptr, FM_ERR_VALUE, "NULL data pointer");
    FM_REQUIRE(x != y, FM_ERR_VALUE, "the fused Kron pair does not run in place");
    FM_REQUIRE((uintptr_t(y) & 15) == 0, FM_ERR_VALUE, "output must be 16-byte aligned");
    KronPairArgs k;
    memset(&k, 0, sizeof(k));
    TileArgs<float> &t = k.t;
    t = blank_args<float>();
    t.logN = KF_LOGN;
    t.O = outer;
    t.J = KF_M;
    t.B = KF_M;
    t.ntiles = outer;
    t.logJ = 7;
    t.x = (const cpx<float> *)x;
    t.in_os = KF_SLAB; t.in_js = 1; t.in_ns = KF_M;
    t.in_gn = 1; t.in_gj = 0;
    t.seqfast_in = 1;
    t.tw = (const cpx<double> *)plan->geo.tw_full;
    t.D = (const cpx<float> *)plan->D;
    t.d_gn = 1; t.d_gj = 0;
    t.conj_d = conj_circulant ? 1 : 0;
    // the slabs are read once and written once: non-temporal loads and
    // stores (the kron workload 0.53 -> 0.50 ms per step)
    t.stream_out = 3;
    k.y = (cpx<float> *)y;
    k.A = (const float *)M;
    const size_t lds = size_t(KF_N) * (KF_N + (KF_N >> 5) + 1) * sizeof(cpx<float>) >
                               size_t(KF_N) * KF_ROW * sizeof(cpx<float>)
                           ? size_t(KF_N) * (KF_N + (KF_N >> 5) + 1) * sizeof(cpx<float>)
                           : size_t(KF_N) * KF_ROW * sizeof(cpx<float>);
    static bool attr = false;
    if (!attr) {
        (void)hipFuncSetAttribute((const void *)kron_conv_dense_kernel<true>,
                                  hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
        (void)hipFuncSetAttribute((const void *)kron_conv_dense_kernel<false>,
                                  hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
        attr = true;
    }
    int dev = 0, ncu = 256;
    if (hipGetDevice(&dev) == hipSuccess)
        (void)hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
    int64_t grid = ncu;
    if (grid > outer) grid = outer;
    // 32-bit element offsets when the whole block stays below 4 GiB
    const bool i32 = double(outer) * KF_SLAB * sizeof(cpx<float>) < 4294967296.0;
    ProfToken tok;
    prof_begin(as_stream(stream), &tok);
    const dim3 gd = dim3(unsigned(grid)), bd = dim3(KF_NT);
    hipStream_t s = as_stream(stream);
    if (i32)
        hipLaunchKernelGGL((kron_conv_dense_kernel<true>), gd, bd, lds, s, k);
    else
        hipLaunchKernelGGL((kron_conv_dense_kernel<false>), gd, bd, lds, s, k);
    FM_CHECK_LAUNCH();
    if (tok.active) {
        const double bytes = 2.0 * double(outer) * KF_SLAB * sizeof(cpx<float>) +
                             double(KF_M) * KF_M * 4.0 + double(KF_N) * 8.0;
        const double flops = 2.0 * double(KF_M) * KF_M * 2.0 * double(KF_N) * double(outer);
        prof_end(as_stream(stream), &tok, "kron_conv_dense_kernel", bytes, flops);
    }
    return FM_SUCCESS;
}
