// FFT engine: LDS-resident batched Stockham FFT tiles and the tile kernel
// used by every Fourier / Circulant / Toeplitz / Bluestein pass.
//
// Method (written from the published Stockham autosort formulation): a
// length-N transform (N = 2^logN) is done as radix-8 stages plus one radix-2
// or radix-4 stage.  Stage with radix R and sub-length p (1, R, R^2, ...):
// butterfly i in [0, N/R) reads X[i + r N/R], multiplies input r by
// W_{pR}^{r k} with k = i mod p, does an R-point DFT in registers and writes
// Y[(i - k) R + k + r p].  Output comes out in natural order; all stages run
// in place on one LDS buffer (read to registers, barrier, write, barrier).
//
// Tile kernel: a tile is B consecutive sequences s = o*J + j of a (O x J)
// sequence space.  Element n of sequence (o, j) is read at
//     x[o*in_os + j*in_js + n*in_ns]
// and written at y[o*out_os + j*out_js + n*out_ns]; this one addressing form
// covers contiguous columns, four-step passes (strided and transposed sides)
// and Kron mode axes.  Per-element work is fused around the transform:
//     load:  zero-pad (global index n*in_gn + j*in_gj >= n_valid), conj,
//            pre-multiply by a chirp, then FFT(sign1)
//     conv:  multiply by the spectrum D[n*d_gn + j*d_gj] (or conj), FFT(sign2)
//     store: twiddle W_L^{+-j n} of an enclosing four-step, crop
//            (n*out_gn + j*out_gj >= m_valid), post-multiply, conj, scale.
#pragma once

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

template <int R, int SIGN, typename T> __device__ __forceinline__ void dft(cpx<T> *v) {
    if constexpr (R == 2) dft2<SIGN>(v);
    else if constexpr (R == 4) dft4<SIGN>(v);
    else dft8<SIGN>(v);
}

template <int SIGN, typename T>
__device__ __forceinline__ cpx<T> twiddle_mul(cpx<T> v, cpx<T> w) {
    // w is the forward twiddle exp(-2 pi i e / N); the inverse uses conj(w)
    return SIGN < 0 ? cmul(v, w) : cmulc(v, w);
}

// One in-place Stockham stage over B sequences of length N held in LDS
// (sequence b at lds + b*S).  `tw` holds W_N^e = exp(-2 pi i e / N), e < N.
template <typename T, int R, int SIGN, int NT, int MAXE>
__device__ __forceinline__ void fft_stage(cpx<T> *lds, int S, int logN, int B,
                                          int logp, const cpx<T> *__restrict__ tw) {
    constexpr int LOGR = R == 2 ? 1 : (R == 4 ? 2 : 3);
    constexpr int MAXB = MAXE / R;
    const int logNR = logN - LOGR;
    const int NR = 1 << logNR;
    const int total = B << logNR;
    const int p = 1 << logp;
    cpx<T> v[MAXB * R];
#pragma unroll
    for (int e = 0; e < MAXB; ++e) {
        const int g = threadIdx.x + e * NT;
        if (g < total) {
            const int b = g >> logNR, i = g & (NR - 1);
            const cpx<T> *src = lds + b * S + i;
#pragma unroll
            for (int r = 0; r < R; ++r) v[e * R + r] = src[r * NR];
        }
    }
    __syncthreads();
#pragma unroll
    for (int e = 0; e < MAXB; ++e) {
        const int g = threadIdx.x + e * NT;
        if (g < total) {
            const int b = g >> logNR, i = g & (NR - 1);
            const int k = i & (p - 1);
            if (logp > 0) {
                // W_{pR}^{r k} = W_N^{r k N / (p R)}
                const int step = k << (logN - logp - LOGR);
#pragma unroll
                for (int r = 1; r < R; ++r)
                    v[e * R + r] = twiddle_mul<SIGN>(v[e * R + r], tw[r * step]);
            }
            dft<R, SIGN>(v + e * R);
            cpx<T> *dst = lds + b * S + ((i - k) << LOGR) + k;
#pragma unroll
            for (int r = 0; r < R; ++r) dst[r * p] = v[e * R + r];
        }
    }
    __syncthreads();
}

template <typename T, int SIGN, int NT, int MAXE>
__device__ void lds_fft(cpx<T> *lds, int S, int logN, int B, const cpx<T> *__restrict__ tw) {
    int logp = 0;
    int rem = logN;
    while (rem >= 3) {
        fft_stage<T, 8, SIGN, NT, MAXE>(lds, S, logN, B, logp, tw);
        logp += 3;
        rem -= 3;
    }
    if (rem == 2) fft_stage<T, 4, SIGN, NT, MAXE>(lds, S, logN, B, logp, tw);
    else if (rem == 1) fft_stage<T, 2, SIGN, NT, MAXE>(lds, S, logN, B, logp, tw);
}

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
    const cpx<T> *tw;               // W_N^e table
    // spectral multiply + second transform (conv mode)
    const cpx<T> *D;
    int64_t d_gn, d_gj;
    int conj_d;
    // enclosing four-step twiddle applied before the store: W_L^{tw_sign * j * n}
    const cpx<T> *twb_hi, *twb_lo;  // W_L^{a * 2^twb_shift}, W_L^{b}
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
};

template <typename T, int NT, int MAXE, bool CONV>
__global__ __launch_bounds__(NT) void fft_tile_kernel(const TileArgs<T> a) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const int B = a.B;
    const int N = 1 << a.logN;
    const int S = N + 1;
    cpx<T> *lds = reinterpret_cast<cpx<T> *>(smem_raw);
    int64_t *seq_in = reinterpret_cast<int64_t *>(lds + B * S);
    int64_t *seq_out = seq_in + B;
    int64_t *seq_j = seq_out + B;

    const int64_t tile = xcd_remap(blockIdx.x, gridDim.x);
    const int64_t s0 = tile * B;
    const int64_t nseq = a.O * a.J;
    for (int b = threadIdx.x; b < B; b += NT) {
        const int64_t s = s0 + b;
        if (s < nseq) {
            const int64_t o = s / a.J, j = s - o * a.J;
            seq_in[b] = o * a.in_os + j * a.in_js;
            seq_out[b] = o * a.out_os + j * a.out_js;
            seq_j[b] = j;
        } else {
            seq_j[b] = -1;
        }
    }
    __syncthreads();

    // ---- load (zero-pad, conj, pre-multiply)
    const int tot = B * N;
    int logB = 0;
    while ((1 << logB) < B) ++logB;
    for (int idx = threadIdx.x; idx < tot; idx += NT) {
        int b, n;
        if (a.seqfast_in) { b = idx & (B - 1); n = idx >> logB; }
        else { n = idx & (N - 1); b = idx >> a.logN; }
        const int64_t j = seq_j[b];
        cpx<T> v = mk(T(0), T(0));
        if (j >= 0) {
            const int64_t gidx = n * a.in_gn + j * a.in_gj;
            if (gidx < a.n_valid) {
                v = a.x[seq_in[b] + n * a.in_ns];
                if (a.conj_in) v = conj(v);
                if (a.pre) v = cmul(v, a.pre[gidx]);
            }
        }
        lds[b * S + n] = v;
    }
    __syncthreads();

    if (a.sign1 < 0) lds_fft<T, -1, NT, MAXE>(lds, S, a.logN, B, a.tw);
    else lds_fft<T, +1, NT, MAXE>(lds, S, a.logN, B, a.tw);

    if constexpr (CONV) {
        for (int idx = threadIdx.x; idx < tot; idx += NT) {
            const int n = idx & (N - 1), b = idx >> a.logN;
            const int64_t j = seq_j[b];
            if (j >= 0) {
                const cpx<T> d = a.D[n * a.d_gn + j * a.d_gj];
                lds[b * S + n] = a.conj_d ? cmulc(lds[b * S + n], d) : cmul(lds[b * S + n], d);
            }
        }
        __syncthreads();
        if (a.sign1 < 0) lds_fft<T, +1, NT, MAXE>(lds, S, a.logN, B, a.tw);
        else lds_fft<T, -1, NT, MAXE>(lds, S, a.logN, B, a.tw);
    }

    // ---- store (four-step twiddle, crop, post-multiply, conj, scale)
    const int lo_mask = (1 << a.twb_shift) - 1;
    for (int idx = threadIdx.x; idx < tot; idx += NT) {
        int b, n;
        if (a.seqfast_out) { b = idx & (B - 1); n = idx >> logB; }
        else { n = idx & (N - 1); b = idx >> a.logN; }
        const int64_t j = seq_j[b];
        if (j < 0) continue;
        const int64_t gidx = n * a.out_gn + j * a.out_gj;
        if (gidx >= a.m_valid) continue;
        cpx<T> v = lds[b * S + n];
        if (a.twb_hi) {
            const int64_t e = j * n;  // < L
            const cpx<T> w = cmul(a.twb_hi[e >> a.twb_shift], a.twb_lo[e & lo_mask]);
            v = a.twb_sign < 0 ? cmul(v, w) : cmulc(v, w);
        }
        if (a.post) v = cmul(v, a.post[gidx]);
        if (a.conj_out) v = conj(v);
        v = scale(v, a.out_scale);
        a.y[seq_out[b] + n * a.out_ns] = v;
    }
}

// Host-side launcher for one tile pass (defined in fm_fft.hip).
template <typename T>
fm_status launch_tile(const TileArgs<T> &a, bool conv, hipStream_t stream);

// Largest tile (B*N elements) the engine supports for one precision.
template <typename T> constexpr int max_tile_elems() {
    return sizeof(T) == 4 ? 16384 : 8192;
}

}  // namespace fm
