/*
 * fastmat_amd — C ABI of the MI355X (gfx950) structured-matvec library.
 *
 * Drop-in boundary for fastmat's C-style transform hook
 * `_forwardC(arrX, arrRes, typeX, typeRes)` / `_backwardC(...)`
 * (fastmat/Matrix.pyx:1819-1829, 1912-1922): every apply entry point takes a
 * read-only input block X and writes a caller-allocated output block Y, both
 * column-major with k right-hand sides (column j of X starts at x + j*ldx
 * elements), exactly the F-contiguous (n, k) arrays fastmat passes.  All
 * pointers are device pointers; `stream` is a hipStream_t (NULL = default
 * stream).  Calls are stream-ordered and never synchronise with the host.
 *
 * Plans (Fourier, Circulant/Toeplitz) are immutable after creation and may be
 * shared between threads, like fastmat matrices (Matrix.pyx:341-358).
 * Multi-pass entry points take a caller-owned workspace whose size is given by
 * the matching *_workspace_size query; apply calls allocate nothing.
 *
 * Element types use the reference's type ids (fastmat/core/types.pxd:45-55).
 * Errors are integer status codes; fm_get_last_error() returns the message of
 * the calling thread's last failure.  The Python layer maps FM_ERR_VALUE to
 * ValueError, FM_ERR_TYPE to TypeError, FM_ERR_NOT_IMPLEMENTED to
 * NotImplementedError, and the others to RuntimeError / MemoryError.
 */
#ifndef FASTMAT_AMD_H
#define FASTMAT_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FM_ABI_VERSION 1

typedef int32_t fm_status;
enum {
    FM_SUCCESS = 0,
    FM_ERR_VALUE = 1,           /* bad size/shape/argument   -> ValueError */
    FM_ERR_TYPE = 2,            /* unsupported element type  -> TypeError */
    FM_ERR_HIP = 3,             /* HIP runtime failure       -> RuntimeError */
    FM_ERR_ALLOC = 4,           /* device allocation failed  -> MemoryError */
    FM_ERR_NOT_IMPLEMENTED = 5  /* valid but unsupported     -> NotImplementedError */
};

/* Element type ids: same order as the reference's TYPES enum
 * (core/types.pxd:45-55). */
enum {
    FM_INT8 = 0,
    FM_INT16 = 1,
    FM_INT32 = 2,
    FM_INT64 = 3,
    FM_FLOAT32 = 4,
    FM_FLOAT64 = 5,
    FM_COMPLEX64 = 6,
    FM_COMPLEX128 = 7
};

/* ------------------------------------------------------------ library */
int fm_abi_version(void);
const char *fm_get_last_error(void);

/* ----------------------------------------------------------- Hadamard
 * Y = H_n X, H_n the unnormalised Sylvester Hadamard matrix of size 2^order.
 * Replaces Hadamard._forwardC / _backwardC (Hadamard.pyx:164-239; the
 * transform is symmetric, so backward == forward).  Any of the 8 element
 * types; integer types wrap modulo 2^bits like the reference's C loop
 * (bit-exact); float types apply the radix-2 stages in the reference's order
 * (distance 1, 2, 4, ...), so float results equal the reference's bit for bit.
 * In-place (x == y, ldx == ldy) is allowed. */
fm_status fm_hadamard_apply(int dtype, int order, const void *x, int64_t ldx,
                            void *y, int64_t ldy, int64_t k, void *stream);

/* Mode-wise Hadamard along a strided axis: for every (o, j) in
 * [0, outer) x [0, inner) the vector v[i] = x[(o*2^order + i)*inner + j] is
 * transformed into y at the same positions (Kron mode application,
 * Kron.pyx:267-303). */
fm_status fm_hadamard_apply_strided(int dtype, int order, const void *x,
                                    void *y, int64_t outer, int64_t inner,
                                    void *stream);

/* Host-side self-check of the Hadamard pass kernel's tile index mapping for a
 * tile of `threads` threads (256, 512 or 1024) holding 2^loge elements each
 * over vectors of 2^logn, with the tile shape derived for `shape_threads`
 * (equal to `threads` in the kernel; 512 with 1024 threads restates the
 * shape error a wrong log2 of the thread count makes): *bad receives how many
 * (vector, element) slots the strided load, the stage exchanges or the
 * strided re-stage store do not hit exactly once (0 = bijective; -1 = no
 * such tile).  No counterpart in the reference; runs on the CPU. */
fm_status fm_wht_tile_check(int logn, int threads, int shape_threads, int loge, int64_t *bad);

/* ------------------------------------------------------------- Fourier
 * Unnormalised DFT of order N, F[j,l] = exp(-2 pi i j l / N); adjoint = F^H.
 * Replaces Fourier.__init__ / _forward / _backward (Fourier.pyx:63-238).
 * precision: FM_COMPLEX64 (FP32 butterflies, FP64-generated twiddles and
 * chirps) or FM_COMPLEX128.  Power-of-two N uses radix kernels (one LDS pass
 * up to the tile size, else a two-pass four-step); any other N uses Bluestein
 * with a power-of-two length L >= 2N-1.  X and Y are complex of `precision`. */
typedef struct fm_fourier_plan fm_fourier_plan;
fm_status fm_fourier_plan_create(int64_t order, int precision, void *stream,
                                 fm_fourier_plan **plan);
fm_status fm_fourier_plan_info(const fm_fourier_plan *plan, int64_t *fft_length,
                               int *bluestein, int *passes);
fm_status fm_fourier_workspace_size(const fm_fourier_plan *plan, int64_t k,
                                    int64_t *bytes);
fm_status fm_fourier_apply(const fm_fourier_plan *plan, int adjoint,
                           const void *x, int64_t ldx, void *y, int64_t ldy,
                           int64_t k, void *workspace, int64_t workspace_bytes,
                           void *stream);
/* Mode-wise DFT along a strided axis (Kron mode application,
 * Kron.pyx:267-303): vectors v[i] = x[(o*N + i)*inner + j]; requires a
 * single-pass plan (N up to the LDS tile size). */
fm_status fm_fourier_apply_strided(const fm_fourier_plan *plan, int adjoint,
                                   const void *x, void *y, int64_t outer,
                                   int64_t inner, void *stream);
fm_status fm_fourier_plan_destroy(fm_fourier_plan *plan);

/* ------------------------------------------------- Circulant / Toeplitz
 * Y = A X with A the top-left rows x cols block of an L x L circulant whose
 * generator t' embeds the structure; the spectrum fft(t')/L is computed at
 * plan creation in FP64 on the device and stored in `precision`.
 *   Circulant(c), c of length n: A[i,j] = c[(i-j) mod n]
 *     (Circulant.pyx:104-137, 414-425); L = n for power-of-two n, otherwise
 *     the smallest power of two >= 2n-1 with t' = [c, 0..., c[1:]].
 *   Toeplitz(vecC, vecR), m x n: tenT = [vecC (m), vecR (n-1)],
 *     A[i,j] = tenT[(i-j) mod (m+n-1)] (Toeplitz.pyx:146-176, 719-732);
 *     L = smallest power of two >= m+n-1, zeros inserted at the split point
 *     (Toeplitz.pyx:357-366).
 * `generator` is a device pointer to complex128 values (c, or tenT).
 * Replaces Partial(Product(F.H, Diag(fft(t')/L), F)) forward/backward
 * (Partial.pyx:268-294, Product.pyx:205-240, Diag.pyx:149-167).
 * Adjoint applies A^H. */
typedef struct fm_conv_plan fm_conv_plan;
fm_status fm_circulant_plan_create(int precision, int64_t n,
                                   const void *generator, void *stream,
                                   fm_conv_plan **plan);
fm_status fm_toeplitz_plan_create(int precision, int64_t m, int64_t n,
                                  const void *generator, void *stream,
                                  fm_conv_plan **plan);
/* Two-level circulant of size (n0 n1) x (n0 n1), n0 and n1 powers of two
 * (2 .. 2^13 for complex64, 2 .. 2^12 for complex128): the multi-level
 * Circulant of Circulant.pyx:138-215 with power-of-two levels, here as three
 * passes (level-0 DFT, level-1 DFT / spectrum / inverse, level-0 inverse)
 * instead of Kron(Fourier) products.  `spectrum` is device complex128,
 * fftn(t) / (n0 n1) in C order (n1 fastest).  Applied with fm_conv_apply /
 * fm_conv_workspace_size; no strided apply, batch or generator gradient
 * (FM_ERR_NOT_IMPLEMENTED). */
fm_status fm_circulant2_plan_create(int precision, int64_t n0, int64_t n1, const void *spectrum,
                                    void *stream, fm_conv_plan **plan);
fm_status fm_conv_plan_info(const fm_conv_plan *plan, int64_t *rows,
                            int64_t *cols, int64_t *fft_length, int *passes);
fm_status fm_conv_workspace_size(const fm_conv_plan *plan, int64_t k,
                                 int64_t *bytes);
fm_status fm_conv_apply(const fm_conv_plan *plan, int adjoint, const void *x,
                        int64_t ldx, void *y, int64_t ldy, int64_t k,
                        void *workspace, int64_t workspace_bytes, void *stream);
/* Allocation-free fm_conv_apply: no counterpart in the reference; for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  The four-step
 * intermediate lives in a 64 MiB scratch array of the library (one per
 * device, held until the process ends) instead of a caller workspace; the
 * columns run in chunks that fit it.  A transform whose single column does
 * not fit (L * element size > 64 MiB: L > 2^23 complex64, L > 2^22
 * complex128) is rejected with FM_ERR_VALUE.  All calls on one device share
 * the scratch: calls that overlap on one device (other streams, other
 * threads) race; calls on one stream are ordered. */
fm_status fm_conv_apply_noalloc(const fm_conv_plan *plan, int adjoint, const void *x,
                                int64_t ldx, void *y, int64_t ldy, int64_t k, void *stream);
/* Mode-wise application along a strided axis (square plans, single pass). */
fm_status fm_conv_apply_strided(const fm_conv_plan *plan, int adjoint,
                                const void *x, void *y, int64_t outer,
                                int64_t inner, void *stream);
/* Gradient of a real loss with respect to the generator, given the forward
 * input X (cols x k) and the upstream gradient G = dLoss/dY (rows x k), in
 * PyTorch's convention for complex tensors (grad = dL/dRe + i dL/dIm):
 *   grad_t'[d] = sum_cols sum_i G[i] conj(X'[(i-d) mod L])   (a correlation)
 * folded back onto the generator (c for Circulant: t'[0], t'[j] and
 * t'[L-n+j] -> c[j]; [vecC, vecR] for Toeplitz).  `grad` receives complex
 * values of `precision` (n entries for Circulant, m+n-1 for Toeplitz). */
fm_status fm_conv_generator_grad_workspace_size(const fm_conv_plan *plan,
                                                int64_t k, int64_t *bytes);
fm_status fm_conv_generator_grad(const fm_conv_plan *plan, const void *x,
                                 int64_t ldx, const void *g, int64_t ldg,
                                 int64_t k, void *grad, void *workspace,
                                 int64_t workspace_bytes, void *stream);
/* The spectrum as a function of a generator held on the device (pattern (a):
 * the operator *is* its generator's spectrum, Circulant.pyx:130
 * `Diag(fft(tenC)/size)`, Toeplitz.pyx:253-255 `fft(tenThat)`).  Rewrites the
 * plan's spectrum with fft(t')/L of `generator` (device pointer, gen_len
 * elements of gen_dtype: FLOAT32, FLOAT64, COMPLEX64 or COMPLEX128; c for a
 * Circulant, [vecC, vecR] for a Toeplitz), embedded as at plan creation and
 * transformed in the plan's precision (FP32 butterflies with FP64-generated
 * twiddles for complex64).  Stream-ordered: applies issued after it on the
 * same stream use the new spectrum.  Allocates nothing (graph-capturable);
 * the workspace size comes from fm_conv_set_generator_workspace_size.  Not
 * for two-level plans (FM_ERR_NOT_IMPLEMENTED).  A plan whose spectrum is
 * rewritten is no longer immutable: share it only with callers that expect
 * that, or work on a copy from fm_conv_plan_clone. */
fm_status fm_conv_set_generator_workspace_size(const fm_conv_plan *plan, int64_t *bytes);
fm_status fm_conv_set_generator(fm_conv_plan *plan, int gen_dtype, const void *generator,
                                void *workspace, int64_t workspace_bytes, void *stream);
/* A copy of a plan with its own spectrum buffer (synchronises `stream`). */
fm_status fm_conv_plan_clone(const fm_conv_plan *plan, void *stream, fm_conv_plan **copy);
fm_status fm_conv_plan_destroy(fm_conv_plan *plan);

/* ------------------------------------------------- uneven operator batch
 * Many Circulant/Toeplitz plans of one precision, each with its own k_i
 * right-hand sides, applied together.  No single reference call does this;
 * it replaces the per-operator loop of forward()/backward() calls fastmat
 * users write (Matrix.pyx:1844-2007 once per operator).  A direction is a
 * program of persistent launches, one per (pass, transform length) class,
 * over device work tables, issued on the caller's stream and on side streams
 * forked from it and joined back (graph-capturable).  The multi-pass
 * operators run in groups whose four-step intermediates fit a budget, so
 * each group's passes find their input in the memory-side cache
 * (fm_batch_set_schedule).  fm_batch_bind() binds one direction to buffers:
 * input i is cols_i x k_i (rows_i x k_i for the adjoint), output the other
 * shape, and uploads the work tables (synchronising `stream`);
 * fm_batch_apply() then only launches.  The workspace holds, per direction
 * and multi-pass operator, the pass-A and pass-B intermediates.  Plans must
 * outlive the batch plan. */
typedef struct fm_batch_plan fm_batch_plan;
fm_status fm_batch_plan_create(int count, const fm_conv_plan *const *plans,
                               const int64_t *ks, fm_batch_plan **plan);
fm_status fm_batch_workspace_size(const fm_batch_plan *plan, int64_t *bytes);
fm_status fm_batch_set_grid(fm_batch_plan *plan, int workgroups_per_cu);
/* Schedule of the directions: persistent grid (workgroups per CU), the
 * budget of one group's intermediates in MB (0: one group) and the number of
 * lanes (streams, 1..4) the groups alternate between.  A change of the
 * grouping unbinds both directions (bind them again). */
fm_status fm_batch_set_schedule(fm_batch_plan *plan, int workgroups_per_cu, int group_mb,
                                int lanes);
fm_status fm_batch_bind(fm_batch_plan *plan, int adjoint, const void *const *inputs,
                        const int64_t *ld_in, void *const *outputs,
                        const int64_t *ld_out, void *workspace,
                        int64_t workspace_bytes, void *stream);
fm_status fm_batch_apply(const fm_batch_plan *plan, int adjoint, void *stream);

/* fm_batch_apply with the direction's outputs redirected: every element the
 * bound direction writes to its output blocks lands shift_bytes further on
 * (the same layout in another buffer; a whole number of elements) and is
 * multiplied by the device scalar *scale (float for complex64 plans, double
 * for complex128; NULL: 1).  The bound buffers themselves are not written.
 * The backward of a loss writes the input gradient, scaled by the loss's
 * incoming gradient, straight into a new tensor this way (Matrix.pyx:1844
 * applies the adjoint per operator; the scaling is torch autograd's). */
fm_status fm_batch_apply_redirected(const fm_batch_plan *plan, int adjoint, int64_t shift_bytes,
                                    const void *scale, void *stream);
/* Residual mode of the forward direction (the least-squares step
 * sum_i ||A_i x_i - t_i||^2 of a training loop): like fm_batch_bind for the
 * forward direction, but output i receives scale * (A_i x_i - t_i) (t_i a
 * rows_i x k_i target block with the output's leading dimension) and every
 * tile adds up |A_i x_i - t_i|^2 of what it stores; after fm_batch_apply,
 * fm_batch_residual_loss writes the sum over all operators (FP64, summed in a
 * fixed order) to the device double *loss.  Gradient reuse
 * (fm_batch_grad_apply_reuse) works as with fm_batch_bind. */
fm_status fm_batch_bind_residual(fm_batch_plan *plan, const void *const *inputs,
                                 const int64_t *ld_in, const void *const *targets,
                                 const int64_t *ld_t, void *const *outputs,
                                 const int64_t *ld_out, double scale, void *workspace,
                                 int64_t workspace_bytes, void *stream);
fm_status fm_batch_residual_loss(const fm_batch_plan *plan, double *loss, void *stream);
/* Both directions of a least-squares step at once: the forward direction in
 * residual mode on inputs x_i and targets t_i (as fm_batch_bind_residual,
 * residual blocks r_i with the targets' leading dimension), the adjoint
 * direction on inputs r_i with outputs gx_i = A_i^H r_i.  Each multi-pass
 * operator's forward column inverse and the adjoint's column transform of
 * its residual run as one pass (the residual is formed between the two
 * transforms and never written: r_i holds the residual of the one-pass
 * operators only), so after fm_batch_apply(plan, 0) (and
 * fm_batch_residual_loss) fm_batch_apply(plan, 1) starts at the adjoint's
 * row pass.  fm_batch_grad_apply_reuse then gives the generator gradients
 * (the forward's and the adjoint's column transforms are in the workspace);
 * fm_batch_grad_apply without reuse is not valid on these bindings.
 * Backward of Sum_i ||A_i x_i - t_i||^2 (Circulant / Toeplitz backward,
 * Circulant.pyx:104-137, Toeplitz.pyx:211-316, applied to the residual). */
fm_status fm_batch_bind_lsq(fm_batch_plan *plan, const void *const *inputs,
                            const int64_t *ld_in, const void *const *targets,
                            const int64_t *ld_t, void *const *residuals,
                            const int64_t *ld_r, void *const *grads_x,
                            const int64_t *ld_gx, double scale, void *workspace,
                            int64_t workspace_bytes, void *stream);
/* Batched generator gradient (pattern (a) over the uneven batch): for every
 * operator i, with x_i its forward input block (k_i x cols_i, leading
 * dimension ldx[i]) and g_i the gradient of the loss with respect to its
 * output (k_i x rows_i, ldg[i]), grads[i] (gen_len_i elements, the plan's
 * generator layout) receives what fm_conv_generator_grad returns for that
 * operator alone.  Per operator: the column transforms (pass A) of x and g,
 * one correlation launch that sums FFT(g rows) conj(FFT(x rows)) over the
 * columns and transforms back, the column inverse, the fold.  The pass-A
 * results live in the directions' workspace, so both directions must have
 * been bound (fm_batch_bind) first.  fm_batch_grad_bind builds the program
 * for one set of device pointers (allocates its tables, synchronises its
 * stream); fm_batch_grad_apply only launches (capturable).
 * fm_batch_grad_apply_reuse skips the pass-A launches and reads the forward
 * direction's and the adjoint direction's pass-A intermediates instead: the
 * forward direction must be bound to the same x blocks and the adjoint to
 * the same g blocks (checked), and both applied since those intermediates
 * were last overwritten (the caller's ordering, as in a training step:
 * forward, then adjoint, then this). */
fm_status fm_batch_grad_workspace_size(const fm_batch_plan *plan, int64_t *bytes);
fm_status fm_batch_grad_bind(fm_batch_plan *plan, const void *const *xs, const int64_t *ldx,
                             const void *const *gs, const int64_t *ldg, void *const *grads,
                             void *workspace, int64_t workspace_bytes, void *stream);
fm_status fm_batch_grad_apply(const fm_batch_plan *plan, void *stream);
fm_status fm_batch_grad_apply_reuse(const fm_batch_plan *plan, void *stream);
/* The backward of a least-squares step bound by fm_batch_bind_lsq, in one
 * call: fm_batch_apply_redirected(plan, 1, shift_bytes, scale) (the input
 * gradients, scaled by the device scalar *scale, written shift_bytes past
 * the bound gradient blocks) and fm_batch_grad_apply_reuse (the generator
 * gradients, also scaled by *scale, written gen_shift_bytes past the blocks
 * fm_batch_grad_bind gave; that binding must be to the step's inputs and
 * residuals).  Neither needs the other's results, so their launches share
 * one fork of the lanes.  The reference differentiates
 * nothing; this is the backward of sum_i ||A_i x_i - t_i||^2 through
 * Matrix.backward (Matrix.pyx:1844) and the generator's spectrum
 * (Circulant.pyx:130, Toeplitz.pyx:253-260). */
fm_status fm_batch_lsq_backward(const fm_batch_plan *plan, int64_t shift_bytes, const void *scale,
                                int64_t gen_shift_bytes, void *stream);
/* Batched fm_conv_set_generator: every operator's spectrum from its device
 * generator gens[i] (element type gen_dtype for all), as one embedding launch
 * and one batched FFT program.  fm_batch_gen_bind builds the program for one
 * set of generator pointers (allocates its tables, synchronises `stream`);
 * fm_batch_gen_apply only launches (capturable), rewriting the spectra of the
 * plans the batch was created from. */
fm_status fm_batch_gen_workspace_size(const fm_batch_plan *plan, int64_t *bytes);
fm_status fm_batch_gen_bind(fm_batch_plan *plan, int gen_dtype, const void *const *gens,
                            void *workspace, int64_t workspace_bytes, void *stream);
fm_status fm_batch_gen_apply(const fm_batch_plan *plan, void *stream);
/* fm_batch_gen_apply followed by fm_batch_apply(plan, 0), with the forward's
 * column transforms of its inputs (the multi-pass operators' first pass,
 * which needs no spectrum) issued on an internal side stream while the
 * spectra are formed, joined back before the passes that read them; stream-
 * ordered and capturable like the two calls it replaces (which it falls back
 * to when the schedule leaves no side stream free).  The forward of a
 * training step whose operators are recomputed from their generators
 * (Circulant.pyx:130, Toeplitz.pyx:253-255). */
fm_status fm_batch_gen_forward_apply(const fm_batch_plan *plan, void *stream);
fm_status fm_batch_plan_destroy(fm_batch_plan *plan);

/* --------------------------------------------------------------- dense
 * Y = A X (adjoint: Y = A^H X) for a dense m x n matrix A stored row-major
 * (A[i*lda + j]); replaces Matrix._forward / _backward (Matrix.pyx:1831-1842,
 * 1924-1935, ndarray.dot).  Supported (a_dtype, x_dtype) pairs:
 * (FLOAT32, FLOAT32), (FLOAT32, COMPLEX64) on FP32 MFMA with FP32
 * accumulation (a complex64 matrix goes as its real 2m x 2n form), and
 * (T, T) for T in INT8, INT16, INT32, INT64, FLOAT64, COMPLEX128: integers
 * accumulate modulo 2^64 and are truncated to their width (the residue
 * numpy's integer dot produces, overflow included), float64 / complex128
 * accumulate in FP64.  Y has the promoted type. */
fm_status fm_dense_apply(int a_dtype, int x_dtype, int adjoint, int64_t m,
                         int64_t n, const void *A, int64_t lda, const void *x,
                         int64_t ldx, void *y, int64_t ldy, int64_t k,
                         void *stream);
/* Mode-wise dense factor along a strided axis (Kron, Kron.pyx:267-303):
 * square n x n A applied to v[i] = x[(o*n + i)*inner + j]. */
fm_status fm_dense_apply_strided(int a_dtype, int x_dtype, int adjoint,
                                 int64_t n, const void *A, int64_t lda,
                                 const void *x, void *y, int64_t outer,
                                 int64_t inner, void *stream);
/* The last two factors of a Kron product in one pass (replaces their two
 * mode passes of Kron._forward / _backward, Kron.pyx:267-303): x is viewed
 * as (outer, 128, 128) complex64 slabs, contiguous, the last axis fastest;
 * y = (C (x) M) x per slab, with C the 128-point Circulant plan (complex64,
 * one level) along the middle axis, its spectrum conjugated when
 * conj_circulant is set, and M a 128 x 128 row-major float32 matrix
 * (ldm = 128) along the last axis.  The adjoint of (C (x) A) for a real A is
 * conj_circulant = 1 with M = A^T.  x and y must not overlap; y 16-byte
 * aligned. */
fm_status fm_kron_conv_dense_apply(const fm_conv_plan *plan, int conj_circulant,
                                   const void *M, int64_t ldm, const void *x,
                                   void *y, int64_t outer, void *stream);

/* ---------------------------------------------------------- elementwise
 * All element types; operands share `dtype`. */
/* y[:,j] = d (.) x[:,j]  (conj_d: conj(d)); Diag._forwardC/_backwardC
 * (Diag.pyx:149-167). */
fm_status fm_diag_apply(int dtype, int conj_d, int64_t n, const void *d,
                        const void *x, int64_t ldx, void *y, int64_t ldy,
                        int64_t k, void *stream);
/* y[i,:] = x[idx[i],:] for i < n_out (row gather; Permutation._forward /
 * _backward, Permutation.pyx:121-125, and Partial's row selection,
 * Partial.pyx:268-294).  idx: device int64. */
fm_status fm_gather_rows(int dtype, int64_t n_out, const int64_t *idx,
                         const void *x, int64_t ldx, void *y, int64_t ldy,
                         int64_t k, void *stream);
/* y = 0 (n_out rows); y[idx[i],:] = x[i,:] for i < n_in (Partial's
 * zero-padded scatter, Partial.pyx:270-276). */
fm_status fm_scatter_rows(int dtype, int64_t n_in, const int64_t *idx,
                          const void *x, int64_t ldx, void *y, int64_t ldy,
                          int64_t n_out, int64_t k, void *stream);
/* y = alpha * op(x) + beta * y over an n x k block, op = conj if conj_x
 * (Sum accumulation Sum.pyx:151-175, Product's scalar Product.pyx:213-238).
 * alpha/beta are given as two doubles each (re, im); for real and integer
 * types only the real parts are used. */
fm_status fm_axpby(int dtype, int64_t n, int64_t k, const double *alpha,
                   const void *x, int64_t ldx, int conj_x, const double *beta,
                   void *y, int64_t ldy, void *stream);

/* ------------------------------------------------------------ solvers
 * One shrinkage step of ISTA / FISTA (algorithms/ISTA.py:152-168,
 * algorithms/FISTA.py:154-178, softThreshold ISTA.py:113-123), fused after
 * the operator's forward and adjoint applies produced g = A^H (A y - b).
 * Over n contiguous elements of a float32 / float64 / complex64 / complex128
 * block, in that precision:
 *   step  = y - L g
 *   x     = step * m / (m + alpha),  m = max(|step| - alpha, 0)   (x in/out)
 *   y_next = x_new + beta (x_new - x_old)  if y_next is not NULL (FISTA;
 *            x_old is x's value on entry; y_next may alias y)
 * beta is beta_table[*step_counter] when beta_table is not NULL (device
 * pointers; step_counter may be NULL for index 0), else the argument `beta`.
 * fm_counter_advance increments *step_counter on the stream, so a captured
 * iteration reads the next beta on each replay. */
fm_status fm_shrink_step(int dtype, int64_t n, const void *y, const void *g, double L,
                         double alpha, void *step, void *x, const double *beta_table,
                         const int64_t *step_counter, double beta, void *y_next,
                         void *stream);
fm_status fm_counter_advance(int64_t *step_counter, void *stream);

/* LFSRCirculant (LFSRCirculant.pyx:159-237): host-side register sequences.
 * Runs the generating register from `start` (masked to `order` bits) until
 * it returns, with the reference's checks (order 1..31, non-zero start, a
 * sequence that repeats its start state; LFSRCirculant.pyx:196-222), and
 * returns the period.  With states/taps non-NULL (capacity >= period) it
 * also writes states[t] = the state after t generator steps and taps[t] =
 * the tap-side state after t steps from 1, the scatter and gather indices
 * of the Hadamard-based apply (LFSRCirculant.pyx:316-396). */
fm_status fm_lfsr_sequence(uint32_t polynomial, uint32_t start, int order, int64_t *period,
                           int64_t *states, int64_t *taps, int64_t capacity);
/* LFSRCirculant apply (LFSRCirculant._core, LFSRCirculant.pyx:316-396):
 * y[:, c] = H-based circular product of x[:, c] for c < k, n = period rows:
 * z[scatter_idx[i], c] = x[i, c] (all other rows of z zero), w = H_order z,
 * y[i, c] = w[gather_idx[i], c].  The index tables hold the generator and
 * tap-side states with the reference's flips for forward or backward
 * (values < 2^order).  work: 2 * 2^order * k elements of dtype (the row-major
 * z and w); every type of the Hadamard transform, integers wrapping. */
fm_status fm_lfsr_apply(int dtype, int order, int64_t n, const int64_t *scatter_idx,
                        const int64_t *gather_idx, const void *x, int64_t ldx, void *y,
                        int64_t ldy, int64_t k, void *work, int64_t work_bytes, void *stream);

/* OMP (algorithms/OMP.pyx:122-240): orthogonal matching pursuit for L
 * problems in float64 / complex128 (dtype), the reference's rank-1
 * pseudo-inverse update kept step for step.  Per step ii (0 <= ii < steps):
 *   phase 0: idx[l] = first argmax_j |g[l, j]| inv_norm[j] (g = A^H r, an
 *            (L, m) block), support[l, ii] = idx[l], e = the one-hot (L, m)
 *            block of idx (the caller then forms c = A e, an (L, n) block)
 *   phase 1: with the new columns c: v1 = P[:ii] c, v2 = c - Acols[:ii]^T v1,
 *            y = <v2, b> / ||v2||^2, x[:ii] -= y v1, x[ii] += y,
 *            P[:ii] -= v1 conj(v2)^T / ||v2||^2, P[ii] = conj(v2) / ||v2||^2,
 *            Acols[ii] = c, r -= y v2
 *   phase 2: x_full (L, m) = 0, x_full[l, support[l, i]] = x[l, i]
 * P and Acols are (L, steps, n) row-major, b / r / c (L, n), x (L, steps),
 * support (L, steps) int64; work as fm_omp_workspace_size reports. */
fm_status fm_omp_workspace_size(int dtype, int64_t L, int64_t n, int steps, int64_t *bytes);
fm_status fm_omp_step(int dtype, int phase, int64_t L, int64_t m, int64_t n, int steps, int ii,
                      const void *g, const double *inv_norm, int64_t *idx, int64_t *support,
                      void *e, const void *c, const void *b, void *P, void *acols, void *r,
                      void *work, int64_t work_bytes, void *x, void *x_full, void *stream);

/* STELA (algorithms/STELA.py:128-267): soft thresholding with exact line
 * search, one call per phase of a step, all on the stream, no allocation.
 * The k x n iterate x, bx, dlt and the k x m residual res / forward f are of
 * dtype_x (FM_FLOAT64 or FM_COMPLEX128, the reference's promote(float64, b));
 * z (k x n, A^H res) is of dtype_z (equal to dtype_x, or FM_COMPLEX128 with a
 * real iterate when the operator is complex).  d[n] = 1 / colNorms^2.
 * partials is a 4 x k x parts float64 scratch; column sums are formed in two
 * deterministic stages (parts blocks per column, then in order).
 *   fm_stela_prox:   bx = soft(d x - z, lambda) / d, dlt = bx - x, and the
 *                    partial sums of ||z - clamp(z - x, +-lambda)||^2 and of
 *                    sum(|bx| - |x|) (STELA.py:179-211)
 *   fm_stela_line:   partial sums of Re<res, f> and ||f||^2 for f = A dlt,
 *                    then per column active = stop > max_error and the step
 *                    gamma = clamp(-(Re<res,f> + lambda sum(|bx|-|x|)) /
 *                    ||f||^2, 0, 1) (STELA.py:213-248)
 *   fm_stela_update: on active columns x += gamma (bx - x), res += gamma f
 *                    (STELA.py:250-257)
 *   fm_stela_select: on active columns z = zn (zn = A^H res, STELA.py:258-261) */
fm_status fm_stela_prox(int dtype_x, int dtype_z, int64_t n, int64_t k, const void *x,
                        const void *z, const double *d, double lambda, void *bx, void *dlt,
                        double *partials, int64_t parts, void *stream);
fm_status fm_stela_line(int dtype_x, int64_t m, int64_t k, const void *res, const void *f,
                        double *partials, int64_t parts, double lambda, double max_error,
                        double *gamma, int32_t *active, void *stream);
fm_status fm_stela_update(int dtype_x, int64_t n, int64_t m, int64_t k, const void *bx, void *x,
                          const void *f, void *res, const double *gamma, const int32_t *active,
                          void *stream);
fm_status fm_stela_select(int dtype_z, int64_t n, int64_t k, void *z, const void *zn,
                          const int32_t *active, void *stream);

/* ------------------------------------------------------------ profiler
 * No counterpart in the reference.  For benchmarks: when enabled, each of the
 * library's kernel launches is bracketed by HIP events on its own stream and
 * tagged with its algorithmic bytes (its share of the transforms' input and
 * output blocks and generators: an operator's (n + m) k elements plus its
 * generator, split evenly over the operator's passes), the bytes the launch
 * itself reads and writes (intermediates included) and its flops.  Launches
 * on a stream being captured into a graph are not timed.
 * fm_profiler_count() synchronises on the recorded events and aggregates per
 * kernel label; fm_profiler_entry() reads one aggregate. */
fm_status fm_profiler_enable(int on);
fm_status fm_profiler_reset(void);
fm_status fm_profiler_count(int *count);
fm_status fm_profiler_entry(int index, char *name, int name_len, int64_t *launches,
                            double *total_ms, double *total_bytes, double *total_flops,
                            double *total_pass_bytes);

#ifdef __cplusplus
}
#endif

#endif /* FASTMAT_AMD_H */
