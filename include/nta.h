/*
 * nta.h -- C ABI of the MI355X analytic-kernel engine (libnta.so).
 *
 * The reference (neural-tangents 0.6.6) is pure Python over JAX/XLA and has
 * no FFI of its own; each entry point below replaces the XLA op sites that the
 * reference's kernel_fn / predict path emits (paths relative to
 * neural_tangents/_src/).
 *
 * Conventions
 *   - Every buffer is device memory owned by the caller; the library never
 *     allocates on these paths.  Pointers are plain float / double pointers.
 *   - `stream` is a hipStream_t passed as void* (NULL = default stream);
 *     launches are stream-ordered and asynchronous.
 *   - Return value: 0 ok; negative = invalid argument / unsupported shape
 *     (NTA_EINVAL, NTA_EUNSUPPORTED); positive = a hipError_t code.
 *     nta_last_error() returns a thread-local message for the last failure.
 *   - Kernel state layout ("pair state"): for a tile of T pairs, an fp32
 *     array [T][H][H'][W][W'] -- the reference's zipped full-spatial layout
 *     (utils.py:324-342) with the batch axes flattened.  Spatial axes are kept
 *     in natural (h, w) order; the reference's is_reversed flag only permutes
 *     axes (linear.py:1343-1349) and is applied by the host on output.
 *   - Pair p of a cross tile is (i, j) = (p / n2, p % n2); a self tile
 *     (n2 == 0) holds the pairs (n, n).
 *   - Per-image diagonals q[n][h][w] are the ReLU normalisers
 *     (stax/requirements.py:1057-1117).
 */
#ifndef NTA_H_
#define NTA_H_

#ifdef __cplusplus
extern "C" {
#endif

#define NTA_OK 0
#define NTA_EINVAL (-1)
#define NTA_EUNSUPPORTED (-2)

/* Flags for the fused stage kernels. */
#define NTA_RELU 1        /* apply the ReLU arc-cosine map after the conv   */
#define NTA_NTK 2         /* propagate the ntk tensor next to nngp          */
#define NTA_CONV 4        /* apply the 3x3-type stencil (else identity)     */
#define NTA_CONV_VALID 8  /* VALID padding for the stencil (default SAME)   */
#define NTA_FLATTEN 16    /* head: Flatten (diagonal trace) instead of GAP  */
#define NTA_CONV_CIRCULAR 128  /* periodic (wrap-around) padding, SAME size */
/* Activation of a stage with NTA_RELU set (default: ABRelu with the
 * act_k1 = (a-b)^2, act_k0 = ab constants of the entry point): */
#define NTA_ACT_ERF 32    /* a Erf(x) + c: act_k1 = a^2, act_k0 = c^2 (the  */
                          /* input scale b is folded into the stage's conv) */
#define NTA_ACT_GELU 64   /* Gelu (act_k1, act_k0 unused)                   */

/* Library version (major*10000 + minor*100 + patch). */
int nta_version(void);

/* Message for the most recent non-zero status on this thread. */
const char* nta_last_error(void);

/*
 * Layer-1 ReLU normaliser from the inputs.
 * q[n][h][w] = w_var/(fh*fw*C) * sum_{a,e,c} x~[n, h+a, w+e, c]^2 + b_var
 * where x~ is x zero-padded for a SAME (or VALID) fh x fw stencil.
 * Replaces: _cov_diag_batch_full_spatial (stax/requirements.py:529-539) +
 * the first Conv's stencil on cov1 (linear.py:1392, 3341-3378) +
 * get_diagonal (stax/requirements.py:1057-1074).
 * x: [n][H][W][C] fp32 (NHWC).  q: [n][Ho][Wo].
 */
int nta_first_diag(const float* x, int n, int H, int W, int C, int fh, int fw,
                   int flags, float w_var, float b_var, float* q,
                   void* stream);

/*
 * Input covariance + first Conv (+ ReLU) on a tile of pairs, fused: the
 * stencil of the input Gram is the Gram of zero-padded fh x fw x C patches.
 * Replaces: _inputs_to_kernel / _cov (stax/requirements.py:585-638, 641-830),
 * Conv.kernel_fn (linear.py:1325-1424) on nngp and ntk=0, and
 * ABRelu.kernel_fn (elementwise.py:423-477) when flags has NTA_RELU.
 * x1: [n1][H][W][C], x2: [n2][H][W][C] (ignored when n2 == 0: self tile of
 * x1).  q1: [n1][Ho][Wo], q2: [n2][Ho][Wo] (layer-1 normalisers; unused
 * without NTA_RELU).  Outputs nngp, ntk: [T][Ho][Ho][Wo][Wo] with
 * T = n1*n2 (or n1 for a self tile); ntk is written only with NTA_NTK.
 * act_k1, act_k0: the activation is ABRelu(a, b) with act_k1 = (a - b)^2,
 * act_k0 = a b (Relu: 1, 0; LeakyRelu(alpha): (1 - alpha)^2, alpha; Abs:
 * 4, -1), i.e. nngp' = act_k1 J + act_k0 c and dot_sigma' = act_k1
 * dot_sigma + act_k0 with the ReLU map's J and dot_sigma.
 */
int nta_first_stage(const float* x1, const float* x2, int n1, int n2, int H,
                    int W, int C, int fh, int fw, int flags, float w_var,
                    float b_var, float act_k1, float act_k0, const float* q1,
                    const float* q2, float* nngp, float* ntk, void* stream);

/*
 * ReLU normaliser of a later stage from the self state of each image:
 * q[n][h][w] = w_var * stencil(R)[n][h][h][w][w] + b_var (NTA_CONV set) or
 * R[n][h][h][w][w] (identity).  R: [n][H][H][W][W].
 * Replaces: get_diagonal (stax/requirements.py:1057-1074) of the Conv output
 * (linear.py:1392).
 */
int nta_stage_diag(const float* R, int n, int H, int W, int fh, int fw,
                   int flags, float w_var, float b_var, float* q,
                   void* stream);

/*
 * One fused stage on a tile of pairs: [Conv] -> [ReLU] -> [AvgPool ph x pw,
 * stride = window, VALID].  ph = pw = 1 means no pooling.
 * Replaces: Conv.kernel_fn (linear.py:1325-1424 via
 * _conv_kernel_full_spatial_shared 3115-3207, 3341-3378), ABRelu.kernel_fn
 * (elementwise.py:423-477) and _Pool.kernel_fn (linear.py:1634-1664,
 * _pool_kernel 3499-3572).
 * in_nngp, in_ntk: [T][H][H][W][W]; out: [T][Ho'][Ho'][Wo'][Wo'].
 * q1: [n1][Hc][Wc], q2: [n2][Hc][Wc] at the conv-output resolution
 * (q2 ignored for a self tile, n2 == 0).  act_k1, act_k0: the ABRelu
 * constants, as in nta_first_stage.
 */
int nta_stage(const float* in_nngp, const float* in_ntk, int n1, int n2,
              int H, int W, int fh, int fw, int ph, int pw, int flags,
              float w_var, float b_var, float act_k1, float act_k0,
              const float* q1, const float* q2, float* out_nngp,
              float* out_ntk, void* stream);

/*
 * nta_stage for Conv(3x3 SAME) -> Relu -> AvgPool(2x2 / 2) on S x S states
 * (S = 8, 16, 32; flags NTA_CONV | NTA_RELU [| NTA_NTK]) that also writes
 * the stencil means of the inputs at every pre-pool position,
 * pre[T][S][S][S][S] = (1/9) sum of the nine taps -- the pre-activations
 * before the w_var, b_var affine map, which the stage's adjoint
 * (nta_relu_bwd: aN, aT) consumes.  NTA_EUNSUPPORTED for other stages.
 * Replaces: the same reference functions as nta_stage; the saved means
 * stand in for what jax.grad keeps of the forward pass of Conv.kernel_fn
 * (linear.py:1387-1398) for its reverse pass.
 */
int nta_stage_keep_pre(const float* in_nngp, const float* in_ntk, int n1,
                       int n2, int S, int flags, float w_var, float b_var,
                       const float* q1, const float* q2, float* out_nngp,
                       float* out_ntk, float* pre_nngp, float* pre_ntk,
                       void* stream);

/*
 * The first two stages of a Myrtle-type network as one launch on a tile of
 * pairs: Conv(3x3, SAME) -> Relu -> Conv(3x3, SAME) -> Relu ->
 * AvgPool(2x2, stride 2) on 32x32 images with 1..4 channels; the first
 * stage's [32][32][32][32] state is never written.  Same result as
 * nta_first_stage followed by nta_stage (within FP32 rounding: the patch
 * Gram is accumulated in another order).
 * Replaces: the same reference functions as those two entry points, for
 * stax.serial(Conv, Relu, Conv, Relu, AvgPool) (stax/combinators.py:40-68).
 * x1: [n1][32][32][C], x2: [n2][32][32][C] (ignored when n2 == 0).
 * qn1a, qn2a: normalisers of the first ReLU as (sqrt(q), 1/sqrt(q)) pairs,
 * [n][32][32][2] (nta_qnorm of the nta_first_diag output); qn1b, qn2b: the
 * same for the second ReLU (nta_qnorm of the nta_stage_diag output).
 * flags: NTA_NTK or 0.  Outputs: [T][16][16][16][16].
 * Returns NTA_EUNSUPPORTED for other image shapes.
 */
int nta_first_pool_stage(const float* x1, const float* x2, int n1, int n2,
                         int H, int W, int C, int flags, float w_var1,
                         float b_var1, const float* qn1a, const float* qn2a,
                         float w_var2, float b_var2, const float* qn1b,
                         const float* qn2b, float* out_nngp, float* out_ntk,
                         void* stream);

/*
 * nta_first_pool_stage that also writes the stencil means of the first
 * stage's outputs at every position, pre[T][32][32][32][32] = (1/9) sum of
 * the nine taps -- the second stage's pre-activations before w_var2,
 * b_var2, which its adjoint (nta_relu_bwd: aN, aT) consumes.
 * Replaces: the same reference functions as nta_first_pool_stage; the
 * saved means stand in for what jax.grad keeps of the forward pass of the
 * second Conv.kernel_fn (linear.py:1387-1398) for its reverse pass.
 */
int nta_first_pool_stage_keep_pre(const float* x1, const float* x2, int n1,
                                  int n2, int H, int W, int C, int flags,
                                  float w_var1, float b_var1,
                                  const float* qn1a, const float* qn2a,
                                  float w_var2, float b_var2,
                                  const float* qn1b, const float* qn2b,
                                  float* out_nngp, float* out_ntk,
                                  float* pre_nngp, float* pre_ntk,
                                  void* stream);

/*
 * The first three stages of a Myrtle-10-type network as one launch on a
 * tile of pairs: [Conv(3x3, SAME) -> Relu] x 3 -> AvgPool(2x2, stride 2) on
 * 32x32 images with 1..4 channels; neither the first nor the second
 * stage's [32][32][32][32] state is written.  Same result as
 * nta_first_stage followed by two nta_stage calls (within FP32 rounding).
 * Replaces: the reference functions of those entry points, for
 * stax.serial(Conv, Relu, Conv, Relu, Conv, Relu, AvgPool)
 * (stax/combinators.py:40-68).
 * Arguments as nta_first_pool_stage, with qn1c, qn2c the (sqrt(q),
 * 1/sqrt(q)) tables of the third ReLU and w_var3, b_var3 its conv's.
 * Outputs: [T][16][16][16][16].  Returns NTA_EUNSUPPORTED for other image
 * shapes.
 */
int nta_first3_pool_stage(const float* x1, const float* x2, int n1, int n2,
                          int H, int W, int C, int flags, float w_var1,
                          float b_var1, const float* qn1a, const float* qn2a,
                          float w_var2, float b_var2, const float* qn1b,
                          const float* qn2b, float w_var3, float b_var3,
                          const float* qn1c, const float* qn2c,
                          float* out_nngp, float* out_ntk, void* stream);

/*
 * out[2k] = sqrt(q[k]), out[2k + 1] = 1 / sqrt(q[k]) (both 0 where
 * q[k] <= 0), for k < n: the per-position factors of the ReLU map
 * (elementwise.py:444-455 divides the covariance by sqrt(q1 q2)).
 */
int nta_qnorm(const float* q, long long n, float* out, void* stream);

/*
 * Readout head: GlobalAvgPool (or Flatten with NTA_FLATTEN) followed by the
 * composed affine map of the trailing Dense layers:
 *   nngp_out = a * m + b,   ntk_out = c * m + d * mt + e
 * with m, mt the spatial means of nngp, ntk.  Writes out_nngp[i*ld + j]
 * (and out_ntk) for the tile pairs; a self tile writes out[n].
 * Replaces: _GlobalPool.kernel_fn (linear.py:1774-1801), Flatten.kernel_fn
 * (linear.py:1868-1899), Dense.kernel_fn (linear.py:902-926).
 */
int nta_head(const float* nngp, const float* ntk, int n1, int n2, int H,
             int W, int flags, double a, double b, double c, double d,
             double e, float* out_nngp, float* out_ntk, long long ld,
             void* stream);

/*
 * Adjoint of one [Conv] -> ReLU -> [AvgPool p x p] stage over a tile of
 * pairs (no counterpart in the reference, which differentiates its XLA graph
 * with jax.grad; this is the reverse of stax/elementwise.py:444-455 with the
 * _sqrt/_arctan2 JVP rules of 1278-1310 and of the pool
 * stax/linear.py:3551-3571).  aN, aT: stencil means of the stage's nngp and
 * ntk inputs [T][S][S][S][S] (aT = NULL: first stage, ntk input 0); gN, gT:
 * pooled cotangents [T][S/p]^4 (NULL: zero).  Writes
 *   gaN: the cotangent of aN (already times w_var), in the stage layout,
 *     or with gram_scale != 0 (cross tiles only) times gram_scale in the
 *     patch-Gram layout [n1][S][S][n2][S][S] = (a, h, w, b, h', w'): the
 *     first stage's Gram adjoint is then two plain GEMMs;
 *   gaT: the cotangent of aT (NULL: not wanted);
 *   gqa, gqb [T][S][S]: the cotangents of the normalisers q1 (at (h, w),
 *     summed over (h', w')) and q2 (at (h', w'), summed over (h, w));
 *   gwb: device doubles, 2 per workgroup (nta_relu_bwd_parts(n1, n2, S)
 *     pairs), the workgroups' partial sums of the d/dw_var and d/db_var
 *     terms (the caller adds the slots: a fixed-order, deterministic
 *     reduction).
 * work: device workspace of nta_relu_bwd_work_bytes(n1, n2, S) bytes.
 * Pairs (a, b) with a + self_off == b (and every pair of a self tile,
 * n2 == 0) are pairs of identical images: their spatial diagonal is treated
 * as rho = 1 exactly.  Pass self_off = INT_MAX for none.  act_k1, act_k0:
 * the ABRelu constants, as in nta_first_stage.  T < 2^24 (and n2 S^3 < 2^31
 * with the patch-Gram layout).
 */
int nta_relu_bwd(const float* aN, const float* aT, int n1, int n2, int S,
                 int pool, int self_off, float w_var, float b_var,
                 float act_k1, float act_k0, const float* q1,
                 const float* q2, const float* gN, const float* gT,
                 float gram_scale, float* gaN, float* gaT, float* gqa,
                 float* gqb, double* gwb, void* work, long long work_bytes,
                 void* stream);

/* Partial-sum slots (pairs of doubles) of an nta_relu_bwd call on a tile of
 * n1 x n2 pairs (n2 == 0: n1 self pairs) of side S; -1 for bad sizes. */
long long nta_relu_bwd_parts(int n1, int n2, int S);

/* Workspace bytes of an nta_relu_bwd call (S^3 floats per pair for sides 8,
 * 16 and 32, 2 S^4 otherwise); -1 for bad sizes. */
long long nta_relu_bwd_work_bytes(int n1, int n2, int S);

/*
 * Accumulates the outputs of an nta_relu_bwd call on the same tile into
 * per-image and per-stage totals, in a fixed order (each argument may be
 * NULL to skip it):
 *   wb_acc[0], wb_acc[1] += the sums of gwb's d/dw_var, d/db_var slots;
 *   gq1_acc [n1][S][S] += gqa summed over the tile's second images (a self
 *     tile, n2 == 0: += gqa + gqb, both normalisers being q1's);
 *   gq2_acc [n2][S][S] += gqb summed over the tile's first images (cross
 *     tiles), after gq1_acc, so the two may be rows of one array.
 * No counterpart in the reference (jax.grad accumulates these itself).
 */
int nta_relu_bwd_accumulate(const float* gqa, const float* gqb,
                            const double* gwb, int n1, int n2, int S,
                            double* wb_acc, float* gq1_acc, float* gq2_acc,
                            void* stream);

/*
 * Diagonal-spatial chain: a network of nstage [Conv fh x fw, stride 1]
 * [-> ReLU] stages without pooling, read out by Flatten and the composed
 * Dense map (a, b, c, d, e as in nta_head), on the spatial diagonal
 * (h = h', w = w') only, where the Conv stencil is an fh x fw window mean.
 * One launch runs every stage of every pair; states stay on chip.
 *   stage_desc (host) [nstage][3]: flags (NTA_CONV, NTA_CONV_VALID,
 *   NTA_RELU), fh, fw;  stage_wb (host) [nstage][4]: w_var, b_var and
 *   the activation constants act_k1, act_k0 (see nta_stage).
 *   q1, q2: device [nstage][n][H][W] normaliser tables (a stage's values
 *   fill the first n * hout * wout floats of its slot).
 * Self mode (n2 == 0): the chains of x1's images against themselves; q1 is
 *   written (the pre-activation of every ReLU stage) and out_nngp[i]
 *   receives the readout of image i (cov1).
 * Cross mode: out_nngp[i * ld + j] (and out_ntk with NTA_NTK in flags);
 *   sym = 1 (x2 == x1, n2 == n1) computes the upper triangle and writes both
 *   (i, j) and (j, i).
 * H * W <= 3840, nstage <= 48.
 * Replaces: Conv.kernel_fn with diagonal_spatial
 * (_conv_kernel_diagonal_spatial, linear.py:3381-3437), ABRelu.kernel_fn
 * (elementwise.py:423-477) on diagonal kernels, Flatten.kernel_fn
 * (linear.py:1868-1899), Dense.kernel_fn (linear.py:902-926) and the
 * diagonal-spatial input covariance (requirements.py:556-582).
 */
int nta_diag_chain(const float* x1, const float* x2, int n1, int n2, int sym,
                   int H, int W, int C, int nstage, const int* stage_desc,
                   const float* stage_wb, float* q1, const float* q2,
                   int flags, double a, double b, double c, double d,
                   double e, float* out_nngp, float* out_ntk, long long ld,
                   void* stream);

/* ------------------------------------------------------------------------
 * Kernel-regression solve (predict.py:1235-1258, cho_factor / cho_solve of
 * K + lambda I, mixed precision).  Row-major matrices, element (i, j) at
 * [i * ld + j]; n, and the leading dimensions, multiples of 64 and 4.
 */

/*
 * The split n1 of the top-level block of nta_chol_f32 for an n x n matrix
 * (n itself when n <= 64; 0 when n is not a positive multiple of 64).
 */
int nta_chol_split(int n);

/*
 * A = K + shift I on the 64x64 tiles on and below the diagonal of an n x n
 * fp32 matrix (the tiles above are not written): the input nta_chol_f32
 * factors, at half the bytes of a full copy.  n a multiple of 64; K, A
 * 16-byte aligned, ldk, lda multiples of 4.
 * Replaces: _add_diagonal_regularizer (predict.py:1224-1232) on the copy
 * that cho_factor (predict.py:1244) consumes.
 */
int nta_copy_lower_shift(const float* K, long long ldk, int n, float shift,
                         float* A, long long lda, void* stream);

/*
 * FP32 Cholesky of A (lower triangle read; A is overwritten with
 * intermediates): right-looking, blocked on 128 columns (leaves factored
 * and inverted in LDS) and 512-column panels, with fp32 MFMA products; the
 * trailing updates overlap the next panel on a second stream of the
 * library's (events order it against `stream`, which the call returns on).
 * Each half of the factor is then inverted by recursive doubling.  W
 * (n x n) need not be initialised.  On return, with L L^T = A and
 * n1 = nta_chol_split(n):
 *   W[0:n1, 0:n1] = L11^-1, W[n1:n, n1:n] = L22^-1 (lower; zeros above the
 *     diagonal inside each diagonal 64x64 block, the rest of the upper
 *     triangles left as it was),
 *   W[n1:n, 0:n1] = L21,
 * the blocks from which the preconditioner (L L^T)^-1 is applied, and WT
 * (n x n, leading dimension ldw) holds W^T on and above its diagonal.
 * *info (device int, zero on entry) receives the 1-based index of the first
 * non-positive pivot.
 * Replaces: jsp.linalg.cho_factor (predict.py:1244).
 */
int nta_chol_f32(float* A, int n, long long lda, float* W, long long ldw,
                 float* WT, int* info, void* stream);

/* Bytes of device workspace nta_pcg64 / nta_chol_precond need (-1 when the
 * arguments are invalid: nrhs must be 1..16). */
long long nta_pcg64_work_bytes(int n, int nrhs);

/*
 * Z = (L L^T)^-1 R in fp32 from the blocks nta_chol_f32 left in W and WT
 * (six row-major products on fp32 MFMA); R, Z [n][nrhs] fp32, nrhs <= 16,
 * n1 = nta_chol_split(n).
 * Replaces: the triangular solves of cho_solve (predict.py:1254).
 */
int nta_chol_precond(const float* W, const float* WT, int n, long long ldw,
                     int n1, const float* R, int nrhs, float* Z, void* work,
                     long long work_bytes, void* stream);

/*
 * Solves (K + lam I) X = B by conjugate gradients preconditioned with the
 * FP32 factor in W, WT (nta_chol_f32 of K + shift I), FP64 vectors, FP64
 * products on fp64 MFMA (K read in fp32, exact) and the FP64 shift lam.  B, X
 * [n][nrhs] fp64, nrhs <= 16.  Iterates until, for every column, the
 * relative residual ||r|| / ||b|| or the estimated relative error
 * ||M^-1 r|| / ||x|| (M = L L^T the FP32 factor, so M^-1 r is the error of
 * x to the factor's accuracy) is <= tol -- tested on the recurrence
 * residual and confirmed on the residual recomputed as B - (K + lam I) X --
 * or the recomputed residual stops improving over 5 checkpoints (every 8
 * iterations) or is NaN (a broken-down factor), or max_iter is reached.
 * *iters receives the iteration count and *resid the largest relative
 * residual of the returned X, recomputed from scratch (NaN if any column's
 * is; host pointers, may be NULL).  Iterations run in batches of
 * 4 without host synchronisation (the kernels of an iteration after
 * convergence return at once), each batch followed by the confirmation on
 * the recomputed residual, which the device runs only once the recurrence
 * has converged; the stream is synchronised once per batch and per
 * checkpoint.
 * Replaces: cho_solve (predict.py:1254) of the regularised train kernel.
 */
int nta_pcg64(const float* K, int n, long long ldk, double lam,
              const float* W, const float* WT, long long ldw, int n1,
              const double* B, int nrhs, double* X, double tol, int max_iter,
              void* work, long long work_bytes, int* iters, double* resid,
              void* stream);

/*
 * nta_pcg64 for a kernel matrix held in FP64 (K [n][ldk] fp64, ldk even):
 * the residual operator (K + lam I) reads the FP64 entries, so the solution
 * is that of the FP64 system even where K is not representable in fp32;
 * W, WT are the FP32 factor of fp32(K) + shift I as for nta_pcg64.
 * Replaces: cho_solve (predict.py:1254) of the regularised train kernel
 * given in float64.
 */
int nta_pcg64_k64(const double* K, int n, long long ldk, double lam,
                  const float* W, const float* WT, long long ldw, int n1,
                  const double* B, int nrhs, double* X, double tol,
                  int max_iter, void* work, long long work_bytes, int* iters,
                  double* resid, void* stream);

/*
 * Y[:, :ncols] = K X[:, :ncols] in FP64 for an fp32 m x n matrix K
 * (row-major, leading dimension ldk, 16-byte aligned) and X [n][ldx] fp64
 * (1 <= ncols <= 16 columns read): K's entries are exact in FP64 and the
 * products accumulate on fp64 MFMA.  Y [m][ldy] fp64 (the first ncols
 * columns written).  m, n and ldk must be multiples of 32, 16 and 4
 * (NTA_EUNSUPPORTED otherwise: pad).  With a device workspace of
 * nta_gemm_f32_f64_work_bytes(m, n) bytes the columns of K are split over
 * several workgroups per row block (for short K: m = 1,024 rows are only
 * 32 row blocks), the parts summed in a fixed order, and X, Y may be
 * narrower than 16 columns; without it (NULL or too small) ncols, ldx and
 * ldy must be 16 and one workgroup per row block runs.
 * Replaces: the test-set mean jnp.tensordot(k_td, k_inv_y, ...)
 * (predict.py:696) with the FP32 test-train kernel and the FP64 solution.
 */
int nta_gemm_f32_f64(const float* K, long long ldk, int m, int n,
                     const double* X, long long ldx, int ncols, double* Y,
                     long long ldy, void* work, long long work_bytes,
                     void* stream);

/* Workspace bytes of nta_gemm_f32_f64 (-1: invalid sizes). */
long long nta_gemm_f32_f64_work_bytes(int m, int n);

/*
 * nta_pcg64 without a workspace argument: the iteration's vectors live in a
 * static device array of the library (one per GPU, held until the process
 * ends; 19,792,896 bytes, enough for n <= 16384 and nrhs <= 16), so the call
 * allocates nothing.  n > 16384 returns NTA_EINVAL.  Calls on one GPU share
 * that array: calls that overlap on one GPU (other streams or threads)
 * race.  Arguments and results as nta_pcg64.
 * No counterpart in the reference: it is for callers that cannot allocate
 * device memory (code under HIP graph capture, serving loops, harnesses
 * that forbid hipMalloc).
 */
int nta_pcg64_noalloc(const float* K, int n, long long ldk, double lam,
                      const float* W, const float* WT, long long ldw, int n1,
                      const double* B, int nrhs, double* X, double tol,
                      int max_iter, int* iters, double* resid, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* NTA_H_ */
