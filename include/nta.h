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
 */
int nta_first_stage(const float* x1, const float* x2, int n1, int n2, int H,
                    int W, int C, int fh, int fw, int flags, float w_var,
                    float b_var, const float* q1, const float* q2,
                    float* nngp, float* ntk, void* stream);

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
 * (q2 ignored for a self tile, n2 == 0).
 */
int nta_stage(const float* in_nngp, const float* in_ntk, int n1, int n2,
              int H, int W, int fh, int fw, int ph, int pw, int flags,
              float w_var, float b_var, const float* q1, const float* q2,
              float* out_nngp, float* out_ntk, void* stream);

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
 * Adjoint of one [Conv] -> ReLU -> [AvgPool p x p] stage, per pre-pool
 * element of a tile (no counterpart in the reference, which differentiates
 * its XLA graph with jax.grad; this is the reverse of stax/elementwise.py:
 * 444-455 with the _sqrt/_arctan2 JVP rules of 1278-1310 and of the pool
 * stax/linear.py:3551-3571).  aN, aT: stencil means of the stage's nngp and
 * ntk inputs [T][S][S][S][S] (aT = NULL: first stage, ntk input 0); gN, gT:
 * pooled cotangents [T][S/p]^4.  Writes the cotangents of aN (gaN, already
 * times w_var), aT, of the normalisers per element (gqa, gqb; to be summed
 * over the partner axes) and the per-element terms of d/dw_var (gwv) and
 * d/db_var (gbv).  Pairs (a, b) with a + self_off == b (and every pair of a
 * self tile, n2 == 0) are pairs of identical images: their spatial diagonal
 * is treated as rho = 1 exactly.  Pass self_off = INT_MAX for none.
 */
int nta_relu_bwd(const float* aN, const float* aT, int n1, int n2, int S,
                 int pool, int self_off, float w_var, float b_var,
                 const float* q1,
                 const float* q2, const float* gN, const float* gT,
                 float* gaN, float* gaT, float* gqa, float* gqb, float* gwv,
                 float* gbv, void* stream);

/* ------------------------------------------------------------------------
 * Kernel-regression solve (predict.py:1235-1258, cho_factor / cho_solve of
 * K + lambda I, mixed precision).  Row-major matrices, element (i, j) at
 * [i * ld + j].
 */

/*
 * In-place FP32 Cholesky of the lower triangle of A (n x n, n a multiple of
 * 64; upper triangle zeroed), with the trailing updates on fp32 MFMA, and
 * W = L^-1 (lower, n x n).  *info (device int, zero on entry) receives the
 * 1-based index of the first non-positive pivot.
 * Replaces: jsp.linalg.cho_factor (predict.py:1244).
 */
int nta_chol_f32(float* A, int n, long long lda, float* W, long long ldw,
                 int* info, void* stream);

/*
 * Y = (A + lam I) X in FP64: A fp32 (upcast exactly), X, Y fp64 [n][nrhs],
 * nrhs <= 16.  The FP64 residual of the refinement.
 * Replaces: the regularised operand of cho_solve (predict.py:1224-1232,
 * 1254).
 */
int nta_symv64(const float* A, int n, long long lda, double lam,
               const double* X, int nrhs, double* Y, void* stream);

/*
 * Y = T X for a triangular fp32 T (upper = 0: lower, j <= i; upper = 1:
 * upper, j >= i), fp32 arithmetic, fp64 X and Y [n][nrhs], nrhs <= 16.
 * The preconditioner M^-1 = W^T W of the refinement.
 * Replaces: the triangular solves of cho_solve (predict.py:1254).
 */
int nta_trmv(const float* W, int n, long long ldw, int upper,
             const double* X, int nrhs, double* Y, void* stream);

/* WT = W^T (n x n fp32). */
int nta_transpose_f32(const float* W, int n, long long ldw, float* WT,
                      long long ldt, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* NTA_H_ */
