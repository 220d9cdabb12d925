// Dense linear algebra of the kernel-regression solve
// (predict.py:1235-1258: cho_factor / cho_solve of K + lambda I).
//
// Mixed precision: the train kernel is factored in FP32 (blocked, lower,
// row-major, block NB = 64) with the trailing updates on fp32 MFMA, the
// factor is inverted (W = L^-1) so the preconditioner M^-1 = W^T W is two
// dense products, and the solution is refined by preconditioned conjugate
// gradients on (K + lambda I) with FP64 residuals (host loop, _mixed.py).
//
// Matrices are row-major; element (i, j) of a matrix with leading dimension
// ld is at [i * ld + j].  Only the lower triangle of A is read or written.
#include "nta_common.h"

namespace nta {

typedef float floatx16 __attribute__((ext_vector_type(16)));

constexpr int NB = 64;

// Factors the NB x NB diagonal block at (k0, k0) in place (lower) and writes
// its inverse into W at (k0, k0).  One wave.  *info receives the 1-based
// global index of the first non-positive pivot (left unchanged on success).
__global__ __launch_bounds__(64) void k_chol_diag(float* __restrict__ A,
                                                  long long lda, int k0,
                                                  float* __restrict__ W,
                                                  long long ldw,
                                                  int* __restrict__ info) {
  __shared__ float a[NB][NB + 1];
  const int lane = threadIdx.x;
  for (int r = 0; r < NB; ++r)
    a[r][lane] = A[(long long)(k0 + r) * lda + k0 + lane];
  __syncthreads();
  bool bad = false;
  for (int j = 0; j < NB; ++j) {
    float d = a[j][j];
    if (!(d > 0.0f) && !bad) {
      if (lane == 0 && *info == 0) *info = k0 + j + 1;
      bad = true;
    }
    const float rd = bad ? 0.0f : 1.0f / sqrtf(d);
    __syncthreads();
    if (lane >= j) a[lane][j] *= rd;       // a[j][j] becomes sqrt(d)
    __syncthreads();
    if (lane > j) {
      const float lij = a[lane][j];
      for (int c = j + 1; c <= lane; ++c) a[lane][c] -= lij * a[c][j];
    }
    __syncthreads();
  }
  // Column `lane` of L^-1 by forward substitution.
  float w[NB];
#pragma unroll
  for (int r = 0; r < NB; ++r) {
    float acc = (r == lane) ? 1.0f : 0.0f;
#pragma unroll
    for (int t = 0; t < r; ++t) acc = fmaf(-a[r][t], w[t], acc);
    w[r] = r >= lane ? acc / a[r][r] : 0.0f;
  }
  for (int r = 0; r < NB; ++r) {
    A[(long long)(k0 + r) * lda + k0 + lane] = lane <= r ? a[r][lane] : 0.0f;
    W[(long long)(k0 + r) * ldw + k0 + lane] = w[r];
  }
}

// Panel: rows i >= k0 + NB of block column k0 become A[i, k] W_kk^T.
// Block (64 rows) per workgroup of 256 threads; thread owns 16 outputs.
__global__ __launch_bounds__(256) void k_chol_panel(
    float* __restrict__ A, long long lda, int n, int k0,
    const float* __restrict__ W, long long ldw) {
  __shared__ float wk[NB][NB + 1];
  __shared__ float pr[NB][NB + 1];
  const int r0 = k0 + NB + blockIdx.x * NB;
  const int t = threadIdx.x;
  for (int e = t; e < NB * NB; e += 256) {
    const int r = e / NB, c = e % NB;
    wk[r][c] = W[(long long)(k0 + r) * ldw + k0 + c];
    pr[r][c] = (r0 + r < n) ? A[(long long)(r0 + r) * lda + k0 + c] : 0.0f;
  }
  __syncthreads();
  const int r = t / 4;            // row of the block
  const int cq = (t % 4) * 16;    // 16 output columns
  float out[16];
#pragma unroll
  for (int u = 0; u < 16; ++u) out[u] = 0.0f;
  for (int s = 0; s < NB; ++s) {
    const float v = pr[r][s];
#pragma unroll
    for (int u = 0; u < 16; ++u) out[u] = fmaf(v, wk[cq + u][s], out[u]);
  }
  if (r0 + r < n) {
#pragma unroll
    for (int u = 0; u < 16; ++u)
      A[(long long)(r0 + r) * lda + k0 + cq + u] = out[u];
  }
}

// Trailing update C_ij -= P_i P_j^T for the lower tiles (bi >= bj) of the
// matrix below/right of block column k0; P = A[:, k0:k0+NB].  One 64x64
// tile per workgroup, 4 waves x 32x32 quadrants, 32 fp32 MFMAs (K = 64).
__global__ __launch_bounds__(256) void k_chol_syrk(float* __restrict__ A,
                                                   long long lda, int n,
                                                   int k0, int m) {
  __shared__ float pi_[NB][NB + 1];
  __shared__ float pj_[NB][NB + 1];
  // blockIdx.x -> (bi, bj), bj <= bi < m, row-major over the lower triangle
  const int tix = blockIdx.x;
  int bi = (int)((sqrtf(8.0f * tix + 1.0f) - 1.0f) * 0.5f);
  while ((bi + 1) * (bi + 2) / 2 <= tix) ++bi;
  while (bi * (bi + 1) / 2 > tix) --bi;
  const int bj = tix - bi * (bi + 1) / 2;
  if (bi >= m) return;
  const int ri = k0 + NB + bi * NB, rj = k0 + NB + bj * NB;
  const int t = threadIdx.x;
  for (int e = t; e < NB * NB; e += 256) {
    const int r = e / NB, c = e % NB;
    pi_[r][c] = (ri + r < n) ? A[(long long)(ri + r) * lda + k0 + c] : 0.0f;
    pj_[r][c] = (rj + r < n) ? A[(long long)(rj + r) * lda + k0 + c] : 0.0f;
  }
  __syncthreads();
  const int lane = t & 63, wave = t >> 6;
  const int qm = (wave >> 1) * 32, qn = (wave & 1) * 32;
  const int col = lane & 31, half = lane >> 5;
  floatx16 acc;
#pragma unroll
  for (int e = 0; e < 16; ++e) acc[e] = 0.0f;
#pragma unroll 8
  for (int kk = 0; kk < NB / 2; ++kk) {
    const float av = pi_[qm + col][2 * kk + half];
    const float bv = pj_[qn + col][2 * kk + half];
    acc = __builtin_amdgcn_mfma_f32_32x32x2f32(av, bv, acc, 0, 0, 0);
  }
#pragma unroll
  for (int e = 0; e < 16; ++e) {
    const int r = ri + qm + 8 * (e >> 2) + 4 * half + (e & 3);
    const int c = rj + qn + col;
    if (r < n && c < n && c <= r) A[(long long)r * lda + c] -= acc[e];
  }
}

// Off-diagonal blocks of W = L^-1 for block row bi:
//   W_ij = -W_ii sum_{k=j}^{i-1} L_ik W_kj,  j < i.
// One workgroup per (bi, j); 256 threads, 16 outputs each.
__global__ __launch_bounds__(256) void k_trtri_row(
    const float* __restrict__ L, long long lda, float* __restrict__ W,
    long long ldw, int n, int bi) {
  __shared__ float la[NB][NB + 1];
  __shared__ float wb[NB][NB + 1];
  const int bj = blockIdx.x;
  const int t = threadIdx.x;
  const int r = t / 4, cq = (t % 4) * 16;
  const int ri = bi * NB, cj = bj * NB;
  float acc[16];
#pragma unroll
  for (int u = 0; u < 16; ++u) acc[u] = 0.0f;
  for (int bk = bj; bk < bi; ++bk) {
    __syncthreads();
    for (int e = t; e < NB * NB; e += 256) {
      const int rr = e / NB, cc = e % NB;
      la[rr][cc] = (ri + rr < n) ? L[(long long)(ri + rr) * lda + bk * NB +
                                     cc] : 0.0f;
      wb[rr][cc] = W[(long long)(bk * NB + rr) * ldw + cj + cc];
    }
    __syncthreads();
    for (int s = 0; s < NB; ++s) {
      const float v = la[r][s];
#pragma unroll
      for (int u = 0; u < 16; ++u) acc[u] = fmaf(v, wb[s][cq + u], acc[u]);
    }
  }
  __syncthreads();
  for (int e = t; e < NB * NB; e += 256) {
    const int rr = e / NB, cc = e % NB;
    la[rr][cc] = W[(long long)(ri + rr) * ldw + ri + cc];   // W_ii
  }
  for (int u = 0; u < 16; ++u) wb[r][cq + u] = acc[u];
  __syncthreads();
  float out[16];
#pragma unroll
  for (int u = 0; u < 16; ++u) out[u] = 0.0f;
  for (int s = 0; s <= r; ++s) {
    const float v = la[r][s];
#pragma unroll
    for (int u = 0; u < 16; ++u) out[u] = fmaf(v, wb[s][cq + u], out[u]);
  }
  if (ri + r < n) {
#pragma unroll
    for (int u = 0; u < 16; ++u)
      W[(long long)(ri + r) * ldw + cj + cq + u] = -out[u];
  }
}

// Y = (A + lam I) X in FP64 with A stored in fp32 (exact upcast);
// X, Y: [n][nrhs] fp64, nrhs <= 16.  One workgroup per 4 rows (a wave per
// row, lanes stride the columns).
__global__ __launch_bounds__(256) void k_symv64(const float* __restrict__ A,
                                                long long lda, int n,
                                                double lam,
                                                const double* __restrict__ X,
                                                int nrhs,
                                                double* __restrict__ Y) {
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int i = blockIdx.x * 4 + wave;
  if (i >= n) return;
  double acc[16];
#pragma unroll
  for (int c = 0; c < 16; ++c) acc[c] = 0.0;
  const float* a = A + (long long)i * lda;
  for (int j = lane; j < n; j += 64) {
    const double aij = (double)a[j];
    const double* x = X + (long long)j * nrhs;
#pragma unroll
    for (int c = 0; c < 16; ++c)
      if (c < nrhs) acc[c] = fma(aij, x[c], acc[c]);
  }
#pragma unroll
  for (int c = 0; c < 16; ++c) {
    if (c < nrhs) {
      double v = acc[c];
      for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
      acc[c] = v;
    }
  }
  if (lane < nrhs) {
    double v = 0.0;
#pragma unroll
    for (int c = 0; c < 16; ++c)
      if (c == lane) v = acc[c];
    Y[(long long)i * nrhs + lane] = v + lam * X[(long long)i * nrhs + lane];
  }
}

// Y = op(W) X in fp32 arithmetic with fp64 vectors, W lower triangular
// (fp32).  trans = 0: Y = W X (row i sums j <= i); trans = 1: Y = W^T X,
// computed from the stored transpose WT (row i of W^T = column i of W, so
// WT is upper triangular: j >= i).
__global__ __launch_bounds__(256) void k_trmv(const float* __restrict__ W,
                                              long long ldw, int n, int upper,
                                              const double* __restrict__ X,
                                              int nrhs,
                                              double* __restrict__ Y) {
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int i = blockIdx.x * 4 + wave;
  if (i >= n) return;
  float acc[16];
#pragma unroll
  for (int c = 0; c < 16; ++c) acc[c] = 0.0f;
  const float* w = W + (long long)i * ldw;
  const int j0 = upper ? i : 0, j1 = upper ? n : i + 1;
  for (int j = j0 + lane; j < j1; j += 64) {
    const float wij = w[j];
    const double* x = X + (long long)j * nrhs;
#pragma unroll
    for (int c = 0; c < 16; ++c)
      if (c < nrhs) acc[c] = fmaf(wij, (float)x[c], acc[c]);
  }
#pragma unroll
  for (int c = 0; c < 16; ++c) {
    if (c < nrhs) {
      float v = acc[c];
      for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
      acc[c] = v;
    }
  }
  if (lane < nrhs) {
    float v = 0.0f;
#pragma unroll
    for (int c = 0; c < 16; ++c)
      if (c == lane) v = acc[c];
    Y[(long long)i * nrhs + lane] = (double)v;
  }
}

// Out-of-place transpose of the lower triangle of W into WT (upper).
__global__ __launch_bounds__(256) void k_transpose(const float* __restrict__ W,
                                                   long long ldw, int n,
                                                   float* __restrict__ WT,
                                                   long long ldt) {
  __shared__ float tile[32][33];
  const int bx = blockIdx.x * 32, by = blockIdx.y * 32;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  for (int r = ty; r < 32; r += 8) {
    const int i = by + r, j = bx + tx;
    tile[r][tx] = (i < n && j < n) ? W[(long long)i * ldw + j] : 0.0f;
  }
  __syncthreads();
  for (int r = ty; r < 32; r += 8) {
    const int i = bx + r, j = by + tx;
    if (i < n && j < n) WT[(long long)i * ldt + j] = tile[tx][r];
  }
}

}  // namespace nta

using namespace nta;

extern "C" int nta_chol_f32(float* A, int n, long long lda, float* W,
                            long long ldw, int* info, void* stream) {
  if (n <= 0 || lda < n || ldw < n || !A || !W || !info)
    return set_error(NTA_EINVAL, "nta_chol_f32: bad n=%d lda=%lld ldw=%lld "
                     "or null pointer", n, lda, ldw);
  if (n % NB != 0)
    return set_error(NTA_EUNSUPPORTED, "nta_chol_f32: n=%d must be a "
                     "multiple of %d (pad the matrix)", n, NB);
  hipStream_t st = (hipStream_t)stream;
  const int nb = n / NB;
  for (int b = 0; b < nb; ++b) {
    const int k0 = b * NB;
    hipLaunchKernelGGL(k_chol_diag, dim3(1), dim3(64), 0, st, A, lda, k0, W,
                       ldw, info);
    const int m = nb - b - 1;
    if (m > 0) {
      hipLaunchKernelGGL(k_chol_panel, dim3(m), dim3(256), 0, st, A, lda, n,
                         k0, W, ldw);
      hipLaunchKernelGGL(k_chol_syrk, dim3(m * (m + 1) / 2), dim3(256), 0,
                         st, A, lda, n, k0, m);
    }
  }
  int rc = check_launch("nta_chol_f32");
  if (rc) return rc;
  // Off-diagonal blocks of W = L^-1, one block row at a time.
  for (int bi = 1; bi < nb; ++bi)
    hipLaunchKernelGGL(k_trtri_row, dim3(bi), dim3(256), 0, st, A, lda, W,
                       ldw, n, bi);
  return check_launch("nta_chol_f32 (inverse)");
}

extern "C" int nta_symv64(const float* A, int n, long long lda, double lam,
                          const double* X, int nrhs, double* Y,
                          void* stream) {
  if (n <= 0 || lda < n || nrhs <= 0 || nrhs > 16 || !A || !X || !Y)
    return set_error(NTA_EINVAL, "nta_symv64: bad n=%d nrhs=%d (1..16) or "
                     "null pointer", n, nrhs);
  hipLaunchKernelGGL(k_symv64, dim3((n + 3) / 4), dim3(256), 0,
                     (hipStream_t)stream, A, lda, n, lam, X, nrhs, Y);
  return check_launch("nta_symv64");
}

extern "C" int nta_trmv(const float* W, int n, long long ldw, int upper,
                        const double* X, int nrhs, double* Y, void* stream) {
  if (n <= 0 || ldw < n || nrhs <= 0 || nrhs > 16 || !W || !X || !Y)
    return set_error(NTA_EINVAL, "nta_trmv: bad n=%d nrhs=%d (1..16) or "
                     "null pointer", n, nrhs);
  hipLaunchKernelGGL(k_trmv, dim3((n + 3) / 4), dim3(256), 0,
                     (hipStream_t)stream, W, ldw, n, upper, X, nrhs, Y);
  return check_launch("nta_trmv");
}

extern "C" int nta_transpose_f32(const float* W, int n, long long ldw,
                                 float* WT, long long ldt, void* stream) {
  if (n <= 0 || ldw < n || ldt < n || !W || !WT)
    return set_error(NTA_EINVAL, "nta_transpose_f32: bad n=%d", n);
  dim3 grid((n + 31) / 32, (n + 31) / 32);
  hipLaunchKernelGGL(k_transpose, grid, dim3(256), 0, (hipStream_t)stream, W,
                     ldw, n, WT, ldt);
  return check_launch("nta_transpose_f32");
}
