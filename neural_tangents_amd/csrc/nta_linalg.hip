// Dense linear algebra of the kernel-regression solve
// (predict.py:1235-1258: cho_factor / cho_solve of K + lambda I).
//
// Mixed precision.  K + shift I is factored in FP32 by a recursive blocked
// Cholesky that also inverts the factor as it goes: for a diagonal block
//   [A11 A21^T; A21 A22] = [L11 0; L21 L22] [L11 0; L21 L22]^T
// with W = L^-1,
//   W11 = node(A11)                       (recursion)
//   S   = A21 W11^T            (= L21)
//   A22 -= S S^T
//   T   = S W11                (into the A21 slot)
//   W22 = node(A22)                       (recursion)
//   W21 = -W22 T
// All products run on one FP32 MFMA GEMM kernel (v_mfma_f32_32x32x2f32,
// 128x128 tiles) that skips the zero blocks of triangular operands; 64x64
// leaves are factored and inverted by one wavefront.  At the top level the
// last two products are skipped: W keeps W11 = L11^-1, W22 = L22^-1 on its
// diagonal blocks and S = L21 below them, and the preconditioner
// M^-1 = (L L^T)^-1 is applied as two block triangular solves built from
// those three blocks.
//
// Refinement: preconditioned conjugate gradients on (K + lambda I) with
// FP64 vectors.  (K + lambda I) p reads K in fp32 (exact in fp64) and
// accumulates in FP64; the preconditioner runs in fp32.  Every column of the
// right-hand side runs its own CG recurrence.
//
// Matrices are row-major; element (i, j) of a matrix with leading dimension
// ld is at [i * ld + j].
#include <math.h>
#include <string.h>

#include <vector>

#include "nta_common.h"

namespace nta {

typedef float floatx16 __attribute__((ext_vector_type(16)));

// ---------------------------------------------------------------- GEMM
// C = alpha A op(B) + beta C on one 128x128 tile per workgroup.
//   A: M x K row-major (element (m, k) at A[m * lda + k]).
//   B: bkm = 0: N x K row-major, op(B) = B^T;  bkm = 1: K x N row-major.
//   tri = 1: B (bkm = 0) is lower triangular (B[n][k] = 0 for k > n);
//   tri = 2: B (bkm = 1) is lower triangular (B[k][n] = 0 for n > k);
//   tri = 3: A is lower triangular (A[m][k] = 0 for k > m).
//   lower = 1: only tiles on or below the diagonal (M == N, SYRK).
// M, N, K are multiples of 64.  The zero triangles must hold zeros: the
// k-range is cut at tile, not element, granularity.
struct GemmJob {
  const float* A;
  const float* B;
  float* C;
  long long lda, ldb, ldc;
  int M, N, K;
  float alpha, beta;
  int bkm, tri, lower;
  int tiles_n, ntiles;
};
struct GemmJobs {
  GemmJob j[2];
};

constexpr int GT = 128;          // tile edge
constexpr int GK = 32;           // k per stage
constexpr int GSR = GK + 4;      // row-layout stride (floats)
constexpr int GSK = GT + 4;      // k-major-layout stride (floats)
constexpr int GBUF = GT * GSR;   // floats per operand buffer (>= GK * GSK)

template <bool BKM>
__device__ __forceinline__ void gemm_tile(const GemmJob& J, int ti, int tj,
                                          float* __restrict__ As,
                                          float* __restrict__ Bs) {
  const int t = threadIdx.x;
  const int lane = t & 63, wave = t >> 6;
  const int wm = wave >> 1, wn = wave & 1;
  const int r = lane & 31, h = lane >> 5;
  const int i0 = ti * GT, j0 = tj * GT;
  int kb = 0, ke = J.K;
  if (J.tri == 1) ke = min(J.K, j0 + GT);
  if (J.tri == 2) kb = j0;
  if (J.tri == 3) ke = min(J.K, i0 + GT);

  floatx16 acc[2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int e = 0; e < 16; ++e) acc[a][b][e] = 0.0f;

  float4 ra[4], rb[4];
  auto gload = [&](int k0) {
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int f = t + 256 * u;
      const int row = f >> 3, c4 = (f & 7) * 4;
      ra[u] = (i0 + row < J.M)
                  ? *(const float4*)(J.A + (long long)(i0 + row) * J.lda +
                                     k0 + c4)
                  : make_float4(0.f, 0.f, 0.f, 0.f);
      if (BKM) {
        const int kr = f >> 5, n4 = (f & 31) * 4;
        rb[u] = (j0 + n4 < J.N)
                    ? *(const float4*)(J.B + (long long)(k0 + kr) * J.ldb +
                                       j0 + n4)
                    : make_float4(0.f, 0.f, 0.f, 0.f);
      } else {
        rb[u] = (j0 + row < J.N)
                    ? *(const float4*)(J.B + (long long)(j0 + row) * J.ldb +
                                       k0 + c4)
                    : make_float4(0.f, 0.f, 0.f, 0.f);
      }
    }
  };
  auto lstore = [&](int buf) {
    float* as = As + buf * GBUF;
    float* bs = Bs + buf * GBUF;
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int f = t + 256 * u;
      const int row = f >> 3, c4 = (f & 7) * 4;
      *(float4*)(as + row * GSR + c4) = ra[u];
      if (BKM) {
        const int kr = f >> 5, n4 = (f & 31) * 4;
        *(float4*)(bs + kr * GSK + n4) = rb[u];
      } else {
        *(float4*)(bs + row * GSR + c4) = rb[u];
      }
    }
  };

  if (kb < ke) {
    gload(kb);
    lstore(0);
    __syncthreads();
  }
  int buf = 0;
  for (int k0 = kb; k0 < ke; k0 += GK) {
    const bool more = k0 + GK < ke;
    if (more) gload(k0 + GK);
    const float* as = As + buf * GBUF;
    const float* bs = Bs + buf * GBUF;
    float af[2][16], bf[2][16];
#pragma unroll
    for (int mi = 0; mi < 2; ++mi) {
      const float* p = as + (64 * wm + 32 * mi + r) * GSR + 16 * h;
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const float4 v = *(const float4*)(p + 4 * q);
        af[mi][4 * q] = v.x;
        af[mi][4 * q + 1] = v.y;
        af[mi][4 * q + 2] = v.z;
        af[mi][4 * q + 3] = v.w;
      }
    }
#pragma unroll
    for (int ni = 0; ni < 2; ++ni) {
      const int n = 64 * wn + 32 * ni + r;
      if (BKM) {
#pragma unroll
        for (int kk = 0; kk < 16; ++kk) bf[ni][kk] = bs[(16 * h + kk) * GSK + n];
      } else {
        const float* p = bs + n * GSR + 16 * h;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const float4 v = *(const float4*)(p + 4 * q);
          bf[ni][4 * q] = v.x;
          bf[ni][4 * q + 1] = v.y;
          bf[ni][4 * q + 2] = v.z;
          bf[ni][4 * q + 3] = v.w;
        }
      }
    }
#pragma unroll
    for (int kk = 0; kk < 16; ++kk)
#pragma unroll
      for (int mi = 0; mi < 2; ++mi)
#pragma unroll
        for (int ni = 0; ni < 2; ++ni)
          acc[mi][ni] = __builtin_amdgcn_mfma_f32_32x32x2f32(
              af[mi][kk], bf[ni][kk], acc[mi][ni], 0, 0, 0);
    if (more) lstore(buf ^ 1);
    __syncthreads();
    buf ^= 1;
  }

#pragma unroll
  for (int mi = 0; mi < 2; ++mi)
#pragma unroll
    for (int ni = 0; ni < 2; ++ni)
#pragma unroll
      for (int e = 0; e < 16; ++e) {
        const int row = i0 + 64 * wm + 32 * mi + 8 * (e >> 2) + 4 * h +
                        (e & 3);
        const int col = j0 + 64 * wn + 32 * ni + r;
        if (row < J.M && col < J.N) {
          float* c = J.C + (long long)row * J.ldc + col;
          float v = J.alpha * acc[mi][ni][e];
          if (J.beta != 0.0f) v = fmaf(J.beta, *c, v);
          *c = v;
        }
      }
}

__global__ __launch_bounds__(256) void k_gemm(GemmJobs jobs) {
  __shared__ __align__(16) float As[2 * GBUF];
  __shared__ __align__(16) float Bs[2 * GBUF];
  int b = blockIdx.x;
  const bool second = b >= jobs.j[0].ntiles;
  const GemmJob J = second ? jobs.j[1] : jobs.j[0];
  if (second) b -= jobs.j[0].ntiles;
  int ti, tj;
  if (J.lower) {
    ti = (int)((sqrtf(8.0f * b + 1.0f) - 1.0f) * 0.5f);
    while ((ti + 1) * (ti + 2) / 2 <= b) ++ti;
    while (ti * (ti + 1) / 2 > b) --ti;
    tj = b - ti * (ti + 1) / 2;
  } else {
    ti = b / J.tiles_n;
    tj = b % J.tiles_n;
  }
  if (J.bkm)
    gemm_tile<true>(J, ti, tj, As, Bs);
  else
    gemm_tile<false>(J, ti, tj, As, Bs);
}

static GemmJob make_job(const float* A, long long lda, const float* B,
                        long long ldb, float* C, long long ldc, int M, int N,
                        int K, float alpha, float beta, int bkm, int tri,
                        int lower) {
  GemmJob j;
  j.A = A;
  j.B = B;
  j.C = C;
  j.lda = lda;
  j.ldb = ldb;
  j.ldc = ldc;
  j.M = M;
  j.N = N;
  j.K = K;
  j.alpha = alpha;
  j.beta = beta;
  j.bkm = bkm;
  j.tri = tri;
  j.lower = lower;
  const int tm = (M + GT - 1) / GT;
  j.tiles_n = (N + GT - 1) / GT;
  j.ntiles = lower ? tm * (tm + 1) / 2 : tm * j.tiles_n;
  return j;
}

static void launch_gemm(const GemmJob& a, const GemmJob* b,
                        hipStream_t st) {
  GemmJobs jobs;
  jobs.j[0] = a;
  if (b)
    jobs.j[1] = *b;
  else {
    jobs.j[1] = a;
    jobs.j[1].ntiles = 0;
  }
  const int grid = jobs.j[0].ntiles + jobs.j[1].ntiles;
  hipLaunchKernelGGL(k_gemm, dim3(grid), dim3(256), 0, st, jobs);
}

// ---------------------------------------------------------------- leaf
// Factors the 64x64 diagonal block of A at (k0, k0) (lower triangle read)
// and writes its inverse factor W = L^-1 (64x64, zeros above the diagonal)
// at (k0, k0) of W.  One wavefront: lane i holds row i of the block in
// registers; each column step broadcasts the scaled column through LDS.
// *info receives the 1-based global index of the first non-positive pivot
// (the factorisation continues with a unit pivot so the output stays
// finite).
__global__ __launch_bounds__(64) void k_leaf64(const float* __restrict__ A,
                                               long long lda,
                                               float* __restrict__ W,
                                               long long ldw, int k0,
                                               int* __restrict__ info) {
  __shared__ __align__(16) float lcol[64];
  __shared__ __align__(16) float Ls[64 * 68];
  __shared__ __align__(16) float dinv[64];
  const int i = threadIdx.x;
  float a[64];
  const float* row = A + (long long)(k0 + i) * lda + k0;
#pragma unroll
  for (int c = 0; c < 64; c += 4) {
    const float4 v = *(const float4*)(row + c);
    a[c] = c <= i ? v.x : 0.0f;
    a[c + 1] = c + 1 <= i ? v.y : 0.0f;
    a[c + 2] = c + 2 <= i ? v.z : 0.0f;
    a[c + 3] = c + 3 <= i ? v.w : 0.0f;
  }
  int bad = 0;
#pragma unroll
  for (int j = 0; j < 64; ++j) {
    float piv = __int_as_float(
        __builtin_amdgcn_readlane(__float_as_int(a[j]), j));
    if (!(piv > 0.0f)) {
      if (bad == 0) bad = k0 + j + 1;
      piv = 1.0f;
    }
    const float rs = 1.0f / sqrtf(piv);
    float l = (i == j) ? sqrtf(piv) : a[j] * rs;
    if (i < j) l = 0.0f;
    a[j] = l;
    lcol[i] = l;
    __syncthreads();
#pragma unroll
    for (int k = j + 1; k < 64; ++k) a[k] = fmaf(-l, lcol[k], a[k]);
    __syncthreads();
  }
  if (bad != 0 && i == 0 && *info == 0) *info = bad;
  // L rows to LDS (zeros above the diagonal), reciprocal diagonal.
#pragma unroll
  for (int c = 0; c < 64; c += 4) {
    float4 v;
    v.x = c <= i ? a[c] : 0.0f;
    v.y = c + 1 <= i ? a[c + 1] : 0.0f;
    v.z = c + 2 <= i ? a[c + 2] : 0.0f;
    v.w = c + 3 <= i ? a[c + 3] : 0.0f;
    *(float4*)(Ls + i * 68 + c) = v;
  }
  float di = 0.0f;
#pragma unroll
  for (int c = 0; c < 64; ++c)
    if (c == i) di = a[c];
  dinv[i] = 1.0f / di;
  __syncthreads();
  // Lane i computes column i of W by forward substitution.
  float w[64];
#pragma unroll
  for (int k = 0; k < 64; ++k) {
    float s = (k == i) ? 1.0f : 0.0f;
#pragma unroll
    for (int m = 0; m < k; ++m) s = fmaf(-Ls[k * 68 + m], w[m], s);
    w[k] = s * dinv[k];
  }
#pragma unroll
  for (int k = 0; k < 64; ++k)
    W[(long long)(k0 + k) * ldw + k0 + i] = w[k];
}

// Split of a block of n rows into n1 + n2 (n1 a multiple of 64; at the top
// level a multiple of 256 when n >= 512).
static int split_of(int n, bool top) {
  if (n <= 64) return n;
  if (top && n >= 512) return 256 * ((n / 256 + 1) / 2);
  return 64 * ((n / 64 + 1) / 2);
}

static void chol_node(float* A, long long lda, float* W, long long ldw, int o,
                      int n, bool top, int* info, hipStream_t st) {
  if (n <= 64) {
    hipLaunchKernelGGL(k_leaf64, dim3(1), dim3(64), 0, st, A, lda, W, ldw, o,
                       info);
    return;
  }
  const int n1 = split_of(n, top), n2 = n - n1, r2 = o + n1;
  chol_node(A, lda, W, ldw, o, n1, false, info, st);
  float* A21 = A + (long long)r2 * lda + o;
  float* A22 = A + (long long)r2 * lda + r2;
  float* W11 = W + (long long)o * ldw + o;
  float* W21 = W + (long long)r2 * ldw + o;
  float* W22 = W + (long long)r2 * ldw + r2;
  // S = A21 W11^T  (W11 lower: B of the NT form is lower)
  launch_gemm(make_job(A21, lda, W11, ldw, W21, ldw, n2, n1, n1, 1.0f, 0.0f,
                       0, 1, 0),
              nullptr, st);
  // A22 -= S S^T (lower tiles), and T = S W11 into the A21 slot.
  const GemmJob syrk = make_job(W21, ldw, W21, ldw, A22, lda, n2, n2, n1,
                                -1.0f, 1.0f, 0, 0, 1);
  if (top) {
    launch_gemm(syrk, nullptr, st);
  } else {
    const GemmJob tt = make_job(W21, ldw, W11, ldw, A21, lda, n2, n1, n1,
                                1.0f, 0.0f, 1, 2, 0);
    launch_gemm(syrk, &tt, st);
  }
  chol_node(A, lda, W, ldw, r2, n2, false, info, st);
  if (!top) {
    // W21 = -W22 T
    launch_gemm(make_job(W22, ldw, A21, lda, W21, ldw, n2, n1, n2, -1.0f,
                         0.0f, 1, 3, 0),
                nullptr, st);
  }
}

// ----------------------------------------------------------- PCG kernels
// Vectors are [n][NR] (NR right-hand sides, padded).  Partial sums are
// indexed [block][n][NR] and summed in a fixed order, so every result is
// deterministic.
constexpr int SV_R = 64;     // rows of a (K p) / N-product tile
constexpr int SV_C = 512;    // columns of the same
constexpr int GTR = 128;     // reduced rows of a T-product tile
constexpr int GTC = 256;     // output columns of a T-product tile
constexpr int VR = 64;       // rows per workgroup of the vector kernels

// Sum of nslot partials [slot][NR] of column c (fixed order).
template <int NR>
__device__ __forceinline__ double sum_slots(const double* __restrict__ s,
                                            int nslot, int c) {
  double v = 0.0;
  for (int k = 0; k < nslot; ++k) v += s[k * NR + c];
  return v;
}

// Tile partial of (K + lambda I) X over 64 rows x 512 columns, FP64.
// mode 0 (CG step): X = z + beta p_old with beta = rz_new / rz_old per
//   column (0 on the first step); row-block-0 tiles store X into p_new;
//   every tile writes its share of p^T (K + lambda I) p into pq_part.
// mode 1: X = p_old (plain product, no stores besides the partials).
template <int NR>
__global__ __launch_bounds__(256) void k_symv_part(
    const float* __restrict__ K, long long ldk, int n, double lam, int mode,
    const float* __restrict__ z, const double* __restrict__ p_old,
    double* __restrict__ p_new, const double* __restrict__ rz_new,
    const double* __restrict__ rz_old, int nslot, int first,
    double* __restrict__ P64, double* __restrict__ pq_part) {
  __shared__ __align__(16) double xs[SV_C * NR];
  __shared__ double beta_s[NR];
  __shared__ double red[SV_R * NR];
  const int t = threadIdx.x;
  const int cb = blockIdx.x, rb = blockIdx.y;
  const int i0 = rb * SV_R, j0 = cb * SV_C;
  const int ncols = min(SV_C, n - j0);
  if (t < NR) {
    double b = 0.0;
    if (mode == 0 && !first) {
      const double num = sum_slots<NR>(rz_new, nslot, t);
      const double den = sum_slots<NR>(rz_old, nslot, t);
      b = den != 0.0 ? num / den : 0.0;
    }
    beta_s[t] = b;
  }
  __syncthreads();
  for (int e = t; e < ncols * NR; e += 256) {
    const int j = e / NR, c = e % NR;
    const long long g = (long long)(j0 + j) * NR + c;
    double x;
    if (mode == 0) {
      x = (double)z[g] + beta_s[c] * p_old[g];
      if (rb == 0) p_new[g] = x;
    } else {
      x = p_old[g];
    }
    xs[j * NR + c] = x;
  }
  __syncthreads();
  const int row = t >> 2, q = t & 3;
  double acc[NR];
#pragma unroll
  for (int c = 0; c < NR; ++c) acc[c] = 0.0;
  const float* krow = K + (long long)(i0 + row) * ldk + j0;
  for (int j = 4 * q; j < ncols; j += 16) {
    const float4 kv = *(const float4*)(krow + j);
    const float kk[4] = {kv.x, kv.y, kv.z, kv.w};
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const double kd = (double)kk[u];
      const double* xr = xs + (j + u) * NR;
#pragma unroll
      for (int c = 0; c < NR; ++c) acc[c] = fma(kd, xr[c], acc[c]);
    }
  }
#pragma unroll
  for (int c = 0; c < NR; ++c) {
    double v = acc[c];
    v += __shfl_xor(v, 1);
    v += __shfl_xor(v, 2);
    acc[c] = v;
  }
  const int gi = i0 + row;
  const bool diag = (gi >= j0 && gi < j0 + ncols);
  if (q == 0) {
#pragma unroll
    for (int c = 0; c < NR; ++c)
      P64[((long long)cb * n + gi) * NR + c] = acc[c];
  }
  if (mode != 0) return;
  // p_i (K_tile p)_i (+ lambda p_i^2 on the diagonal tile), per column.
  if (q == 0) {
#pragma unroll
    for (int c = 0; c < NR; ++c) {
      const long long g = (long long)gi * NR + c;
      const double pi = (double)z[g] + beta_s[c] * p_old[g];
      double v = pi * acc[c];
      if (diag) v += lam * pi * pi;
      red[row * NR + c] = v;
    }
  }
  __syncthreads();
  if (t < NR) {
    double s = 0.0;
    for (int r = 0; r < SV_R; ++r) s += red[r * NR + t];
    pq_part[((long long)rb * gridDim.x + cb) * NR + t] = s;
  }
}

// q = sum of the (K p) partials + lambda p; alpha = rz / pq;
// x += alpha p; r -= alpha q; r32 = r; rr_part = per-block sum of r^2.
template <int NR>
__global__ __launch_bounds__(256) void k_pcg_q(
    int n, double lam, int ncb, const double* __restrict__ P64,
    const double* __restrict__ pq_part, int ntiles,
    const double* __restrict__ rz, int nslot, const double* __restrict__ p,
    double* __restrict__ x, double* __restrict__ r, float* __restrict__ r32,
    double* __restrict__ rr_part) {
  __shared__ double alpha_s[NR];
  __shared__ double red[256 * NR > VR * NR ? 256 * NR : VR * NR];
  const int t = threadIdx.x;
  // pq: thread-strided partial sums, then a fixed-order tree.
  {
    double s[NR];
#pragma unroll
    for (int c = 0; c < NR; ++c) s[c] = 0.0;
    for (int k = t; k < ntiles; k += 256)
#pragma unroll
      for (int c = 0; c < NR; ++c) s[c] += pq_part[(long long)k * NR + c];
#pragma unroll
    for (int c = 0; c < NR; ++c) red[t * NR + c] = s[c];
  }
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (t < w)
#pragma unroll
      for (int c = 0; c < NR; ++c) red[t * NR + c] += red[(t + w) * NR + c];
    __syncthreads();
  }
  if (t < NR) {
    const double pq = red[t];
    const double rzc = sum_slots<NR>(rz, nslot, t);
    alpha_s[t] = pq != 0.0 ? rzc / pq : 0.0;
  }
  __syncthreads();
  const int i0 = blockIdx.x * VR;
  for (int e = t; e < VR * NR; e += 256) {
    const int i = i0 + e / NR, c = e % NR;
    double v = 0.0;
    double rr = 0.0;
    if (i < n) {
      const long long g = (long long)i * NR + c;
      double qv = 0.0;
      for (int k = 0; k < ncb; ++k) qv += P64[((long long)k * n + i) * NR + c];
      const double pv = p[g];
      qv = fma(lam, pv, qv);
      x[g] = fma(alpha_s[c], pv, x[g]);
      v = fma(-alpha_s[c], qv, r[g]);
      r[g] = v;
      r32[g] = (float)v;
      rr = v * v;
    }
    red[e] = rr;
  }
  __syncthreads();
  if (t < NR) {
    double s = 0.0;
    for (int k = 0; k < VR; ++k) s += red[k * NR + t];
    rr_part[(long long)blockIdx.x * NR + t] = s;
  }
}

// Y partials of A X (fp32) for a row-major block A (m x k), tiles of 64 rows
// x 512 reduced columns: P[cb][i][c] = sum_{j in cb} A[i][j] X[j][c].
// tri: A lower triangular (tiles above the diagonal are zero).
template <int NR>
__global__ __launch_bounds__(256) void k_gemv_n(
    const float* __restrict__ A, long long lda, int m, int k, int tri,
    const float* __restrict__ X, float* __restrict__ P) {
  __shared__ __align__(16) float xs[SV_C * NR];
  const int t = threadIdx.x;
  const int cb = blockIdx.x, rb = blockIdx.y;
  const int i0 = rb * SV_R, j0 = cb * SV_C;
  const int ncols = min(SV_C, k - j0);
  const int row = t >> 2, q = t & 3;
  float acc[NR];
#pragma unroll
  for (int c = 0; c < NR; ++c) acc[c] = 0.0f;
  if (!(tri && j0 > i0 + SV_R - 1)) {
    for (int e = t; e < ncols * NR; e += 256)
      xs[e] = X[(long long)j0 * NR + e];
    __syncthreads();
    const float* arow = A + (long long)(i0 + row) * lda + j0;
    for (int j = 4 * q; j < ncols; j += 16) {
      const float4 av = *(const float4*)(arow + j);
      const float aa[4] = {av.x, av.y, av.z, av.w};
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const float* xr = xs + (j + u) * NR;
#pragma unroll
        for (int c = 0; c < NR; ++c) acc[c] = fmaf(aa[u], xr[c], acc[c]);
      }
    }
#pragma unroll
    for (int c = 0; c < NR; ++c) {
      float v = acc[c];
      v += __shfl_xor(v, 1);
      v += __shfl_xor(v, 2);
      acc[c] = v;
    }
  }
  if (q == 0 && i0 + row < m) {
#pragma unroll
    for (int c = 0; c < NR; ++c)
      P[((long long)cb * m + i0 + row) * NR + c] = acc[c];
  }
}

// Partials of A^T Y (fp32) for a row-major block A (k rows x m columns):
// tiles of 128 reduced rows x 256 output columns,
// P[rb][j][c] = sum_{i in rb} A[i][j] Y[i][c].  tri: A lower triangular.
template <int NR>
__global__ __launch_bounds__(256) void k_gemv_t(
    const float* __restrict__ A, long long lda, int k, int m, int tri,
    const float* __restrict__ Y, float* __restrict__ P) {
  __shared__ __align__(16) float ys[GTR * NR];
  __shared__ __align__(16) float red[4 * GTC * NR];
  const int t = threadIdx.x;
  const int lane = t & 63, wave = t >> 6;
  const int cb = blockIdx.x, rb = blockIdx.y;
  const int j0 = cb * GTC, i0 = rb * GTR;
  const int nrows = min(GTR, k - i0);
  const int jc = j0 + 4 * lane;
  float acc[4][NR];
#pragma unroll
  for (int u = 0; u < 4; ++u)
#pragma unroll
    for (int c = 0; c < NR; ++c) acc[u][c] = 0.0f;
  const bool zero = tri && j0 > i0 + GTR - 1;
  if (!zero) {
    for (int e = t; e < nrows * NR; e += 256)
      ys[e] = Y[(long long)i0 * NR + e];
    __syncthreads();
    if (jc < m) {
      for (int i = wave; i < nrows; i += 4) {
        const float4 av = *(const float4*)(A + (long long)(i0 + i) * lda + jc);
        const float* yr = ys + i * NR;
#pragma unroll
        for (int c = 0; c < NR; ++c) {
          const float y = yr[c];
          acc[0][c] = fmaf(av.x, y, acc[0][c]);
          acc[1][c] = fmaf(av.y, y, acc[1][c]);
          acc[2][c] = fmaf(av.z, y, acc[2][c]);
          acc[3][c] = fmaf(av.w, y, acc[3][c]);
        }
      }
    }
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
      for (int c = 0; c < NR; ++c)
        red[(wave * GTC + 4 * lane + u) * NR + c] = acc[u][c];
    __syncthreads();
  }
  for (int e = t; e < GTC * NR; e += 256) {
    const int j = e / NR;
    if (j0 + j >= m) continue;
    float v = 0.0f;
    if (!zero)
      v = red[e] + red[GTC * NR + e] + red[2 * GTC * NR + e] +
          red[3 * GTC * NR + e];
    P[((long long)rb * m + j0) * NR + e] = v;
  }
}

// out[i][c] = s1 sum_k P[k][i][c] + s2 V[i][c] (V optional), i < len.
// With r64 given, also rz partials: rzp[slot0 + block][c] =
// sum_i r64[i][c] out[i][c] (r64 indexed from the same origin as out).
template <int NR>
__global__ __launch_bounds__(256) void k_reduce(
    const float* __restrict__ P, int np, int len, float s1,
    const float* __restrict__ V, float s2, float* __restrict__ out,
    const double* __restrict__ r64, double* __restrict__ rzp, int slot0) {
  __shared__ double red[VR * NR];
  const int t = threadIdx.x;
  const int i0 = blockIdx.x * VR;
  for (int e = t; e < VR * NR; e += 256) {
    const int i = i0 + e / NR, c = e % NR;
    double rz = 0.0;
    if (i < len) {
      const long long g = (long long)i * NR + c;
      float v = 0.0f;
      for (int k = 0; k < np; ++k) v += P[(long long)k * len * NR + g];
      v *= s1;
      if (V) v = fmaf(s2, V[g], v);
      out[g] = v;
      if (r64) rz = r64[g] * (double)v;
    }
    red[e] = rz;
  }
  if (!r64) return;
  __syncthreads();
  if (t < NR) {
    double s = 0.0;
    for (int k = 0; k < VR; ++k) s += red[k * NR + t];
    rzp[(long long)(slot0 + blockIdx.x) * NR + t] = s;
  }
}

// b, r <- B (nrhs columns, zero padded to NR); r32 = b; x = p0 = p1 = 0;
// bb_part = per-block sum of b^2.
template <int NR>
__global__ __launch_bounds__(256) void k_pcg_init(
    int n, const double* __restrict__ B, int nrhs, double* __restrict__ b,
    double* __restrict__ r, float* __restrict__ r32, double* __restrict__ x,
    double* __restrict__ p0, double* __restrict__ p1,
    double* __restrict__ bb_part) {
  __shared__ double red[VR * NR];
  const int t = threadIdx.x;
  const int i0 = blockIdx.x * VR;
  for (int e = t; e < VR * NR; e += 256) {
    const int i = i0 + e / NR, c = e % NR;
    double v = 0.0;
    if (i < n) {
      const long long g = (long long)i * NR + c;
      v = c < nrhs ? B[(long long)i * nrhs + c] : 0.0;
      b[g] = v;
      r[g] = v;
      r32[g] = (float)v;
      x[g] = 0.0;
      p0[g] = 0.0;
      p1[g] = 0.0;
    }
    red[e] = v * v;
  }
  __syncthreads();
  if (t < NR) {
    double s = 0.0;
    for (int k = 0; k < VR; ++k) s += red[k * NR + t];
    bb_part[(long long)blockIdx.x * NR + t] = s;
  }
}

// True residual d = b - (K + lambda I) x from the (K x) partials, and its
// per-block sum of squares in rr_part.  With r given, the CG residual is
// replaced by d (r = d, r32 = d); with X given, x (nrhs columns) is copied
// out.
template <int NR>
__global__ __launch_bounds__(256) void k_pcg_resid(
    int n, double lam, int ncb, const double* __restrict__ P64,
    const double* __restrict__ b, const double* __restrict__ x,
    double* __restrict__ r, float* __restrict__ r32, int nrhs,
    double* __restrict__ X, double* __restrict__ rr_part) {
  __shared__ double red[VR * NR];
  const int t = threadIdx.x;
  const int i0 = blockIdx.x * VR;
  for (int e = t; e < VR * NR; e += 256) {
    const int i = i0 + e / NR, c = e % NR;
    double rr = 0.0;
    if (i < n) {
      const long long g = (long long)i * NR + c;
      double qv = 0.0;
      for (int k = 0; k < ncb; ++k) qv += P64[((long long)k * n + i) * NR + c];
      qv = fma(lam, x[g], qv);
      const double d = b[g] - qv;
      rr = d * d;
      if (r) {
        r[g] = d;
        r32[g] = (float)d;
      }
      if (X && c < nrhs) X[(long long)i * nrhs + c] = x[g];
    }
    red[e] = rr;
  }
  __syncthreads();
  if (t < NR) {
    double s = 0.0;
    for (int k = 0; k < VR; ++k) s += red[k * NR + t];
    rr_part[(long long)blockIdx.x * NR + t] = s;
  }
}

// fp32 [n][nrhs] <-> padded [n][NR]
template <int NR>
__global__ __launch_bounds__(256) void k_pad32(int n, const float* __restrict__ src,
                                               int nrhs, float* __restrict__ dst,
                                               int to_padded) {
  const long long e = (long long)blockIdx.x * 256 + threadIdx.x;
  if (e >= (long long)n * NR) return;
  const long long i = e / NR;
  const int c = (int)(e % NR);
  if (to_padded)
    dst[e] = c < nrhs ? src[i * nrhs + c] : 0.0f;
  else if (c < nrhs)
    dst[i * nrhs + c] = src[e];
}

// ------------------------------------------------------------ host side
struct Ws {
  double *b, *x, *r, *p0, *p1, *P64, *pq, *rzp0, *rzp1, *rr, *bb;
  float *r32, *z32, *y1, *u2, *y2, *v1, *P32;
  size_t bytes;
};

static int pick_nr(int nrhs) {
  if (nrhs <= 4) return 4;
  if (nrhs <= 8) return 8;
  if (nrhs <= 12) return 12;
  return 16;
}

static Ws carve(char* base, int n, int NR) {
  Ws w;
  size_t off = 0;
  auto take = [&](size_t bytes) {
    char* p = base ? base + off : nullptr;
    off += (bytes + 255) & ~(size_t)255;
    return p;
  };
  const size_t v64 = (size_t)n * NR * sizeof(double);
  const size_t v32 = (size_t)n * NR * sizeof(float);
  const int ncb = (n + SV_C - 1) / SV_C;
  const int nslot = (n + VR - 1) / VR;
  const int ntiles = (n / SV_R) * ncb;
  const int np_max = (n + GTR - 1) / GTR > ncb ? (n + GTR - 1) / GTR : ncb;
  w.b = (double*)take(v64);
  w.x = (double*)take(v64);
  w.r = (double*)take(v64);
  w.p0 = (double*)take(v64);
  w.p1 = (double*)take(v64);
  w.P64 = (double*)take(v64 * ncb);
  w.pq = (double*)take((size_t)ntiles * NR * sizeof(double));
  w.rzp0 = (double*)take((size_t)nslot * NR * sizeof(double));
  w.rzp1 = (double*)take((size_t)nslot * NR * sizeof(double));
  w.rr = (double*)take((size_t)nslot * NR * sizeof(double));
  w.bb = (double*)take((size_t)nslot * NR * sizeof(double));
  w.r32 = (float*)take(v32);
  w.z32 = (float*)take(v32);
  w.y1 = (float*)take(v32);
  w.u2 = (float*)take(v32);
  w.y2 = (float*)take(v32);
  w.v1 = (float*)take(v32);
  w.P32 = (float*)take(v32 * np_max);
  w.bytes = off;
  return w;
}

template <int NR>
static void gemv_n(const float* A, long long lda, int m, int k, int tri,
                   const float* X, float* P, hipStream_t st) {
  dim3 grid((k + SV_C - 1) / SV_C, (m + SV_R - 1) / SV_R);
  hipLaunchKernelGGL(k_gemv_n<NR>, grid, dim3(256), 0, st, A, lda, m, k, tri,
                     X, P);
}
template <int NR>
static void gemv_t(const float* A, long long lda, int k, int m, int tri,
                   const float* Y, float* P, hipStream_t st) {
  dim3 grid((m + GTC - 1) / GTC, (k + GTR - 1) / GTR);
  hipLaunchKernelGGL(k_gemv_t<NR>, grid, dim3(256), 0, st, A, lda, k, m, tri,
                     Y, P);
}
template <int NR>
static void reduce(const float* P, int np, int len, float s1, const float* V,
                   float s2, float* out, const double* r64, double* rzp,
                   int slot0, hipStream_t st) {
  hipLaunchKernelGGL(k_reduce<NR>, dim3((len + VR - 1) / VR), dim3(256), 0,
                     st, P, np, len, s1, V, s2, out, r64, rzp, slot0);
}

// z = (L L^T)^-1 r (fp32) from the blocks in W; rz partials into rzp when
// r64 is given.
template <int NR>
static void precond(const float* W, long long ldw, int n, int n1,
                    const float* r32, float* z32, const double* r64,
                    double* rzp, const Ws& w, hipStream_t st) {
  const int n2 = n - n1;
  const float* W11 = W;
  const int npn1 = (n1 + SV_C - 1) / SV_C, npt1 = (n1 + GTR - 1) / GTR;
  gemv_n<NR>(W11, ldw, n1, n1, 1, r32, w.P32, st);
  reduce<NR>(w.P32, npn1, n1, 1.0f, nullptr, 0.0f, w.y1, nullptr, nullptr, 0,
             st);
  const float* v1 = w.y1;
  if (n2 > 0) {
    const float* S = W + (long long)n1 * ldw;
    const float* W22 = S + n1;
    const int npn2 = (n2 + SV_C - 1) / SV_C, npt2 = (n2 + GTR - 1) / GTR;
    const long long off2 = (long long)n1 * NR;
    gemv_n<NR>(S, ldw, n2, n1, 0, w.y1, w.P32, st);
    reduce<NR>(w.P32, npn1, n2, -1.0f, r32 + off2, 1.0f, w.u2, nullptr,
               nullptr, 0, st);
    gemv_n<NR>(W22, ldw, n2, n2, 1, w.u2, w.P32, st);
    reduce<NR>(w.P32, npn2, n2, 1.0f, nullptr, 0.0f, w.y2, nullptr, nullptr,
               0, st);
    gemv_t<NR>(W22, ldw, n2, n2, 1, w.y2, w.P32, st);
    reduce<NR>(w.P32, npt2, n2, 1.0f, nullptr, 0.0f, z32 + off2,
               r64 ? r64 + off2 : nullptr, rzp, n1 / VR, st);
    gemv_t<NR>(S, ldw, n2, n1, 0, z32 + off2, w.P32, st);
    reduce<NR>(w.P32, npt2, n1, -1.0f, w.y1, 1.0f, w.v1, nullptr, nullptr, 0,
               st);
    v1 = w.v1;
  }
  gemv_t<NR>(W11, ldw, n1, n1, 1, v1, w.P32, st);
  reduce<NR>(w.P32, npt1, n1, 1.0f, nullptr, 0.0f, z32, r64, rzp, 0, st);
}

template <int NR>
static void symv(const float* K, long long ldk, int n, double lam, int mode,
                 const float* z, const double* p_old, double* p_new,
                 const double* rz_new, const double* rz_old, int first,
                 const Ws& w, hipStream_t st) {
  const int ncb = (n + SV_C - 1) / SV_C;
  dim3 grid(ncb, n / SV_R);
  hipLaunchKernelGGL(k_symv_part<NR>, grid, dim3(256), 0, st, K, ldk, n, lam,
                     mode, z, p_old, p_new, rz_new, rz_old, (n + VR - 1) / VR,
                     first, w.P64, w.pq);
}

template <int NR>
static int pcg_run(const float* K, int n, long long ldk, double lam,
                   const float* W, long long ldw, int n1, const double* B,
                   int nrhs, double* X, double tol, int max_iter, char* work,
                   int* iters, double* resid, hipStream_t st) {
  const Ws w = carve(work, n, NR);
  const int ncb = (n + SV_C - 1) / SV_C;
  const int nslot = (n + VR - 1) / VR;
  const int ntiles = (n / SV_R) * ncb;
  std::vector<double> host((size_t)nslot * NR);
  auto col_sums = [&](const double* dev, double* out) -> int {
    hipError_t e = hipMemcpyAsync(host.data(), dev,
                                  host.size() * sizeof(double),
                                  hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess)
      return set_error((int)e, "nta_pcg64: %s", hipGetErrorString(e));
    for (int c = 0; c < NR; ++c) {
      double s = 0.0;
      for (int k = 0; k < nslot; ++k) s += host[(size_t)k * NR + c];
      out[c] = s;
    }
    return 0;
  };
  hipLaunchKernelGGL(k_pcg_init<NR>, dim3(nslot), dim3(256), 0, st, n, B,
                     nrhs, w.b, w.r, w.r32, w.x, w.p0, w.p1, w.bb);
  double bn[NR], rn[NR];
  int rc = col_sums(w.bb, bn);
  if (rc) return rc;
  for (int c = 0; c < NR; ++c) bn[c] = sqrt(bn[c]);
  precond<NR>(W, ldw, n, n1, w.r32, w.z32, w.r, w.rzp0, w, st);
  double* pbuf[2] = {w.p0, w.p1};
  double* rzp[2] = {w.rzp0, w.rzp1};
  // Relative residual (max over the columns) from per-column sums.
  auto rel_of = [&](const double* ss) {
    double v = 0.0;
    for (int c = 0; c < nrhs; ++c)
      if (bn[c] > 0.0) v = fmax(v, sqrt(ss[c]) / bn[c]);
    return v;
  };
  // r <- b - A x recomputed in FP64 (residual replacement).
  auto true_residual = [&](double* r, float* r32, double* X) -> int {
    symv<NR>(K, ldk, n, lam, 1, nullptr, w.x, nullptr, nullptr, nullptr, 0, w,
             st);
    hipLaunchKernelGGL(k_pcg_resid<NR>, dim3(nslot), dim3(256), 0, st, n,
                       lam, ncb, w.P64, w.b, w.x, r, r32, nrhs, X, w.rr);
    return col_sums(w.rr, rn);
  };
  double best = INFINITY;
  int stall = 0, it = 0;
  for (it = 1; it <= max_iter; ++it) {
    const double* p_old = pbuf[(it + 1) & 1];
    double* p_new = pbuf[it & 1];
    symv<NR>(K, ldk, n, lam, 0, w.z32, p_old, p_new, rzp[(it + 1) & 1],
             rzp[it & 1], it == 1, w, st);
    hipLaunchKernelGGL(k_pcg_q<NR>, dim3(nslot), dim3(256), 0, st, n, lam,
                       ncb, w.P64, w.pq, ntiles, rzp[(it + 1) & 1], nslot,
                       p_new, w.x, w.r, w.r32, w.rr);
    rc = col_sums(w.rr, rn);
    if (rc) return rc;
    double rel = rel_of(rn);
    // The recurrence drifts from b - A x on ill-conditioned systems: every
    // 10 iterations, and before accepting convergence, r is recomputed.
    if (!(rel > tol) || it % 10 == 0) {
      rc = true_residual(w.r, w.r32, nullptr);
      if (rc) return rc;
      rel = rel_of(rn);
      if (!(rel > tol)) break;
    }
    if (rel < best) {
      best = rel;
      stall = 0;
    } else if (++stall >= 20) {
      break;
    }
    precond<NR>(W, ldw, n, n1, w.r32, w.z32, w.r, rzp[it & 1], w, st);
  }
  if (it > max_iter) it = max_iter;
  // True residual of the returned x.
  rc = true_residual(nullptr, nullptr, X);
  if (rc) return rc;
  if (iters) *iters = it;
  if (resid) *resid = rel_of(rn);
  return check_launch("nta_pcg64");
}

template <int NR>
static int precond_run(const float* W, int n, long long ldw, int n1,
                       const float* R, int nrhs, float* Z, char* work,
                       hipStream_t st) {
  const Ws w = carve(work, n, NR);
  const long long tot = (long long)n * NR;
  const unsigned g = (unsigned)((tot + 255) / 256);
  hipLaunchKernelGGL(k_pad32<NR>, dim3(g), dim3(256), 0, st, n, R, nrhs,
                     w.r32, 1);
  precond<NR>(W, ldw, n, n1, w.r32, w.z32, nullptr, nullptr, w, st);
  hipLaunchKernelGGL(k_pad32<NR>, dim3(g), dim3(256), 0, st, n, w.z32, nrhs,
                     Z, 0);
  return check_launch("nta_chol_precond");
}

}  // namespace nta

using namespace nta;

extern "C" int nta_chol_split(int n) {
  if (n <= 0 || n % 64 != 0) return 0;
  return n <= 64 ? n : split_of(n, true);
}

extern "C" int nta_chol_f32(float* A, int n, long long lda, float* W,
                            long long ldw, int* info, void* stream) {
  if (n <= 0 || lda < n || ldw < n || !A || !W || !info)
    return set_error(NTA_EINVAL, "nta_chol_f32: bad n=%d lda=%lld ldw=%lld "
                     "or null pointer", n, lda, ldw);
  if (n % 64 != 0 || lda % 4 != 0 || ldw % 4 != 0)
    return set_error(NTA_EUNSUPPORTED, "nta_chol_f32: n=%d, lda and ldw must "
                     "be multiples of 64, 4 and 4 (pad the matrix)", n);
  chol_node(A, lda, W, ldw, 0, n, true, info, (hipStream_t)stream);
  return check_launch("nta_chol_f32");
}

#define NTA_NR_DISPATCH(NRV, FN, ...)  \
  switch (NRV) {                         \
    case 4: return FN<4>(__VA_ARGS__);   \
    case 8: return FN<8>(__VA_ARGS__);   \
    case 12: return FN<12>(__VA_ARGS__); \
    default: return FN<16>(__VA_ARGS__); \
  }

extern "C" long long nta_pcg64_work_bytes(int n, int nrhs) {
  if (n <= 0 || nrhs <= 0 || nrhs > 16) return -1;
  return (long long)carve(nullptr, n, pick_nr(nrhs)).bytes;
}

extern "C" int nta_chol_precond(const float* W, int n, long long ldw, int n1,
                                const float* R, int nrhs, float* Z,
                                void* work, long long work_bytes,
                                void* stream) {
  if (n <= 0 || n % 64 != 0 || ldw < n || ldw % 4 != 0 || nrhs <= 0 ||
      nrhs > 16 || !W || !R || !Z || !work)
    return set_error(NTA_EINVAL, "nta_chol_precond: bad n=%d ldw=%lld "
                     "nrhs=%d or null pointer", n, ldw, nrhs);
  if (n1 != nta_chol_split(n))
    return set_error(NTA_EINVAL, "nta_chol_precond: n1=%d is not the split "
                     "%d of nta_chol_f32", n1, nta_chol_split(n));
  if (work_bytes < nta_pcg64_work_bytes(n, nrhs))
    return set_error(NTA_EINVAL, "nta_chol_precond: work_bytes=%lld < %lld",
                     work_bytes, nta_pcg64_work_bytes(n, nrhs));
  hipStream_t st = (hipStream_t)stream;
  NTA_NR_DISPATCH(pick_nr(nrhs), precond_run, W, n, ldw, n1, R, nrhs, Z,
                  (char*)work, st);
}

extern "C" int nta_pcg64(const float* K, int n, long long ldk, double lam,
                         const float* W, long long ldw, int n1,
                         const double* B, int nrhs, double* X, double tol,
                         int max_iter, void* work, long long work_bytes,
                         int* iters, double* resid, void* stream) {
  if (n <= 0 || n % 64 != 0 || ldk < n || ldk % 4 != 0 || ldw < n ||
      ldw % 4 != 0 || nrhs <= 0 || nrhs > 16 || max_iter < 1 || !K || !W ||
      !B || !X || !work)
    return set_error(NTA_EINVAL, "nta_pcg64: bad n=%d ldk=%lld ldw=%lld "
                     "nrhs=%d max_iter=%d or null pointer", n, ldk, ldw, nrhs,
                     max_iter);
  if (n1 != nta_chol_split(n))
    return set_error(NTA_EINVAL, "nta_pcg64: n1=%d is not the split %d of "
                     "nta_chol_f32", n1, nta_chol_split(n));
  if (work_bytes < nta_pcg64_work_bytes(n, nrhs))
    return set_error(NTA_EINVAL, "nta_pcg64: work_bytes=%lld < %lld",
                     work_bytes, nta_pcg64_work_bytes(n, nrhs));
  hipStream_t st = (hipStream_t)stream;
  NTA_NR_DISPATCH(pick_nr(nrhs), pcg_run, K, n, ldk, lam, W, ldw, n1, B,
                  nrhs, X, tol, max_iter, (char*)work, iters, resid, st);
}
