// Dense linear algebra of the kernel-regression solve
// (predict.py:1235-1258: cho_factor / cho_solve of K + lambda I).
//
// Mixed precision.  K + shift I is factored in FP32 by a right-looking
// blocked Cholesky on 128-column blocks (blocked factorisation below): a
// one-workgroup leaf factors and inverts each diagonal block in LDS, so the
// triangular solve of the block column under it is a product with the
// leaf's inverse, and the trailing updates are products too.  All products
// run on one FP32 MFMA GEMM kernel (v_mfma_f32_32x32x2f32, 64x64 tiles)
// that skips the zero blocks of triangular operands.  The
// factor's inverse is then assembled inside each half of the matrix by
// recursive doubling, so W keeps W11 = L11^-1, W22 = L22^-1 on its diagonal
// blocks and L21 below them, and the preconditioner M^-1 = (L L^T)^-1 is
// applied as two block triangular solves built from those three blocks.
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

#include <algorithm>
#include <mutex>
#include <type_traits>
#include <vector>

#include "nta_common.h"

namespace nta {

typedef float floatx4 __attribute__((ext_vector_type(4)));
typedef float floatx16 __attribute__((ext_vector_type(16)));

// ---------------------------------------------------------------- GEMM
// C = alpha A op(B) + beta C on one TILE x TILE tile per workgroup.
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
  int tiles_m, tiles_n;    // tile rows and columns of the output
  int ntiles;              // output tiles of one batch member (or range)
  int batch;               // members of identical shape at fixed strides
  long long sA, sB, sC;    // element strides between members
};
struct GemmJobs {
  GemmJob j[2];
};

constexpr int GK = 32;           // k per stage
constexpr int GSR = GK + 4;      // row-layout stride (floats)

// One TILE x TILE output tile (TILE = 128: 4 waves of 64x64, 2x2 MFMA
// tiles each; TILE = 64: 4 waves of 32x32) with 32-deep k stages
// double-buffered in LDS.
template <int TILE, bool BKM>
__device__ __forceinline__ void gemm_tile(const GemmJob& J, int ti, int tj,
                                          float* __restrict__ As,
                                          float* __restrict__ Bs) {
  constexpr int MT = TILE / 64;                 // MFMA tiles per wave edge
  constexpr int NLD = TILE / 32;                // float4 loads per operand
  constexpr int GSK = TILE + 4;                 // k-major stride
  constexpr int GBUF = TILE * GSR;              // floats per operand buffer
  const int t = threadIdx.x;
  const int lane = t & 63, wave = t >> 6;
  const int wm = wave >> 1, wn = wave & 1;
  const int r = lane & 31, h = lane >> 5;
  const int i0 = ti * TILE, j0 = tj * TILE;
  int kb = 0, ke = J.K;
  if (J.tri == 1) ke = min(J.K, j0 + TILE);
  if (J.tri == 2) kb = j0;
  if (J.tri == 3) ke = min(J.K, i0 + TILE);

  floatx16 acc[MT][MT];
#pragma unroll
  for (int a = 0; a < MT; ++a)
#pragma unroll
    for (int b = 0; b < MT; ++b)
#pragma unroll
      for (int e = 0; e < 16; ++e) acc[a][b][e] = 0.0f;

  // Two register sets of global loads: stage s + 2 is fetched while stage s
  // is multiplied and stage s + 1 goes from registers to LDS, so a load has
  // two stages of MFMA work to arrive.
  float4 ra[2][NLD], rb[2][NLD];
  auto gload = [&](int k0, int set) {
#pragma unroll
    for (int u = 0; u < NLD; ++u) {
      const int f = t + 256 * u;
      const int row = f >> 3, c4 = (f & 7) * 4;
      ra[set][u] = (i0 + row < J.M)
                       ? *(const float4*)(J.A + (long long)(i0 + row) * J.lda +
                                          k0 + c4)
                       : make_float4(0.f, 0.f, 0.f, 0.f);
      if (BKM) {
        constexpr int NQ = TILE / 4;            // float4 per k-row
        const int kr = f / NQ, n4 = (f % NQ) * 4;
        rb[set][u] = (j0 + n4 < J.N)
                         ? *(const float4*)(J.B + (long long)(k0 + kr) * J.ldb +
                                            j0 + n4)
                         : make_float4(0.f, 0.f, 0.f, 0.f);
      } else {
        rb[set][u] = (j0 + row < J.N)
                         ? *(const float4*)(J.B + (long long)(j0 + row) * J.ldb +
                                            k0 + c4)
                         : make_float4(0.f, 0.f, 0.f, 0.f);
      }
    }
  };
  auto lstore = [&](int buf, int set) {
    float* as = As + buf * GBUF;
    float* bs = Bs + buf * GBUF;
#pragma unroll
    for (int u = 0; u < NLD; ++u) {
      const int f = t + 256 * u;
      const int row = f >> 3, c4 = (f & 7) * 4;
      *(float4*)(as + row * GSR + c4) = ra[set][u];
      if (BKM) {
        constexpr int NQ = TILE / 4;
        const int kr = f / NQ, n4 = (f % NQ) * 4;
        *(float4*)(bs + kr * GSK + n4) = rb[set][u];
      } else {
        *(float4*)(bs + row * GSR + c4) = rb[set][u];
      }
    }
  };

  const int nst = kb < ke ? (ke - kb) / GK : 0;
  if (nst > 0) {
    gload(kb, 0);
    if (nst > 1) gload(kb + GK, 1);
    lstore(0, 0);
    __syncthreads();
  }
  // The register set and LDS buffer of a stage alternate; the loop is
  // unrolled by two so that both are compile-time (no dynamic indexing of
  // the register arrays).
  auto stage = [&](int st, auto CUR) {
    constexpr int cur = decltype(CUR)::value;
    if (st + 2 < nst) gload(kb + (st + 2) * GK, cur);
    const float* as = As + cur * GBUF;
    const float* bs = Bs + cur * GBUF;
    float af[MT][16], bf[MT][16];
#pragma unroll
    for (int mi = 0; mi < MT; ++mi) {
      const float* p = as + (32 * MT * wm + 32 * mi + r) * GSR + 16 * h;
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
    for (int ni = 0; ni < MT; ++ni) {
      const int n = 32 * MT * wn + 32 * ni + r;
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
      for (int mi = 0; mi < MT; ++mi)
#pragma unroll
        for (int ni = 0; ni < MT; ++ni)
          acc[mi][ni] = __builtin_amdgcn_mfma_f32_32x32x2f32(
              af[mi][kk], bf[ni][kk], acc[mi][ni], 0, 0, 0);
    if (st + 1 < nst) lstore(cur ^ 1, cur ^ 1);
    __syncthreads();
  };
  for (int st = 0; st < nst; st += 2) {
    stage(st, std::integral_constant<int, 0>());
    if (st + 1 < nst) stage(st + 1, std::integral_constant<int, 1>());
  }

#pragma unroll
  for (int mi = 0; mi < MT; ++mi)
#pragma unroll
    for (int ni = 0; ni < MT; ++ni)
#pragma unroll
      for (int e = 0; e < 16; ++e) {
        const int row = i0 + 32 * MT * wm + 32 * mi + 8 * (e >> 2) + 4 * h +
                        (e & 3);
        const int col = j0 + 32 * MT * wn + 32 * ni + r;
        if (row < J.M && col < J.N) {
          float* c = J.C + (long long)row * J.ldc + col;
          float v = J.alpha * acc[mi][ni][e];
          if (J.beta != 0.0f) v = fmaf(J.beta, *c, v);
          *c = v;
        }
      }
}

// Tile bt of a two-job launch (workgroups loop over the tiles when the grid
// is capped below the tile count).
template <int TILE>
__device__ __forceinline__ void gemm_jobs_tile(const GemmJobs& jobs, int bt,
                                               float* __restrict__ As,
                                               float* __restrict__ Bs) {
  int b = bt;
  const int n0 = jobs.j[0].ntiles * jobs.j[0].batch;
  const bool second = b >= n0;
  GemmJob J = second ? jobs.j[1] : jobs.j[0];
  if (second) b -= n0;
  const int bi = b / J.ntiles;
  b -= bi * J.ntiles;
  if (bi) {
    J.A += bi * J.sA;
    J.B += bi * J.sB;
    J.C += bi * J.sC;
  }
  int ti, tj;
  if (J.lower) {
    ti = (int)((sqrtf(8.0f * b + 1.0f) - 1.0f) * 0.5f);
    while ((ti + 1) * (ti + 2) / 2 <= b) ++ti;
    while (ti * (ti + 1) / 2 > b) --ti;
    tj = b - ti * (ti + 1) / 2;
  } else {
    // Triangular operands give tiles of unequal k-length: dispatch the
    // longest first (tri 1: long for large tj; tri 3: for large ti), so the
    // in-order dispatch balances the CUs.
    const int tm = J.tiles_m;
    if (J.tri == 1) {
      tj = J.tiles_n - 1 - b / tm;
      ti = b % tm;
    } else if (J.tri == 3) {
      ti = tm - 1 - b / J.tiles_n;
      tj = b % J.tiles_n;
    } else if (J.tri == 2) {
      tj = b / tm;
      ti = b % tm;
    } else {
      ti = b / J.tiles_n;
      tj = b % J.tiles_n;
    }
  }
  if (J.bkm)
    gemm_tile<TILE, true>(J, ti, tj, As, Bs);
  else
    gemm_tile<TILE, false>(J, ti, tj, As, Bs);
  __syncthreads();                     // LDS free for the next tile
}

template <int TILE>
__global__ __launch_bounds__(256) void k_gemm(GemmJobs jobs) {
  __shared__ __align__(16) float As[2 * TILE * GSR];
  __shared__ __align__(16) float Bs[2 * TILE * GSR];
  const int total = jobs.j[0].ntiles * jobs.j[0].batch +
                    jobs.j[1].ntiles * jobs.j[1].batch;
  for (int bt = blockIdx.x; bt < total; bt += gridDim.x)
    gemm_jobs_tile<TILE>(jobs, bt, As, Bs);
}

// The tile size of every launch: 64x64 tiles keep two or more workgroups
// on a CU (a 128x128 tile's accumulators and prefetch registers fill a CU
// with one), which measured faster on every shape the factorisation uses.
constexpr int kTile = 64;

static GemmJob make_job(const float* A, long long lda, const float* B,
                        long long ldb, float* C, long long ldc, int M, int N,
                        int K, float alpha, float beta, int bkm, int tri,
                        int lower) {
  constexpr int tile = kTile;
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
  const int tm = (M + tile - 1) / tile;
  j.tiles_m = tm;
  j.tiles_n = (N + tile - 1) / tile;
  j.ntiles = lower ? tm * (tm + 1) / 2 : tm * j.tiles_n;
  j.batch = 1;
  j.sA = j.sB = j.sC = 0;
  return j;
}

// The job repeated for `batch` members whose operands lie sA, sB, sC
// elements apart.
static GemmJob batched(GemmJob j, int batch, long long sA, long long sB,
                       long long sC) {
  j.batch = batch;
  j.sA = sA;
  j.sB = sB;
  j.sC = sC;
  return j;
}

static void launch_gemm(const GemmJob& a, const GemmJob* b, hipStream_t st) {
  GemmJobs jobs;
  jobs.j[0] = a;
  if (b)
    jobs.j[1] = *b;
  else {
    jobs.j[1] = a;
    jobs.j[1].ntiles = 0;
  }
  int grid = jobs.j[0].ntiles * jobs.j[0].batch +
             jobs.j[1].ntiles * jobs.j[1].batch;
  if (grid <= 0) return;
  hipLaunchKernelGGL(k_gemm<kTile>, dim3(grid), dim3(256), 0, st, jobs);
}

// ---------------------------------------------------------------- leaves
// A 64x64 diagonal block is factored and inverted in LDS by one wavefront,
// in 2 x 2 blocks of 32:
//   1. the 64 x 32 panel [A11; A21] is factored in registers (lane i holds
//      row i), giving L11 (lanes 0-31) and L21 (lanes 32-63) together;
//   2. A22 -= L21 L21^T on v_mfma_f32_32x32x2f32;
//   3. A22 is factored the same way (lanes 0-31);
//   4. L11^-1 and L22^-1: their 16x16 diagonal blocks by forward
//      substitution (one column of one block per lane), the blocks below
//      them by two 16x16 MFMA products;
//   5. W21 = -L22^-1 (L21 L11^-1) on MFMA.
// k_leaf64 runs this on a 64 block (the last block when n is an odd
// multiple of 64); the 128 blocks have their own leaf (leaf128_body).
constexpr int LP = 33;                      // LDS row stride of 32x32 tiles
constexpr int LQ = 36;                      // 16-byte aligned LDS rows
constexpr int XS = 65;                      // LDS row stride of 64x64 blocks

// Right-looking factorisation of the 32 columns held in a[] (lane i: row i
// of the panel; lanes < 32 are the diagonal block), four columns at a time.
// Inside a block of four the few entries L[j + v][j + u] the lanes need are
// broadcast with v_readlane; the block's columns then go through LDS once
// (lane k writes L[k][j .. j + 3] as a float4) and every lane applies the
// rank-4 update to its columns k >= j + 4.  Eight LDS exchanges per panel
// instead of 32 keep the dependent chain short.  lc: 2 x 32 float4 of LDS.
// Pivots use the hardware reciprocal square root: the factor only
// preconditions.  Returns the 1-based column of the first non-positive
// pivot (0 if none).
__device__ __forceinline__ int panel32(float (&a)[32], int i,
                                       float4* __restrict__ lc) {
  int bad = 0;
#pragma unroll
  for (int j = 0; j < 32; j += 4) {
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int c = j + u;
      float piv = __int_as_float(
          __builtin_amdgcn_readlane(__float_as_int(a[c]), c));
      if (!(piv > 0.0f)) {
        if (bad == 0) bad = c + 1;
        piv = 1.0f;
      }
      const float rs = __builtin_amdgcn_rsqf(piv);
      float l = (i == c) ? piv * rs : a[c] * rs;
      if (i < c) l = 0.0f;
      a[c] = l;
#pragma unroll
      for (int v = u + 1; v < 4; ++v) {
        const float lk = __int_as_float(
            __builtin_amdgcn_readlane(__float_as_int(l), j + v));
        a[j + v] = fmaf(-l, lk, a[j + v]);
      }
    }
    if (j + 4 < 32) {
      float4* buf = lc + 32 * ((j >> 2) & 1);
      if (i < 32) buf[i] = make_float4(a[j], a[j + 1], a[j + 2], a[j + 3]);
      __builtin_amdgcn_s_waitcnt(0xc07f);    // lgkmcnt(0): writes visible
      __builtin_amdgcn_wave_barrier();
#pragma unroll
      for (int k = j + 4; k < 32; ++k) {
        const float4 v = buf[k];
        float x = a[k];
        x = fmaf(-a[j], v.x, x);
        x = fmaf(-a[j + 1], v.y, x);
        x = fmaf(-a[j + 2], v.z, x);
        x = fmaf(-a[j + 3], v.w, x);
        a[k] = x;
      }
      __builtin_amdgcn_wave_barrier();
    }
  }
  return bad;
}

// D (32x32, MFMA layout: register e of lane l is row 8 (e / 4) + 4 (l / 32)
// + e % 4, column l % 32) = X Y over K, for LDS operands with row strides SX,
// SY; TA / TB: read X / Y transposed (element (r, c) of the operand at
// [c * S + r]).
template <int SX, int SY, int K, bool TA, bool TB>
__device__ __forceinline__ floatx16 mm32(const float* __restrict__ X,
                                         const float* __restrict__ Y,
                                         int lane) {
  const int r = lane & 31, h = (lane >> 5) & 1;
  floatx16 acc;
#pragma unroll
  for (int e = 0; e < 16; ++e) acc[e] = 0.0f;
#pragma unroll
  for (int kk = 0; kk < K / 2; ++kk) {
    const int k = 2 * kk + h;
    const float av = TA ? X[k * SX + r] : X[r * SX + k];
    const float bv = TB ? Y[r * SY + k] : Y[k * SY + r];
    acc = __builtin_amdgcn_mfma_f32_32x32x2f32(av, bv, acc, 0, 0, 0);
  }
  return acc;
}

__device__ __forceinline__ int mfma_row(int e, int lane) {
  return 8 * (e >> 2) + 4 * ((lane >> 5) & 1) + (e & 3);
}

struct LeafLds {
  float L11T[32 * LQ], L22T[32 * LQ];       // L11, L22 transposed
  float L21[32 * LP], W11[32 * LP], W22[32 * LP], T[32 * LP];
  float dg[64];
  float4 lc[2 * 32];
};

// Factors the SPD block held in X (64x64, LDS, row stride XS; lower triangle
// read) and overwrites X with W = L^-1 (lower; zeros above the diagonal).
// The work runs on threads 0-63 (wavefront 0); every thread of the
// workgroup must call (the phases are separated by workgroup barriers).
// Returns, on wavefront 0, the 1-based index of the first non-positive
// pivot in the block (0 if none); the factorisation continues with a unit
// pivot so the output stays finite.
__device__ __forceinline__ int chol_inv64(float* __restrict__ X, int t,
                                          LeafLds& s) {
  const bool act = t < 64;
  const int i = t & 63, r = i & 31, h = i >> 5;
  float a[32];
  int bad = 0;
  if (act) {
#pragma unroll
    for (int c = 0; c < 32; ++c)
      a[c] = (h == 1 || c <= i) ? X[i * XS + c] : 0.0f;
    bad = panel32(a, i, s.lc);
    if (h == 0) {
#pragma unroll
      for (int c = 0; c < 32; ++c) s.L11T[c * LQ + r] = a[c];
    } else {
#pragma unroll
      for (int c = 0; c < 32; ++c) s.L21[r * LP + c] = a[c];
    }
  }
  __syncthreads();
  if (act) {
    if (h == 0) s.dg[r] = __builtin_amdgcn_rcpf(s.L11T[r * LQ + r]);
    const floatx16 d = mm32<LP, LP, 32, false, true>(s.L21, s.L21, i);
#pragma unroll
    for (int e = 0; e < 16; ++e)
      X[(32 + mfma_row(e, i)) * XS + 32 + r] -= d[e];
  }
  __syncthreads();
  if (act) {
#pragma unroll
    for (int c = 0; c < 32; ++c)
      a[c] = (h == 0 && c <= r) ? X[(32 + r) * XS + 32 + c] : 0.0f;
    const int bad2 = panel32(a, h ? 64 : r, s.lc);
    if (bad == 0 && bad2) bad = 32 + bad2;
    if (h == 0) {
#pragma unroll
      for (int c = 0; c < 32; ++c) s.L22T[c * LQ + r] = a[c];
    }
  }
  __syncthreads();
  if (act && h == 0) s.dg[32 + r] = __builtin_amdgcn_rcpf(s.L22T[r * LQ + r]);
  __syncthreads();
  if (act) {
    // The 16x16 diagonal blocks of L11 and L22 are inverted per lane:
    // lane i takes column c = i % 16 of block b = i / 16 (b = 0, 1: upper-
    // left / lower-right block of L11; 2, 3: of L22), right-looking: once
    // w[k] is final, column k (a row of the transposed copy, broadcast
    // float4s that do not depend on w) updates the rows below.
    const int b = i >> 4, c = i & 15, o = 16 * (b & 1);
    const float* Lt = (b < 2 ? s.L11T : s.L22T) + o * LQ + o;
    const float* dv = s.dg + 32 * (b >> 1) + o;
    float w[16];
#pragma unroll
    for (int k = 0; k < 16; ++k) w[k] = (k == c) ? 1.0f : 0.0f;
#pragma unroll
    for (int k = 0; k < 16; ++k) {
      w[k] *= dv[k];
#pragma unroll
      for (int m = (k + 1) & ~3; m < 16; m += 4) {
        const float4 lv = *(const float4*)(Lt + k * LQ + m);
        if (m > k) w[m] = fmaf(-lv.x, w[k], w[m]);
        if (m + 1 > k) w[m + 1] = fmaf(-lv.y, w[k], w[m + 1]);
        if (m + 2 > k) w[m + 2] = fmaf(-lv.z, w[k], w[m + 2]);
        if (m + 3 > k) w[m + 3] = fmaf(-lv.w, w[k], w[m + 3]);
      }
    }
    float* Wb = b < 2 ? s.W11 : s.W22;
#pragma unroll
    for (int k = 0; k < 16; ++k) {
      Wb[(o + k) * LP + o + c] = w[k];
      if (o == 0) Wb[k * LP + 16 + c] = 0.0f;   // upper-right 16 block
    }
  }
  __syncthreads();
  // The lower-left 16 blocks: with L = [P 0; Q R], L^-1 has -R^-1 Q P^-1
  // there; Y = Q P^-1 (into T), then -R^-1 Y, on v_mfma_f32_16x16x4f32
  // (lane l: A[l % 16][l / 16], B[l / 16][l % 16]; register v of the result
  // is row 4 (l / 16) + v, column l % 16).
  if (act) {
    const int q = i & 15, kq = i >> 4;
#pragma unroll
    for (int u = 0; u < 2; ++u) {
      const float* Lt = u ? s.L22T : s.L11T;
      const float* Wu = u ? s.W22 : s.W11;
      floatx4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
      for (int kk = 0; kk < 4; ++kk) {
        const int k = 4 * kk + kq;
        acc = __builtin_amdgcn_mfma_f32_16x16x4f32(
            Lt[k * LQ + 16 + q], Wu[k * LP + q], acc, 0, 0, 0);
      }
#pragma unroll
      for (int v = 0; v < 4; ++v)
        s.T[(16 * u + 4 * kq + v) * LP + q] = acc[v];
    }
  }
  __syncthreads();
  if (act) {
    const int q = i & 15, kq = i >> 4;
#pragma unroll
    for (int u = 0; u < 2; ++u) {
      float* Wu = u ? s.W22 : s.W11;
      floatx4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
      for (int kk = 0; kk < 4; ++kk) {
        const int k = 4 * kk + kq;
        acc = __builtin_amdgcn_mfma_f32_16x16x4f32(
            Wu[(16 + q) * LP + 16 + k], s.T[(16 * u + k) * LP + q], acc, 0,
            0, 0);
      }
#pragma unroll
      for (int v = 0; v < 4; ++v)
        Wu[(16 + 4 * kq + v) * LP + q] = -acc[v];
    }
  }
  __syncthreads();
  if (act) {
    // T = L21 W11
    const floatx16 d = mm32<LP, LP, 32, false, false>(s.L21, s.W11, i);
#pragma unroll
    for (int e = 0; e < 16; ++e) s.T[mfma_row(e, i) * LP + r] = d[e];
  }
  __syncthreads();
  if (act) {
    const floatx16 w21 = mm32<LP, LP, 32, false, false>(s.W22, s.T, i);
#pragma unroll
    for (int e = i; e < 32 * 32; e += 64) {
      const int rr = e >> 5, cc = e & 31;
      X[rr * XS + cc] = s.W11[rr * LP + cc];
      X[rr * XS + 32 + cc] = 0.0f;
      X[(32 + rr) * XS + 32 + cc] = s.W22[rr * LP + cc];
    }
#pragma unroll
    for (int e = 0; e < 16; ++e)
      X[(32 + mfma_row(e, i)) * XS + r] = -w21[e];
  }
  __syncthreads();
  return bad;
}

// Global (row stride ld) <-> LDS (row stride XS) copies of a 64x64 block by
// NT threads.
template <int NT>
__device__ __forceinline__ void block64_in(float* __restrict__ X,
                                           const float* __restrict__ G,
                                           long long ld, int t) {
#pragma unroll
  for (int f = t; f < 64 * 16; f += NT) {
    const int row = f >> 4, c4 = (f & 15) * 4;
    const float4 v = *(const float4*)(G + (long long)row * ld + c4);
    float* x = X + row * XS + c4;
    x[0] = v.x;
    x[1] = v.y;
    x[2] = v.z;
    x[3] = v.w;
  }
}

// Writes a 64x64 LDS block to global memory: all of it (full), or only its
// 32-blocks on and below the diagonal (the upper-right one is skipped).
template <int NT>
__device__ __forceinline__ void block64_out(float* __restrict__ G,
                                            long long ld,
                                            const float* __restrict__ X,
                                            bool full, int t) {
#pragma unroll
  for (int f = t; f < 64 * 16; f += NT) {
    const int row = f >> 4, c4 = (f & 15) * 4;
    if (!full && c4 >= 32 && row < 32) continue;
    const float* x = X + row * XS + c4;
    *(float4*)(G + (long long)row * ld + c4) =
        make_float4(x[0], x[1], x[2], x[3]);
  }
}

// Factors and inverts the 64x64 diagonal block of A at (k0, k0) (lower
// triangle read) into W = L^-1 at (k0, k0) of W (lower, with zeros in the
// upper-right 32x32 block).  *info receives the 1-based global index of the
// first non-positive pivot.
__global__ __launch_bounds__(64) void k_leaf64(const float* __restrict__ A,
                                               long long lda,
                                               float* __restrict__ W,
                                               long long ldw, int k0,
                                               int* __restrict__ info) {
  __shared__ float X[64 * XS];
  __shared__ __align__(16) LeafLds s;
  const int t = threadIdx.x;
  block64_in<64>(X, A + (long long)k0 * lda + k0, lda, t);
  __syncthreads();
  const int bad = chol_inv64(X, t, s);
  if (t == 0 && bad != 0 && *info == 0) *info = k0 + bad;
  block64_out<64>(W + (long long)k0 * ldw + k0, ldw, X, false, t);
  // zeros in the upper-right 32-block (triangular products read whole
  // 64-tiles)
  for (int f = t; f < 32 * 8; f += 64)
    *(float4*)(W + (long long)(k0 + (f >> 3)) * ldw + k0 + 32 + (f & 7) * 4) =
        make_float4(0.0f, 0.0f, 0.0f, 0.0f);
}

// ------------------------------------------------ 128x128 leaf
// The 128x128 diagonal block at (k0, k0) in one workgroup of four
// wavefronts, factored first and inverted afterwards:
//   1. four 32-column panels, right-looking.  A panel's rows are held one
//      per lane, up to 64 rows by the leading wavefront (its first 32 lanes
//      are the diagonal block and carry the pivot chain, as in panel32) and
//      the rows below by a second wavefront, which takes each four-column
//      block's pivots and columns from LDS after a workgroup barrier;
//      between panels the trailing lower 32x32 tiles are updated on MFMA by
//      all four wavefronts;
//   2. W = L^-1: the eight 16x16 diagonal blocks by forward substitution
//      (two wavefronts, one column per lane), then recursive doubling
//      [W_l 0; L_rl W_l] -> [W_l 0; -W_r (L_rl W_l) W_r] at 16 (four joins,
//      one per wavefront), 32 (two) and 64 (one, its four 32x32 tiles on
//      the four wavefronts), each product on MFMA from LDS.
// Only the pivot chains of step 1 are serial; the inverse, on one
// wavefront inside each half when the halves were inverted on the way,
// now runs after the factor on all four.  Writes W = L^-1 at (k0, k0): its
// lower 32-blocks and zeros in the upper-right 32-blocks of its two diagonal
// 64-blocks (the upper-right 64-block is left untouched); A is only read.
constexpr int XL = 129;                     // LDS row stride of the block
constexpr int LT = 20;                      // 16-byte aligned 16-blocks
struct Leaf2Lds {
  float X[128 * XL];                        // the block, then L, then W
  float T[64 * 65];                         // join products
  float Lt[8][16 * LT];                     // diagonal 16-blocks, L^T
  float dg[128];                            // 1 / L_ii
  float4 lc[2][32];                         // a panel block's columns
  float rs[2][4];                           // and its pivots' 1/sqrt
};

// One 32-column panel (columns [c0, c0 + 32), rows [c0, 128)) held in a[]
// by wavefronts 0 (rows c0 + lane) and 1 (rows c0 + 64 + lane); every
// thread of the workgroup calls (barrier per four-column block).  Returns
// on wavefront 0 the 1-based column of the first non-positive pivot in the
// panel (0 if none).
__device__ __forceinline__ int panel32_2w(float (&a)[32], int w, int lane,
                                          Leaf2Lds& s) {
  int bad = 0;
#pragma unroll
  for (int j = 0; j < 32; j += 4) {
    const int pb = (j >> 2) & 1;
    float4* buf = s.lc[pb];
    if (w == 0) {
      float rsv[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int c = j + u;
        float piv = __int_as_float(
            __builtin_amdgcn_readlane(__float_as_int(a[c]), c));
        if (!(piv > 0.0f)) {                  // scale 1: output finite
          if (bad == 0) bad = c + 1;
          piv = 1.0f;
        }
        // lane c holds the pivot, so a[c] r is its sqrt; the lanes above
        // the diagonal compute values that are never read (only the
        // write-back masks them)
        const float r = __builtin_amdgcn_rsqf(piv);
        rsv[u] = r;
        const float l = a[c] * r;
        a[c] = l;
#pragma unroll
        for (int v = u + 1; v < 4; ++v) {
          const float lk = __int_as_float(
              __builtin_amdgcn_readlane(__float_as_int(l), j + v));
          a[j + v] = fmaf(-l, lk, a[j + v]);
        }
      }
      if (lane < 32) buf[lane] = make_float4(a[j], a[j + 1], a[j + 2], a[j + 3]);
      if (lane == 0) {
        s.rs[pb][0] = rsv[0];
        s.rs[pb][1] = rsv[1];
        s.rs[pb][2] = rsv[2];
        s.rs[pb][3] = rsv[3];
      }
    }
    __syncthreads();
    if (w == 1) {
      // the block's four columns for rows below the leading wavefront
      float l[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        float x = a[j + u];
        const float4 bu = buf[j + u];               // L[j + u][j .. j + 3]
        if (u > 0) x = fmaf(-l[0], bu.x, x);
        if (u > 1) x = fmaf(-l[1], bu.y, x);
        if (u > 2) x = fmaf(-l[2], bu.z, x);
        l[u] = x * s.rs[pb][u];
        a[j + u] = l[u];
      }
    }
    if (w < 2 && j + 4 < 32) {
      float4 bv[28];                          // all reads in flight at once
#pragma unroll
      for (int k = j + 4; k < 32; ++k) bv[k - j - 4] = buf[k];
#pragma unroll
      for (int k = j + 4; k < 32; ++k) {
        const float4 v = bv[k - j - 4];
        float x = a[k];
        x = fmaf(-a[j], v.x, x);
        x = fmaf(-a[j + 1], v.y, x);
        x = fmaf(-a[j + 2], v.z, x);
        x = fmaf(-a[j + 3], v.w, x);
        a[k] = x;
      }
    }
  }
  return bad;
}

__device__ __forceinline__ void leaf128_body(
    const float* __restrict__ A, long long lda, float* __restrict__ W,
    long long ldw, int k0, int* __restrict__ info, Leaf2Lds& s) {
  const int t = threadIdx.x, w = t >> 6, lane = t & 63;
  float* X = s.X;
  const float* A0 = A + (long long)k0 * lda + k0;
  // The block's lower triangle (and the float4 overhang into the diagonal
  // 32-blocks, masked later); zeros in the upper-right 32-blocks of the
  // diagonal 64-blocks, the zero blocks of W_l, W_r in the join at 64.  All
  // loads of a thread are issued before its stores.
  {
    float4 v[16];
#pragma unroll
    for (int q = 0; q < 16; ++q) {
      const int f = t + 256 * q, row = f >> 5, c4 = (f & 31) * 4;
      v[q] = c4 <= row ? *(const float4*)(A0 + (long long)row * lda + c4)
                       : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    }
#pragma unroll
    for (int q = 0; q < 16; ++q) {
      const int f = t + 256 * q, row = f >> 5, c4 = (f & 31) * 4;
      const bool zero_blk = (row >> 5) + 1 == (c4 >> 5) && !(row & 32);
      if (c4 <= row || zero_blk) {
        float* x = X + row * XL + c4;
        x[0] = v[q].x;
        x[1] = v[q].y;
        x[2] = v[q].z;
        x[3] = v[q].w;
      }
    }
  }
  __syncthreads();
  // 1. factor
  int bad = 0;
#pragma unroll 1
  for (int p = 0; p < 4; ++p) {
    const int c0 = 32 * p;
    const int row = c0 + 64 * w + lane;             // this lane's row
    const bool act = w < 2 && row < 128;
    const int i = row - c0;                         // row within the panel
    float a[32];
#pragma unroll
    for (int c = 0; c < 32; ++c)
      a[c] = (act && c <= i) ? X[row * XL + c0 + c] : 0.0f;
    const int b = panel32_2w(a, w, lane, s);
    if (w == 0 && bad == 0 && b != 0) bad = c0 + b;
    if (act) {
#pragma unroll
      for (int c = 0; c < 32; ++c)
        X[row * XL + c0 + c] = c <= i ? a[c] : 0.0f;
      if (w == 0 && lane < 32) {
        // the panel's two diagonal 16-blocks transposed for the inverse
        // (only the entries below the diagonal are read there), and the
        // reciprocal pivot (a non-positive pivot was reported and left in
        // place: 1 keeps the inverse finite)
        float* lt = s.Lt[2 * p + (lane >> 4)] + (lane & 15);
        if (lane < 16) {                      // static register indices
#pragma unroll
          for (int k = 0; k < 16; ++k) lt[k * LT] = a[k];
        } else {
#pragma unroll
          for (int k = 0; k < 16; ++k) lt[k * LT] = a[16 + k];
        }
        const float d = X[row * XL + row];
        s.dg[row] = d > 0.0f ? __builtin_amdgcn_rcpf(d) : 1.0f;
      }
    }
    __syncthreads();
    if (p < 3) {
      // A[c0 + 32:, c0 + 32:] -= L L^T on its lower 32x32 tiles
      const int nb = 3 - p;
      for (int tt = w; tt < nb * (nb + 1) / 2; tt += 4) {
        int ti = 0;
        while ((ti + 1) * (ti + 2) / 2 <= tt) ++ti;
        const int tj = tt - ti * (ti + 1) / 2;
        const int R0 = c0 + 32 + 32 * ti, C0 = c0 + 32 + 32 * tj;
        const floatx16 d = mm32<XL, XL, 32, false, true>(
            X + R0 * XL + c0, X + C0 * XL + c0, lane);
#pragma unroll
        for (int e = 0; e < 16; ++e)
          X[(R0 + mfma_row(e, lane)) * XL + C0 + (lane & 31)] -= d[e];
      }
      __syncthreads();
    }
  }
  if (t == 0 && bad != 0 && *info == 0) *info = k0 + bad;
  // 2. W = L^-1 (the transposed diagonal 16-blocks and the reciprocal
  // pivots were written with the panels)
  if (w < 2) {
    // lane: column c of block b's inverse, right-looking on the transposed
    // copy (the float4 reads do not depend on the solve)
    const int b = 4 * w + (lane >> 4), c = lane & 15;
    const float* Lt = s.Lt[b];
    const float* dv = s.dg + 16 * b;
    float v[16];
#pragma unroll
    for (int k = 0; k < 16; ++k) v[k] = (k == c) ? 1.0f : 0.0f;
#pragma unroll
    for (int k = 0; k < 16; ++k) {
      v[k] *= dv[k];
#pragma unroll
      for (int m = (k + 1) & ~3; m < 16; m += 4) {
        const float4 lv = *(const float4*)(Lt + k * LT + m);
        if (m > k) v[m] = fmaf(-lv.x, v[k], v[m]);
        if (m + 1 > k) v[m + 1] = fmaf(-lv.y, v[k], v[m + 1]);
        if (m + 2 > k) v[m + 2] = fmaf(-lv.z, v[k], v[m + 2]);
        if (m + 3 > k) v[m + 3] = fmaf(-lv.w, v[k], v[m + 3]);
      }
    }
#pragma unroll
    for (int k = 0; k < 16; ++k) X[(16 * b + k) * XL + 16 * b + c] = v[k];
  }
  __syncthreads();
  {
    // joins at 16 (wavefront w: the 32-block w) on v_mfma_f32_16x16x4f32
    // (lane l: A[l % 16][k], B[k][l % 16] with k = 4 kk + l / 16; register v
    // of the result is row 4 (l / 16) + v, column l % 16)
    const int q = lane & 15, kq = lane >> 4, o = 32 * w;
    float* Tq = s.T + w * (16 * 17);
    floatx4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int kk = 0; kk < 4; ++kk) {
      const int k = 4 * kk + kq;
      acc = __builtin_amdgcn_mfma_f32_16x16x4f32(
          X[(o + 16 + q) * XL + o + k], X[(o + k) * XL + o + q], acc, 0, 0,
          0);
    }
#pragma unroll
    for (int v = 0; v < 4; ++v) Tq[(4 * kq + v) * 17 + q] = acc[v];
    __builtin_amdgcn_s_waitcnt(0xc07f);
    __builtin_amdgcn_wave_barrier();
    acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int kk = 0; kk < 4; ++kk) {
      const int k = 4 * kk + kq;
      acc = __builtin_amdgcn_mfma_f32_16x16x4f32(
          X[(o + 16 + q) * XL + o + 16 + k], Tq[k * 17 + q], acc, 0, 0, 0);
    }
#pragma unroll
    for (int v = 0; v < 4; ++v)
      X[(o + 16 + 4 * kq + v) * XL + o + q] = -acc[v];
  }
  __syncthreads();
  if (w < 2) {
    // joins at 32 (wavefront w: the 64-block w)
    const int o = 64 * w;
    float* Th = s.T + w * (32 * 33);
    const floatx16 d = mm32<XL, XL, 32, false, false>(
        X + (o + 32) * XL + o, X + o * XL + o, lane);
#pragma unroll
    for (int e = 0; e < 16; ++e)
      Th[mfma_row(e, lane) * 33 + (lane & 31)] = d[e];
    __builtin_amdgcn_s_waitcnt(0xc07f);
    __builtin_amdgcn_wave_barrier();
    const floatx16 g = mm32<XL, 33, 32, false, false>(
        X + (o + 32) * XL + o + 32, Th, lane);
#pragma unroll
    for (int e = 0; e < 16; ++e)
      X[(o + 32 + mfma_row(e, lane)) * XL + o + (lane & 31)] = -g[e];
  }
  __syncthreads();
  // 3. W out: its diagonal 64-blocks W_l, W_r now (the lower 32-blocks,
  // and zeros in the upper-right ones: triangular products read whole
  // 64-tiles), the join at 64 straight from its product registers
  float* W0 = W + (long long)k0 * ldw + k0;
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    const int f = t + 256 * q;                 // float4 slot; block q
    const int blk = f >> 8, e = f & 255;
    int bi, bj;
    if (blk < 6) {                             // lower blocks of W_l, W_r
      const int h = blk >= 3, bb = blk - 3 * h;  // (0,0) (1,0) (1,1)
      bi = 2 * h + (bb > 0);
      bj = 2 * h + (bb > 1);
    } else {                                   // (0,1) of W_l, W_r
      bi = 2 * (blk - 6);
      bj = bi + 1;
    }
    const int row = 32 * bi + (e >> 3), c4 = 32 * bj + (e & 7) * 4;
    const float* x = X + row * XL + c4;
    *(float4*)(W0 + (long long)row * ldw + c4) =
        blk < 6 ? make_float4(x[0], x[1], x[2], x[3])
                : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  }
  {
    // the join at 64: tile (ib, jb) of the 64x64 products per wavefront
    const int ib = w >> 1, jb = w & 1;
    const floatx16 d = mm32<XL, XL, 64, false, false>(
        X + (64 + 32 * ib) * XL, X + 32 * jb, lane);
#pragma unroll
    for (int e = 0; e < 16; ++e)
      s.T[(32 * ib + mfma_row(e, lane)) * 65 + 32 * jb + (lane & 31)] = d[e];
    __syncthreads();
    const floatx16 g = mm32<XL, 65, 64, false, false>(
        X + (64 + 32 * ib) * XL + 64, s.T + 32 * jb, lane);
#pragma unroll
    for (int e = 0; e < 16; ++e)
      W0[(long long)(64 + 32 * ib + mfma_row(e, lane)) * ldw + 32 * jb +
         (lane & 31)] = -g[e];
  }
}

__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(1, 1)))
void k_leaf128(const float* __restrict__ A,
                                                 long long lda,
                                                 float* __restrict__ W,
                                                 long long ldw, int k0,
                                                 int* __restrict__ info) {
  __shared__ __align__(16) Leaf2Lds s;
  leaf128_body(A, lda, W, ldw, k0, info, s);
}

// ------------------------------------------------ blocked factorisation
// Right-looking blocked Cholesky on 128-column blocks (the last one 64 when
// n is an odd multiple of 64), two levels deep, on two streams.  Outer
// panels of 512 columns; inside a panel, step k (block k at rows
// [o, o + b), trailing rows r0 = o + b):
//   leaf:   W_kk = L_kk^-1 from A_kk (one workgroup);
//   panel:  L_ik = A_ik W_kk^T for the rows below (the inverse makes the
//           triangular solve a GEMM), into W's block column k;
//   inner:  A[r0:n, r0:panel end] -= L_{., k} L_{r0:panel end, k}^T.
// After a panel (columns [O, E0)), its L updates the rest with K = 512
// (four times the arithmetic intensity of per-128 updates, whose
// read-modify-write of the trailing matrix would be memory-bound):
//   head: the next panel's first 128 columns, on the caller's stream (the
//         next leaf needs them), after the previous bulk update (which also
//         writes them);
//   mid:  the rest of the next panel's columns, on the side stream (the
//         next panel's first inner update waits for it);
//   bulk: everything right of the next panel (lower tiles), on the side
//         stream, overlapping the next panel's factorisation.
// The chain of leaves, panel products and inner updates is the critical
// path; the n^3 / 3 flops of the bulk updates fill the GPU beside it.  The
// updates use 64x64 tiles: a 128x128-tile workgroup fills a CU's registers
// (one wave per SIMD), and on these K = 512 updates the smaller tiles, two
// or more workgroups per CU, are faster (7168^2 lower x 512: 94 against 62
// TFLOP/s on one MI355X).  The factor's inverse is built by recursive
// doubling inside each half (chol_invert_half): the top half on a third
// stream as soon as it is factored, the bottom half after the
// factorisation.
// Events: the last panel product, the last mid update, the last two bulk
// updates, the side stream's end, the end of the top half's factorisation
// and of its inversion (on the third stream).
enum { kEvPanel, kEvMid, kEvBulk0, kEvBulk1, kEvSideEnd, kEvSplit, kEvInvEnd,
       kCholEvents };
struct CholStreams {
  hipStream_t side = nullptr, inv = nullptr;
  hipEvent_t ev[kCholEvents] = {};
};
constexpr int kCholDevices = 64;
static std::mutex g_chol_mu;
static CholStreams g_chol[kCholDevices];

// The side and inversion streams and the events of the current device
// (created on first use, default priority, not synchronising with the null
// stream), or nullptr after setting the error.
static CholStreams* chol_streams() {
  int dev = -1;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess || dev < 0 || dev >= kCholDevices) {
    set_error(e != hipSuccess ? (int)e : NTA_EINVAL,
              "nta_chol_f32: no usable device (%d)", dev);
    return nullptr;
  }
  std::lock_guard<std::mutex> lock(g_chol_mu);
  CholStreams& c = g_chol[dev];
  if (!c.side) {
    hipStream_t s = nullptr, iv = nullptr;
    hipEvent_t ev[kCholEvents] = {};
    e = hipStreamCreateWithFlags(&s, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&iv, hipStreamNonBlocking);
    for (int k = 0; k < kCholEvents && e == hipSuccess; ++k)
      e = hipEventCreateWithFlags(&ev[k], hipEventDisableTiming);
    if (e != hipSuccess) {
      set_error((int)e, "nta_chol_f32: stream/event setup: %s",
                hipGetErrorString(e));
      return nullptr;
    }
    for (int k = 0; k < kCholEvents; ++k) c.ev[k] = ev[k];
    c.inv = iv;
    c.side = s;
  }
  return &c;
}

// Block offsets of an n x n factorisation (n a positive multiple of 64).
static std::vector<int> chol_offsets(int n) {
  std::vector<int> off;
  for (int o = 0; o < n; o += 128) off.push_back(o);
  off.push_back(n);
  return off;                 // block k is [off[k], off[k + 1])
}

static int chol_split_of(int n) {
  const std::vector<int> off = chol_offsets(n);
  const int m = (int)off.size() - 1;
  return m == 1 ? n : off[m / 2];
}

// Transpose of the lower blocks of W's first ncols columns (32x32 tiles on
// or below the diagonal, rows < n) into WT (WT's tiles strictly below the
// diagonal are left untouched; the preconditioner never reads them).  Grid:
// (ncols / 32, n / 32) tiles, rounded up.
__global__ __launch_bounds__(256) void k_transpose_lower(
    const float* __restrict__ W, long long ldw, int n, int ncols,
    float* __restrict__ WT, long long ldt) {
  __shared__ float tile[32][33];
  const int bx = blockIdx.x * 32, by = blockIdx.y * 32;   // W tile (by, bx)
  if (bx > by || bx >= ncols) return;
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

constexpr int kOuter = 512;

static void chol_invert_half(float* W, long long ldw,
                             const std::vector<int>& off, int kb, int ke,
                             float* tmp, long long ldt, hipStream_t st,
                             int smax);

// split > 0 (a panel boundary): once columns [0, split) are factored, the
// top half [0, split) is inverted on the inversion stream (scratch: A's
// rows above split, which the rest of the factorisation no longer touches)
// beside the rest of the factorisation; kEvInvEnd marks its end.
static int chol_factor(float* A, long long lda, float* W, long long ldw,
                       float* WT, int n, int* info, hipStream_t st,
                       CholStreams& cs, int split) {
  hipStream_t sd = cs.side;
  hipEvent_t* ev = cs.ev;
  hipError_t e = hipSuccess;
  auto wait = [&](hipStream_t waiter, int id) {
    if (e == hipSuccess) e = hipStreamWaitEvent(waiter, ev[id], 0);
  };
  auto record = [&](int id, hipStream_t recorder) {
    if (e == hipSuccess) e = hipEventRecord(ev[id], recorder);
  };
  // A[r:n, c0:c1] -= L[r:n, kc:kc+K] L[c0:c1, kc:kc+K]^T (lower: only the
  // tiles on and below the diagonal, for c0 == r, c1 == n)
  auto update = [&](int r, int c0, int c1, int kc, int K, int lower,
                    hipStream_t s) {
    const float* Lr = W + (long long)r * ldw + kc;
    const float* Lc = W + (long long)c0 * ldw + kc;
    launch_gemm(make_job(Lr, ldw, Lc, ldw, A + (long long)r * lda + c0, lda,
                         n - r, c1 - c0, K, -1.0f, 1.0f, 0, 0, lower),
                nullptr, s);
  };
  const std::vector<int> off = chol_offsets(n);
  const int m = (int)off.size() - 1;
  int bulk_prev = -1;                 // event of the previous bulk update
  int nbulk = 0;
  bool mid_pending = false;           // the panel's first update waits
  int k = 0;
  while (k < m && e == hipSuccess) {
    const int O = off[k];
    const int ke = std::min(m, k + kOuter / 128);   // inner blocks [k, ke)
    const int E0 = off[ke];                          // end of the panel
    for (; k < ke; ++k) {
      const int o = off[k], b = off[k + 1] - o, r0 = o + b, rem = n - r0;
      if (b == 128)
        hipLaunchKernelGGL(k_leaf128, dim3(1), dim3(256), 0, st, A, lda, W,
                           ldw, o, info);
      else
        hipLaunchKernelGGL(k_leaf64, dim3(1), dim3(64), 0, st, A, lda, W,
                           ldw, o, info);
      if (rem == 0) break;
      // panel: W[r0:n, o:o+b] = A[r0:n, o:o+b] W_kk^T (W_kk lower)
      launch_gemm(make_job(A + (long long)r0 * lda + o, lda,
                           W + (long long)o * ldw + o, ldw,
                           W + (long long)r0 * ldw + o, ldw, rem, b, b, 1.0f,
                           0.0f, 0, 1, 0),
                  nullptr, st);
      if (r0 < E0) {
        // the columns [r0, E0) take the previous panel's mid update first
        if (mid_pending) wait(st, kEvMid);
        mid_pending = false;
        update(r0, r0, E0, o, b, 0, st);
      }
    }
    if (E0 == split) {
      record(kEvSplit, st);
      wait(cs.inv, kEvSplit);
      if (e == hipSuccess) {
        chol_invert_half(W, ldw, off, 0, ke, A, lda, cs.inv, m);
        // W11 and L21 (columns [0, split)) are final: their transposes too
        const dim3 grid((split + 31) / 32, (n + 31) / 32);
        hipLaunchKernelGGL(k_transpose_lower, grid, dim3(256), 0, cs.inv, W,
                           ldw, n, split, WT, ldw);
        record(kEvInvEnd, cs.inv);
      }
    }
    if (E0 >= n || e != hipSuccess) break;
    const int E1 = off[std::min(m, ke + kOuter / 128)];   // next panel end
    const int H1 = std::min(E1, E0 + 128);                // its first block
    const int Wd = E0 - O;
    // side stream: mid and bulk updates by this panel's L[., O:E0]
    record(kEvPanel, st);
    wait(sd, kEvPanel);
    if (H1 < E1) {
      update(H1, H1, E1, O, Wd, 0, sd);
      record(kEvMid, sd);
      mid_pending = true;
    }
    int bulk_now = -1;
    if (E1 < n) {
      update(E1, E1, n, O, Wd, 1, sd);
      bulk_now = kEvBulk0 + (nbulk++ & 1);
      record(bulk_now, sd);
    }
    // the next panel's first block columns, after the previous bulk update
    if (bulk_prev >= 0) wait(st, bulk_prev);
    bulk_prev = bulk_now;
    update(E0, E0, H1, O, Wd, 0, st);
  }
  // every side-stream update is complete before the stream goes on
  record(kEvSideEnd, sd);
  wait(st, kEvSideEnd);
  if (e != hipSuccess)
    return set_error((int)e, "nta_chol_f32: %s", hipGetErrorString(e));
  return 0;
}

// Inverse of the factor inside blocks [kb, ke) (rows/columns off[kb] ..
// off[ke]), from the block inverses on W's diagonal and the L blocks below
// them, by recursive doubling: groups of s blocks are joined pairwise,
//   [W_l 0; L_rl W_r] -> [W_l 0; -W_r (L_rl W_l) W_r],
// the product T = L_rl W_l into scratch rows of `tmp` (leading dimension
// ldt), then -W_r T over L_rl in W.  All joins of one size in a level run
// as one batched launch (the one shorter join at the end, if any, as the
// second job of that launch).
static void chol_invert_half(float* W, long long ldw,
                             const std::vector<int>& off, int kb, int ke,
                             float* tmp, long long ldt, hipStream_t st,
                             int smax) {
  const int nb = ke - kb;
  for (int s = 1; s < nb && s < smax; s *= 2) {
    // joins: left blocks [kb + 2 g s, + s), right [kb + (2 g + 1) s, + s)
    int full = 0;
    while (kb + (2 * full + 2) * s <= ke &&
           off[kb + (2 * full + 2) * s] - off[kb + (2 * full + 1) * s] ==
               s * 128 &&
           off[kb + (2 * full + 1) * s] - off[kb + 2 * full * s] == s * 128)
      ++full;
    const int g1 = full;                          // a shorter last join?
    const bool extra = kb + (2 * g1 + 1) * s < ke;
    const long long stW = (long long)(2 * s * 128) * (ldw + 1);
    GemmJob tj[2], wj[2];
    int nj = 0;
    // T of join g in scratch rows [2 g s 128, + rs): the batch members
    // and the shorter last join never overlap.
    auto jobs_for = [&](int g, int cnt, int ls, int rs) {
      const int lo = off[kb + 2 * g * s], ro = lo + ls;
      float* Tg = tmp + (long long)(2 * g * s * 128) * ldt;
      GemmJob t = make_job(W + (long long)ro * ldw + lo, ldw,
                           W + (long long)lo * ldw + lo, ldw, Tg, ldt, rs, ls,
                           ls, 1.0f, 0.0f, 1, 2, 0);
      GemmJob w = make_job(W + (long long)ro * ldw + ro, ldw, Tg, ldt,
                           W + (long long)ro * ldw + lo, ldw, rs, ls, rs,
                           -1.0f, 0.0f, 1, 3, 0);
      tj[nj] = batched(t, cnt, stW, stW, (long long)(2 * s * 128) * ldt);
      wj[nj] = batched(w, cnt, stW, (long long)(2 * s * 128) * ldt, stW);
      ++nj;
    };
    if (full > 0) jobs_for(0, full, s * 128, s * 128);
    if (extra) {
      const int lo = off[kb + 2 * g1 * s], ro = off[kb + (2 * g1 + 1) * s];
      const int re = off[kb + (2 * g1 + 2) * s <= ke ? kb + (2 * g1 + 2) * s
                                                     : ke];
      jobs_for(g1, 1, ro - lo, re - ro);
    }
    if (nj == 0) break;
    launch_gemm(tj[0], nj > 1 ? &tj[1] : nullptr, st);
    launch_gemm(wj[0], nj > 1 ? &wj[1] : nullptr, st);
  }
}

// ----------------------------------------------------------- PCG kernels
// Vectors are [n][16] (up to 16 right-hand sides, zero padded): one MFMA
// 16x16 tile spans the right-hand sides.  Every product is row-complete
// (a workgroup owns whole output rows), so there are no partial-sum
// buffers; per-workgroup dot-product partials are summed in a fixed order.
constexpr int RHS = 16;
constexpr int VROWS = 256;        // rows per workgroup of the vector kernels

typedef double doublex4 __attribute__((ext_vector_type(4)));

// Column range of the k-loop of row stripe [i0, i0 + 16) for a lower (tri =
// 1: columns <= row) or upper (tri = 2: columns >= row) triangular operand.
__device__ __forceinline__ void tri_range(int tri, int i0, int k, int& kb,
                                          int& ke) {
  kb = 0;
  ke = k;
  if (tri == 1) ke = min(k, i0 + 16);
  if (tri == 2) kb = i0;
}

// One 16-column chunk of a row-major fp32 operand and the matching 16 rows
// of a [k][RHS] right-hand side: lane l holds A[row l%16][4 (l/16) + kk]
// and X[4 (l/16) + kk][rhs l%16] for MFMA kk (the reduction index is
// permuted identically in both operands).
// TK: element type of the matrix operand (float, or double for an FP64
// kernel matrix in the residual operator).
template <typename TK, typename TX>
struct Chunk {
  TK a[4];
  TX x[4];
};

template <typename TK, typename TX>
__device__ __forceinline__ void load_chunk(Chunk<TK, TX>& c, bool ok,
                                           const TK* __restrict__ arow,
                                           const TX* __restrict__ xcol,
                                           int j0) {
  if (ok) {
    if constexpr (std::is_same<TK, float>::value) {
      const float4 v = *(const float4*)(arow + j0);
      c.a[0] = v.x;
      c.a[1] = v.y;
      c.a[2] = v.z;
      c.a[3] = v.w;
    } else {
      const double2 v0 = *(const double2*)(arow + j0);
      const double2 v1 = *(const double2*)(arow + j0 + 2);
      c.a[0] = v0.x;
      c.a[1] = v0.y;
      c.a[2] = v1.x;
      c.a[3] = v1.y;
    }
#pragma unroll
    for (int kk = 0; kk < 4; ++kk) c.x[kk] = xcol[(long long)(j0 + kk) * RHS];
  } else {
#pragma unroll
    for (int kk = 0; kk < 4; ++kk) {
      c.a[kk] = (TK)0;
      c.x[kk] = (TX)0;
    }
  }
}

template <typename TK>
__device__ __forceinline__ void mfma_chunk(doublex4& acc,
                                           const Chunk<TK, double>& c) {
#pragma unroll
  for (int kk = 0; kk < 4; ++kk)
    acc = __builtin_amdgcn_mfma_f64_16x16x4f64((double)c.a[kk], c.x[kk], acc,
                                               0, 0, 0);
}

__device__ __forceinline__ void mfma_chunk(floatx4& acc,
                                           const Chunk<float, float>& c) {
#pragma unroll
  for (int kk = 0; kk < 4; ++kk)
    acc = __builtin_amdgcn_mfma_f32_16x16x4f32(c.a[kk], c.x[kk], acc, 0, 0,
                                               0);
}

// acc += A[16 rows, chunks c0, c0 + step, ...] X over the chunk indices
// c0 + step * i < c1 (chunks of 16 columns), software-pipelined: the loads
// of the next group of 4 chunks are in flight while the current group's
// MFMAs run.
template <typename TK, typename TX, typename ACC>
__device__ __forceinline__ void row_stripe_product(
    ACC& acc, const TK* __restrict__ arow, const TX* __restrict__ xcol,
    int c0, int c1, int step) {
  Chunk<TK, TX> cur[4], nxt[4];
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    const int c = c0 + step * u;
    load_chunk(cur[u], c < c1, arow, xcol, 16 * c);
  }
  for (int cb = c0; cb < c1; cb += 4 * step) {
    const int cn = cb + 4 * step;
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int c = cn + step * u;
      load_chunk(nxt[u], c < c1, arow, xcol, 16 * c);
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) mfma_chunk(acc, cur[u]);
#pragma unroll
    for (int u = 0; u < 4; ++u) cur[u] = nxt[u];
  }
}

// (K + lam I) P in FP64 for 32 rows per workgroup of 2 kSvSets waves:
// wave w computes rows [i0 + 16 (w & 1), +16) over the 16-column chunks
// c = w >> 1 (mod kSvSets) on v_mfma_f64_16x16x4f64 (K read in fp32 --
// exact in fp64 -- or in fp64); the kSvSets partial tiles of a row stripe
// are added through LDS in a fixed order.
// mode 0: Q = (K + lam I) P, part[wg][c] = sum_rows P Q.
// mode 2: Q = K P for a rectangular K (rows = 32 gridDim.x, n columns):
//         neither P's rows of Q's index nor part are touched; with
//         gridDim.y > 1, part y of the 16-column chunks into Q + y (32
//         gridDim.x) RHS.
// mode 1: D = B - (K + lam I) P; R = D, R32 = D (when R is given),
//         part[wg][c] = sum_rows D^2.
constexpr int kSvSets = 4;
constexpr int kSvThreads = 128 * kSvSets;
template <typename TK>
__global__ __launch_bounds__(kSvThreads) void k_symv_rows(
    const TK* __restrict__ K, long long ldk, int n, double lam, int mode,
    const double* __restrict__ P, double* __restrict__ Q,
    const double* __restrict__ B, double* __restrict__ R,
    float* __restrict__ R32, double* __restrict__ part,
    const int* __restrict__ done) {
  if (done && *done) return;           // converged: the rest is a no-op
  __shared__ doublex4 red[kSvSets][2][64];
  __shared__ double dots[2][RHS];
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  const int s = w & 1, qd = w >> 1;
  const int r = lane & 15, g = lane >> 4;
  const int i0 = blockIdx.x * 32 + 16 * s;
  // mode 2 may split the columns over gridDim.y parts (partial products
  // into Q's part blockIdx.y)
  const int nch = n / 16, per = (nch + gridDim.y - 1) / gridDim.y;
  const int cb = blockIdx.y * per, ce = min(nch, cb + per);
  if (gridDim.y > 1) Q += (long long)blockIdx.y * gridDim.x * 32 * RHS;
  doublex4 acc = {0.0, 0.0, 0.0, 0.0};
  row_stripe_product(acc, K + (long long)(i0 + r) * ldk + 4 * g,
                     P + (long long)(4 * g) * RHS + r, cb + qd, ce,
                     kSvSets);
  red[qd][s][lane] = acc;
  __syncthreads();
  if (qd == 0) {
    double d = 0.0;
    // v_mfma_f64_16x16x4f64 output: register v of lane l is row
    // 4 v + l / 16, column l % 16 (the f32 16x16x4 layout is row
    // 4 (l / 16) + v).
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const long long e = (long long)(i0 + 4 * v + g) * RHS + r;
      double kp = acc[v];
#pragma unroll
      for (int q = 1; q < kSvSets; ++q) kp += red[q][s][lane][v];
      if (mode == 2) {
        Q[e] = kp;
        continue;
      }
      const double pv = P[e];
      const double qv = fma(lam, pv, kp);
      if (mode == 0) {
        Q[e] = qv;
        d += pv * qv;
      } else {
        const double dv = B[e] - qv;
        if (R) {
          R[e] = dv;
          R32[e] = (float)dv;
        }
        d += dv * dv;
      }
    }
    d += __shfl_xor(d, 16);
    d += __shfl_xor(d, 32);
    if (g == 0) dots[s][r] = d;
  }
  __syncthreads();
  if (part && t < RHS)
    part[(long long)blockIdx.x * RHS + t] = dots[0][t] + dots[1][t];
}

// Out = sa (A X) + sv V (fp32) for a row-major block A (m x k) on
// v_mfma_f32_16x16x4f32: a workgroup of 16 waves owns 16 output rows, the
// waves take the 16-column chunks round-robin and are added through LDS in
// a fixed order.  tri: A is lower (1) or upper (2) triangular.  With r64
// given, rzp[wg][c] = sum_rows r64 * Out and zzp[wg][c] = sum_rows Out^2
// (fp64).
constexpr int GV_WAVES = 16;
__global__ __launch_bounds__(64 * GV_WAVES) void k_gemv_rows(
    const float* __restrict__ A, long long lda, int m, int k, int tri,
    const float* __restrict__ X, float* __restrict__ Out, float sa,
    const float* __restrict__ V, float sv, const double* __restrict__ r64,
    double* __restrict__ rzp, double* __restrict__ zzp,
    const int* __restrict__ done) {
  if (done && *done) return;
  __shared__ floatx4 red[GV_WAVES][64];
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  const int r = lane & 15, g = lane >> 4;
  const int i0 = blockIdx.x * 16;
  int kb, ke;
  tri_range(tri, i0, k, kb, ke);
  floatx4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
  row_stripe_product(acc, A + (long long)(i0 + r) * lda + 4 * g,
                     X + (long long)(4 * g) * RHS + r, kb / 16 + w, ke / 16,
                     GV_WAVES);
  red[w][lane] = acc;
  __syncthreads();
  if (w != 0) return;
  double d = 0.0, zz = 0.0;
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    float o = acc[v];
#pragma unroll
    for (int u = 1; u < GV_WAVES; ++u) o += red[u][lane][v];
    const long long e = (long long)(i0 + 4 * g + v) * RHS + r;
    o *= sa;
    if (V) o = fmaf(sv, V[e], o);
    Out[e] = o;
    if (r64) {
      d += r64[e] * (double)o;
      zz += (double)o * (double)o;
    }
  }
  if (!r64) return;
  d += __shfl_xor(d, 16);
  d += __shfl_xor(d, 32);
  zz += __shfl_xor(zz, 16);
  zz += __shfl_xor(zz, 32);
  if (g == 0) {
    rzp[(long long)blockIdx.x * RHS + r] = d;
    if (zzp) zzp[(long long)blockIdx.x * RHS + r] = zz;
  }
}

// Iteration state of the refinement, in device memory, written only by
// the single-block kernels below (the vector kernels read it).
struct PcgState {
  double alpha[RHS];   // step of the current iteration
  double beta[RHS];    // direction update of the next iteration
  double rz[RHS];      // r . z of the current residual
  double bn[RHS];      // ||b|| per column
  double rel[RHS];     // ||r|| / ||b|| (recurrence or recomputed residual)
  double est[RHS];     // ||M^-1 r|| / ||x||: estimated relative error of x
  int done;            // 1: converged, every later kernel is a no-op
  int iters;           // iteration that set done (or the last checked)
  int restart;         // 1: the next direction is z (beta = 0)
  int skip;            // 1: skip the rest of a residual recomputation
  int notdone;         // 1 - done (gate of the speculative verification)
  int pad;
};

// Per-column sums of nslot partials [slot][RHS] by a 256-thread block:
// thread t adds slots t/16, t/16 + 16, ... of column t%16 (the loop
// unrolled by four so the loads are in flight together), then a fixed order
// over the 16 strands.  Result in out[RHS] (shared); ends with a barrier.
__device__ __forceinline__ void block_slot_sums(const double* __restrict__ s,
                                                int nslot,
                                                double* __restrict__ scratch,
                                                double* __restrict__ out) {
  const int t = threadIdx.x;
  const int c = t % RHS, strand = t / RHS;
  constexpr int ST = 256 / RHS;
  double v0 = 0.0, v1 = 0.0, v2 = 0.0, v3 = 0.0;
  int k = strand;
  for (; k + 3 * ST < nslot; k += 4 * ST) {
    v0 += s[(k) * RHS + c];
    v1 += s[(k + ST) * RHS + c];
    v2 += s[(k + 2 * ST) * RHS + c];
    v3 += s[(k + 3 * ST) * RHS + c];
  }
  for (; k < nslot; k += ST) v0 += s[k * RHS + c];
  scratch[t] = (v0 + v1) + (v2 + v3);
  __syncthreads();
  if (t < RHS) {
    double a = 0.0;
    for (int q = 0; q < ST; ++q) a += scratch[q * RHS + t];
    out[t] = a;
  }
  __syncthreads();
}

// Start of the iteration: ||b|| from the bb partials, r . z of the first
// preconditioned residual; the first direction is z.
__global__ __launch_bounds__(256) void k_pcg_init(
    const double* __restrict__ bb, int nbb, const double* __restrict__ rz,
    int nrz, PcgState* __restrict__ st) {
  __shared__ double scratch[256], sb[RHS], sr[RHS];
  block_slot_sums(bb, nbb, scratch, sb);
  block_slot_sums(rz, nrz, scratch, sr);
  const int t = threadIdx.x;
  if (t < RHS) {
    st->bn[t] = sqrt(sb[t]);
    st->rz[t] = sr[t];
    st->alpha[t] = 0.0;
    st->beta[t] = 0.0;
    st->rel[t] = 1.0;
    st->est[t] = INFINITY;
  }
  if (t == 0) {
    st->done = 0;
    st->notdone = 1;
    st->iters = 0;
    st->restart = 1;
    st->skip = 0;
  }
}

// P_new = Z + beta P_old (beta = 0 after a restart).
__global__ __launch_bounds__(256) void k_pcg_p(
    int n, const float* __restrict__ Z, const double* __restrict__ P_old,
    double* __restrict__ P_new, const PcgState* __restrict__ st) {
  if (st->done) return;
  const int t = threadIdx.x;
  const bool first = st->restart != 0;
  const double beta = first ? 0.0 : st->beta[t % RHS];
  const long long e0 = (long long)blockIdx.x * VROWS * RHS;
  for (int e = t; e < VROWS * RHS; e += 256) {
    const long long g = e0 + e;
    if (g >= (long long)n * RHS) break;
    P_new[g] = first ? (double)Z[g] : fma(beta, P_old[g], (double)Z[g]);
  }
}

// alpha = (r . z) / (p . A p) per column (every workgroup sums the
// product's npq partials pq itself, in a fixed order); X += alpha P;
// R -= alpha Q; R32 = R; partials per workgroup of rr = sum_rows R^2,
// rzx = sum_rows R . Z (Z: the previous preconditioned residual, for the
// Polak-Ribiere beta) and xx = sum_rows X^2.
__global__ __launch_bounds__(256) void k_pcg_xr(
    int n, PcgState* __restrict__ st, const double* __restrict__ pq,
    int npq, const double* __restrict__ P, const double* __restrict__ Q,
    const float* __restrict__ Z, double* __restrict__ X,
    double* __restrict__ R, float* __restrict__ R32, double* __restrict__ rr,
    double* __restrict__ rzx, double* __restrict__ xx) {
  if (st->done) return;
  __shared__ double red[256], den[RHS];
  block_slot_sums(pq, npq, red, den);
  const int t = threadIdx.x;
  const int c = t % RHS;
  const double alpha = den[c] != 0.0 ? st->rz[c] / den[c] : 0.0;
  if (blockIdx.x == 0 && t < RHS) st->alpha[t] = alpha;
  const long long e0 = (long long)blockIdx.x * VROWS * RHS;
  double s = 0.0, x = 0.0, q = 0.0;
  for (int e = t; e < VROWS * RHS; e += 256) {
    const long long g = e0 + e;
    if (g >= (long long)n * RHS) break;
    const double xv = fma(alpha, P[g], X[g]);
    X[g] = xv;
    const double rv = fma(-alpha, Q[g], R[g]);
    R[g] = rv;
    R32[g] = (float)rv;
    s += rv * rv;
    x += rv * (double)Z[g];
    q += xv * xv;
  }
  double* outs[3] = {rr, rzx, xx};
  const double vals[3] = {s, x, q};
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    __syncthreads();
    red[t] = vals[k];
    __syncthreads();
    if (t < RHS) {
      double v = 0.0;
      for (int m = t; m < 256; m += RHS) v += red[m];
      outs[k][(long long)blockIdx.x * RHS + t] = v;
    }
  }
}

// Convergence test and the next direction's beta, after the
// preconditioner has produced z+ = M^-1 r+ (rz partials r+ . z+, zz
// partials z+ . z+).  rr: partials of r+ . r+ (nrr slots); rzx: of
// r+ . z (nullptr after a recomputed residual, which restarts the
// direction); xx: of x . x.  Every column c < nrhs with b != 0 must have
// rel <= tol or est <= tol.
__global__ __launch_bounds__(256) void k_pcg_check(
    const double* __restrict__ rr, int nrr, const double* __restrict__ rzx,
    int nrzx, const double* __restrict__ xx, int nxx,
    const double* __restrict__ rz, const double* __restrict__ zz, int nrz,
    int nrhs, double tol, int it, PcgState* __restrict__ st) {
  if (rzx ? st->done : st->skip) return;
  __shared__ double scratch[256], srr[RHS], srx[RHS], sxx[RHS], srz[RHS],
      szz[RHS];
  __shared__ int ok[RHS];
  block_slot_sums(rr, nrr, scratch, srr);
  if (rzx) block_slot_sums(rzx, nrzx, scratch, srx);
  block_slot_sums(xx, nxx, scratch, sxx);
  block_slot_sums(rz, nrz, scratch, srz);
  block_slot_sums(zz, nrz, scratch, szz);
  const int t = threadIdx.x;
  if (t < RHS) {
    const double bn = st->bn[t];
    const double rel = bn > 0.0 ? sqrt(srr[t]) / bn : 0.0;
    const double est = sxx[t] > 0.0 ? sqrt(szz[t] / sxx[t])
                                    : (bn > 0.0 ? INFINITY : 0.0);
    st->rel[t] = rel;
    st->est[t] = est;
    ok[t] = (t >= nrhs || bn == 0.0 || rel <= tol || est <= tol) ? 1 : 0;
    const double rz_old = st->rz[t];
    st->beta[t] = (rzx && rz_old != 0.0) ? (srz[t] - srx[t]) / rz_old : 0.0;
    st->rz[t] = srz[t];
  }
  __syncthreads();
  if (t == 0) {
    int all = 1;
    for (int c = 0; c < RHS; ++c) all &= ok[c];
    st->done = all;
    st->notdone = !all;
    if (rzx) st->iters = it;     // a recomputation keeps the iteration count
    st->restart = rzx ? 0 : 1;
  }
}

// The test on a recomputed residual before its preconditioning: when every
// column's relative residual is <= tol the solve has converged (skip = 1,
// done = 1) and the preconditioner and k_pcg_check of the recomputation
// return at once; otherwise skip = 0, done = 0 and they run.  With a gate
// that is set (the recurrence has not converged: the speculative
// recomputation did not run), skip = 1 and nothing else changes.
__global__ __launch_bounds__(256) void k_pcg_rcheck(
    const double* __restrict__ rr, int nrr, int nrhs, double tol,
    const int* __restrict__ gate, PcgState* __restrict__ st) {
  if (gate && *gate) {
    if (threadIdx.x == 0) st->skip = 1;
    return;
  }
  __shared__ double scratch[256], srr[RHS];
  __shared__ int ok[RHS];
  block_slot_sums(rr, nrr, scratch, srr);
  const int t = threadIdx.x;
  double rel = 0.0;
  if (t < RHS) {
    const double bn = st->bn[t];
    rel = bn > 0.0 ? sqrt(srr[t]) / bn : 0.0;
    ok[t] = (t >= nrhs || bn == 0.0 || rel <= tol) ? 1 : 0;
  }
  __syncthreads();
  int all = 1;
  for (int c = 0; c < RHS; ++c) all &= ok[c];
  if (all && t < RHS) st->rel[t] = rel;
  if (t == 0) {
    st->skip = all;
    st->done = all;     // a failed test overrides the recurrence's result
    if (all) {
      st->notdone = 0;
      st->restart = 1;
    }
  }
}

// B, R <- rhs (nrhs columns, zero padded); R32 = B; X = P0 = 0;
// bb[wg][c] = sum_rows B^2.  With X_out given instead, copies X (nrhs
// columns) out.
__global__ __launch_bounds__(256) void k_pcg_io(
    int n, int nrhs, const double* __restrict__ rhs, double* __restrict__ B,
    double* __restrict__ R, float* __restrict__ R32, double* __restrict__ X,
    double* __restrict__ P0, double* __restrict__ bb,
    double* __restrict__ X_out) {
  __shared__ double red[256];
  const int t = threadIdx.x;
  const int c = t % RHS;
  const long long e0 = (long long)blockIdx.x * VROWS * RHS;
  double s = 0.0;
  for (int e = t; e < VROWS * RHS; e += 256) {
    const long long g = e0 + e;
    if (g >= (long long)n * RHS) break;
    const long long i = g / RHS;
    if (X_out) {
      if (c < nrhs) X_out[i * nrhs + c] = X[g];
      continue;
    }
    const double v = c < nrhs ? rhs[i * nrhs + c] : 0.0;
    B[g] = v;
    R[g] = v;
    R32[g] = (float)v;
    X[g] = 0.0;
    P0[g] = 0.0;
    s += v * v;
  }
  if (X_out) return;
  red[t] = s;
  __syncthreads();
  if (t < RHS) {
    double v = 0.0;
    for (int k = t; k < 256; k += RHS) v += red[k];
    bb[(long long)blockIdx.x * RHS + t] = v;
  }
}

// fp32 [n][nrhs] <-> padded [n][RHS]
__global__ __launch_bounds__(256) void k_pad32(int n, const float* __restrict__ src,
                                               int nrhs, float* __restrict__ dst,
                                               int to_padded) {
  const long long e = (long long)blockIdx.x * 256 + threadIdx.x;
  if (e >= (long long)n * RHS) return;
  const long long i = e / RHS;
  const int c = (int)(e % RHS);
  if (to_padded)
    dst[e] = c < nrhs ? src[i * nrhs + c] : 0.0f;
  else if (c < nrhs)
    dst[i * nrhs + c] = src[e];
}

// ------------------------------------------------------------ host side
struct Ws {
  double *b, *x, *r, *p0, *p1, *q, *pq, *rz, *zz, *rr, *bb, *rzx, *xx;
  PcgState* st;
  float *r32, *z32, *y1, *u2, *y2, *v1;
  size_t bytes;
};

static Ws carve(char* base, int n) {
  Ws w;
  size_t off = 0;
  auto take = [&](size_t bytes) {
    char* p = base ? base + off : nullptr;
    off += (bytes + 255) & ~(size_t)255;
    return p;
  };
  const size_t v64 = (size_t)n * RHS * sizeof(double);
  const size_t v32 = (size_t)n * RHS * sizeof(float);
  const size_t slots = (size_t)((n + 15) / 16) * RHS * sizeof(double);
  w.b = (double*)take(v64);
  w.x = (double*)take(v64);
  w.r = (double*)take(v64);
  w.p0 = (double*)take(v64);
  w.p1 = (double*)take(v64);
  w.q = (double*)take(v64);
  w.pq = (double*)take(slots);
  w.rz = (double*)take(slots);
  w.zz = (double*)take(slots);
  w.rr = (double*)take(slots);
  w.bb = (double*)take(slots);
  w.rzx = (double*)take(slots);
  w.xx = (double*)take(slots);
  w.st = (PcgState*)take(sizeof(PcgState));
  w.r32 = (float*)take(v32);
  w.z32 = (float*)take(v32);
  w.y1 = (float*)take(v32);
  w.u2 = (float*)take(v32);
  w.y2 = (float*)take(v32);
  w.v1 = (float*)take(v32);
  w.bytes = off;
  return w;
}

static void gemv(const float* A, long long lda, int m, int k, int tri,
                 const float* X, float* out, float sa, const float* V,
                 float sv, const double* r64, double* rzp, double* zzp,
                 const int* done, hipStream_t st) {
  hipLaunchKernelGGL(k_gemv_rows, dim3(m / 16), dim3(64 * GV_WAVES), 0, st,
                     A, lda, m, k, tri, X, out, sa, V, sv, r64, rzp, zzp,
                     done);
}

// z = (L L^T)^-1 r (fp32) from W (W11, L21, W22) and WT = W^T; with r64
// given, rz and zz partials (r . z and z . z, one slot per 16 rows, n / 16
// slots).  done: the refinement's convergence flag (nullptr: always run).
static void precond(const float* W, const float* WT, long long ld, int n,
                    int n1, const float* r32, float* z32, const double* r64,
                    double* rzp, double* zzp, const int* done, const Ws& w,
                    hipStream_t st) {
  const int n2 = n - n1;
  const long long o2 = (long long)n1 * RHS;
  if (n2 == 0) {
    gemv(W, ld, n, n, 1, r32, w.y1, 1.0f, nullptr, 0.0f, nullptr, nullptr,
         nullptr, done, st);
    gemv(WT, ld, n, n, 2, w.y1, z32, 1.0f, nullptr, 0.0f, r64, rzp, zzp,
         done, st);
    return;
  }
  const float* S = W + (long long)n1 * ld;            // L21 (n2 x n1)
  const float* W22 = S + n1;
  const float* ST = WT + n1;                           // L21^T (n1 x n2)
  const float* W22T = WT + (long long)n1 * ld + n1;
  const long long s2 = (long long)(n1 / 16) * RHS;     // slots of rows n1..
  gemv(W, ld, n1, n1, 1, r32, w.y1, 1.0f, nullptr, 0.0f, nullptr, nullptr,
       nullptr, done, st);                             // y1 = W11 r1
  gemv(S, ld, n2, n1, 0, w.y1, w.u2, -1.0f, r32 + o2, 1.0f, nullptr,
       nullptr, nullptr, done, st);                    // u2 = r2 - S y1
  gemv(W22, ld, n2, n2, 1, w.u2, w.y2, 1.0f, nullptr, 0.0f, nullptr,
       nullptr, nullptr, done, st);                    // y2 = W22 u2
  gemv(W22T, ld, n2, n2, 2, w.y2, z32 + o2, 1.0f, nullptr, 0.0f,
       r64 ? r64 + o2 : nullptr, rzp ? rzp + s2 : nullptr,
       zzp ? zzp + s2 : nullptr, done, st);            // z2 = W22^T y2
  gemv(ST, ld, n1, n2, 0, z32 + o2, w.v1, -1.0f, w.y1, 1.0f, nullptr,
       nullptr, nullptr, done, st);                    // v1 = y1 - S^T z2
  gemv(WT, ld, n1, n1, 2, w.v1, z32, 1.0f, nullptr, 0.0f, r64, rzp, zzp,
       done, st);                                      // z1 = W11^T v1
}

// The refinement loop.  Iterations are launched in batches without host
// synchronisation: every kernel of an iteration returns at once when the
// device-side state says the solve has converged, so a batch that passes
// convergence costs a few empty launches.  Each batch is followed by a
// recomputation of the residual as b - A x and its test (the recurrence
// drifts on ill-conditioned systems) that the device runs only if the
// recurrence converged, and by the solution's copy out; then the host reads
// the state (one synchronisation per batch).  Every second batch without
// convergence the residual is recomputed unconditionally (a checkpoint:
// drift, stall test), and the direction restarts.
// Stop: every column's relative residual ||r|| / ||b|| or estimated
// relative error ||M^-1 r|| / ||x|| <= tol (M^-1 r is the error x* - x to
// the accuracy of the FP32 factor), or no progress of the recomputed
// residual's test quantity over 5 checkpoints.
constexpr int kPcgBatch = 4;

template <typename TK>
static int pcg_run(const TK* K, int n, long long ldk, double lam,
                   const float* W, const float* WT, long long ldw, int n1,
                   const double* Bin, int nrhs, double* Xout, double tol,
                   int max_iter, char* work, int* iters, double* resid,
                   hipStream_t st) {
  const Ws w = carve(work, n);
  const int nvec = (n + VROWS - 1) / VROWS;    // vector-kernel workgroups
  const int nsym = n / 32;                     // k_symv_rows workgroups
  const int nrz = n / 16;                      // rz / zz slots
  const int* done = &w.st->done;
  PcgState hs;
  auto read_state = [&]() -> int {
    hipError_t e = hipMemcpyAsync(&hs, w.st, sizeof(PcgState),
                                  hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess)
      return set_error((int)e, "nta_pcg64: %s", hipGetErrorString(e));
    return 0;
  };
  // Largest value over the columns with b != 0; NaN if any is NaN (a
  // broken-down preconditioner must never read as converged).
  auto worst = [&](const double* v) {
    double m = 0.0;
    for (int c = 0; c < nrhs; ++c) {
      if (!(hs.bn[c] > 0.0)) continue;
      if (v[c] != v[c]) return v[c];
      m = fmax(m, v[c]);
    }
    return m;
  };
  // r <- b - A x (FP64), z <- M^-1 r, then the test on the recomputed
  // residual (restarts the direction).  Speculative: the device runs it only
  // if the recurrence has converged (gate notdone).
  // r <- b - A x and the test of its norm alone (verify), then, unless
  // that test passed, its preconditioning and the full test (finish).
  auto verify = [&](bool speculative) {
    const int* gate = speculative ? &w.st->notdone : nullptr;
    hipLaunchKernelGGL(k_symv_rows<TK>, dim3(nsym), dim3(kSvThreads), 0, st,
                       K, ldk, n, lam, 1, w.x, nullptr, w.b, w.r, w.r32,
                       w.rr, gate);
    hipLaunchKernelGGL(k_pcg_rcheck, dim3(1), dim3(256), 0, st, w.rr, nsym,
                       nrhs, tol, gate, w.st);
  };
  auto finish = [&]() {
    precond(W, WT, ldw, n, n1, w.r32, w.z32, w.r, w.rz, w.zz, &w.st->skip,
            w, st);
    hipLaunchKernelGGL(k_pcg_check, dim3(1), dim3(256), 0, st, w.rr, nsym,
                       nullptr, 0, w.xx, nvec, w.rz, w.zz, nrz, nrhs, tol,
                       0, w.st);
  };
  auto recompute = [&]() {
    verify(false);
    finish();
  };
  auto copy_out = [&]() {
    hipLaunchKernelGGL(k_pcg_io, dim3(nvec), dim3(256), 0, st, n, nrhs,
                       nullptr, nullptr, nullptr, nullptr, w.x, nullptr,
                       nullptr, Xout);
  };
  hipLaunchKernelGGL(k_pcg_io, dim3(nvec), dim3(256), 0, st, n, nrhs, Bin,
                     w.b, w.r, w.r32, w.x, w.p0, w.bb, nullptr);
  precond(W, WT, ldw, n, n1, w.r32, w.z32, w.r, w.rz, w.zz, nullptr, w, st);
  hipLaunchKernelGGL(k_pcg_init, dim3(1), dim3(256), 0, st, w.bb, nvec,
                     w.rz, nrz, w.st);
  double* pbuf[2] = {w.p0, w.p1};
  double best = INFINITY;
  int stall = 0, it = 0, batches = 0, rc = 0;
  bool converged = false, copied = false;
  while (it < max_iter && !converged) {
    const int nb = max_iter - it < kPcgBatch ? max_iter - it : kPcgBatch;
    for (int k = 0; k < nb; ++k) {
      ++it;
      const double* p_old = pbuf[(it + 1) & 1];
      double* p_new = pbuf[it & 1];
      hipLaunchKernelGGL(k_pcg_p, dim3(nvec), dim3(256), 0, st, n, w.z32,
                         p_old, p_new, w.st);
      hipLaunchKernelGGL(k_symv_rows<TK>, dim3(nsym), dim3(kSvThreads), 0,
                         st, K, ldk, n, lam, 0, p_new, w.q, nullptr, nullptr,
                         nullptr, w.pq, done);
      hipLaunchKernelGGL(k_pcg_xr, dim3(nvec), dim3(256), 0, st, n, w.st,
                         w.pq, nsym, p_new, w.q, w.z32, w.x, w.r, w.r32,
                         w.rr, w.rzx, w.xx);
      precond(W, WT, ldw, n, n1, w.r32, w.z32, w.r, w.rz, w.zz, done, w,
              st);
      hipLaunchKernelGGL(k_pcg_check, dim3(1), dim3(256), 0, st, w.rr, nvec,
                         w.rzx, nvec, w.xx, nvec, w.rz, w.zz, nrz, nrhs, tol,
                         it, w.st);
    }
    ++batches;
    // If the recurrence converged in this batch, the device confirms it on
    // the recomputed residual and copies the solution out before the one
    // synchronisation of the batch.
    verify(true);
    copy_out();
    rc = read_state();
    if (rc) return rc;
    if (hs.done) {                    // confirmed (iteration that converged)
      it = hs.iters;
      converged = copied = true;
      break;
    }
    if (!hs.skip) {
      // The recurrence converged but the recomputed residual alone does
      // not meet tol: its preconditioning and the full test (the copy-out
      // above stays valid: x is unchanged).
      finish();
      rc = read_state();
      if (rc) return rc;
      if (hs.done) {
        it = hs.iters;
        converged = copied = true;
        break;
      }
    }
    if (batches % 2 != 0) continue;
    // a periodic checkpoint on the recomputed residual (the recurrence
    // drifts on ill-conditioned systems), also the stall test
    recompute();
    rc = read_state();
    if (rc) return rc;
    if (hs.done) {
      it = hs.iters;
      converged = true;
      break;
    }
    const double wr = worst(hs.rel);
    if (wr != wr) break;              // NaN residual: give up at once
    const double q = fmin(wr, worst(hs.est));
    if (q < best) {
      best = q;
      stall = 0;
    } else if (++stall >= 5) {
      break;
    }
  }
  if (!converged) {
    // The returned x's recomputed residual (the test above may not have
    // run for the last iterations).
    recompute();
    rc = read_state();
    if (rc) return rc;
  }
  if (!copied) copy_out();
  if (iters) *iters = it < max_iter ? it : max_iter;
  if (resid) *resid = worst(hs.rel);
  return check_launch("nta_pcg64");
}

// A = K + shift I on the 64x64 tiles on and below the diagonal (the tiles
// above are not touched: the factorisation reads only the lower triangle).
// Grid (n / 64, n / 64); one tile per workgroup of 256 threads.
__global__ __launch_bounds__(256) void k_copy_lower_shift(
    const float* __restrict__ K, long long ldk, int n, float shift,
    float* __restrict__ A, long long lda) {
  const int bx = blockIdx.x * 64, by = blockIdx.y * 64;   // tile (by, bx)
  if (bx > by) return;
  const int t = threadIdx.x;
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int f = t + 256 * q, row = by + (f >> 4), c = bx + (f & 15) * 4;
    float4 v = *(const float4*)(K + (long long)row * ldk + c);
    if (bx == by) {
      const int d = row - c;                  // diagonal inside this float4
      if (d == 0) v.x += shift;
      if (d == 1) v.y += shift;
      if (d == 2) v.z += shift;
      if (d == 3) v.w += shift;
    }
    *(float4*)(A + (long long)row * lda + c) = v;
  }
}

}  // namespace nta

using namespace nta;

extern "C" int nta_chol_split(int n) {
  if (n <= 0 || n % 64 != 0) return 0;
  return chol_split_of(n);
}

extern "C" int nta_copy_lower_shift(const float* K, long long ldk, int n,
                                    float shift, float* A, long long lda,
                                    void* stream) {
  if (n <= 0 || ldk < n || lda < n || !K || !A)
    return set_error(NTA_EINVAL, "nta_copy_lower_shift: bad n=%d ldk=%lld "
                     "lda=%lld or null pointer", n, ldk, lda);
  if (n % 64 != 0 || ldk % 4 != 0 || lda % 4 != 0 ||
      ((uintptr_t)K & 15) != 0 || ((uintptr_t)A & 15) != 0)
    return set_error(NTA_EUNSUPPORTED, "nta_copy_lower_shift: n=%d, ldk and "
                     "lda must be multiples of 64, 4 and 4 and the matrices "
                     "16-byte aligned", n);
  const dim3 grid(n / 64, n / 64);
  hipLaunchKernelGGL(k_copy_lower_shift, grid, dim3(256), 0,
                     (hipStream_t)stream, K, ldk, n, shift, A, lda);
  return check_launch("nta_copy_lower_shift");
}

extern "C" int nta_chol_f32(float* A, int n, long long lda, float* W,
                            long long ldw, float* WT, int* info,
                            void* stream) {
  if (n <= 0 || lda < n || ldw < n || !A || !W || !WT || !info)
    return set_error(NTA_EINVAL, "nta_chol_f32: bad n=%d lda=%lld ldw=%lld "
                     "or null pointer", n, lda, ldw);
  if (n % 64 != 0 || lda % 4 != 0 || ldw % 4 != 0)
    return set_error(NTA_EUNSUPPORTED, "nta_chol_f32: n=%d, lda and ldw must "
                     "be multiples of 64, 4 and 4 (pad the matrix)", n);
  hipStream_t st = (hipStream_t)stream;
  CholStreams* cs = chol_streams();
  if (!cs) return NTA_EINVAL;
  const std::vector<int> off = chol_offsets(n);
  const int m = (int)off.size() - 1;
  const int mh = m / 2;
  // W11 = L11^-1, W22 = L22^-1; the block below them keeps L21.  The
  // products' scratch is A (no longer read once factored).  When the split
  // is a panel boundary, the top half is inverted on the third stream while
  // the bottom half is factored (its scratch rows [0, n1) of A are done
  // with by then), and the bottom half after the factorisation, its scratch
  // in rows [n1, n).
  const bool overlap = m > 1 && mh % (kOuter / 128) == 0;
  int rc = chol_factor(A, lda, W, ldw, WT, n, info, st, *cs,
                       overlap ? off[mh] : -1);
  if (rc) return rc;
  if (overlap) {
    chol_invert_half(W, ldw, off, mh, m, A + (long long)off[mh] * lda, lda,
                     st, m);
    hipError_t e = hipStreamWaitEvent(st, cs->ev[kEvInvEnd], 0);
    if (e != hipSuccess)
      return set_error((int)e, "nta_chol_f32: %s", hipGetErrorString(e));
  } else if (m > 1) {
    // equal powers of two of 128-blocks: each level's joins of both halves
    // run as one batch (no join of size < mh crosses the boundary)
    const bool pow2 = (mh & (mh - 1)) == 0;
    if (m == 2 * mh && n % 128 == 0 && pow2) {
      chol_invert_half(W, ldw, off, 0, m, A, lda, st, mh);
    } else {
      chol_invert_half(W, ldw, off, 0, mh, A, lda, st, m);
      chol_invert_half(W, ldw, off, mh, m, A, lda, st, m);
    }
  }
  // WT = W^T: with the overlap, columns [0, n1) were transposed on the
  // inversion stream; W22 remains
  const int t0 = overlap ? off[mh] : 0, nt = n - t0;
  const dim3 grid((nt + 31) / 32, (nt + 31) / 32);
  hipLaunchKernelGGL(k_transpose_lower, grid, dim3(256), 0, st,
                     W + (long long)t0 * ldw + t0, ldw, nt, nt,
                     WT + (long long)t0 * ldw + t0, ldw);
  return check_launch("nta_chol_f32");
}

// Y[row][c] (c < ncols, leading dimension ldy) = sum over the parts of the
// column-split products ([nparts][rows][16]), in a fixed order.
__global__ __launch_bounds__(256) void k_sum_parts(
    const double* __restrict__ parts, int nparts, long long rows,
    double* __restrict__ Y, long long ldy, int ncols) {
  const long long e = (long long)blockIdx.x * 256 + threadIdx.x;
  const long long count = rows * RHS;
  if (e >= count) return;
  const long long row = e / RHS;
  const int c = (int)(e - row * RHS);
  if (c >= ncols) return;
  double v = parts[e];
  for (int q = 1; q < nparts; ++q) v += parts[(long long)q * count + e];
  Y[row * ldy + c] = v;
}

// Xp[k][c] = X[k][c] (c < ncols, leading dimension ldx), zero for c >= ncols:
// the [n][16] operand layout of the row kernel.
__global__ __launch_bounds__(256) void k_pad_cols(
    const double* __restrict__ X, long long ldx, int ncols, long long n,
    double* __restrict__ Xp) {
  const long long e = (long long)blockIdx.x * 256 + threadIdx.x;
  if (e >= n * RHS) return;
  const long long k = e / RHS;
  const int c = (int)(e - k * RHS);
  Xp[e] = c < ncols ? X[k * ldx + c] : 0.0;
}

// Column parts of nta_gemm_f32_f64: enough row-block x part workgroups to
// cover the CUs twice (one part when the rows alone do).
static int gemm_f32_f64_parts(int m, int n) {
  int parts = (512 + m / 32 - 1) / (m / 32);
  parts = std::min(parts, std::max(1, n / 256));    // >= 16 chunks a part
  return std::max(1, std::min(parts, 16));
}

extern "C" long long nta_gemm_f32_f64_work_bytes(int m, int n) {
  if (m <= 0 || n <= 0 || m % 32 != 0 || n % 16 != 0) return -1;
  const int parts = gemm_f32_f64_parts(m, n);
  return ((long long)parts * m + n) * RHS * (long long)sizeof(double);
}

extern "C" int nta_gemm_f32_f64(const float* K, long long ldk, int m, int n,
                                const double* X, long long ldx, int ncols,
                                double* Y, long long ldy, void* work,
                                long long work_bytes, void* stream) {
  if (m <= 0 || n <= 0 || ldk < n || !K || !X || !Y || ncols < 1 ||
      ncols > RHS || ldx < ncols || ldy < ncols)
    return set_error(NTA_EINVAL, "nta_gemm_f32_f64: bad m=%d n=%d ldk=%lld "
                     "ncols=%d ldx=%lld ldy=%lld or null pointer", m, n,
                     ldk, ncols, ldx, ldy);
  if (m % 32 != 0 || n % 16 != 0 || ldk % 4 != 0 ||
      ((uintptr_t)K & 15) != 0)
    return set_error(NTA_EUNSUPPORTED, "nta_gemm_f32_f64: m=%d, n=%d and "
                     "ldk=%lld must be multiples of 32, 16 and 4 and K "
                     "16-byte aligned (pad the matrix)", m, n, ldk);
  hipStream_t st = (hipStream_t)stream;
  const bool dense = ncols == RHS && ldx == RHS && ldy == RHS;
  const long long need = nta_gemm_f32_f64_work_bytes(m, n);
  if (!work || work_bytes < need) {
    if (!dense)
      return set_error(NTA_EINVAL, "nta_gemm_f32_f64: ncols=%d, ldx=%lld, "
                       "ldy=%lld need a workspace of %lld bytes", ncols,
                       ldx, ldy, need);
    hipLaunchKernelGGL(k_symv_rows<float>, dim3(m / 32), dim3(kSvThreads), 0,
                       st, K, ldk, n, 0.0, 2, X, Y, nullptr, nullptr,
                       nullptr, nullptr, nullptr);
    return check_launch("nta_gemm_f32_f64");
  }
  const int parts = gemm_f32_f64_parts(m, n);
  double* pw = (double*)work;
  const double* Xp = X;
  if (!(ncols == RHS && ldx == RHS)) {
    double* xw = pw + (long long)parts * m * RHS;
    const long long cx = (long long)n * RHS;
    hipLaunchKernelGGL(k_pad_cols, dim3((unsigned)((cx + 255) / 256)),
                       dim3(256), 0, st, X, ldx, ncols, (long long)n, xw);
    Xp = xw;
  }
  hipLaunchKernelGGL(k_symv_rows<float>, dim3(m / 32, parts),
                     dim3(kSvThreads), 0, st, K, ldk, n, 0.0, 2, Xp, pw,
                     nullptr, nullptr, nullptr, nullptr, nullptr);
  const long long count = (long long)m * RHS;
  hipLaunchKernelGGL(k_sum_parts, dim3((unsigned)((count + 255) / 256)),
                     dim3(256), 0, st, pw, parts, (long long)m, Y, ldy,
                     ncols);
  return check_launch("nta_gemm_f32_f64");
}

extern "C" long long nta_pcg64_work_bytes(int n, int nrhs) {
  if (n <= 0 || nrhs <= 0 || nrhs > RHS) return -1;
  return (long long)carve(nullptr, n).bytes;
}

static int check_solve_args(const char* who, int n, long long ldw, int n1,
                            int nrhs, long long work_bytes) {
  if (n <= 0 || n % 64 != 0 || ldw < n || ldw % 4 != 0 || nrhs <= 0 ||
      nrhs > RHS)
    return set_error(NTA_EINVAL, "%s: bad n=%d ldw=%lld nrhs=%d", who, n, ldw,
                     nrhs);
  if (n1 != nta_chol_split(n))
    return set_error(NTA_EINVAL, "%s: n1=%d is not the split %d of "
                     "nta_chol_f32", who, n1, nta_chol_split(n));
  if (work_bytes < nta_pcg64_work_bytes(n, nrhs))
    return set_error(NTA_EINVAL, "%s: work_bytes=%lld < %lld", who,
                     work_bytes, nta_pcg64_work_bytes(n, nrhs));
  return 0;
}

extern "C" int nta_chol_precond(const float* W, const float* WT, int n,
                                long long ldw, int n1, const float* R,
                                int nrhs, float* Z, void* work,
                                long long work_bytes, void* stream) {
  int rc = check_solve_args("nta_chol_precond", n, ldw, n1, nrhs,
                            work_bytes);
  if (rc) return rc;
  if (!W || !WT || !R || !Z || !work)
    return set_error(NTA_EINVAL, "nta_chol_precond: null pointer");
  hipStream_t st = (hipStream_t)stream;
  const Ws w = carve((char*)work, n);
  const unsigned g = (unsigned)(((long long)n * RHS + 255) / 256);
  hipLaunchKernelGGL(k_pad32, dim3(g), dim3(256), 0, st, n, R, nrhs, w.r32,
                     1);
  precond(W, WT, ldw, n, n1, w.r32, w.z32, nullptr, nullptr, nullptr,
          nullptr, w, st);
  hipLaunchKernelGGL(k_pad32, dim3(g), dim3(256), 0, st, n, w.z32, nrhs, Z,
                     0);
  return check_launch("nta_chol_precond");
}

extern "C" int nta_pcg64(const float* K, int n, long long ldk, double lam,
                         const float* W, const float* WT, long long ldw,
                         int n1, const double* B, int nrhs, double* X,
                         double tol, int max_iter, void* work,
                         long long work_bytes, int* iters, double* resid,
                         void* stream) {
  int rc = check_solve_args("nta_pcg64", n, ldw, n1, nrhs, work_bytes);
  if (rc) return rc;
  if (ldk < n || ldk % 4 != 0 || max_iter < 1 || !K || !W || !WT || !B ||
      !X || !work)
    return set_error(NTA_EINVAL, "nta_pcg64: bad ldk=%lld max_iter=%d or "
                     "null pointer", ldk, max_iter);
  return pcg_run<float>(K, n, ldk, lam, W, WT, ldw, n1, B, nrhs, X, tol,
                        max_iter, (char*)work, iters, resid,
                        (hipStream_t)stream);
}

extern "C" int nta_pcg64_k64(const double* K, int n, long long ldk,
                             double lam, const float* W, const float* WT,
                             long long ldw, int n1, const double* B,
                             int nrhs, double* X, double tol, int max_iter,
                             void* work, long long work_bytes, int* iters,
                             double* resid, void* stream) {
  int rc = check_solve_args("nta_pcg64_k64", n, ldw, n1, nrhs, work_bytes);
  if (rc) return rc;
  if (ldk < n || ldk % 2 != 0 || max_iter < 1 || !K || !W || !WT || !B ||
      !X || !work)
    return set_error(NTA_EINVAL, "nta_pcg64_k64: bad ldk=%lld max_iter=%d "
                     "or null pointer", ldk, max_iter);
  return pcg_run<double>(K, n, ldk, lam, W, WT, ldw, n1, B, nrhs, X, tol,
                         max_iter, (char*)work, iters, resid,
                         (hipStream_t)stream);
}

// ---------------------------------------------- allocation-free refinement
// nta_pcg64 with its workspace in a file-scope device array ("scratch"), for
// callers that cannot allocate.  One copy per GPU; its device address is
// taken with hipGetSymbolAddress once per device (the array's name is only a
// host-side stand-in) and kept in g_scratch.  Calls on one GPU share it, so
// overlapping calls on one GPU race; every call initialises the parts it
// reads (k_pcg_io, the first preconditioner application), so nothing of a
// previous call is seen.
namespace nta {

constexpr int kScratchMaxN = 16384;
constexpr size_t kScratchBytes =
    6 * (size_t)kScratchMaxN * RHS * sizeof(double) +         // v64
    7 * (size_t)(kScratchMaxN / 16) * RHS * sizeof(double) +  // slots
    (sizeof(PcgState) + 255) / 256 * 256 +                    // state
    6 * (size_t)kScratchMaxN * RHS * sizeof(float);           // v32

__device__ __align__(256) char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static void* g_scratch[kMaxDevices];

// Device address of this device's scratch (set up on the first call on the
// current device), or nullptr after setting the error.
static void* scratch_address() {
  int dev = -1;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) {
    set_error((int)e, "nta_pcg64_noalloc: hipGetDevice: %s",
              hipGetErrorString(e));
    return nullptr;
  }
  if (dev < 0 || dev >= kMaxDevices) {
    set_error(NTA_EINVAL, "nta_pcg64_noalloc: device %d out of range", dev);
    return nullptr;
  }
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch[dev]) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      set_error(e != hipSuccess ? (int)e : NTA_EINVAL,
                "nta_pcg64_noalloc: hipGetSymbolAddress(scratch): %s",
                hipGetErrorString(e));
      return nullptr;
    }
    g_scratch[dev] = p;
  }
  return g_scratch[dev];
}

}  // namespace nta

extern "C" int nta_pcg64_noalloc(const float* K, int n, long long ldk,
                                 double lam, const float* W, const float* WT,
                                 long long ldw, int n1, const double* B,
                                 int nrhs, double* X, double tol,
                                 int max_iter, int* iters, double* resid,
                                 void* stream) {
  if (n > kScratchMaxN)
    return set_error(NTA_EINVAL, "nta_pcg64_noalloc: n=%d exceeds the "
                     "static workspace's limit n <= %d (use nta_pcg64)", n,
                     kScratchMaxN);
  const long long need = nta_pcg64_work_bytes(n, nrhs);
  if (need > (long long)kScratchBytes)
    return set_error(NTA_EINVAL, "nta_pcg64_noalloc: n=%d nrhs=%d needs "
                     "%lld bytes of workspace, the static one holds %lld", n,
                     nrhs, need, (long long)kScratchBytes);
  void* ws = scratch_address();
  if (!ws) return NTA_EINVAL;
  return nta_pcg64(K, n, ldk, lam, W, WT, ldw, n1, B, nrhs, X, tol, max_iter,
                   ws, (long long)kScratchBytes, iters, resid, stream);
}
