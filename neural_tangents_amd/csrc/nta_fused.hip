// Shape-specialised kernels for the Myrtle path (3x3 SAME Conv + ReLU
// [+ 2x2 AvgPool], square images of side 8/16/32).
//
// k_stage_fast: one fused Conv -> ReLU [-> AvgPool] stage on a tile of pairs.
//   A workgroup owns one diagonal (h - h' = const, in output planes) of JPW
//   pairs and sweeps it.  The 3x3 stencil is separable into a 3-tap sum along
//   the (w, w') diagonal of each (h, h') plane (done once per loaded plane
//   through an LDS exchange) and a 3-tap sum along the (h, h') diagonal (a
//   register sliding window over consecutive planes of the sweep), so every
//   input plane is read from HBM once per diagonal that needs it.  Each thread
//   owns a 2x2 (w, w') block, which makes the 2x2x2x2 pool register-local.
//
// k_first_fast: input covariance + first Conv + ReLU.  The stencil of the
//   input Gram is the Gram of zero-padded 3x3xC patches (K = 9 * 4 with C
//   padded to 4), computed per (h, h') plane with fp32 MFMA
//   (32x32x2 for S = 32, 16x16x4 for S = 16); the ReLU is the epilogue.
//
// Both remap blockIdx so that consecutive logical blocks (the diagonals or
// rows of one pair) land on the same XCD: dispatch is round-robin over the 8
// XCDs, each with its own L2.
#include "nta_common.h"

namespace nta {

typedef float floatx4 __attribute__((ext_vector_type(4)));
typedef float floatx16 __attribute__((ext_vector_type(16)));
typedef _Float16 half4v __attribute__((ext_vector_type(4)));
typedef _Float16 half8v __attribute__((ext_vector_type(8)));

__device__ __forceinline__ float2 qnorm(float q) {
  float sq = fsqrt(fmaxf(q, 0.0f));
  float rq = q > 0.0f ? 1.0f / sq : 0.0f;
  return make_float2(sq, rq);
}

__device__ __forceinline__ int xcd_remap(int b, int nblk) {
  return (b & 7) * (nblk >> 3) + (b >> 3);
}

// Buffer resource over [base, base + bytes): loads at offsets >= bytes
// return 0, which zero-fills planes outside the grid without branches.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t make_rsrc(const float* base,
                                                            unsigned bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<float*>(base), 0,
                                           (int)bytes, 0x00020000);
}

__device__ __forceinline__ float2 buf_load2(__amdgpu_buffer_rsrc_t r,
                                            unsigned voff, unsigned soff) {
  return __builtin_bit_cast(
      float2, __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0));
}

constexpr unsigned kOutside = 0x40000000u;   // byte offset past any state

// State written by a stage is read back only by the next launch, after
// gigabytes of other traffic: streaming (non-temporal) stores keep it from
// displacing the planes the current sweep re-reads from L2.
#ifndef NTA_NT_STORE
#define NTA_NT_STORE 0
#endif
__device__ __forceinline__ void st_state(float* p, float v) {
#if NTA_NT_STORE
  __builtin_nontemporal_store(v, p);
#else
  *p = v;
#endif
}

// ReLU contribution of one pre-activation element, accumulated for the
// pool: acc_n += sq1 sq2 (sqrt(1 - rho^2) + (pi - theta) rho) and
// acc_t += (pi - theta) t; both sums are scaled by 1 / (2 pi) by the caller.
template <bool NTK>
__device__ __forceinline__ void relu_acc(float cv, float sqp, float rqp,
                                         float t, float& acc_n,
                                         float& acc_t) {
  const float rho = cv * rqp;
  const float a = fminf(fabsf(rho), 1.0f);
  const float u = 1.0f - a;
  const float th = acos_pos5(a, u);
  const float pm = rho < 0.0f ? th : kPi - th;
  const float sn = fsqrt(fmaf(-a, a, 1.0f));
  acc_n = fmaf(fmaf(pm, rho, sn), sqp, acc_n);
  if constexpr (NTK) acc_t = fmaf(pm, t, acc_t);
}

// Issues the buffer loads of K planes of one tensor: plane k is
// (hh[k], hp[k]); planes outside [0, S)^2 (and every plane of a tail pair,
// outside the resource) read as zero, and so does every plane when live is
// false (the loads are issued but touch no memory).  Each thread loads its
// 2x2 (w, w')
// block (2W.., 2Wp..): r0 = row 2W, r1 = row 2W + 1.
template <int S, int K>
__device__ __forceinline__ void issue_planes(__amdgpu_buffer_rsrc_t rs,
                                             const int* hh, const int* hp,
                                             unsigned tofs, float2 (&r0)[K],
                                             float2 (&r1)[K],
                                             bool live = true) {
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const bool ok = live && (unsigned)hh[k] < (unsigned)S &&
                    (unsigned)hp[k] < (unsigned)S;
    const unsigned soff = ok ? (unsigned)((hh[k] * S + hp[k]) * S * S) * 4u
                             : kOutside;
    r0[k] = buf_load2(rs, tofs, soff);
    r1[k] = buf_load2(rs, tofs + S * 4, soff);
  }
}

// (w, w')-diagonal 3-tap sum of K loaded planes through an LDS exchange:
// ws[k][e], e = 2c + d, is element (2W + c, 2Wp + d) of
// X[w-1, w'-1] + X[w, w'] + X[w+1, w'+1].  Two barriers; every thread of the
// block calls it the same number of times.
template <int S, int K>
__device__ __forceinline__ void exchange(const float2 (&r0)[K],
                                         const float2 (&r1)[K],
                                         float* __restrict__ lds, int lofs,
                                         float (&ws)[K][4]) {
  constexpr int LS = S + 2;
  constexpr int PL = LS * LS;
#pragma unroll
  for (int k = 0; k < K; ++k) {
    float* d = lds + k * PL + lofs;
    d[0] = r0[k].x;
    d[1] = r0[k].y;
    d[LS] = r1[k].x;
    d[LS + 1] = r1[k].y;
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const float* s0 = lds + k * PL + lofs;
    const float v00 = r0[k].x, v01 = r0[k].y, v10 = r1[k].x, v11 = r1[k].y;
    const float a_mm = s0[-LS - 1];       // (2W-1, 2Wp-1)
    const float a_m0 = s0[-LS];           // (2W-1, 2Wp)
    const float a_0m = s0[-1];            // (2W,   2Wp-1)
    const float a_12 = s0[LS + 2];        // (2W+1, 2Wp+2)
    const float a_21 = s0[2 * LS + 1];    // (2W+2, 2Wp+1)
    const float a_22 = s0[2 * LS + 2];    // (2W+2, 2Wp+2)
    ws[k][0] = a_mm + v00 + v11;
    ws[k][1] = a_m0 + v01 + a_12;
    ws[k][2] = a_0m + v10 + a_21;
    ws[k][3] = v00 + v11 + a_22;
  }
  __syncthreads();
}

template <int S, int K>
__device__ __forceinline__ void wpass(__amdgpu_buffer_rsrc_t rs,
                                      const int* hh, const int* hp,
                                      float* __restrict__ lds, unsigned tofs,
                                      int lofs, float (&ws)[K][4]) {
  float2 r0[K], r1[K];
  issue_planes<S, K>(rs, hh, hp, tofs, r0, r1);
  exchange<S, K>(r0, r1, lds, lofs, ws);
}

// (w, w')-diagonal 3-tap sums of one LDS plane written by the thread block
// (zero border): ws[e], e = 2c + d, as in exchange(), with the thread's own
// elements read back from LDS (paired 8-byte reads).
template <int S, int LS = S + 2>
__device__ __forceinline__ void read_diag3(const float* __restrict__ plane,
                                           int lofs, float (&ws)[4]) {
  const float* s = plane + lofs;
  const float2 ra = *reinterpret_cast<const float2*>(s - LS - 1);
  const float2 rb = *reinterpret_cast<const float2*>(s - 1);
  const float v01 = s[1];
  const float v10 = s[LS];
  const float2 re = *reinterpret_cast<const float2*>(s + LS + 1);
  const float2 rf = *reinterpret_cast<const float2*>(s + 2 * LS + 1);
  ws[0] = ra.x + rb.y + re.x;    // (2W-1, 2Wp-1) + (2W, 2Wp) + (2W+1, 2Wp+1)
  ws[1] = ra.y + v01 + re.y;     // (2W-1, 2Wp) + (2W, 2Wp+1) + (2W+1, 2Wp+2)
  ws[2] = rb.x + v10 + rf.x;     // (2W, 2Wp-1) + (2W+1, 2Wp) + (2W+2, 2Wp+1)
  ws[3] = rb.y + re.x + rf.y;    // (2W, 2Wp) + (2W+1, 2Wp+1) + (2W+2, 2Wp+2)
}

// Slots of the 16x16 pooled stage: the 15 pooled diagonals D = -7..7 (8 - |D|
// steps each, 64 in all) in four groups of 16 steps, one workgroup each.
__device__ __forceinline__ int stage16_nsweep(int slot) {
  return slot == 0 ? 3 : 4;
}
__device__ __forceinline__ int stage16_diag(int slot, int k) {
  // 0: {0, 4, -4}; 1: {1, -1, 7, -7}; 2: {2, 3, 5, 6}; 3: {-2, -3, -5, -6}
  if (slot == 0) return k == 0 ? 0 : (k == 1 ? 4 : -4);
  if (slot == 1) return k == 0 ? 1 : (k == 1 ? -1 : (k == 2 ? 7 : -7));
  const int v = k == 0 ? 2 : (k == 1 ? 3 : (k == 2 ? 5 : 6));
  return slot == 2 ? v : -v;
}

// PF selects how the planes of a step are fetched (POOL == 2):
//   0: issued at the step that uses them;
//   1: prefetched into registers one step ahead, after the step's exchange;
//   2: as 1, but issued right after the step's planes are written to LDS
//      (both tensors share one exchange: two barriers per step);
//   3: as 2 with two steps in flight (two register sets, even/odd steps).
// KP (POOL == 2): the stencil means of the inputs at every pre-pool
// position, (1/9) sum of the nine taps -- the pre-activations before the
// w_var, b_var affine map -- also go to pre_n, pre_t ([T][S][S][S][S]),
// for the adjoint of the stage.
template <int S, int POOL, bool NTK, int PF, int SL, bool KP = false>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void k_stage_fast(
    const float* __restrict__ in_n, const float* __restrict__ in_t, int n1,
    int n2, int ndiag, int ngroups, int nblk, float wv9, float b_var,
    const float* __restrict__ q1, const float* __restrict__ q2,
    float* __restrict__ out_n, float* __restrict__ out_t,
    float* __restrict__ pre_n, float* __restrict__ pre_t) {
  static_assert(!KP || POOL == 2, "pre-activations of pooled stages only");
  constexpr int H2 = S / 2;
  constexpr int TPP = H2 * H2;            // threads per plane
  constexpr int JPW = 256 / TPP;          // pairs per workgroup
  constexpr int LS = S + 2;
  constexpr int PL = LS * LS;
  constexpr int NT = NTK ? 2 : 1;
  constexpr int NB = POOL == 2 ? (PF >= 2 ? 6 * NT : 6) : 1;  // LDS planes
  constexpr int SO = S / POOL;
  constexpr long long PER_IN = (long long)S * S * S * S;
  constexpr long long PER_OUT = (long long)SO * SO * SO * SO;
  __shared__ __align__(16) float lds[JPW * NB * PL];
  __shared__ float2 lq[JPW * 2 * S * S];

  const int L = xcd_remap(blockIdx.x, nblk);
  if (L >= ndiag * ngroups) return;       // whole block exits together
  const int g = L / ndiag;
  const int di = L - g * ndiag;
  const int job = threadIdx.x / TPP;
  const int t = threadIdx.x - job * TPP;
  const int W = t / H2;
  const int Wp = t - W * H2;
  const bool self = n2 == 0;
  const long long T = self ? (long long)n1 : (long long)n1 * n2;
  const long long p = (long long)g * JPW + job;
  const bool valid = p < T;
  const long long pc = valid ? p : 0;
  const long long i = self ? pc : pc / n2;
  const long long j = self ? pc : pc - i * n2;

  for (int k = threadIdx.x; k < JPW * NB * PL; k += 256) lds[k] = 0.0f;
  float2* lq1 = lq + job * 2 * S * S;
  float2* lq2 = lq1 + S * S;
  {
    const float* qa = q1 + i * (S * S);
    const float* qb = (self ? q1 : q2) + j * (S * S);
    for (int k = t; k < S * S; k += TPP) {
      lq1[k] = qnorm(valid ? qa[k] : 0.0f);
      lq2[k] = qnorm(valid ? qb[k] : 0.0f);
    }
  }
  __syncthreads();
  float* L0 = lds + job * NB * PL;
  // One resource per workgroup over its JPW pairs (uniform, so the loads
  // take it from SGPRs); pairs past the tile end fall outside its size.
  const long long p0 = (long long)g * JPW;
  const long long left = T - p0;
  const unsigned nbytes =
      (unsigned)((left < JPW ? left : JPW) * PER_IN * 4);
  const __amdgpu_buffer_rsrc_t rn = make_rsrc(in_n + p0 * PER_IN, nbytes);
  const __amdgpu_buffer_rsrc_t rt =
      make_rsrc((NTK ? in_t : in_n) + p0 * PER_IN, nbytes);
  const unsigned tofs =
      (unsigned)(job * PER_IN + (2 * W) * S + 2 * Wp) * 4u;
  const int lofs = (2 * W + 1) * LS + 2 * Wp + 1;
  const float4* lq1v = reinterpret_cast<const float4*>(lq1);
  const float4* lq2v = reinterpret_cast<const float4*>(lq2);

  if constexpr (POOL == 2) {
    // SL > 0: the workgroup sweeps the diagonals of slot di (shared
    // normaliser tables, fewer workgroups); else diagonal D = di - (H2 - 1).
    const int nsw = SL ? stage16_nsweep(di) : 1;
    for (int sw = 0; sw < nsw; ++sw) {
    const int D = SL ? stage16_diag(di, sw) : di - (H2 - 1);
    const int Hs = D > 0 ? D : 0;
    const int He = D < 0 ? H2 - 1 + D : H2 - 1;
    const int d0 = 2 * D, dp = 2 * D + 1, dm = 2 * D - 1;
    float c0a[NT][4], c0b[NT][4], cp[NT][4], cm[NT][4];
    {
      const int h0 = 2 * Hs;
      const int hh[4] = {h0 - 1, h0, h0, h0 - 1};
      const int hp[4] = {h0 - 1 - d0, h0 - d0, h0 - dp, h0 - 1 - dm};
#pragma unroll
      for (int tt = 0; tt < NT; ++tt) {
        float ws[4][4];
        wpass<S, 4>(tt ? rt : rn, hh, hp, L0, tofs, lofs, ws);
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          c0a[tt][e] = ws[0][e];
          c0b[tt][e] = ws[1][e];
          cp[tt][e] = ws[2][e];
          cm[tt][e] = ws[3][e];
        }
      }
    }
    // Planes of step H: d0 @ 2H+1, 2H+2; dp @ 2H+1, 2H+2; dm @ 2H, 2H+1.
    auto step_planes = [&](int H, int (&hh)[6], int (&hp)[6]) {
      hh[0] = 2 * H + 1; hp[0] = 2 * H + 1 - d0;
      hh[1] = 2 * H + 2; hp[1] = 2 * H + 2 - d0;
      hh[2] = 2 * H + 1; hp[2] = 2 * H + 1 - dp;
      hh[3] = 2 * H + 2; hp[3] = 2 * H + 2 - dp;
      hh[4] = 2 * H;     hp[4] = 2 * H - dm;
      hh[5] = 2 * H + 1; hp[5] = 2 * H + 1 - dm;
    };
    // One step: emits pooled plane (H, H - D).  With PF >= 2 the planes of
    // the step arrive in (r0, r1) and the loads of step Hn (if any) are
    // issued into the same registers once they are in LDS.
    auto step = [&](int H, float2 (&r0)[NT][6], float2 (&r1)[NT][6],
                    int Hn) {
      const int Hp = H - D;
      if constexpr (PF >= 2) {
        __syncthreads();             // the previous step's reads are done
#pragma unroll
        for (int tt = 0; tt < NT; ++tt)
#pragma unroll
          for (int k = 0; k < 6; ++k) {
            float* d = L0 + (tt * 6 + k) * PL + lofs;
            d[0] = r0[tt][k].x;
            d[1] = r0[tt][k].y;
            d[LS] = r1[tt][k].x;
            d[LS + 1] = r1[tt][k].y;
          }
        if (PF == 3 || Hn <= He) {
          // PF 3 issues past the sweep end too (as empty loads) so that the
          // count of loads in flight is the same on every path.
          int hh[6], hp[6];
          step_planes(Hn, hh, hp);
#pragma unroll
          for (int tt = 0; tt < NT; ++tt)
            issue_planes<S, 6>(tt ? rt : rn, hh, hp, tofs, r0[tt], r1[tt],
                               Hn <= He);
        }
        __syncthreads();
      }
      float pre[NT][4][4];   // [tensor][P00, P11, P10, P01][element]
#pragma unroll
      for (int tt = 0; tt < NT; ++tt) {
        float ws[6][4];
        if constexpr (PF >= 2) {
#pragma unroll
          for (int k = 0; k < 6; ++k)
            read_diag3<S>(L0 + (tt * 6 + k) * PL, lofs, ws[k]);
        } else if constexpr (PF == 1) {
          exchange<S, 6>(r0[tt], r1[tt], L0, lofs, ws);
        } else {
          int hh[6], hp[6];
          step_planes(H, hh, hp);
          wpass<S, 6>(tt ? rt : rn, hh, hp, L0, tofs, lofs, ws);
        }
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          const float s0 = c0b[tt][e] + ws[0][e];
          pre[tt][0][e] = c0a[tt][e] + s0;
          pre[tt][1][e] = s0 + ws[1][e];
          pre[tt][2][e] = cp[tt][e] + ws[2][e] + ws[3][e];
          pre[tt][3][e] = cm[tt][e] + ws[4][e] + ws[5][e];
          c0a[tt][e] = ws[0][e];
          c0b[tt][e] = ws[1][e];
          cp[tt][e] = ws[3][e];
          cm[tt][e] = ws[5][e];
        }
      }
      if constexpr (PF == 1) {
        if (Hn <= He) {
          int hh[6], hp[6];
          step_planes(Hn, hh, hp);
#pragma unroll
          for (int tt = 0; tt < NT; ++tt)
            issue_planes<S, 6>(tt ? rt : rn, hh, hp, tofs, r0[tt], r1[tt]);
        }
      }
      if constexpr (KP) {
        // plane pl at rows (hr, hc) of (i, j); element e at w = 2W + e / 2,
        // w' = 2Wp + e % 2
        if (valid) {
          const int hr[4] = {2 * H, 2 * H + 1, 2 * H + 1, 2 * H};
          const int hc[4] = {2 * Hp, 2 * Hp + 1, 2 * Hp, 2 * Hp + 1};
          constexpr float kNinth = 1.0f / 9.0f;
#pragma unroll
          for (int tt = 0; tt < NT; ++tt) {
            float* dst = (tt ? pre_t : pre_n) + p * PER_IN;
#pragma unroll
            for (int pl = 0; pl < 4; ++pl) {
              float* row = dst + ((hr[pl] * S + hc[pl]) * S + 2 * W) * S +
                           2 * Wp;
              *reinterpret_cast<float2*>(row) = make_float2(
                  pre[tt][pl][0] * kNinth, pre[tt][pl][1] * kNinth);
              *reinterpret_cast<float2*>(row + S) = make_float2(
                  pre[tt][pl][2] * kNinth, pre[tt][pl][3] * kNinth);
            }
          }
        }
      }
      // (sq, rq) at rows 2H, 2H+1 of image i and 2Hp, 2Hp+1 of image j,
      // columns (2W, 2W+1) and (2Wp, 2Wp+1).
      const float4 qa0 = lq1v[((2 * H) * S + 2 * W) >> 1];
      const float4 qa1 = lq1v[((2 * H + 1) * S + 2 * W) >> 1];
      const float4 qb0 = lq2v[((2 * Hp) * S + 2 * Wp) >> 1];
      const float4 qb1 = lq2v[((2 * Hp + 1) * S + 2 * Wp) >> 1];
      // plane -> (row of i, row of j): P00 (0,0) P11 (1,1) P10 (1,0) P01 (0,1)
      const float4 qa_of[4] = {qa0, qa1, qa1, qa0};
      const float4 qb_of[4] = {qb0, qb1, qb0, qb1};
      float acc_n = 0.0f, acc_t = 0.0f;
#pragma unroll
      for (int pl = 0; pl < 4; ++pl) {
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          const float4 A = qa_of[pl], B = qb_of[pl];
          const float sa = (e >> 1) ? A.z : A.x, ra = (e >> 1) ? A.w : A.y;
          const float sb = (e & 1) ? B.z : B.x, rb = (e & 1) ? B.w : B.y;
          const float cv = fmaf(wv9, pre[0][pl][e], b_var);
          const float tv = NTK ? fmaf(wv9, pre[NT - 1][pl][e], cv) : 0.0f;
          relu_acc<NTK>(cv, sa * sb, ra * rb, tv, acc_n, acc_t);
        }
      }
      if (valid) {
        const long long o =
            p * PER_OUT + ((long long)(H * SO + Hp) * SO + W) * SO + Wp;
        constexpr float kScale = kInv2Pi / 16.0f;
        st_state(out_n + o, acc_n * kScale);
        if constexpr (NTK) st_state(out_t + o, acc_t * kScale);
      }
    };
    float2 na0[NT][6], na1[NT][6];     // planes in flight (register set A)
    if constexpr (PF != 0) {
      int hh[6], hp[6];
      step_planes(Hs, hh, hp);
#pragma unroll
      for (int tt = 0; tt < NT; ++tt)
        issue_planes<S, 6>(tt ? rt : rn, hh, hp, tofs, na0[tt], na1[tt]);
    }
    if constexpr (PF == 3) {
      // Two steps in flight: set A carries even steps, set B odd ones.
      float2 nb0[NT][6], nb1[NT][6];
      {
        int hh[6], hp[6];
        step_planes(Hs + 1, hh, hp);
#pragma unroll
        for (int tt = 0; tt < NT; ++tt)
          issue_planes<S, 6>(tt ? rt : rn, hh, hp, tofs, nb0[tt], nb1[tt],
                             Hs + 1 <= He);
      }
      int H = Hs;
      for (; H + 1 <= He; H += 2) {
        step(H, na0, na1, H + 2);
        step(H + 1, nb0, nb1, H + 3);
      }
      if (H <= He) step(H, na0, na1, H + 2);
      __builtin_amdgcn_s_waitcnt(0x0f70);    // vmcnt(0): drain the empty loads
    } else {
      for (int H = Hs; H <= He; ++H) step(H, na0, na1, H + 1);
    }
    }
  } else {
    const int d = di - (S - 1);
    const int hs = d > 0 ? d : 0;
    const int he = d < 0 ? S - 1 + d : S - 1;
    float cA[NT][4], cB[NT][4];
    {
      const int hh[1] = {hs};
      const int hp[1] = {hs - d};
#pragma unroll
      for (int tt = 0; tt < NT; ++tt) {
        float ws[1][4];
        wpass<S, 1>(tt ? rt : rn, hh, hp, L0, tofs, lofs, ws);
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          cA[tt][e] = 0.0f;        // plane (hs-1, hs-1-d) is always outside
          cB[tt][e] = ws[0][e];
        }
      }
    }
    for (int h = hs; h <= he; ++h) {
      const int hq = h - d;
      const int hh[1] = {h + 1};
      const int hp[1] = {h + 1 - d};
      float pre[NT][4];
#pragma unroll
      for (int tt = 0; tt < NT; ++tt) {
        float ws[1][4];
        wpass<S, 1>(tt ? rt : rn, hh, hp, L0, tofs, lofs, ws);
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          pre[tt][e] = cA[tt][e] + cB[tt][e] + ws[0][e];
          cA[tt][e] = cB[tt][e];
          cB[tt][e] = ws[0][e];
        }
      }
      const float4 A = lq1v[(h * S + 2 * W) >> 1];
      const float4 B = lq2v[(hq * S + 2 * Wp) >> 1];
      float vn[4], vt[4];
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const float sa = (e >> 1) ? A.z : A.x, ra = (e >> 1) ? A.w : A.y;
        const float sb = (e & 1) ? B.z : B.x, rb = (e & 1) ? B.w : B.y;
        const float cv = fmaf(wv9, pre[0][e], b_var);
        const float tv = NTK ? fmaf(wv9, pre[NT - 1][e], cv) : 0.0f;
        float an = 0.0f, at = 0.0f;
        relu_acc<NTK>(cv, sa * sb, ra * rb, tv, an, at);
        vn[e] = an * kInv2Pi;
        vt[e] = at * kInv2Pi;
      }
      if (valid) {
        const long long o = p * PER_OUT + ((long long)(h * S + hq) * S +
                                           2 * W) * S + 2 * Wp;
        *reinterpret_cast<float2*>(out_n + o) = make_float2(vn[0], vn[1]);
        *reinterpret_cast<float2*>(out_n + o + S) = make_float2(vn[2], vn[3]);
        if constexpr (NTK) {
          *reinterpret_cast<float2*>(out_t + o) = make_float2(vt[0], vt[1]);
          *reinterpret_cast<float2*>(out_t + o + S) =
              make_float2(vt[2], vt[3]);
        }
      }
    }
  }
}

// Stencil alone (no activation, no pool): out = w9 (3x3 shared-offset
// stencil sum of in) + b over T states [S]^4 -- the Conv of a stage without
// its ReLU, e.g. the stencil means the adjoint recomputes and its
// self-adjoint transpose.  A workgroup sweeps diagonal d = h - h' of JPW
// states as the no-pool path of k_stage_fast does.
template <int S>
__global__ __launch_bounds__(256) void k_stencil_fast(
    const float* __restrict__ in, long long T, int ngroups, int nblk,
    float w9, float b_var, float* __restrict__ out) {
  constexpr int H2 = S / 2;
  constexpr int TPP = H2 * H2;
  constexpr int JPW = 256 / TPP;
  constexpr int LS = S + 2;
  constexpr int PL = LS * LS;
  constexpr long long PER = (long long)S * S * S * S;
  constexpr int ndiag = 2 * S - 1;
  __shared__ __align__(16) float lds[JPW * PL];
  const int L = xcd_remap(blockIdx.x, nblk);
  if (L >= ndiag * ngroups) return;       // whole block exits together
  const int g = L / ndiag;
  const int di = L - g * ndiag;
  const int job = threadIdx.x / TPP;
  const int t = threadIdx.x - job * TPP;
  const int W = t / H2;
  const int Wp = t - W * H2;
  const long long p = (long long)g * JPW + job;
  const bool valid = p < T;
  for (int k = threadIdx.x; k < JPW * PL; k += 256) lds[k] = 0.0f;
  __syncthreads();
  float* L0 = lds + job * PL;
  const long long p0 = (long long)g * JPW;
  const long long left = T - p0;
  const unsigned nbytes = (unsigned)((left < JPW ? left : JPW) * PER * 4);
  const __amdgpu_buffer_rsrc_t rn = make_rsrc(in + p0 * PER, nbytes);
  const unsigned tofs = (unsigned)(job * PER + (2 * W) * S + 2 * Wp) * 4u;
  const int lofs = (2 * W + 1) * LS + 2 * Wp + 1;
  const int d = di - (S - 1);
  const int hs = d > 0 ? d : 0;
  const int he = d < 0 ? S - 1 + d : S - 1;
  float cA[4], cB[4];
  {
    const int hh[1] = {hs};
    const int hp[1] = {hs - d};
    float ws[1][4];
    wpass<S, 1>(rn, hh, hp, L0, tofs, lofs, ws);
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      cA[e] = 0.0f;               // plane (hs-1, hs-1-d) is always outside
      cB[e] = ws[0][e];
    }
  }
  for (int h = hs; h <= he; ++h) {
    const int hq = h - d;
    const int hh[1] = {h + 1};
    const int hp[1] = {h + 1 - d};
    float ws[1][4], v[4];
    wpass<S, 1>(rn, hh, hp, L0, tofs, lofs, ws);
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      v[e] = fmaf(w9, cA[e] + cB[e] + ws[0][e], b_var);
      cA[e] = cB[e];
      cB[e] = ws[0][e];
    }
    if (valid) {
      const long long o =
          p * PER + ((long long)(h * S + hq) * S + 2 * W) * S + 2 * Wp;
      *reinterpret_cast<float2*>(out + o) = make_float2(v[0], v[1]);
      *reinterpret_cast<float2*>(out + o + S) = make_float2(v[2], v[3]);
    }
  }
}

int launch_stencil_fast(const float* in, long long T, int S, float w_var,
                        float b_var, float* out, hipStream_t st) {
  const float w9 = w_var / 9.0f;
  auto go = [&](auto sc) -> int {
    constexpr int SS = decltype(sc)::value;
    constexpr int JPW = 256 / ((SS / 2) * (SS / 2));
    const long long ngroups = (T + JPW - 1) / JPW;
    const long long nb = ngroups * (2 * SS - 1);
    if (nb > 0x7ffffff0LL)
      return set_error(NTA_EINVAL, "nta_stage: tile too large (%lld)", T);
    const int nblk = (int)((nb + 7) / 8 * 8);
    hipLaunchKernelGGL(k_stencil_fast<SS>, dim3(nblk), dim3(256), 0, st, in,
                       T, (int)ngroups, nblk, w9, b_var, out);
    return check_launch("nta_stage");
  };
  switch (S) {
    case 32: return go(std::integral_constant<int, 32>());
    case 16: return go(std::integral_constant<int, 16>());
    case 8: return go(std::integral_constant<int, 8>());
  }
  return kNotHandled;
}

// ------------------------------------------------------------ first stage

// Loads image n of x ([N][S][S][C], C <= 4) into a zero-padded,
// channel-planar LDS image xp[4][S+2][S+2] (channels >= C are zero).
template <int S>
__device__ __forceinline__ void load_padded(const float* __restrict__ x,
                                            long long n, int C, bool valid,
                                            float* __restrict__ xp) {
  constexpr int SP = S + 2;
  for (int k = threadIdx.x; k < 4 * SP * SP; k += blockDim.x) {
    const int c = k / (SP * SP);
    const int r = k - c * (SP * SP);
    const int hh = r / SP - 1;
    const int ww = r - (r / SP) * SP - 1;
    float v = 0.0f;
    if (valid && c < C && (unsigned)hh < (unsigned)S &&
        (unsigned)ww < (unsigned)S)
      v = x[((n * S + hh) * S + ww) * C + c];
    xp[k] = v;
  }
}

template <bool NTK>
__device__ __forceinline__ void first32_epilogue(
    const floatx16& acc, const float (&sq1r)[16], const float (&rq1r)[16],
    float2 qb, float scale, float b_var, float* __restrict__ on,
    float* __restrict__ ot, int half, int col) {
  constexpr int S = 32;
#pragma unroll
  for (int r = 0; r < 16; ++r) {
    const int m = 8 * (r >> 2) + 4 * half + (r & 3);
    const float cv = fmaf(scale, acc[r], b_var);
    const float rho = cv * (rq1r[r] * qb.y);
    const float a = fminf(fabsf(rho), 1.0f);
    const float u = 1.0f - a;
    const float th = acos_pos5(a, u);
    const float pm = rho < 0.0f ? th : kPi - th;
    const float sn = fsqrt(fmaf(-a, a, 1.0f));
    // sq1r already carries the 1 / (2 pi) factor.
    st_state(on + m * S + col, fmaf(pm, rho, sn) * (sq1r[r] * qb.x));
    if constexpr (NTK) st_state(ot + m * S + col, (cv * kInv2Pi) * pm);
  }
}

// RELU = false: the stencil of the input Gram alone (scale G + b_var into
// out_n; q1, q2 and out_t unused), e.g. for the adjoint's recomputation.
template <bool NTK, bool RELU = true>
__global__ __launch_bounds__(256) void k_first32(
    const float* __restrict__ x1, const float* __restrict__ x2, int n1,
    int n2, int C, int nblk, float scale, float b_var,
    const float* __restrict__ q1, const float* __restrict__ q2,
    float* __restrict__ out_n, float* __restrict__ out_t) {
  constexpr int S = 32, SP = S + 2, PLN = SP * SP;
  constexpr long long PER = (long long)S * S * S * S;
  // xp1 is dead once the A operand is in registers; its space then holds
  // the (sq, rq) normalisers of image j.
  __shared__ float xp1[4 * PLN];
  __shared__ float xp2[4 * PLN];
  __shared__ float2 lq1[S];
  float2* lq2 = reinterpret_cast<float2*>(xp1);
  const bool self = n2 == 0;
  const long long T = self ? (long long)n1 : (long long)n1 * n2;
  const int L = xcd_remap(blockIdx.x, nblk);
  if ((long long)L >= T * S) return;
  const long long p = L / S;
  const int h = L - (int)(p * S);
  const long long i = self ? p : p / n2;
  const long long j = self ? p : p - i * n2;
  const float* xb = self ? x1 : x2;
  const float* qbp = self ? q1 : q2;
  load_padded<S>(x1, i, C, true, xp1);
  load_padded<S>(xb, j, C, true, xp2);
  if (RELU && threadIdx.x < S) {
    float2 v = qnorm(q1[(i * S + h) * S + threadIdx.x]);
    lq1[threadIdx.x] = make_float2(v.x * kInv2Pi, v.y);
  }
  __syncthreads();

  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int col = lane & 31;
  const int half = lane >> 5;
  // K = 9 taps x 4 channels; step kk covers tap kk >> 1 and channels
  // 2 (kk & 1) + {0, 1} (lane half).
  float a[18];
  {
    const float* pa = xp1 + half * PLN + h * SP + col;
#pragma unroll
    for (int kk = 0; kk < 18; ++kk) {
      const int tap = kk >> 1;
      a[kk] = pa[2 * (kk & 1) * PLN + (tap / 3) * SP + tap % 3];
    }
  }
  float sq1r[16], rq1r[16];
#pragma unroll
  for (int r = 0; r < 16; ++r) {
    const int m = 8 * (r >> 2) + 4 * half + (r & 3);
    sq1r[r] = RELU ? lq1[m].x : 0.0f;
    rq1r[r] = RELU ? lq1[m].y : 0.0f;
  }
  if constexpr (RELU) {
    __syncthreads();                    // xp1 reads done: reuse as lq2
    for (int k = threadIdx.x; k < S * S; k += 256)
      lq2[k] = qnorm(qbp[j * (S * S) + k]);
    __syncthreads();
  }

  const float* pb0 = xp2 + half * PLN + col;
  auto gram = [&](int hq) {
    floatx16 acc;
#pragma unroll
    for (int r = 0; r < 16; ++r) acc[r] = 0.0f;
    const float* pb = pb0 + hq * SP;
#pragma unroll
    for (int kk = 0; kk < 18; ++kk) {
      const int tap = kk >> 1;
      const float bv = pb[2 * (kk & 1) * PLN + (tap / 3) * SP + tap % 3];
      acc = __builtin_amdgcn_mfma_f32_32x32x2f32(a[kk], bv, acc, 0, 0, 0);
    }
    return acc;
  };
  // Planes hq = wave, wave + 4, ...; the Gram of the next plane is issued
  // before the ReLU epilogue of the current one so MFMA and VALU overlap.
  floatx16 cur = gram(wave);
  for (int hq = wave; hq < S; hq += 4) {
    floatx16 nxt;
    if (hq + 4 < S) nxt = gram(hq + 4);
    float* on = out_n + p * PER + (long long)(h * S + hq) * (S * S);
    if constexpr (RELU) {
      const float2 qb = lq2[hq * S + col];
      float* ot = NTK ? out_t + p * PER + (long long)(h * S + hq) * (S * S)
                      : nullptr;
      first32_epilogue<NTK>(cur, sq1r, rq1r, qb, scale, b_var, on, ot, half,
                            col);
    } else {
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        const int m = 8 * (r >> 2) + 4 * half + (r & 3);
        on[m * S + col] = fmaf(scale, cur[r], b_var);
      }
    }
    cur = nxt;
  }
}

template <bool NTK>
__global__ __launch_bounds__(256) void k_first16(
    const float* __restrict__ x1, const float* __restrict__ x2, int n1,
    int n2, int C, int nblk, float scale, float b_var,
    const float* __restrict__ q1, const float* __restrict__ q2,
    float* __restrict__ out_n, float* __restrict__ out_t) {
  constexpr int S = 16, SP = S + 2, PLN = SP * SP;
  constexpr long long PER = (long long)S * S * S * S;
  __shared__ float xp1[4 * PLN];
  __shared__ float xp2[4 * PLN];
  __shared__ float2 lq2[S * S];
  __shared__ float2 lq1[S];
  const bool self = n2 == 0;
  const long long T = self ? (long long)n1 : (long long)n1 * n2;
  const int L = xcd_remap(blockIdx.x, nblk);
  if ((long long)L >= T * S) return;
  const long long p = L / S;
  const int h = L - (int)(p * S);
  const long long i = self ? p : p / n2;
  const long long j = self ? p : p - i * n2;
  const float* xb = self ? x1 : x2;
  const float* qbp = self ? q1 : q2;
  load_padded<S>(x1, i, C, true, xp1);
  load_padded<S>(xb, j, C, true, xp2);
  for (int k = threadIdx.x; k < S * S; k += 256)
    lq2[k] = qnorm(qbp[j * (S * S) + k]);
  if (threadIdx.x < S) lq1[threadIdx.x] = qnorm(q1[(i * S + h) * S +
                                                   threadIdx.x]);
  __syncthreads();

  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int col = lane & 15;
  const int ch = lane >> 4;           // channel of this lane's K slice
  // K = 9 taps x 4 channels; step kk = tap, the 4 channels across lanes.
  float a[9];
  const float* pa = xp1 + ch * PLN + h * SP + col;
#pragma unroll
  for (int kk = 0; kk < 9; ++kk) a[kk] = pa[(kk / 3) * SP + kk % 3];
  float sq1r[4], rq1r[4];
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int m = 4 * ch + r;
    sq1r[r] = lq1[m].x;
    rq1r[r] = lq1[m].y;
  }
  for (int hq = wave; hq < S; hq += 4) {
    floatx4 acc;
#pragma unroll
    for (int r = 0; r < 4; ++r) acc[r] = 0.0f;
    const float* pb = xp2 + ch * PLN + hq * SP + col;
#pragma unroll
    for (int kk = 0; kk < 9; ++kk) {
      const float bv = pb[(kk / 3) * SP + kk % 3];
      acc = __builtin_amdgcn_mfma_f32_16x16x4f32(a[kk], bv, acc, 0, 0, 0);
    }
    const float2 qb = lq2[hq * S + col];
    float* on = out_n + p * PER + (long long)(h * S + hq) * (S * S);
    float* ot = NTK ? out_t + p * PER + (long long)(h * S + hq) * (S * S)
                    : nullptr;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int m = 4 * ch + r;
      const float cv = fmaf(scale, acc[r], b_var);
      float ds;
      const float v = relu_map_norm(cv, sq1r[r] * qb.x, rq1r[r] * qb.y, &ds);
      st_state(on + m * S + col, v);
      if constexpr (NTK) st_state(ot + m * S + col, cv * ds);
    }
  }
}

// ------------------------------------------- first two stages, fused

// Conv(3x3 SAME) -> ReLU -> Conv(3x3 SAME) -> ReLU -> AvgPool(2x2) on 32x32
// images, with the first stage's state never written to memory.  The
// decomposition is k_stage_fast's (a sweep in H along each pooled diagonal
// H - H' = D of one pair), with several diagonals of 128 steps in all per
// workgroup (f12_diag) so that the images are prepared once per 128 steps;
// the six first-stage planes a step consumes are computed in place from the
// two images held in LDS:
//   plane (h, h') = ReLU(scale1 * P_i(h) P_j(h')^T + b1),
// P_n(h) the [32 w][9 taps x 4 channels] patch matrix of image n at row h
// (the Gram of zero-padded patches equals the input-covariance stencil).
// Each wave owns one 16x16 (w, w') quadrant of every plane.  The Gram runs
// on the FP16 matrix pipe with FP32 accuracy: each image is scaled by a
// power of two and split as hi + lo (two FP16 numbers), and the three
// products hi hi + hi lo + lo hi (v_mfma_f32_16x16x32_f16, two per product
// for the 9 x 4 patch entries) accumulate exactly-multiplied terms in FP32;
// the dropped lo lo term is below 2^-22 of each product.  That is 6 MFMAs
// of 16 cycles per plane quadrant where FP32 MFMAs need 9 of 32, and the
// fused kernel's matrix time (it does not overlap the vector ALU) drops
// 3x.  The ReLU is the epilogue.  The planes then go through the LDS
// exchange of the second stage (nngp first, ntk second).  Planes on odd
// first-stage diagonals are computed by both workgroups that need them
// (1.5x the first-stage work), which costs far less than the 16 MiB per
// pair of state the unfused pair of stages writes and reads back.
constexpr int kXS = 34;               // padded image row (positions)
// Rows past the image that the last sweep steps' operand reads reach (their
// planes lie past the sweep end and are dropped): the image array carries
// them as zeroed padding, so every read stays inside it.
constexpr int kXhTail = 4 * kXS * 4;  // halves
constexpr int kLS2 = 36;              // exchange plane row (4 rows = 16 banks)
// Vector instructions scheduled between consecutive MFMAs of the next step's
// patch Grams (0: leave the order to the compiler).
#ifndef NTA_F12_SCHED
#define NTA_F12_SCHED 16
#endif
constexpr int kPL2 = 34 * kLS2;

// Writes this lane's 4 x K first-stage values (rows w0 + e, column wq of
// each plane, as offsets wr + e * kLS2) and reads back the (w, w')-diagonal
// 3-tap sums of the thread's 2x2 block.  Every thread calls it the same
// number of times.
template <int K>
__device__ __forceinline__ void s1_exchange(float* __restrict__ lp,
                                            const float (&v)[K][4], int wr,
                                            int lofs, float (&ws)[K][4]) {
  __syncthreads();                     // the previous reads are done
#pragma unroll
  for (int k = 0; k < K; ++k)
#pragma unroll
    for (int e = 0; e < 4; ++e) lp[k * kPL2 + wr + e * kLS2] = v[k][e];
  __syncthreads();
#pragma unroll
  for (int k = 0; k < K; ++k) read_diag3<32, kLS2>(lp + k * kPL2, lofs, ws[k]);
}

// Workgroups ("slots") per pair and the diagonals each sweeps.  With 16:
// slot 0 sweeps D = 0, slot s > 0 sweeps s and s - 16 (16 steps).  With 8:
// slot 0 sweeps 0, 8, -8, slot s > 0 sweeps s, s - 16, s + 8, s - 8 (32
// steps).  With 4, 2 (default) or 1: the diagonals of 2, 4 or 8 of those
// groups (64, 128 or 256 steps), so the images are prepared once per 128
// steps by default (measured: 4 slots 0.7% slower, 1 the same, 8 3% slower).
#ifndef NTA_F12_SLOTS
#define NTA_F12_SLOTS 2
#endif
constexpr int kF12Slots = NTA_F12_SLOTS;
static_assert(kF12Slots == 1 || kF12Slots == 2 || kF12Slots == 4 ||
              kF12Slots == 8 || kF12Slots == 16,
              "1, 2, 4, 8 or 16 slots per pair");
__device__ __forceinline__ int f12_nsweep8(int slot) {
  return slot == 0 ? 3 : 4;
}
__device__ __forceinline__ int f12_diag8(int slot, int sw) {
  if (slot == 0) return sw == 0 ? 0 : (sw == 1 ? 8 : -8);
  return sw == 0 ? slot : (sw == 1 ? slot - 16 : (sw == 2 ? slot + 8
                                                          : slot - 8));
}
// With fewer slots, slot t sweeps the diagonals of the 8 / kF12Slots
// consecutive 8-slot groups starting at t * 8 / kF12Slots.
__device__ __forceinline__ int f12_nsweep(int slot) {
  if (kF12Slots == 16) return slot == 0 ? 1 : 2;
  constexpr int G = 8 / (kF12Slots < 8 ? kF12Slots : 8);
  int n = 0;
  for (int g = 0; g < G; ++g) n += f12_nsweep8(slot * G + g);
  return n;
}
__device__ __forceinline__ int f12_diag(int slot, int sw) {
  if (kF12Slots == 16) return sw == 0 ? slot : slot - 16;
  constexpr int G = 8 / (kF12Slots < 8 ? kF12Slots : 8);
  int g = slot * G;
  while (sw >= f12_nsweep8(g)) sw -= f12_nsweep8(g++);
  return f12_diag8(g, sw);
}

// KP: the stencil means of the first stage's outputs at every position,
// (1/9) sum of the nine taps -- the second stage's pre-activations before
// its w_var, b_var affine map -- also go to pre_n, pre_t
// ([T][32][32][32][32]), for the adjoint of the second stage.
template <bool NTK, bool KP = false>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2)))
void k_first_pool32(const float* __restrict__ x1, const float* __restrict__ x2,
                    int n1, int n2, int C, int nblk, float scale1, float b1,
                    const float2* __restrict__ qa1,
                    const float2* __restrict__ qa2, float wv9, float b2,
                    const float2* __restrict__ qb1,
                    const float2* __restrict__ qb2, float* __restrict__ out_n,
                    float* __restrict__ out_t, float* __restrict__ pre_n,
                    float* __restrict__ pre_t) {
  constexpr int S = 32, H2 = 16, SO = 16;
  constexpr long long PER_OUT = (long long)SO * SO * SO * SO;
  constexpr long long PER_IN = (long long)S * S * S * S;
  constexpr int NT = NTK ? 2 : 1;
  // Images i, j as FP16 hi/lo pairs: xh[img][hl][1 + h][1 + w][c] (zero
  // padded, channels padded to 4).
  __shared__ __align__(16) _Float16 xh[2 * 2 * kXS * kXS * 4 + kXhTail];
  __shared__ __align__(16) float lp[6 * kPL2];        // exchange planes
  __shared__ unsigned smax[2];
  const bool self = n2 == 0;
  const long long T = self ? (long long)n1 : (long long)n1 * n2;
  // Workgroup (p, slot): slot 0 sweeps diagonal 0, slot s > 0 sweeps s and
  // then s - 16: 16 steps each, and the images are prepared once per pair
  // and 16 diagonals instead of once per diagonal.
  const int L = xcd_remap(blockIdx.x, nblk);
  if ((long long)L >= T * kF12Slots) return;   // whole block exits together
  const long long p = L / kF12Slots;
  const int slot = L - (int)(p * kF12Slots);
  const long long i = self ? p : p / n2;
  const long long j = self ? p : p - i * n2;
  const int t = threadIdx.x;
  // Second-stage role: the thread's 2x2 (w, w') block at (2W, 2Wp).
  const int W = t / H2;
  const int Wp = t - W * H2;
  const int lofs = (2 * W + 1) * kLS2 + 2 * Wp + 1;
  // First-stage role: quadrant (bw, bwp) of the wave; MFMA output rows
  // w0 .. w0 + 3 (lane group c4), column wq.
  const int lane = t & 63, wave = t >> 6;
  const int m16 = lane & 15, c4 = lane >> 4;
  const int w0 = 16 * (wave >> 1) + 4 * c4;
  const int wq = 16 * (wave & 1) + m16;
  const int wr = (w0 + 1) * kLS2 + wq + 1;

  static_assert((6 * kPL2) % 4 == 0, "exchange planes in float4 units");
  for (int k = t; k < 6 * kPL2 / 4; k += 256)
    reinterpret_cast<float4*>(lp)[k] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  if (t < 2) smax[t] = 0u;
  const float* xi = x1 + i * (S * S * C);
  const float* xj = (self ? x1 : x2) + j * (S * S * C);
  __syncthreads();
  {
    // Per-image power-of-two scale 2^-e with max |x| 2^-e in [0.5, 1):
    // exact, and it keeps the FP16 parts in range.
    float mi = 0.0f, mj = 0.0f;
    for (int k = t; k < S * S * C; k += 256) {
      mi = fmaxf(mi, fabsf(xi[k]));
      mj = fmaxf(mj, fabsf(xj[k]));
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      mi = fmaxf(mi, __shfl_xor(mi, o));
      mj = fmaxf(mj, __shfl_xor(mj, o));
    }
    if (lane == 0) {
      atomicMax(&smax[0], __float_as_uint(mi));
      atomicMax(&smax[1], __float_as_uint(mj));
    }
  }
  __syncthreads();
  int ei, ej;
  frexpf(__uint_as_float(smax[0]), &ei);
  frexpf(__uint_as_float(smax[1]), &ej);
  // x = 2^e (hi + lo) with hi = fp16(x 2^-e), lo = fp16(x 2^-e - hi): the
  // patch Gram is 2^(ei+ej) (hi_i hi_j + hi_i lo_j + lo_i hi_j) up to
  // |lo_i lo_j| <= 2^-22 |x_i x_j|, each FP16 product exact in the FP32
  // accumulator.
  const float scale_g = ldexpf(scale1, ei + ej);
  {
    // One padded position (4 channels: one 8-byte hi and one lo store) per
    // thread and pass.
    for (int k = t; k < 2 * kXS * kXS; k += 256) {
      const int img = k >= kXS * kXS;
      const int pos = k - img * (kXS * kXS);
      const int hh = pos / kXS - 1;
      const int ww = pos - (pos / kXS) * kXS - 1;
      const float sc = ldexpf(1.0f, img ? -ej : -ei);
      const float* src = (img ? xj : xi) + (hh * S + ww) * C;
      const bool in = (unsigned)hh < (unsigned)S && (unsigned)ww < (unsigned)S;
      half4v hv4, lv4;
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        const float v = (in && c < C) ? src[c] * sc : 0.0f;
        const _Float16 hv = (_Float16)v;
        hv4[c] = hv;
        lv4[c] = (_Float16)(v - (float)hv);
      }
      *reinterpret_cast<half4v*>(xh + ((img * 2) * (kXS * kXS) + pos) * 4) =
          hv4;
      *reinterpret_cast<half4v*>(xh + ((img * 2 + 1) * (kXS * kXS) + pos) *
                                          4) = lv4;
    }
    for (int k = t; k < kXhTail / 4; k += 256)
      *reinterpret_cast<half4v*>(xh + 2 * 2 * kXS * kXS * 4 + 4 * k) =
          half4v{0, 0, 0, 0};
  }
  __syncthreads();
  // MFMA operands (v_mfma_f32_16x16x32_f16, 8 halves per lane; A and B
  // share the K order).  The 9 x 4 patch entries (tap (ty, tx), 4 channels)
  // are laid out so that every lane reads contiguous halves:
  //   step 0: lane group c4 < 3 carries taps (c4, 0) and (c4, 1), 16
  //     contiguous bytes of its row;
  //   step 1: lane group c4 < 3 carries tap (c4, 2); image i also reads the
  //     next 8 bytes, image j has zeros there, so they add nothing.
  // Lane group 3 reads zeros: padded column 33 and the next row's column 0
  // (step 0), padded column 0 (step 1; its upper halves meet image j's
  // zeros).  Byte offsets from the image base, before the row term.
  constexpr int kRowB = kXS * 4 * 2;            // bytes per padded row
  constexpr int kHlB = kXS * kXS * 4 * 2;       // bytes of one hi or lo image
  int oa[2], ob[2];
#pragma unroll
  for (int st = 0; st < 2; ++st) {
    const int tx = st == 0 ? 0 : 2;
    oa[st] = c4 < 3 ? c4 * kRowB + (16 * (wave >> 1) + m16 + tx) * 8
                    : (st == 0 ? (kXS - 1) * 8 : 0);
    ob[st] = c4 < 3 ? c4 * kRowB + (16 * (wave & 1) + m16 + tx) * 8
                    : (st == 0 ? (kXS - 1) * 8 : 0);
  }
  const char* xhb = reinterpret_cast<const char*>(xh);
  // Operands of image img (0: i, 1: j), part hl, padded row base `row`.
  auto opnd8 = [&](int img, int hl, int row) {
    const char* b = xhb + (img * 2 + hl) * kHlB + row * kRowB + (img ? ob[0] : oa[0]);
    const half4v lo = *reinterpret_cast<const half4v*>(b);
    const half4v hi = *reinterpret_cast<const half4v*>(b + 8);
    return half8v{lo[0], lo[1], lo[2], lo[3], hi[0], hi[1], hi[2], hi[3]};
  };
  auto opnd4 = [&](int img, int hl, int row) {
    const char* b = xhb + (img * 2 + hl) * kHlB + row * kRowB + (img ? ob[1] : oa[1]);
    const half4v lo = *reinterpret_cast<const half4v*>(b);
    const _Float16 z = 0;
    if (img == 1) return half8v{lo[0], lo[1], lo[2], lo[3], z, z, z, z};
    const half4v hi = *reinterpret_cast<const half4v*>(b + 8);
    return half8v{lo[0], lo[1], lo[2], lo[3], hi[0], hi[1], hi[2], hi[3]};
  };
  const float2* qi = qa1 + i * (S * S);
  const float2* qj = (self ? qa1 : qa2) + j * (S * S);
  const float2* ri = qb1 + i * (S * S);
  const float2* rj = (self ? qb1 : qb2) + j * (S * S);

  // This wave's quadrant of first-stage plane (hh, hp): nngp in vn, ntk in
  // vt (zero outside the grid: the second conv's zero padding), both
  // 2 pi times the first stage's state (the factor is folded into the
  // second conv's weight, wv2 = wv9 / (2 pi)).
  auto s1 = [&](int hh, int hp, float (&vn)[4], float (&vt)[4]) {
    const bool ok = (unsigned)hh < (unsigned)S && (unsigned)hp < (unsigned)S;
    if (!ok) hh = hp = 0;              // uniform; the values are dropped
    floatx4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
    if (ok) {
      const half8v ah = opnd8(0, 0, hh), al = opnd8(0, 1, hh);
      const half8v bh = opnd8(1, 0, hp), bl = opnd8(1, 1, hp);
      acc = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah, bh, acc, 0, 0, 0);
      acc = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah, bl, acc, 0, 0, 0);
      acc = __builtin_amdgcn_mfma_f32_16x16x32_f16(al, bh, acc, 0, 0, 0);
      const half8v ch = opnd4(0, 0, hh), cl = opnd4(0, 1, hh);
      const half8v dh = opnd4(1, 0, hp), dl = opnd4(1, 1, hp);
      acc = __builtin_amdgcn_mfma_f32_16x16x32_f16(ch, dh, acc, 0, 0, 0);
      acc = __builtin_amdgcn_mfma_f32_16x16x32_f16(ch, dl, acc, 0, 0, 0);
      acc = __builtin_amdgcn_mfma_f32_16x16x32_f16(cl, dh, acc, 0, 0, 0);
    }
    const float4 q01 = *reinterpret_cast<const float4*>(qi + hh * S + w0);
    const float4 q23 = *reinterpret_cast<const float4*>(qi + hh * S + w0 + 2);
    const float2 qb = qj[hp * S + wq];
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const float4 Q = e < 2 ? q01 : q23;
      const float sa = (e & 1) ? Q.z : Q.x, ra = (e & 1) ? Q.w : Q.y;
      const float cv = fmaf(scale_g, acc[e], b1);
      const float rho = cv * (ra * qb.y);
      const float a = fminf(fabsf(rho), 1.0f);
      const float u = 1.0f - a;
      const float th = acos_pos5(a, u);
      const float pm = rho < 0.0f ? th : kPi - th;
      const float sn = fsqrt(fmaf(-a, a, 1.0f));
      vn[e] = ok ? fmaf(pm, rho, sn) * (sa * qb.x) : 0.0f;
      vt[e] = ok ? cv * pm : 0.0f;
    }
  };

  const float wv2 = wv9 * kInv2Pi;
  const int nsweep = f12_nsweep(slot);
  for (int sw = 0; sw < nsweep; ++sw) {
  const int D = f12_diag(slot, sw);
  const int Hs = D > 0 ? D : 0;
  const int He = D < 0 ? H2 - 1 + D : H2 - 1;
  const int d0 = 2 * D, dp = 2 * D + 1, dm = 2 * D - 1;

  // Loop-carried state of the sweep.  Steps alternate between two copies
  // (the loop is unrolled by two) so the carried values stay in place.
  struct Carry {
    float c0a[NT][4], c0b[NT][4], cp[NT][4], cm[NT][4];  // (h, h') window
    float4 qal[3], qah[3];   // first-stage (sq, rq), image i rows 2H + r
    float2 qbv[3];           // first-stage (sq, rq), image j rows 2Hp + r
    float4 r0, r1, s0, s1;   // second-stage (sq, rq) rows 2H(+1), 2Hp(+1)
    floatx4 acc[6];          // patch Grams of the step's six planes
  };
  Carry A, B;
  {
    const int h0 = 2 * Hs;
    const int hh[4] = {h0 - 1, h0, h0, h0 - 1};
    const int hp[4] = {h0 - 1 - d0, h0 - d0, h0 - dp, h0 - 1 - dm};
    float vn[4][4], vt[4][4];
#pragma unroll
    for (int k = 0; k < 4; ++k) s1(hh[k], hp[k], vn[k], vt[k]);
    auto init = [&](const float (&v)[4][4], int tt) {
      float ws[4][4];
      s1_exchange<4>(lp, v, wr, lofs, ws);
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        A.c0a[tt][e] = ws[0][e];
        A.c0b[tt][e] = ws[1][e];
        A.cp[tt][e] = ws[2][e];
        A.cm[tt][e] = ws[3][e];
      }
    };
    init(vn, 0);
    if constexpr (NTK) init(vt, NT - 1);
  }

  // Plane k of step H is (2H + ar[k], 2Hp + br[k]) (d0 @ 2H+1, 2H+2; dp @
  // 2H+1, 2H+2; dm @ 2H, 2H+1).  The normalisers of the rows a step adds
  // are loaded one step ahead (row indices clamped into the table: values
  // of rows past the grid are never used).
  constexpr int ar[6] = {1, 2, 1, 2, 0, 1};
  constexpr int br[6] = {1, 2, 0, 1, 1, 2};
  auto ld_a = [&](int row, float4& lo, float4& hi) {
    row = row < S ? row : S - 1;
    lo = *reinterpret_cast<const float4*>(qi + row * S + w0);
    hi = *reinterpret_cast<const float4*>(qi + row * S + w0 + 2);
  };
  auto ld_b = [&](int row) { return qj[(row < S ? row : S - 1) * S + wq]; };
  auto ld_r = [&](const float2* tab, int row, int col) {
    return *reinterpret_cast<const float4*>(
        tab + (row < S ? row : S - 1) * S + col);
  };
  // First-stage patch Grams of the six planes of step Hn (split FP16, 36
  // MFMAs).  Padded rows 2Hn + r, r = 0..2 (+ 0..2 taps); past the sweep
  // end (Hn = He + 1) they read the rows below, the lo image, the other
  // image or the LDS words after them: values of planes that are dropped
  // (every plane has its own accumulator, so they never reach a plane that
  // is used).
  auto gram6 = [&](int Hn, floatx4 (&acc)[6]) {
    const int Hpn = Hn - D;
#pragma unroll
    for (int k = 0; k < 6; ++k) acc[k] = {0.0f, 0.0f, 0.0f, 0.0f};
    {
      half8v ah[3], al[3], bh[3], bl[3];
#pragma unroll
      for (int r = 0; r < 3; ++r) {
        ah[r] = opnd8(0, 0, 2 * Hn + r);
        bh[r] = opnd8(1, 0, 2 * Hpn + r);
      }
#pragma unroll
      for (int k = 0; k < 6; ++k)
        acc[k] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah[ar[k]], bh[br[k]],
                                                        acc[k], 0, 0, 0);
#pragma unroll
      for (int r = 0; r < 3; ++r) bl[r] = opnd8(1, 1, 2 * Hpn + r);
#pragma unroll
      for (int k = 0; k < 6; ++k)
        acc[k] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah[ar[k]], bl[br[k]],
                                                        acc[k], 0, 0, 0);
#pragma unroll
      for (int r = 0; r < 3; ++r) al[r] = opnd8(0, 1, 2 * Hn + r);
#pragma unroll
      for (int k = 0; k < 6; ++k)
        acc[k] = __builtin_amdgcn_mfma_f32_16x16x32_f16(al[ar[k]], bh[br[k]],
                                                        acc[k], 0, 0, 0);
    }
    {
      half8v ah[3], al[3], bh[3], bl[3];
#pragma unroll
      for (int r = 0; r < 3; ++r) {
        ah[r] = opnd4(0, 0, 2 * Hn + r);
        bh[r] = opnd4(1, 0, 2 * Hpn + r);
      }
#pragma unroll
      for (int k = 0; k < 6; ++k)
        acc[k] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah[ar[k]], bh[br[k]],
                                                        acc[k], 0, 0, 0);
#pragma unroll
      for (int r = 0; r < 3; ++r) bl[r] = opnd4(1, 1, 2 * Hpn + r);
#pragma unroll
      for (int k = 0; k < 6; ++k)
        acc[k] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah[ar[k]], bl[br[k]],
                                                        acc[k], 0, 0, 0);
#pragma unroll
      for (int r = 0; r < 3; ++r) al[r] = opnd4(0, 1, 2 * Hn + r);
#pragma unroll
      for (int k = 0; k < 6; ++k)
        acc[k] = __builtin_amdgcn_mfma_f32_16x16x32_f16(al[ar[k]], bh[br[k]],
                                                        acc[k], 0, 0, 0);
    }
  };
  {
    const int Hp = Hs - D;
#pragma unroll
    for (int r = 0; r < 3; ++r) {
      ld_a(2 * Hs + r, A.qal[r], A.qah[r]);
      A.qbv[r] = ld_b(2 * Hp + r);
    }
    A.r0 = ld_r(ri, 2 * Hs, 2 * W);
    A.r1 = ld_r(ri, 2 * Hs + 1, 2 * W);
    A.s0 = ld_r(rj, 2 * Hp, 2 * Wp);
    A.s1 = ld_r(rj, 2 * Hp + 1, 2 * Wp);
    gram6(Hs, A.acc);
  }

  // First-stage ReLU of the step's planes (MASK: some plane is outside the
  // grid and reads as zero).
  auto relu1 = [&](const Carry& in, int H, auto mask, float (&vn)[6][4],
                   float (&vt)[6][4]) {
    constexpr bool MASK = decltype(mask)::value;
    const int Hp = H - D;
#pragma unroll
    for (int k = 0; k < 6; ++k) {
      const bool ok = 2 * H + ar[k] < S && 2 * Hp + br[k] < S;
      const float2 qb = in.qbv[br[k]];
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const float4 Q = e < 2 ? in.qal[ar[k]] : in.qah[ar[k]];
        const float sa = (e & 1) ? Q.z : Q.x, ra = (e & 1) ? Q.w : Q.y;
        const float cv = fmaf(scale_g, in.acc[k][e], b1);
        const float rho = cv * (ra * qb.y);
        const float a = fminf(fabsf(rho), 1.0f);
        const float th = acos_pos5(a, 1.0f - a);
        const float pm = rho < 0.0f ? th : kPi - th;
        const float sn = fsqrt(fmaf(-a, a, 1.0f));
        const float n = fmaf(pm, rho, sn) * (sa * qb.x);
        const float t = cv * pm;
        vn[k][e] = MASK ? (ok ? n : 0.0f) : n;
        vt[k][e] = MASK ? (ok ? t : 0.0f) : t;
      }
    }
  };

  auto step = [&](int H, const Carry& in, Carry& out) {
    const int Hp = H - D;
    // Next step's normaliser rows, in flight during this step.
    out.qal[0] = in.qal[2];
    out.qah[0] = in.qah[2];
    out.qbv[0] = in.qbv[2];
#pragma unroll
    for (int r = 1; r < 3; ++r) {
      ld_a(2 * H + 2 + r, out.qal[r], out.qah[r]);
      out.qbv[r] = ld_b(2 * Hp + 2 + r);
    }
    out.r0 = ld_r(ri, 2 * H + 2, 2 * W);
    out.r1 = ld_r(ri, 2 * H + 3, 2 * W);
    out.s0 = ld_r(rj, 2 * Hp + 2, 2 * Wp);
    out.s1 = ld_r(rj, 2 * Hp + 3, 2 * Wp);

    float vn[6][4], vt[6][4];
    if (2 * H + 2 < S && 2 * Hp + 2 < S)
      relu1(in, H, std::false_type(), vn, vt);
    else
      relu1(in, H, std::true_type(), vn, vt);

    float pre[NT][4][4];   // [tensor][P00, P11, P10, P01][element]
    auto advance = [&](const float (&v)[6][4], int tt) {
      float ws[6][4];
      s1_exchange<6>(lp, v, wr, lofs, ws);
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const float z0 = in.c0b[tt][e] + ws[0][e];
        pre[tt][0][e] = in.c0a[tt][e] + z0;
        pre[tt][1][e] = z0 + ws[1][e];
        pre[tt][2][e] = in.cp[tt][e] + ws[2][e] + ws[3][e];
        pre[tt][3][e] = in.cm[tt][e] + ws[4][e] + ws[5][e];
        out.c0a[tt][e] = ws[0][e];
        out.c0b[tt][e] = ws[1][e];
        out.cp[tt][e] = ws[3][e];
        out.cm[tt][e] = ws[5][e];
      }
    };
    advance(vn, 0);
    if constexpr (NTK) advance(vt, NT - 1);
    if constexpr (KP) {
      // plane pl at rows (hr, hc) of (i, j), element e at w = 2W + e / 2,
      // w' = 2Wp + e % 2; the first stage's values are 2 pi too large
      const int hr[4] = {2 * H, 2 * H + 1, 2 * H + 1, 2 * H};
      const int hc[4] = {2 * Hp, 2 * Hp + 1, 2 * Hp, 2 * Hp + 1};
      constexpr float kMean = kInv2Pi / 9.0f;
#pragma unroll
      for (int tt = 0; tt < NT; ++tt) {
        float* dst = (tt ? pre_t : pre_n) + p * PER_IN;
#pragma unroll
        for (int pl = 0; pl < 4; ++pl) {
          float* row = dst + ((hr[pl] * S + hc[pl]) * S + 2 * W) * S + 2 * Wp;
          *reinterpret_cast<float2*>(row) =
              make_float2(pre[tt][pl][0] * kMean, pre[tt][pl][1] * kMean);
          *reinterpret_cast<float2*>(row + S) =
              make_float2(pre[tt][pl][2] * kMean, pre[tt][pl][3] * kMean);
        }
      }
    }
    // The next step's matrix work overlaps this step's second-stage ReLU
    // (past the sweep end it computes planes that are never used).
    gram6(H + 1, out.acc);
    // Second stage: (sq, rq) at rows 2H, 2H+1 of image i and 2Hp, 2Hp+1 of
    // image j, columns (2W, 2W+1) and (2Wp, 2Wp+1).  The pool sums
    // sq_i sq_j J over the 4 x 4 block as sum_x sq_i[x] sum_y sq_j[y] J.
    float acc_n = 0.0f, acc_t = 0.0f;
#pragma unroll
    for (int x = 0; x < 4; ++x) {          // row of i: (2H + (x >> 1), 2W + (x & 1))
      const float4 Ai = (x >> 1) ? in.r1 : in.r0;
      const float sa = (x & 1) ? Ai.z : Ai.x, ra = (x & 1) ? Ai.w : Ai.y;
      float inner = 0.0f;
#pragma unroll
      for (int y = 0; y < 4; ++y) {        // (2Hp + (y >> 1), 2Wp + (y & 1))
        const float4 Bj = (y >> 1) ? in.s1 : in.s0;
        const float sb = (y & 1) ? Bj.z : Bj.x, rb = (y & 1) ? Bj.w : Bj.y;
        // plane index of (row of i, row of j) = ((x >> 1), (y >> 1)):
        // P00 (0,0) P11 (1,1) P10 (1,0) P01 (0,1); element (x & 1, y & 1)
        const int pl = (x >> 1) == (y >> 1) ? (x >> 1) : 2 + (y >> 1);
        const int e = 2 * (x & 1) + (y & 1);
        const float cv = fmaf(wv2, pre[0][pl][e], b2);
        const float tv = NTK ? fmaf(wv2, pre[NT - 1][pl][e], cv) : 0.0f;
        const float rho = cv * (ra * rb);
        const float a = fminf(fabsf(rho), 1.0f);
        const float th = acos_pos5(a, 1.0f - a);
        const float pm = rho < 0.0f ? th : kPi - th;
        const float sn = fsqrt(fmaf(-a, a, 1.0f));
        inner = fmaf(fmaf(pm, rho, sn), sb, inner);
        if constexpr (NTK) acc_t = fmaf(pm, tv, acc_t);
      }
      acc_n = fmaf(inner, sa, acc_n);
    }
#if NTA_F12_SCHED
    // One MFMA per NTA_F12_SCHED vector instructions of the second-stage
    // ReLU: the wave keeps issuing while the matrix pipe runs.
#pragma unroll
    for (int m = 0; m < 36; ++m) {
      __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);
      __builtin_amdgcn_sched_group_barrier(0x002, NTA_F12_SCHED, 0);
    }
#endif
    const long long o =
        p * PER_OUT + ((long long)(H * SO + Hp) * SO + W) * SO + Wp;
    constexpr float kScale = kInv2Pi / 16.0f;
    st_state(out_n + o, acc_n * kScale);
    if constexpr (NTK) st_state(out_t + o, acc_t * kScale);
  };

  int H = Hs;
  for (; H + 1 <= He; H += 2) {
    step(H, A, B);
    step(H + 1, B, A);
  }
  if (H <= He) step(H, A, B);
  }
}

// ------------------------------------------- first three stages, fused

// Conv(3x3 SAME) -> ReLU -> Conv(3x3 SAME) -> ReLU -> Conv(3x3 SAME) ->
// ReLU -> AvgPool(2x2) on 32x32 images (the first three layers of
// Myrtle-10), with neither the first nor the second stage's 32^4 state
// ever written.  The sweep is k_first_pool32's, one stage deeper: a step of
// pooled row H on pooled diagonal D needs third-stage pre-activations on
// the input diagonals 2D - 1, 2D, 2D + 1 (rows 2H, 2H + 1), hence six new
// second-stage planes per step, at the same offsets (2H + ar, 2Hp + br) as
// the first-stage planes of k_first_pool32, and six new first-stage planes
// one row further on (2H + 1 + ar, 2Hp + 1 + br).  Each conv's (h, h') 3-tap
// sum is a register window along the diagonal (stage 1: two w-summed planes
// carried per input diagonal; stage 2: as the first stage of
// k_first_pool32), each (w, w') 3-tap sum an LDS exchange.  The second
// stage's output is computed per thread on its 2 x 2 (w, w') block, kept
// 2 pi times larger like the first stage's (the factor moves into the next
// conv's weight), and zero on planes outside the grid (the third conv's
// zero padding).
template <bool NTK>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2)))
void k_first3_pool32(const float* __restrict__ x1,
                     const float* __restrict__ x2, int n1, int n2, int C,
                     int nblk, float scale1, float b1,
                     const float2* __restrict__ qa1,
                     const float2* __restrict__ qa2, float wv9b, float b2,
                     const float2* __restrict__ qb1,
                     const float2* __restrict__ qb2, float wv9c, float b3,
                     const float2* __restrict__ qc1,
                     const float2* __restrict__ qc2, float* __restrict__ out_n,
                     float* __restrict__ out_t) {
  constexpr int S = 32, H2 = 16, SO = 16;
  constexpr long long PER_OUT = (long long)SO * SO * SO * SO;
  constexpr int NT = NTK ? 2 : 1;
  __shared__ __align__(16) _Float16 xh[2 * 2 * kXS * kXS * 4 + kXhTail];
  __shared__ __align__(16) float lp[6 * kPL2];        // exchange planes
  __shared__ unsigned smax[2];
  const bool self = n2 == 0;
  const long long T = self ? (long long)n1 : (long long)n1 * n2;
  const int L = xcd_remap(blockIdx.x, nblk);
  if ((long long)L >= T * kF12Slots) return;   // whole block exits together
  const long long p = L / kF12Slots;
  const int slot = L - (int)(p * kF12Slots);
  const long long i = self ? p : p / n2;
  const long long j = self ? p : p - i * n2;
  const int t = threadIdx.x;
  // Block role (stages 2 and 3): the thread's 2x2 (w, w') block at
  // (2W, 2Wp).
  const int W = t / H2;
  const int Wp = t - W * H2;
  const int lofs = (2 * W + 1) * kLS2 + 2 * Wp + 1;
  // MFMA role (stage 1): quadrant of the wave, rows w0 .. w0 + 3, column wq.
  const int lane = t & 63, wave = t >> 6;
  const int m16 = lane & 15, c4 = lane >> 4;
  const int w0 = 16 * (wave >> 1) + 4 * c4;
  const int wq = 16 * (wave & 1) + m16;
  const int wr = (w0 + 1) * kLS2 + wq + 1;

  static_assert((6 * kPL2) % 4 == 0, "exchange planes in float4 units");
  for (int k = t; k < 6 * kPL2 / 4; k += 256)
    reinterpret_cast<float4*>(lp)[k] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  if (t < 2) smax[t] = 0u;
  const float* xi = x1 + i * (S * S * C);
  const float* xj = (self ? x1 : x2) + j * (S * S * C);
  __syncthreads();
  {
    float mi = 0.0f, mj = 0.0f;
    for (int k = t; k < S * S * C; k += 256) {
      mi = fmaxf(mi, fabsf(xi[k]));
      mj = fmaxf(mj, fabsf(xj[k]));
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      mi = fmaxf(mi, __shfl_xor(mi, o));
      mj = fmaxf(mj, __shfl_xor(mj, o));
    }
    if (lane == 0) {
      atomicMax(&smax[0], __float_as_uint(mi));
      atomicMax(&smax[1], __float_as_uint(mj));
    }
  }
  __syncthreads();
  int ei, ej;
  frexpf(__uint_as_float(smax[0]), &ei);
  frexpf(__uint_as_float(smax[1]), &ej);
  const float scale_g = ldexpf(scale1, ei + ej);
  for (int k = t; k < 2 * kXS * kXS; k += 256) {
    const int img = k >= kXS * kXS;
    const int pos = k - img * (kXS * kXS);
    const int hh = pos / kXS - 1;
    const int ww = pos - (pos / kXS) * kXS - 1;
    const float sc = ldexpf(1.0f, img ? -ej : -ei);
    const float* src = (img ? xj : xi) + (hh * S + ww) * C;
    const bool in = (unsigned)hh < (unsigned)S && (unsigned)ww < (unsigned)S;
    half4v hv4, lv4;
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      const float v = (in && c < C) ? src[c] * sc : 0.0f;
      const _Float16 hv = (_Float16)v;
      hv4[c] = hv;
      lv4[c] = (_Float16)(v - (float)hv);
    }
    *reinterpret_cast<half4v*>(xh + ((img * 2) * (kXS * kXS) + pos) * 4) =
        hv4;
    *reinterpret_cast<half4v*>(xh + ((img * 2 + 1) * (kXS * kXS) + pos) *
                                        4) = lv4;
  }
  for (int k = t; k < kXhTail / 4; k += 256)
    *reinterpret_cast<half4v*>(xh + 2 * 2 * kXS * kXS * 4 + 4 * k) =
        half4v{0, 0, 0, 0};
  __syncthreads();
  constexpr int kRowB = kXS * 4 * 2;
  constexpr int kHlB = kXS * kXS * 4 * 2;
  int oa[2], ob[2];
#pragma unroll
  for (int st = 0; st < 2; ++st) {
    const int tx = st == 0 ? 0 : 2;
    oa[st] = c4 < 3 ? c4 * kRowB + (16 * (wave >> 1) + m16 + tx) * 8
                    : (st == 0 ? (kXS - 1) * 8 : 0);
    ob[st] = c4 < 3 ? c4 * kRowB + (16 * (wave & 1) + m16 + tx) * 8
                    : (st == 0 ? (kXS - 1) * 8 : 0);
  }
  const char* xhb = reinterpret_cast<const char*>(xh);
  // Operands of image row `row` (rows past the image belong to dropped
  // planes and read the zeroed tail of xh).
  auto opnd8 = [&](int img, int hl, int row) {
    const char* b = xhb + (img * 2 + hl) * kHlB + row * kRowB +
                    (img ? ob[0] : oa[0]);
    const half4v lo = *reinterpret_cast<const half4v*>(b);
    const half4v hi = *reinterpret_cast<const half4v*>(b + 8);
    return half8v{lo[0], lo[1], lo[2], lo[3], hi[0], hi[1], hi[2], hi[3]};
  };
  auto opnd4 = [&](int img, int hl, int row) {
    const char* b = xhb + (img * 2 + hl) * kHlB + row * kRowB +
                    (img ? ob[1] : oa[1]);
    const half4v lo = *reinterpret_cast<const half4v*>(b);
    const _Float16 z = 0;
    if (img == 1) return half8v{lo[0], lo[1], lo[2], lo[3], z, z, z, z};
    const half4v hi = *reinterpret_cast<const half4v*>(b + 8);
    return half8v{lo[0], lo[1], lo[2], lo[3], hi[0], hi[1], hi[2], hi[3]};
  };
  const float2* qi = qa1 + i * (S * S);
  const float2* qj = (self ? qa1 : qa2) + j * (S * S);
  const float2* ri = qb1 + i * (S * S);
  const float2* rj = (self ? qb1 : qb2) + j * (S * S);
  const float2* ui = qc1 + i * (S * S);
  const float2* uj = (self ? qc1 : qc2) + j * (S * S);
  auto ld_r = [&](const float2* tab, int row, int col) {
    row = row < 0 ? 0 : (row < S ? row : S - 1);
    return *reinterpret_cast<const float4*>(tab + row * S + col);
  };

  // One first-stage plane (hh, hp), this wave's quadrant (MFMA layout),
  // 2 pi times the first stage's state; zero outside the grid.
  auto s1 = [&](int hh, int hp, float (&vn)[4], float (&vt)[4]) {
    const bool ok = (unsigned)hh < (unsigned)S && (unsigned)hp < (unsigned)S;
    if (!ok) hh = hp = 0;
    floatx4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
    if (ok) {
      const half8v ah = opnd8(0, 0, hh), al = opnd8(0, 1, hh);
      const half8v bh = opnd8(1, 0, hp), bl = opnd8(1, 1, hp);
      acc = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah, bh, acc, 0, 0, 0);
      acc = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah, bl, acc, 0, 0, 0);
      acc = __builtin_amdgcn_mfma_f32_16x16x32_f16(al, bh, acc, 0, 0, 0);
      const half8v ch = opnd4(0, 0, hh), cl = opnd4(0, 1, hh);
      const half8v dh = opnd4(1, 0, hp), dl = opnd4(1, 1, hp);
      acc = __builtin_amdgcn_mfma_f32_16x16x32_f16(ch, dh, acc, 0, 0, 0);
      acc = __builtin_amdgcn_mfma_f32_16x16x32_f16(ch, dl, acc, 0, 0, 0);
      acc = __builtin_amdgcn_mfma_f32_16x16x32_f16(cl, dh, acc, 0, 0, 0);
    }
    const float4 q01 = *reinterpret_cast<const float4*>(qi + hh * S + w0);
    const float4 q23 = *reinterpret_cast<const float4*>(qi + hh * S + w0 + 2);
    const float2 qb = qj[hp * S + wq];
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const float4 Q = e < 2 ? q01 : q23;
      const float sa = (e & 1) ? Q.z : Q.x, ra = (e & 1) ? Q.w : Q.y;
      const float cv = fmaf(scale_g, acc[e], b1);
      const float rho = cv * (ra * qb.y);
      const float a = fminf(fabsf(rho), 1.0f);
      const float th = acos_pos5(a, 1.0f - a);
      const float pm = rho < 0.0f ? th : kPi - th;
      const float sn = fsqrt(fmaf(-a, a, 1.0f));
      vn[e] = ok ? fmaf(pm, rho, sn) * (sa * qb.x) : 0.0f;
      vt[e] = ok ? cv * pm : 0.0f;
    }
  };

  // Second-stage ReLU of one plane (row hh of image i, hp of image j) from
  // its pre-activation sums (three w-summed first-stage planes, nngp and
  // ntk) on the thread's 2x2 block; tables A (image i, row hh: (sq, rq) of
  // columns 2W, 2W + 1) and B (image j, row hp).  2 pi times the state,
  // zero outside the grid.
  const float wv2 = wv9b * kInv2Pi, wv3 = wv9c * kInv2Pi;
  auto relu2 = [&](bool ok, float4 A, float4 B, const float (&pn)[4],
                   const float (&pt)[4], float (&vn)[4], float (&vt)[4]) {
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int c = e >> 1, d = e & 1;
      const float sa = c ? A.z : A.x, ra = c ? A.w : A.y;
      const float sb = d ? B.z : B.x, rb = d ? B.w : B.y;
      const float cv = fmaf(wv2, pn[e], b2);
      const float rho = cv * (ra * rb);
      const float a = fminf(fabsf(rho), 1.0f);
      const float th = acos_pos5(a, 1.0f - a);
      const float pm = rho < 0.0f ? th : kPi - th;
      const float sn = fsqrt(fmaf(-a, a, 1.0f));
      vn[e] = ok ? fmaf(pm, rho, sn) * (sa * sb) : 0.0f;
      if constexpr (NTK) {
        const float tv = fmaf(wv2, pt[e], cv);
        vt[e] = ok ? tv * pm : 0.0f;
      } else {
        vt[e] = 0.0f;
      }
    }
  };
  // (w, w')-diagonal 3-tap sums of K planes held in the block layout.
  auto b_exchange = [&](auto kc, const float (*v)[4], float (*ws)[4]) {
    constexpr int K = decltype(kc)::value;
    __syncthreads();
#pragma unroll
    for (int k = 0; k < K; ++k) {
      float* d = lp + k * kPL2 + lofs;
      d[0] = v[k][0];
      d[1] = v[k][1];
      d[kLS2] = v[k][2];
      d[kLS2 + 1] = v[k][3];
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < K; ++k) read_diag3<32, kLS2>(lp + k * kPL2, lofs, ws[k]);
  };
  using K4 = std::integral_constant<int, 4>;
  using K6 = std::integral_constant<int, 6>;

  // Plane offsets: second-stage plane k of step H is (2H + ar[k],
  // 2Hp + br[k]), first-stage plane k is one row further, (2H + 1 + ar[k],
  // 2Hp + 1 + br[k]).  Diagonals: k = 0, 1 on 2D, 2, 3 on 2D + 1, 4, 5 on
  // 2D - 1.
  constexpr int ar[6] = {1, 2, 1, 2, 0, 1};
  constexpr int br[6] = {1, 2, 0, 1, 1, 2};

  const int nsweep = f12_nsweep(slot);
  for (int sw = 0; sw < nsweep; ++sw) {
  const int D = f12_diag(slot, sw);
  const int Hs = D > 0 ? D : 0;
  const int He = D < 0 ? H2 - 1 + D : H2 - 1;

  // Loop-carried state of the sweep (updated in place; the normaliser
  // tables are loaded where they are used, which keeps the carried state
  // within the register budget of two waves per SIMD).
  struct Carry {
    float c1[NT][6][4];      // w-summed first-stage planes (see above)
    float c2[NT][4][4];      // w-summed second-stage planes
    floatx4 acc[6];          // patch Grams of the step's six planes
  };
  Carry A;
  {
    // Sweep start (H = Hs, Hp = Hs - D): first-stage planes on diagonal 2D
    // rows 2Hs-2 .. 2Hs+1 (I0..I3), on 2D+1 rows 2Hs-1 .. 2Hs+1 (J0..J2),
    // on 2D-1 rows 2Hs-2 .. 2Hs (K0..K2); second-stage planes (2Hs-1,
    // 2Hps-1), (2Hs, 2Hps), (2Hs, 2Hps-1), (2Hs-1, 2Hps) from them.
    const int h0 = 2 * Hs, g0 = 2 * (Hs - D);
    const int hh[10] = {h0 - 2, h0 - 1, h0, h0 + 1,  h0 - 1, h0, h0 + 1,
                        h0 - 2, h0 - 1, h0};
    const int hp[10] = {g0 - 2, g0 - 1, g0, g0 + 1,  g0 - 2, g0 - 1, g0,
                        g0 - 1, g0, g0 + 1};
    float w1[NT][10][4];
#pragma unroll
    for (int half = 0; half < 2; ++half) {
      float vn[5][4], vt[5][4];
#pragma unroll
      for (int k = 0; k < 5; ++k) s1(hh[5 * half + k], hp[5 * half + k],
                                     vn[k], vt[k]);
      float ws[5][4];
      s1_exchange<5>(lp, vn, wr, lofs, ws);
#pragma unroll
      for (int k = 0; k < 5; ++k)
#pragma unroll
        for (int e = 0; e < 4; ++e) w1[0][5 * half + k][e] = ws[k][e];
      if constexpr (NTK) {
        s1_exchange<5>(lp, vt, wr, lofs, ws);
#pragma unroll
        for (int k = 0; k < 5; ++k)
#pragma unroll
          for (int e = 0; e < 4; ++e) w1[NT - 1][5 * half + k][e] = ws[k][e];
      }
    }
    // c1: I2, I3, J1, J2, K1, K2
    constexpr int cidx[6] = {2, 3, 5, 6, 8, 9};
#pragma unroll
    for (int tt = 0; tt < NT; ++tt)
#pragma unroll
      for (int k = 0; k < 6; ++k)
#pragma unroll
        for (int e = 0; e < 4; ++e) A.c1[tt][k][e] = w1[tt][cidx[k]][e];
    // second-stage planes: rows (2Hs-1, 2Hps-1) = I0+I1+I2, (2Hs, 2Hps) =
    // I1+I2+I3, (2Hs, 2Hps-1) = J0+J1+J2, (2Hs-1, 2Hps) = K0+K1+K2
    constexpr int sa0[4] = {0, 1, 4, 7};
    const int rh[4] = {h0 - 1, h0, h0, h0 - 1};
    const int rp[4] = {g0 - 1, g0, g0 - 1, g0};
    float pn[4][4], pt[4][4], vn[4][4], vt[4][4];
#pragma unroll
    for (int k = 0; k < 4; ++k)
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int b0 = sa0[k];
        pn[k][e] = w1[0][b0][e] + w1[0][b0 + 1][e] + w1[0][b0 + 2][e];
        pt[k][e] = w1[NT - 1][b0][e] + w1[NT - 1][b0 + 1][e] +
                   w1[NT - 1][b0 + 2][e];
      }
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const bool ok = (unsigned)rh[k] < (unsigned)S &&
                      (unsigned)rp[k] < (unsigned)S;
      relu2(ok, ld_r(ri, rh[k], 2 * W), ld_r(rj, rp[k], 2 * Wp), pn[k],
            pt[k], vn[k], vt[k]);
    }
    float ws[4][4];
    b_exchange(K4(), vn, ws);
#pragma unroll
    for (int k = 0; k < 4; ++k)
#pragma unroll
      for (int e = 0; e < 4; ++e) A.c2[0][k][e] = ws[k][e];
    if constexpr (NTK) {
      b_exchange(K4(), vt, ws);
#pragma unroll
      for (int k = 0; k < 4; ++k)
#pragma unroll
        for (int e = 0; e < 4; ++e) A.c2[NT - 1][k][e] = ws[k][e];
    }
  }

  auto ld_a = [&](int row, float4& lo, float4& hi) {
    row = row < S ? row : S - 1;
    lo = *reinterpret_cast<const float4*>(qi + row * S + w0);
    hi = *reinterpret_cast<const float4*>(qi + row * S + w0 + 2);
  };
  auto ld_b = [&](int row) { return qj[(row < S ? row : S - 1) * S + wq]; };
  // Patch Grams of the six first-stage planes of step Hn: image rows
  // 2Hn + 1 + r (r = 0..2) of i and 2Hpn + 1 + r of j.
  auto gram6 = [&](int Hn, floatx4 (&acc)[6]) {
    const int Hpn = Hn - D;
#pragma unroll
    for (int k = 0; k < 6; ++k) acc[k] = {0.0f, 0.0f, 0.0f, 0.0f};
    {
      half8v ah[3], al[3], bh[3], bl[3];
#pragma unroll
      for (int r = 0; r < 3; ++r) {
        ah[r] = opnd8(0, 0, 2 * Hn + 1 + r);
        bh[r] = opnd8(1, 0, 2 * Hpn + 1 + r);
      }
#pragma unroll
      for (int k = 0; k < 6; ++k)
        acc[k] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah[ar[k]], bh[br[k]],
                                                        acc[k], 0, 0, 0);
#pragma unroll
      for (int r = 0; r < 3; ++r) bl[r] = opnd8(1, 1, 2 * Hpn + 1 + r);
#pragma unroll
      for (int k = 0; k < 6; ++k)
        acc[k] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah[ar[k]], bl[br[k]],
                                                        acc[k], 0, 0, 0);
#pragma unroll
      for (int r = 0; r < 3; ++r) al[r] = opnd8(0, 1, 2 * Hn + 1 + r);
#pragma unroll
      for (int k = 0; k < 6; ++k)
        acc[k] = __builtin_amdgcn_mfma_f32_16x16x32_f16(al[ar[k]], bh[br[k]],
                                                        acc[k], 0, 0, 0);
    }
    {
      half8v ah[3], al[3], bh[3], bl[3];
#pragma unroll
      for (int r = 0; r < 3; ++r) {
        ah[r] = opnd4(0, 0, 2 * Hn + 1 + r);
        bh[r] = opnd4(1, 0, 2 * Hpn + 1 + r);
      }
#pragma unroll
      for (int k = 0; k < 6; ++k)
        acc[k] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah[ar[k]], bh[br[k]],
                                                        acc[k], 0, 0, 0);
#pragma unroll
      for (int r = 0; r < 3; ++r) bl[r] = opnd4(1, 1, 2 * Hpn + 1 + r);
#pragma unroll
      for (int k = 0; k < 6; ++k)
        acc[k] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah[ar[k]], bl[br[k]],
                                                        acc[k], 0, 0, 0);
#pragma unroll
      for (int r = 0; r < 3; ++r) al[r] = opnd4(0, 1, 2 * Hn + 1 + r);
#pragma unroll
      for (int k = 0; k < 6; ++k)
        acc[k] = __builtin_amdgcn_mfma_f32_16x16x32_f16(al[ar[k]], bh[br[k]],
                                                        acc[k], 0, 0, 0);
    }
  };
  gram6(Hs, A.acc);

  // First-stage ReLU of the step's planes (MASK: some plane is outside the
  // grid).
  auto relu1 = [&](const Carry& in, int H, auto mask, float (&vn)[6][4],
                   float (&vt)[6][4]) {
    constexpr bool MASK = decltype(mask)::value;
    const int Hp = H - D;
    float4 qal[3], qah[3];
    float2 qbv[3];
#pragma unroll
    for (int r = 0; r < 3; ++r) {
      ld_a(2 * H + 1 + r, qal[r], qah[r]);
      qbv[r] = ld_b(2 * Hp + 1 + r);
    }
#pragma unroll
    for (int k = 0; k < 6; ++k) {
      const bool ok = 2 * H + 1 + ar[k] < S && 2 * Hp + 1 + br[k] < S;
      const float2 qb = qbv[br[k]];
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const float4 Q = e < 2 ? qal[ar[k]] : qah[ar[k]];
        const float sa = (e & 1) ? Q.z : Q.x, ra = (e & 1) ? Q.w : Q.y;
        const float cv = fmaf(scale_g, in.acc[k][e], b1);
        const float rho = cv * (ra * qb.y);
        const float a = fminf(fabsf(rho), 1.0f);
        const float th = acos_pos5(a, 1.0f - a);
        const float pm = rho < 0.0f ? th : kPi - th;
        const float sn = fsqrt(fmaf(-a, a, 1.0f));
        const float n = fmaf(pm, rho, sn) * (sa * qb.x);
        const float tv = cv * pm;
        vn[k][e] = MASK ? (ok ? n : 0.0f) : n;
        vt[k][e] = MASK ? (ok ? tv : 0.0f) : tv;
      }
    }
  };

  auto step = [&](int H, Carry& cy) {
    const int Hp = H - D;
    float vn[6][4], vt[6][4];
    if (2 * H + 3 < S && 2 * Hp + 3 < S)
      relu1(cy, H, std::false_type(), vn, vt);
    else
      relu1(cy, H, std::true_type(), vn, vt);

    // Stage 1 -> 2: w-sums of the new planes, h-window sums with the
    // carried ones, then the second ReLU on the six second-stage planes.
    float p2[NT][6][4];
    auto adv1 = [&](const float (&v)[6][4], int tt) {
      float ws[6][4];
      s1_exchange<6>(lp, v, wr, lofs, ws);
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        p2[tt][0][e] = cy.c1[tt][0][e] + cy.c1[tt][1][e] + ws[0][e];
        p2[tt][1][e] = cy.c1[tt][1][e] + ws[0][e] + ws[1][e];
        p2[tt][2][e] = cy.c1[tt][2][e] + cy.c1[tt][3][e] + ws[2][e];
        p2[tt][3][e] = cy.c1[tt][3][e] + ws[2][e] + ws[3][e];
        p2[tt][4][e] = cy.c1[tt][4][e] + cy.c1[tt][5][e] + ws[4][e];
        p2[tt][5][e] = cy.c1[tt][5][e] + ws[4][e] + ws[5][e];
#pragma unroll
        for (int k = 0; k < 6; ++k) cy.c1[tt][k][e] = ws[k][e];
      }
    };
    adv1(vn, 0);
    if constexpr (NTK) adv1(vt, NT - 1);
    float v2n[6][4], v2t[6][4];
    {
      float4 r2i[3], r2j[3];
#pragma unroll
      for (int r = 0; r < 3; ++r) {
        r2i[r] = ld_r(ri, 2 * H + r, 2 * W);
        r2j[r] = ld_r(rj, 2 * Hp + r, 2 * Wp);
      }
#pragma unroll
      for (int k = 0; k < 6; ++k) {
        const bool ok = 2 * H + ar[k] < S && 2 * Hp + br[k] < S;
        relu2(ok, r2i[ar[k]], r2j[br[k]], p2[0][k], p2[NT - 1][k], v2n[k],
              v2t[k]);
      }
    }
    // Stage 2 -> 3: w-sums, h-window sums: P00, P11, P10, P01.
    float pre[NT][4][4];
    auto adv2 = [&](const float (*v)[4], int tt) {
      float ws[6][4];
      b_exchange(K6(), v, ws);
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const float z0 = cy.c2[tt][1][e] + ws[0][e];
        pre[tt][0][e] = cy.c2[tt][0][e] + z0;
        pre[tt][1][e] = z0 + ws[1][e];
        pre[tt][2][e] = cy.c2[tt][2][e] + ws[2][e] + ws[3][e];
        pre[tt][3][e] = cy.c2[tt][3][e] + ws[4][e] + ws[5][e];
        cy.c2[tt][0][e] = ws[0][e];
        cy.c2[tt][1][e] = ws[1][e];
        cy.c2[tt][2][e] = ws[3][e];
        cy.c2[tt][3][e] = ws[5][e];
      }
    };
    adv2(v2n, 0);
    if constexpr (NTK) adv2(v2t, NT - 1);
    // The next step's matrix work overlaps the third stage (past the sweep
    // end it computes planes that are never used).
    gram6(H + 1, cy.acc);
    // Third stage + pool, as the second stage of k_first_pool32.
    float4 r3i[2], r3j[2];
#pragma unroll
    for (int r = 0; r < 2; ++r) {
      r3i[r] = ld_r(ui, 2 * H + r, 2 * W);
      r3j[r] = ld_r(uj, 2 * Hp + r, 2 * Wp);
    }
    float acc_n = 0.0f, acc_t = 0.0f;
#pragma unroll
    for (int x = 0; x < 4; ++x) {
      const float4 Ai = r3i[x >> 1];
      const float sa = (x & 1) ? Ai.z : Ai.x, ra = (x & 1) ? Ai.w : Ai.y;
      float inner = 0.0f;
#pragma unroll
      for (int y = 0; y < 4; ++y) {
        const float4 Bj = r3j[y >> 1];
        const float sb = (y & 1) ? Bj.z : Bj.x, rb = (y & 1) ? Bj.w : Bj.y;
        const int pl = (x >> 1) == (y >> 1) ? (x >> 1) : 2 + (y >> 1);
        const int e = 2 * (x & 1) + (y & 1);
        const float cv = fmaf(wv3, pre[0][pl][e], b3);
        const float tv = NTK ? fmaf(wv3, pre[NT - 1][pl][e], cv) : 0.0f;
        const float rho = cv * (ra * rb);
        const float a = fminf(fabsf(rho), 1.0f);
        const float th = acos_pos5(a, 1.0f - a);
        const float pm = rho < 0.0f ? th : kPi - th;
        const float sn = fsqrt(fmaf(-a, a, 1.0f));
        inner = fmaf(fmaf(pm, rho, sn), sb, inner);
        if constexpr (NTK) acc_t = fmaf(pm, tv, acc_t);
      }
      acc_n = fmaf(inner, sa, acc_n);
    }
    const long long o =
        p * PER_OUT + ((long long)(H * SO + Hp) * SO + W) * SO + Wp;
    constexpr float kScale = kInv2Pi / 16.0f;
    st_state(out_n + o, acc_n * kScale);
    if constexpr (NTK) st_state(out_t + o, acc_t * kScale);
  };

  for (int H = Hs; H <= He; ++H) step(H, A);
  }
}

__global__ __launch_bounds__(256) void k_qnorm(const float* __restrict__ q,
                                               long long n,
                                               float2* __restrict__ out) {
  for (long long k = (long long)blockIdx.x * 256 + threadIdx.x; k < n;
       k += (long long)gridDim.x * 256)
    out[k] = qnorm(q[k]);
}

// ---------------------------------------------------------------- launchers

static inline int pad8(long long n) { return (int)((n + 7) / 8 * 8); }

template <int S, int POOL, int PF, int SL = 0>
static int launch_stage_s(const float* in_n, const float* in_t, int n1,
                          int n2, bool ntk, float wv9, float b_var,
                          const float* q1, const float* q2, float* out_n,
                          float* out_t, hipStream_t st,
                          float* pre_n = nullptr, float* pre_t = nullptr) {
  static_assert(SL == 0 || (S == 16 && POOL == 2 && SL == 4),
                "diagonal slots are laid out for the 16x16 pooled stage");
  constexpr int JPW = 256 / ((S / 2) * (S / 2));
  const long long T = n2 == 0 ? (long long)n1 : (long long)n1 * n2;
  const int ndiag = SL ? SL : (POOL == 2 ? S - 1 : 2 * S - 1);
  const long long ngroups = (T + JPW - 1) / JPW;
  if (ngroups * ndiag > 0x7ffffff0LL)
    return set_error(NTA_EINVAL, "nta_stage: tile too large (%lld pairs)", T);
  const int nblk = pad8(ngroups * ndiag);
  if constexpr (POOL == 2) {
    if (pre_n) {
      if (ntk)
        hipLaunchKernelGGL((k_stage_fast<S, POOL, true, PF, SL, true>),
                           dim3(nblk), dim3(256), 0, st, in_n, in_t, n1, n2,
                           ndiag, (int)ngroups, nblk, wv9, b_var, q1, q2,
                           out_n, out_t, pre_n, pre_t);
      else
        hipLaunchKernelGGL((k_stage_fast<S, POOL, false, PF, SL, true>),
                           dim3(nblk), dim3(256), 0, st, in_n, in_t, n1, n2,
                           ndiag, (int)ngroups, nblk, wv9, b_var, q1, q2,
                           out_n, out_t, pre_n, pre_t);
      return check_launch("nta_stage_keep_pre");
    }
  }
  if (ntk)
    hipLaunchKernelGGL((k_stage_fast<S, POOL, true, PF, SL>), dim3(nblk),
                       dim3(256),
                       0, st, in_n, in_t, n1, n2, ndiag, (int)ngroups, nblk,
                       wv9, b_var, q1, q2, out_n, out_t, nullptr, nullptr);
  else
    hipLaunchKernelGGL((k_stage_fast<S, POOL, false, PF, SL>), dim3(nblk),
                       dim3(256),
                       0, st, in_n, in_t, n1, n2, ndiag, (int)ngroups, nblk,
                       wv9, b_var, q1, q2, out_n, out_t, nullptr, nullptr);
  return check_launch("nta_stage");
}

// Plane fetch schedule of the 32x32 pooled stage (see k_stage_fast).
#ifndef NTA_STAGE_PREFETCH
#define NTA_STAGE_PREFETCH 2
#endif
constexpr int kPrefetch = NTA_STAGE_PREFETCH;
// 16x16: loads at the step that uses them (three workgroups per CU hide
// them better than a register prefetch with two).
#ifndef NTA_STAGE16_PREFETCH
#define NTA_STAGE16_PREFETCH 0
#endif
constexpr int kPrefetch16 = NTA_STAGE16_PREFETCH;
// 16x16: four workgroups per group of pairs, each sweeping several
// diagonals (0: one diagonal per workgroup).
#ifndef NTA_STAGE16_SLOTS
#define NTA_STAGE16_SLOTS 4
#endif
constexpr int kSlots16 = NTA_STAGE16_SLOTS;

// The fused kernels implement the ReLU map only (k1 = 1, k0 = 0); other
// ABRelu slopes run the generic kernels.
int launch_stage_fast(const float* in_nngp, const float* in_ntk, int n1,
                      int n2, int S, int pool, int flags, float w_var,
                      float b_var, float k1, float k0, const float* q1,
                      const float* q2, float* out_nngp, float* out_ntk,
                      hipStream_t stream) {
  if (k1 != 1.0f || k0 != 0.0f ||
      (flags & (NTA_ACT_ERF | NTA_ACT_GELU | NTA_CONV_CIRCULAR)))
    return kNotHandled;
  const bool ntk = flags & NTA_NTK;
  const float wv9 = w_var / 9.0f;
  if (pool == 2) {
    switch (S) {
      case 32: return launch_stage_s<32, 2, kPrefetch>(in_nngp, in_ntk, n1, n2, ntk, wv9,
                                            b_var, q1, q2, out_nngp, out_ntk,
                                            stream);
      case 16: return launch_stage_s<16, 2, kPrefetch16, kSlots16>(in_nngp, in_ntk, n1, n2, ntk, wv9,
                                            b_var, q1, q2, out_nngp, out_ntk,
                                            stream);
      case 8: return launch_stage_s<8, 2, 1>(in_nngp, in_ntk, n1, n2, ntk, wv9,
                                          b_var, q1, q2, out_nngp, out_ntk,
                                          stream);
    }
  } else if (pool == 1) {
    switch (S) {
      case 32: return launch_stage_s<32, 1, 0>(in_nngp, in_ntk, n1, n2, ntk, wv9,
                                            b_var, q1, q2, out_nngp, out_ntk,
                                            stream);
      case 16: return launch_stage_s<16, 1, 0>(in_nngp, in_ntk, n1, n2, ntk, wv9,
                                            b_var, q1, q2, out_nngp, out_ntk,
                                            stream);
      case 8: return launch_stage_s<8, 1, 0>(in_nngp, in_ntk, n1, n2, ntk, wv9,
                                          b_var, q1, q2, out_nngp, out_ntk,
                                          stream);
    }
  }
  return kNotHandled;
}

int launch_first_fast(const float* x1, const float* x2, int n1, int n2, int S,
                      int C, int flags, float w_var, float b_var, float k1,
                      float k0, const float* q1, const float* q2,
                      float* nngp, float* ntk, hipStream_t stream) {
  if (k1 != 1.0f || k0 != 0.0f ||
      (flags & (NTA_ACT_ERF | NTA_ACT_GELU | NTA_CONV_CIRCULAR)))
    return kNotHandled;
  if (C > 4 || (S != 32 && S != 16)) return kNotHandled;
  const bool relu = flags & NTA_RELU;
  const bool want_ntk = flags & NTA_NTK;
  if (!relu && (want_ntk || S != 32)) return kNotHandled;
  const long long T = n2 == 0 ? (long long)n1 : (long long)n1 * n2;
  if (T * S > 0x7ffffff0LL)
    return set_error(NTA_EINVAL, "nta_first_stage: tile too large (%lld "
                     "pairs)", T);
  const int nblk = pad8(T * S);
  const float scale = w_var / (9.0f * (float)C);
  if (!relu) {
    hipLaunchKernelGGL((k_first32<false, false>), dim3(nblk), dim3(256), 0,
                       stream, x1, x2, n1, n2, C, nblk, scale, b_var, q1, q2,
                       nngp, ntk);
    return check_launch("nta_first_stage");
  }
  if (S == 32) {
    if (want_ntk)
      hipLaunchKernelGGL(k_first32<true>, dim3(nblk), dim3(256), 0, stream,
                         x1, x2, n1, n2, C, nblk, scale, b_var, q1, q2, nngp,
                         ntk);
    else
      hipLaunchKernelGGL(k_first32<false>, dim3(nblk), dim3(256), 0, stream,
                         x1, x2, n1, n2, C, nblk, scale, b_var, q1, q2, nngp,
                         ntk);
  } else {
    if (want_ntk)
      hipLaunchKernelGGL(k_first16<true>, dim3(nblk), dim3(256), 0, stream,
                         x1, x2, n1, n2, C, nblk, scale, b_var, q1, q2, nngp,
                         ntk);
    else
      hipLaunchKernelGGL(k_first16<false>, dim3(nblk), dim3(256), 0, stream,
                         x1, x2, n1, n2, C, nblk, scale, b_var, q1, q2, nngp,
                         ntk);
  }
  return check_launch("nta_first_stage");
}

// The pooled stage of launch_stage_fast that also writes the inputs'
// stencil means (pre_n, pre_t).
static int launch_stage_keep_pre(const float* in_n, const float* in_t,
                                 int n1, int n2, int S, bool ntk, float w_var,
                                 float b_var, const float* q1,
                                 const float* q2, float* out_n, float* out_t,
                                 float* pre_n, float* pre_t, hipStream_t st) {
  const float wv9 = w_var / 9.0f;
  switch (S) {
    case 32: return launch_stage_s<32, 2, kPrefetch>(
        in_n, in_t, n1, n2, ntk, wv9, b_var, q1, q2, out_n, out_t, st, pre_n,
        pre_t);
    case 16: return launch_stage_s<16, 2, kPrefetch16, kSlots16>(
        in_n, in_t, n1, n2, ntk, wv9, b_var, q1, q2, out_n, out_t, st, pre_n,
        pre_t);
    case 8: return launch_stage_s<8, 2, 1>(
        in_n, in_t, n1, n2, ntk, wv9, b_var, q1, q2, out_n, out_t, st, pre_n,
        pre_t);
  }
  return kNotHandled;
}

}  // namespace nta

using namespace nta;

// nta_stage for Conv(3x3 SAME) -> ReLU -> AvgPool(2x2 / 2) on S x S states
// (S = 8, 16, 32), which also writes the stencil means of the inputs at
// every pre-pool position (the pre-activations before w_var, b_var) for the
// stage's adjoint.
extern "C" int nta_stage_keep_pre(const float* in_nngp, const float* in_ntk,
                                  int n1, int n2, int S, int flags,
                                  float w_var, float b_var, const float* q1,
                                  const float* q2, float* out_nngp,
                                  float* out_ntk, float* pre_nngp,
                                  float* pre_ntk, void* stream) {
  if (n1 < 0 || n2 < 0)
    return set_error(NTA_EINVAL, "nta_stage_keep_pre: bad n1=%d n2=%d", n1,
                     n2);
  if ((S != 8 && S != 16 && S != 32) || !(flags & NTA_CONV) ||
      !(flags & NTA_RELU) ||
      (flags & (NTA_CONV_VALID | NTA_CONV_CIRCULAR | NTA_ACT_ERF |
                NTA_ACT_GELU)))
    return set_error(NTA_EUNSUPPORTED, "nta_stage_keep_pre: needs a 3x3 "
                     "SAME Conv + Relu stage on 8x8, 16x16 or 32x32 (S=%d, "
                     "flags=%d)", S, flags);
  if (n1 == 0) return NTA_OK;
  const bool ntk = flags & NTA_NTK;
  if (!in_nngp || !out_nngp || !pre_nngp || !q1 || (n2 > 0 && !q2) ||
      (ntk && (!in_ntk || !out_ntk || !pre_ntk)))
    return set_error(NTA_EINVAL, "nta_stage_keep_pre: null pointer for a "
                     "requested operand");
  return launch_stage_keep_pre(in_nngp, in_ntk, n1, n2, S, ntk, w_var, b_var,
                               q1, q2, out_nngp, out_ntk, pre_nngp, pre_ntk,
                               (hipStream_t)stream);
}

// Conv(3x3 SAME) -> ReLU -> Conv(3x3 SAME) -> ReLU -> AvgPool(2x2 / 2) on
// 32x32 images with C <= 4 channels, as one kernel (k_first_pool32); with
// pre_nngp given, the second stage's stencil means go there too.
static int first_pool_stage(const char* name, const float* x1,
                            const float* x2, int n1, int n2, int H, int W,
                            int C, int flags, float w_var1, float b_var1,
                            const float* qn1a, const float* qn2a,
                            float w_var2, float b_var2, const float* qn1b,
                            const float* qn2b, float* out_nngp,
                            float* out_ntk, float* pre_nngp, float* pre_ntk,
                            bool keep_pre, void* stream) {
  if (n1 < 0 || n2 < 0)
    return set_error(NTA_EINVAL, "%s: bad n1=%d n2=%d", name, n1, n2);
  if (H != 32 || W != 32 || C < 1 || C > 4)
    return set_error(NTA_EUNSUPPORTED, "%s: needs 32x32 images with 1..4 "
                     "channels (got %dx%dx%d)", name, H, W, C);
  if (n1 == 0) return NTA_OK;
  const bool ntk = flags & NTA_NTK;
  if (!x1 || (n2 > 0 && !x2) || !qn1a || !qn1b ||
      (n2 > 0 && (!qn2a || !qn2b)) || !out_nngp || (ntk && !out_ntk) ||
      (keep_pre && (!pre_nngp || (ntk && !pre_ntk))))
    return set_error(NTA_EINVAL, "%s: null pointer for a requested operand",
                     name);
  const long long T = n2 == 0 ? (long long)n1 : (long long)n1 * n2;
  if (T * kF12Slots > 0x7ffffff0LL)
    return set_error(NTA_EINVAL, "%s: tile too large (%lld pairs)", name, T);
  const int nblk = (int)((T * kF12Slots + 7) / 8 * 8);
  const float scale1 = w_var1 / (9.0f * (float)C);
  const float wv9 = w_var2 / 9.0f;
  const float2* a1 = reinterpret_cast<const float2*>(qn1a);
  const float2* a2 = reinterpret_cast<const float2*>(qn2a);
  const float2* c1 = reinterpret_cast<const float2*>(qn1b);
  const float2* c2 = reinterpret_cast<const float2*>(qn2b);
  const hipStream_t st = (hipStream_t)stream;
#define NTA_F12_LAUNCH(NTK_, KP_)                                           \
  hipLaunchKernelGGL((k_first_pool32<NTK_, KP_>), dim3(nblk), dim3(256), 0, \
                     st, x1, x2, n1, n2, C, nblk, scale1, b_var1, a1, a2,   \
                     wv9, b_var2, c1, c2, out_nngp, out_ntk, pre_nngp,      \
                     pre_ntk)
  if (keep_pre) {
    if (ntk) NTA_F12_LAUNCH(true, true);
    else NTA_F12_LAUNCH(false, true);
  } else {
    if (ntk) NTA_F12_LAUNCH(true, false);
    else NTA_F12_LAUNCH(false, false);
  }
#undef NTA_F12_LAUNCH
  return check_launch(name);
}

extern "C" int nta_first_pool_stage(
    const float* x1, const float* x2, int n1, int n2, int H, int W, int C,
    int flags, float w_var1, float b_var1, const float* qn1a,
    const float* qn2a, float w_var2, float b_var2, const float* qn1b,
    const float* qn2b, float* out_nngp, float* out_ntk, void* stream) {
  return first_pool_stage("nta_first_pool_stage", x1, x2, n1, n2, H, W, C,
                          flags, w_var1, b_var1, qn1a, qn2a, w_var2, b_var2,
                          qn1b, qn2b, out_nngp, out_ntk, nullptr, nullptr,
                          false, stream);
}

extern "C" int nta_first_pool_stage_keep_pre(
    const float* x1, const float* x2, int n1, int n2, int H, int W, int C,
    int flags, float w_var1, float b_var1, const float* qn1a,
    const float* qn2a, float w_var2, float b_var2, const float* qn1b,
    const float* qn2b, float* out_nngp, float* out_ntk, float* pre_nngp,
    float* pre_ntk, void* stream) {
  return first_pool_stage("nta_first_pool_stage_keep_pre", x1, x2, n1, n2, H,
                          W, C, flags, w_var1, b_var1, qn1a, qn2a, w_var2,
                          b_var2, qn1b, qn2b, out_nngp, out_ntk, pre_nngp,
                          pre_ntk, true, stream);
}

// Conv(3x3 SAME) -> ReLU, three times, -> AvgPool(2x2 / 2) on 32x32 images
// with C <= 4 channels, as one kernel (k_first3_pool32).
extern "C" int nta_first3_pool_stage(
    const float* x1, const float* x2, int n1, int n2, int H, int W, int C,
    int flags, float w_var1, float b_var1, const float* qn1a,
    const float* qn2a, float w_var2, float b_var2, const float* qn1b,
    const float* qn2b, float w_var3, float b_var3, const float* qn1c,
    const float* qn2c, float* out_nngp, float* out_ntk, void* stream) {
  if (n1 < 0 || n2 < 0)
    return set_error(NTA_EINVAL, "nta_first3_pool_stage: bad n1=%d n2=%d",
                     n1, n2);
  if (H != 32 || W != 32 || C < 1 || C > 4)
    return set_error(NTA_EUNSUPPORTED, "nta_first3_pool_stage: needs 32x32 "
                     "images with 1..4 channels (got %dx%dx%d)", H, W, C);
  if (n1 == 0) return NTA_OK;
  const bool ntk = flags & NTA_NTK;
  if (!x1 || (n2 > 0 && !x2) || !qn1a || !qn1b || !qn1c ||
      (n2 > 0 && (!qn2a || !qn2b || !qn2c)) || !out_nngp ||
      (ntk && !out_ntk))
    return set_error(NTA_EINVAL, "nta_first3_pool_stage: null pointer for a "
                     "requested operand");
  const long long T = n2 == 0 ? (long long)n1 : (long long)n1 * n2;
  if (T * kF12Slots > 0x7ffffff0LL)
    return set_error(NTA_EINVAL, "nta_first3_pool_stage: tile too large "
                     "(%lld pairs)", T);
  const int nblk = (int)((T * kF12Slots + 7) / 8 * 8);
  const float scale1 = w_var1 / (9.0f * (float)C);
  const float wv9b = w_var2 / 9.0f, wv9c = w_var3 / 9.0f;
  auto f2 = [](const float* q) { return reinterpret_cast<const float2*>(q); };
  if (ntk)
    hipLaunchKernelGGL(k_first3_pool32<true>, dim3(nblk), dim3(256), 0,
                       (hipStream_t)stream, x1, x2, n1, n2, C, nblk, scale1,
                       b_var1, f2(qn1a), f2(qn2a), wv9b, b_var2, f2(qn1b),
                       f2(qn2b), wv9c, b_var3, f2(qn1c), f2(qn2c), out_nngp,
                       out_ntk);
  else
    hipLaunchKernelGGL(k_first3_pool32<false>, dim3(nblk), dim3(256), 0,
                       (hipStream_t)stream, x1, x2, n1, n2, C, nblk, scale1,
                       b_var1, f2(qn1a), f2(qn2a), wv9b, b_var2, f2(qn1b),
                       f2(qn2b), wv9c, b_var3, f2(qn1c), f2(qn2c), out_nngp,
                       out_ntk);
  return check_launch("nta_first3_pool_stage");
}

// (sqrt(q), 1 / sqrt(q)) per element, (0, 0) where q <= 0.
extern "C" int nta_qnorm(const float* q, long long n, float* out,
                         void* stream) {
  if (n < 0 || (n > 0 && (!q || !out)))
    return set_error(NTA_EINVAL, "nta_qnorm: bad n=%lld or null pointer", n);
  if (n == 0) return NTA_OK;
  const long long blocks = (n + 255) / 256;
  const unsigned grid = (unsigned)(blocks < 65536 ? blocks : 65536);
  hipLaunchKernelGGL(k_qnorm, dim3(grid), dim3(256), 0, (hipStream_t)stream,
                     q, n, reinterpret_cast<float2*>(out));
  return check_launch("nta_qnorm");
}
