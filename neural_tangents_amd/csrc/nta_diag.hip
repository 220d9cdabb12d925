// Diagonal-spatial path: networks of [Conv] -> [Relu] stages without
// pooling, read out by Flatten (+ Dense).  Flatten needs only the spatial
// diagonal of the kernel (h = h', w = w'), and on that diagonal the shared-
// offset Conv stencil is a plain fh x fw window mean
// (_conv_kernel_diagonal_spatial, stax/linear.py:3381-3437) and the ReLU map
// pairs q1[i, h, w] with q2[j, h, w] (requirements.py:1077-1117 with
// diagonal_spatial), so each pair carries an [H, W] state instead of
// [H, H, W, W].
//
// One workgroup runs the whole chain of one pair with its (nngp, ntk) state
// in LDS; only the images, the per-image normalisers and the two head values
// touch HBM.  Self chains (n2 == 0) record the pre-activation of every ReLU
// stage as the normaliser q of the image.
#include "nta_common.h"

namespace nta {

constexpr int kDiagMaxStages = 48;    // the stage table travels in kernargs
constexpr int kDiagMaxHW = 3840;       // 4 planes + reduction <= 64 KiB LDS

struct DiagStage {
  int conv, fh, fw, lo_h, lo_w, hin, win, hout, wout, relu, act, circ;
  float w_var, b_var, k1, k0;
};
struct DiagPlan {
  int nstage, H, W, C, ntk;
  float a, b, c, d, e;
  long long qslot1, qslot2;        // floats per stage slot of q1, q2
  DiagStage s[kDiagMaxStages];
};

__device__ __forceinline__ void decode_upper(long long p, int n, int& i,
                                             int& j) {
  // p-th entry of the row-major upper triangle (j >= i) of an n x n matrix.
  const long long T = (long long)n * (n + 1) / 2;
  const long long r = T - 1 - p;               // lower-triangle index
  long long row = (long long)((sqrt(8.0 * (double)r + 1.0) - 1.0) * 0.5);
  while ((row + 1) * (row + 2) / 2 <= r) ++row;
  while (row * (row + 1) / 2 > r) --row;
  const long long col = r - row * (row + 1) / 2;
  i = n - 1 - (int)row;
  j = n - 1 - (int)col;
}

__global__ __launch_bounds__(256) void k_diag_chain(
    DiagPlan P, const float* __restrict__ x1, const float* __restrict__ x2,
    int n1, int n2, int sym, float* __restrict__ q1,
    const float* __restrict__ q2, float* __restrict__ out_n,
    float* __restrict__ out_t, long long ld, long long npairs) {
  // [sn0 | sn1 | st0 | st1] planes of H * W floats, then the reduction.
  extern __shared__ float dyn[];
  const int HWp = P.H * P.W;
  float* sn[2] = {dyn, dyn + HWp};
  float* st[2] = {dyn + 2 * HWp, dyn + 3 * HWp};
  float* red0 = dyn + 4 * HWp;
  float* red1 = red0 + 256;
  const int t = threadIdx.x;
  const bool self = n2 == 0;
  const bool ntk = P.ntk && !self;
  for (long long p = blockIdx.x; p < npairs; p += gridDim.x) {
    int i, j;
    if (self) {
      i = j = (int)p;
    } else if (sym) {
      decode_upper(p, n1, i, j);
    } else {
      i = (int)(p / n2);
      j = (int)(p - (long long)i * n2);
    }
    // Input covariance on the diagonal: (1/C) sum_c x1[i,h,w,c] x2[j,h,w,c]
    // (requirements.py:556-582).
    const int HW = P.H * P.W;
    const float* xa = x1 + (long long)i * HW * P.C;
    const float* xb = (self ? x1 : x2) + (long long)j * HW * P.C;
    const float invc = 1.0f / (float)P.C;
    for (int e = t; e < HW; e += 256) {
      float acc = 0.0f;
      for (int c = 0; c < P.C; ++c)
        acc = fmaf(xa[(long long)e * P.C + c], xb[(long long)e * P.C + c],
                   acc);
      sn[0][e] = acc * invc;
      st[0][e] = 0.0f;
    }
    __syncthreads();
    int cur = 0;
    bool have_t = false;               // the ntk state is 0 before a Conv
    for (int si = 0; si < P.nstage; ++si) {
      const DiagStage S = P.s[si];
      const int nout = S.hout * S.wout;
      const float inv_f = 1.0f / (float)(S.fh * S.fw);
      for (int e = t; e < nout; e += 256) {
        const int h = e / S.wout, w = e - (e / S.wout) * S.wout;
        float cv, tv = 0.0f;
        if (S.conv) {
          float an = 0.0f, at = 0.0f;
          for (int a = 0; a < S.fh; ++a) {
            int hh = h + a - S.lo_h;
            if (S.circ) hh = ((hh % S.hin) + S.hin) % S.hin;
            if ((unsigned)hh >= (unsigned)S.hin) continue;
            for (int b = 0; b < S.fw; ++b) {
              int ww = w + b - S.lo_w;
              if (S.circ) ww = ((ww % S.win) + S.win) % S.win;
              if ((unsigned)ww >= (unsigned)S.win) continue;
              an += sn[cur][hh * S.win + ww];
              if (ntk && have_t) at += st[cur][hh * S.win + ww];
            }
          }
          // nngp = W^2 conv(nngp) + b^2; ntk = W^2 conv(ntk) + nngp
          // (linear.py:1396-1398)
          cv = fmaf(an * inv_f, S.w_var, S.b_var);
          if (ntk) tv = fmaf(S.w_var, at * inv_f, cv);
        } else {
          cv = sn[cur][e];
          if (ntk) tv = have_t ? st[cur][e] : 0.0f;
        }
        if (S.relu) {
          float qa, qb;
          if (self) {
            qa = qb = cv;
            q1[si * P.qslot1 + (long long)i * nout + e] = cv;
          } else {
            qa = q1[si * P.qslot1 + (long long)i * nout + e];
            qb = q2[si * P.qslot2 + (long long)j * nout + e];
          }
          float ds;
          // elementwise.py:444-455
          const float nv = act_map(S.act, cv, qa, qb, S.k1, S.k0, &ds);
          tv *= ds;
          cv = nv;
        }
        sn[cur ^ 1][e] = cv;
        if (ntk) st[cur ^ 1][e] = tv;
      }
      __syncthreads();
      cur ^= 1;
      if (S.conv) have_t = true;
    }
    // Flatten: mean over the spatial diagonal (linear.py:1868-1899), then
    // the composed Dense affine map.
    const DiagStage& L = P.s[P.nstage - 1];
    const int nout = L.hout * L.wout;
    float mn = 0.0f, mt = 0.0f;
    for (int e = t; e < nout; e += 256) {
      mn += sn[cur][e];
      if (ntk) mt += st[cur][e];
    }
    red0[t] = mn;
    red1[t] = mt;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
      if (t < w) {
        red0[t] += red0[t + w];
        red1[t] += red1[t + w];
      }
      __syncthreads();
    }
    if (t == 0) {
      const float m = red0[0] / (float)nout;
      const float vn = fmaf(P.a, m, P.b);
      if (self) {
        out_n[i] = vn;
      } else {
        out_n[(long long)i * ld + j] = vn;
        if (sym) out_n[(long long)j * ld + i] = vn;
        if (ntk) {
          const float m2 = red1[0] / (float)nout;
          const float vt = fmaf(P.c, m, fmaf(P.d, m2, P.e));
          out_t[(long long)i * ld + j] = vt;
          if (sym) out_t[(long long)j * ld + i] = vt;
        }
      }
    }
    __syncthreads();
  }
}

}  // namespace nta

using namespace nta;

extern "C" int nta_diag_chain(const float* x1, const float* x2, int n1,
                              int n2, int sym, int H, int W, int C,
                              int nstage, const int* stage_desc,
                              const float* stage_wb, float* q1,
                              const float* q2, int flags, double a, double b,
                              double c, double d, double e, float* out_nngp,
                              float* out_ntk, long long ld, void* stream) {
  if (n1 < 0 || n2 < 0 || H <= 0 || W <= 0 || C <= 0 || nstage <= 0 ||
      nstage > kDiagMaxStages || H * W > kDiagMaxHW || !stage_desc ||
      !stage_wb)
    return set_error(NTA_EINVAL, "nta_diag_chain: bad n1=%d n2=%d H=%d W=%d "
                     "C=%d nstage=%d (<= %d stages, H*W <= %d)", n1, n2, H,
                     W, C, nstage, kDiagMaxStages, kDiagMaxHW);
  if (sym && n2 != n1)
    return set_error(NTA_EINVAL, "nta_diag_chain: sym needs n2 == n1");
  if (n1 == 0 || (n2 == 0 && sym)) return NTA_OK;
  const bool self = n2 == 0;
  if (!x1 || (!self && !x2) || !out_nngp || (!self && (flags & NTA_NTK) &&
                                             !out_ntk))
    return set_error(NTA_EINVAL, "nta_diag_chain: null pointer");
  DiagPlan P;
  P.nstage = nstage;
  P.H = H;
  P.W = W;
  P.C = C;
  P.ntk = (flags & NTA_NTK) ? 1 : 0;
  P.a = (float)a;
  P.b = (float)b;
  P.c = (float)c;
  P.d = (float)d;
  P.e = (float)e;
  // q tables: [nstage][n][H][W] slots (a stage's values fill the first
  // n * hout * wout floats of its slot); q2 holds n2 images.
  P.qslot1 = (long long)n1 * H * W;
  P.qslot2 = (long long)(self ? n1 : n2) * H * W;
  int h = H, w = W;
  bool any_relu = false;
  for (int s = 0; s < nstage; ++s) {
    DiagStage& S = P.s[s];
    const int f = stage_desc[3 * s];
    S.conv = (f & NTA_CONV) ? 1 : 0;
    S.relu = (f & NTA_RELU) ? 1 : 0;
    S.act = f & (NTA_ACT_ERF | NTA_ACT_GELU);
    S.circ = (S.conv && (f & NTA_CONV_CIRCULAR)) ? 1 : 0;
    S.fh = S.conv ? stage_desc[3 * s + 1] : 1;
    S.fw = S.conv ? stage_desc[3 * s + 2] : 1;
    S.w_var = stage_wb[4 * s];
    S.b_var = stage_wb[4 * s + 1];
    S.k1 = stage_wb[4 * s + 2];
    S.k0 = stage_wb[4 * s + 3];
    S.hin = h;
    S.win = w;
    if (S.fh <= 0 || S.fw <= 0)
      return set_error(NTA_EINVAL, "nta_diag_chain: stage %d filter %dx%d",
                       s, S.fh, S.fw);
    if (S.conv && (f & NTA_CONV_VALID)) {
      if (S.fh > h || S.fw > w)
        return set_error(NTA_EINVAL, "nta_diag_chain: VALID %dx%d filter on "
                         "%dx%d", S.fh, S.fw, h, w);
      S.lo_h = S.lo_w = 0;
      h = h - S.fh + 1;
      w = w - S.fw + 1;
    } else {
      S.lo_h = (S.fh - 1) / 2;       // XLA SAME, stride 1
      S.lo_w = (S.fw - 1) / 2;
    }
    S.hout = h;
    S.wout = w;
    any_relu |= S.relu != 0;
  }
  if (any_relu && (!q1 || (!self && !q2)))
    return set_error(NTA_EINVAL, "nta_diag_chain: q tables are required");
  const long long npairs = self ? (long long)n1
                                : sym ? (long long)n1 * (n1 + 1) / 2
                                      : (long long)n1 * n2;
  const unsigned grid =
      (unsigned)(npairs < (1LL << 20) ? npairs : (1LL << 20));
  const size_t lds = (size_t)(4 * H * W + 512) * sizeof(float);
  hipLaunchKernelGGL(k_diag_chain, dim3(grid), dim3(256), lds,
                     (hipStream_t)stream, P, x1, x2, n1, n2, sym, q1, q2,
                     out_nngp, out_ntk, ld, npairs);
  return check_launch("nta_diag_chain");
}
