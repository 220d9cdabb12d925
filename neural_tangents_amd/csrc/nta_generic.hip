// Generic (any size / filter) kernels of the analytic-kernel path and the
// C-ABI entry points declared in include/nta.h.
//
// These kernels map one thread to one output element and read their operands
// straight from global memory; they serve every shape the engine accepts.
// The shape-specialised LDS kernels in nta_fused.hip take over the Myrtle
// shapes.
#include <stdarg.h>
#include <stdio.h>

#include "nta_common.h"

namespace nta {

static thread_local char g_err[512] = "";

int set_error(int code, const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
  return code;
}

int check_launch(const char* what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    return set_error((int)e, "%s: launch failed: %s", what,
                     hipGetErrorString(e));
  }
  return NTA_OK;
}

static inline unsigned grid_for(uint64_t total, int block) {
  uint64_t g = (total + block - 1) / block;
  if (g > (1u << 20)) g = 1u << 20;
  if (g == 0) g = 1;
  return (unsigned)g;
}

struct StencilGeom {
  int fh, fw;    // filter
  int oh, ow;    // input offset of tap 0 (SAME: -(f-1)/2, VALID: 0)
  int Hc, Wc;    // conv output size
  int circ;      // CIRCULAR: taps wrap around instead of reading zeros
};

// Input index of a tap, or -1 for a zero (padding) tap.
__device__ __forceinline__ int tap_index(int v, int n, int circ) {
  if (circ) {
    v %= n;
    return v < 0 ? v + n : v;
  }
  return (unsigned)v < (unsigned)n ? v : -1;
}

static StencilGeom make_geom(int H, int W, int fh, int fw, int flags) {
  StencilGeom g;
  g.fh = fh;
  g.fw = fw;
  if (flags & NTA_CONV_VALID) {
    g.oh = 0;
    g.ow = 0;
    g.Hc = H - fh + 1;
    g.Wc = W - fw + 1;
  } else {
    g.oh = -((fh - 1) / 2);
    g.ow = -((fw - 1) / 2);
    g.Hc = H;
    g.Wc = W;
  }
  g.circ = (flags & NTA_CONV_CIRCULAR) ? 1 : 0;
  return g;
}

// ----------------------------------------------------------------- kernels

__global__ void k_first_diag(const float* __restrict__ x, int n, int H, int W,
                             int C, StencilGeom g, float scale, float b_var,
                             float* __restrict__ q) {
  uint64_t total = (uint64_t)n * g.Hc * g.Wc;
  for (uint64_t idx = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
       idx < total; idx += (uint64_t)gridDim.x * blockDim.x) {
    int w = (int)(idx % g.Wc);
    int h = (int)((idx / g.Wc) % g.Hc);
    int img = (int)(idx / ((uint64_t)g.Wc * g.Hc));
    const float* xi = x + (uint64_t)img * H * W * C;
    float acc = 0.f;
    for (int k = 0; k < g.fh; ++k) {
      int hh = tap_index(h + k + g.oh, H, g.circ);
      if (hh < 0) continue;
      for (int l = 0; l < g.fw; ++l) {
        int ww = tap_index(w + l + g.ow, W, g.circ);
        if (ww < 0) continue;
        const float* px = xi + ((uint64_t)hh * W + ww) * C;
        for (int c = 0; c < C; ++c) acc = fmaf(px[c], px[c], acc);
      }
    }
    q[idx] = fmaf(scale, acc, b_var);
  }
}

__global__ void k_first_stage(const float* __restrict__ x1,
                              const float* __restrict__ x2, int n1, int n2,
                              int H, int W, int C, StencilGeom g, int flags,
                              float scale, float b_var, float k1, float k0,
                              const float* __restrict__ q1,
                              const float* __restrict__ q2,
                              float* __restrict__ nngp,
                              float* __restrict__ ntk) {
  const bool self = (n2 == 0);
  const uint64_t T = self ? (uint64_t)n1 : (uint64_t)n1 * n2;
  const int Hc = g.Hc, Wc = g.Wc;
  const uint64_t per = (uint64_t)Hc * Hc * Wc * Wc;
  const uint64_t total = T * per;
  for (uint64_t idx = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
       idx < total; idx += (uint64_t)gridDim.x * blockDim.x) {
    uint64_t r = idx % per;
    uint64_t p = idx / per;
    int w2 = (int)(r % Wc);
    int w1 = (int)((r / Wc) % Wc);
    int h2 = (int)((r / ((uint64_t)Wc * Wc)) % Hc);
    int h1 = (int)(r / ((uint64_t)Wc * Wc * Hc));
    int i = self ? (int)p : (int)(p / n2);
    int j = self ? (int)p : (int)(p % n2);
    const float* xa = x1 + (uint64_t)i * H * W * C;
    const float* xb = (self ? x1 : x2) + (uint64_t)j * H * W * C;
    float acc = 0.f;
    for (int k = 0; k < g.fh; ++k) {
      int ha = tap_index(h1 + k + g.oh, H, g.circ);
      int hb = tap_index(h2 + k + g.oh, H, g.circ);
      if (ha < 0 || hb < 0) continue;
      for (int l = 0; l < g.fw; ++l) {
        int wa = tap_index(w1 + l + g.ow, W, g.circ);
        int wb = tap_index(w2 + l + g.ow, W, g.circ);
        if (wa < 0 || wb < 0) continue;
        const float* pa = xa + ((uint64_t)ha * W + wa) * C;
        const float* pb = xb + ((uint64_t)hb * W + wb) * C;
        for (int c = 0; c < C; ++c) acc = fmaf(pa[c], pb[c], acc);
      }
    }
    float v = fmaf(scale, acc, b_var);
    float t = v;   // ntk after the first conv equals nngp (ntk_in = 0)
    if (flags & NTA_RELU) {
      float qa = q1[((uint64_t)i * Hc + h1) * Wc + w1];
      float qb = (self ? q1 : q2)[((uint64_t)j * Hc + h2) * Wc + w2];
      float ds;
      v = act_map(flags, v, qa, qb, k1, k0, &ds);
      t *= ds;
    }
    nngp[idx] = v;
    if (flags & NTA_NTK) ntk[idx] = t;
  }
}

__global__ void k_stage_diag(const float* __restrict__ R, int n, int H, int W,
                             StencilGeom g, int conv, float scale,
                             float b_var, float* __restrict__ q) {
  const int Hc = g.Hc, Wc = g.Wc;
  uint64_t total = (uint64_t)n * Hc * Wc;
  const uint64_t per = (uint64_t)H * H * W * W;
  for (uint64_t idx = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
       idx < total; idx += (uint64_t)gridDim.x * blockDim.x) {
    int w = (int)(idx % Wc);
    int h = (int)((idx / Wc) % Hc);
    int img = (int)(idx / ((uint64_t)Wc * Hc));
    const float* r = R + (uint64_t)img * per;
    float out;
    if (conv) {
      float acc = 0.f;
      for (int k = 0; k < g.fh; ++k) {
        int hh = tap_index(h + k + g.oh, H, g.circ);
        if (hh < 0) continue;
        for (int l = 0; l < g.fw; ++l) {
          int ww = tap_index(w + l + g.ow, W, g.circ);
          if (ww < 0) continue;
          acc += r[(((uint64_t)hh * H + hh) * W + ww) * W + ww];
        }
      }
      out = fmaf(scale, acc, b_var);
    } else {
      out = r[(((uint64_t)h * H + h) * W + w) * W + w];
    }
    q[idx] = out;
  }
}

__global__ void k_stage(const float* __restrict__ in_nngp,
                        const float* __restrict__ in_ntk, int n1, int n2,
                        int H, int W, StencilGeom g, int ph, int pw, int flags,
                        float scale, float w_var, float b_var, float k1,
                        float k0, const float* __restrict__ q1,
                        const float* __restrict__ q2,
                        float* __restrict__ out_nngp,
                        float* __restrict__ out_ntk) {
  const bool self = (n2 == 0);
  const bool conv = flags & NTA_CONV;
  const bool relu = flags & NTA_RELU;
  const bool want_ntk = flags & NTA_NTK;
  const uint64_t T = self ? (uint64_t)n1 : (uint64_t)n1 * n2;
  const int Hc = conv ? g.Hc : H, Wc = conv ? g.Wc : W;
  const int Ho = Hc / ph, Wo = Wc / pw;
  const uint64_t per_in = (uint64_t)H * H * W * W;
  const uint64_t per_out = (uint64_t)Ho * Ho * Wo * Wo;
  const uint64_t total = T * per_out;
  const float inv_pool = 1.0f / (float)(ph * ph * pw * pw);
  for (uint64_t idx = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
       idx < total; idx += (uint64_t)gridDim.x * blockDim.x) {
    uint64_t r = idx % per_out;
    uint64_t p = idx / per_out;
    int ow2 = (int)(r % Wo);
    int ow1 = (int)((r / Wo) % Wo);
    int oh2 = (int)((r / ((uint64_t)Wo * Wo)) % Ho);
    int oh1 = (int)(r / ((uint64_t)Wo * Wo * Ho));
    int i = self ? (int)p : (int)(p / n2);
    int j = self ? (int)p : (int)(p % n2);
    const float* sn = in_nngp + p * per_in;
    const float* st = want_ntk ? in_ntk + p * per_in : nullptr;
    float acc_n = 0.f, acc_t = 0.f;
    for (int a = 0; a < ph; ++a)
      for (int b = 0; b < ph; ++b)
        for (int c = 0; c < pw; ++c)
          for (int d = 0; d < pw; ++d) {
            int h1 = oh1 * ph + a, h2 = oh2 * ph + b;
            int w1 = ow1 * pw + c, w2 = ow2 * pw + d;
            float v, t = 0.f;
            if (conv) {
              float sv = 0.f, stt = 0.f;
              for (int k = 0; k < g.fh; ++k) {
                int ha = tap_index(h1 + k + g.oh, H, g.circ);
                int hb = tap_index(h2 + k + g.oh, H, g.circ);
                if (ha < 0 || hb < 0) continue;
                for (int l = 0; l < g.fw; ++l) {
                  int wa = tap_index(w1 + l + g.ow, W, g.circ);
                  int wb = tap_index(w2 + l + g.ow, W, g.circ);
                  if (wa < 0 || wb < 0) continue;
                  uint64_t o = (((uint64_t)ha * H + hb) * W + wa) * W + wb;
                  sv += sn[o];
                  if (want_ntk) stt += st[o];
                }
              }
              v = fmaf(scale * w_var, sv, b_var);
              if (want_ntk) t = fmaf(scale * w_var, stt, v);
            } else {
              uint64_t o = (((uint64_t)h1 * H + h2) * W + w1) * W + w2;
              v = sn[o];
              if (want_ntk) t = st[o];
            }
            if (relu) {
              float qa = q1[((uint64_t)i * Hc + h1) * Wc + w1];
              float qb = (self ? q1 : q2)[((uint64_t)j * Hc + h2) * Wc + w2];
              float ds;
              v = act_map(flags, v, qa, qb, k1, k0, &ds);
              t *= ds;
            }
            acc_n += v;
            acc_t += t;
          }
    out_nngp[idx] = acc_n * inv_pool;
    if (want_ntk) out_ntk[idx] = acc_t * inv_pool;
  }
}

// One workgroup per pair: spatial mean (or diagonal trace) + affine head.
__global__ void k_head(const float* __restrict__ nngp,
                       const float* __restrict__ ntk, int n1, int n2, int H,
                       int W, int flags, double a, double b, double c,
                       double d, double e, float* __restrict__ out_nngp,
                       float* __restrict__ out_ntk, long long ld) {
  const bool self = (n2 == 0);
  const bool flat = flags & NTA_FLATTEN;
  const bool want_ntk = flags & NTA_NTK;
  const uint64_t p = blockIdx.x;
  const uint64_t per = (uint64_t)H * H * W * W;
  const float* sn = nngp + p * per;
  const float* st = want_ntk ? ntk + p * per : nullptr;
  double sn_acc = 0.0, st_acc = 0.0;
  uint64_t count = flat ? (uint64_t)H * W : per;
  for (uint64_t k = threadIdx.x; k < count; k += blockDim.x) {
    uint64_t o;
    if (flat) {
      int h = (int)(k / W), w = (int)(k % W);
      o = (((uint64_t)h * H + h) * W + w) * W + w;
    } else {
      o = k;
    }
    sn_acc += sn[o];
    if (want_ntk) st_acc += st[o];
  }
  __shared__ double red[2][256];
  red[0][threadIdx.x] = sn_acc;
  red[1][threadIdx.x] = st_acc;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
      red[0][threadIdx.x] += red[0][threadIdx.x + s];
      red[1][threadIdx.x] += red[1][threadIdx.x + s];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    double m = red[0][0] / (double)count;
    double mt = red[1][0] / (double)count;
    uint64_t o;
    if (self) {
      o = p;
    } else {
      uint64_t i = p / n2, j = p % n2;
      o = i * (uint64_t)ld + j;
    }
    out_nngp[o] = (float)(a * m + b);
    if (want_ntk) out_ntk[o] = (float)(c * m + d * mt + e);
  }
}

}  // namespace nta

using namespace nta;

// ----------------------------------------------------------------- C ABI

extern "C" int nta_version(void) { return 100; }

extern "C" const char* nta_last_error(void) { return nta::g_err; }

static int check_filter(int H, int W, int fh, int fw, int flags,
                        const char* who) {
  if (H <= 0 || W <= 0)
    return set_error(NTA_EINVAL, "%s: spatial sizes must be positive, got "
                     "H=%d W=%d", who, H, W);
  if (fh <= 0 || fw <= 0)
    return set_error(NTA_EINVAL, "%s: filter sizes must be positive, got "
                     "%dx%d", who, fh, fw);
  if ((flags & NTA_CONV_VALID) && (flags & NTA_CONV_CIRCULAR))
    return set_error(NTA_EINVAL, "%s: VALID and CIRCULAR padding both set",
                     who);
  if ((flags & NTA_CONV_VALID) && (fh > H || fw > W))
    return set_error(NTA_EINVAL, "%s: VALID %dx%d filter larger than %dx%d "
                     "input", who, fh, fw, H, W);
  return NTA_OK;
}

extern "C" int nta_first_diag(const float* x, int n, int H, int W, int C,
                              int fh, int fw, int flags, float w_var,
                              float b_var, float* q, void* stream) {
  int rc = check_filter(H, W, fh, fw, flags, "nta_first_diag");
  if (rc) return rc;
  if (n < 0 || C <= 0 || (n > 0 && (!x || !q)))
    return set_error(NTA_EINVAL, "nta_first_diag: bad n=%d C=%d or null "
                     "pointer", n, C);
  if (n == 0) return NTA_OK;
  StencilGeom g = make_geom(H, W, fh, fw, flags);
  float scale = w_var / (float)(fh * fw * C);
  uint64_t total = (uint64_t)n * g.Hc * g.Wc;
  hipLaunchKernelGGL(k_first_diag, dim3(grid_for(total, 256)), dim3(256), 0,
                     (hipStream_t)stream, x, n, H, W, C, g, scale, b_var, q);
  return check_launch("nta_first_diag");
}

extern "C" int nta_first_stage(const float* x1, const float* x2, int n1,
                               int n2, int H, int W, int C, int fh, int fw,
                               int flags, float w_var, float b_var,
                               float act_k1, float act_k0, const float* q1,
                               const float* q2, float* nngp, float* ntk,
                               void* stream) {
  int rc = check_filter(H, W, fh, fw, flags, "nta_first_stage");
  if (rc) return rc;
  if (n1 < 0 || n2 < 0 || C <= 0)
    return set_error(NTA_EINVAL, "nta_first_stage: bad n1=%d n2=%d C=%d", n1,
                     n2, C);
  if (n1 == 0) return NTA_OK;
  if (!x1 || !nngp || (n2 > 0 && !x2) || ((flags & NTA_NTK) && !ntk) ||
      ((flags & NTA_RELU) && (!q1 || (n2 > 0 && !q2))))
    return set_error(NTA_EINVAL, "nta_first_stage: null pointer for a "
                     "requested operand");
  if (!(flags & NTA_CONV_VALID) && fh == 3 && fw == 3 && H == W) {
    int rc2 = launch_first_fast(x1, x2, n1, n2, H, C, flags, w_var, b_var,
                                act_k1, act_k0, q1, q2, nngp, ntk,
                                (hipStream_t)stream);
    if (rc2 != kNotHandled) return rc2;
  }
  StencilGeom g = make_geom(H, W, fh, fw, flags);
  float scale = w_var / (float)(fh * fw * C);
  uint64_t T = n2 == 0 ? (uint64_t)n1 : (uint64_t)n1 * n2;
  uint64_t total = T * g.Hc * g.Hc * g.Wc * g.Wc;
  hipLaunchKernelGGL(k_first_stage, dim3(grid_for(total, 256)), dim3(256), 0,
                     (hipStream_t)stream, x1, x2, n1, n2, H, W, C, g, flags,
                     scale, b_var, act_k1, act_k0, q1, q2, nngp, ntk);
  return check_launch("nta_first_stage");
}

extern "C" int nta_stage_diag(const float* R, int n, int H, int W, int fh,
                              int fw, int flags, float w_var, float b_var,
                              float* q, void* stream) {
  int rc = check_filter(H, W, fh, fw, flags, "nta_stage_diag");
  if (rc) return rc;
  if (n < 0 || (n > 0 && (!R || !q)))
    return set_error(NTA_EINVAL, "nta_stage_diag: bad n=%d or null pointer",
                     n);
  if (n == 0) return NTA_OK;
  StencilGeom g = make_geom(H, W, fh, fw, flags);
  int conv = (flags & NTA_CONV) != 0;
  if (!conv) {
    g.Hc = H;
    g.Wc = W;
  }
  float scale = w_var / (float)(fh * fw);
  uint64_t total = (uint64_t)n * g.Hc * g.Wc;
  hipLaunchKernelGGL(k_stage_diag, dim3(grid_for(total, 256)), dim3(256), 0,
                     (hipStream_t)stream, R, n, H, W, g, conv, scale, b_var,
                     q);
  return check_launch("nta_stage_diag");
}

extern "C" int nta_stage(const float* in_nngp, const float* in_ntk, int n1,
                         int n2, int H, int W, int fh, int fw, int ph, int pw,
                         int flags, float w_var, float b_var, float act_k1,
                         float act_k0, const float* q1, const float* q2,
                         float* out_nngp, float* out_ntk, void* stream) {
  int rc = check_filter(H, W, fh, fw, flags, "nta_stage");
  if (rc) return rc;
  if (n1 < 0 || n2 < 0 || ph <= 0 || pw <= 0)
    return set_error(NTA_EINVAL, "nta_stage: bad n1=%d n2=%d pool=%dx%d", n1,
                     n2, ph, pw);
  if (n1 == 0) return NTA_OK;
  if (!in_nngp || !out_nngp || ((flags & NTA_NTK) && (!in_ntk || !out_ntk)) ||
      ((flags & NTA_RELU) && (!q1 || (n2 > 0 && !q2))))
    return set_error(NTA_EINVAL, "nta_stage: null pointer for a requested "
                     "operand");
  StencilGeom g = make_geom(H, W, fh, fw, flags);
  int Hc = (flags & NTA_CONV) ? g.Hc : H, Wc = (flags & NTA_CONV) ? g.Wc : W;
  if (Hc < ph || Wc < pw)
    return set_error(NTA_EINVAL, "nta_stage: pool window %dx%d larger than "
                     "%dx%d", ph, pw, Hc, Wc);
  int Ho = Hc / ph, Wo = Wc / pw;
  if ((flags & NTA_CONV) && !(flags & (NTA_RELU | NTA_NTK | NTA_CONV_VALID |
                                        NTA_CONV_CIRCULAR)) &&
      fh == 3 && fw == 3 && H == W && ph == 1 && pw == 1) {
    const long long T = n2 == 0 ? (long long)n1 : (long long)n1 * n2;
    int rc2 = launch_stencil_fast(in_nngp, T, H, w_var, b_var, out_nngp,
                                  (hipStream_t)stream);
    if (rc2 != kNotHandled) return rc2;
  }
  if ((flags & NTA_CONV) && (flags & NTA_RELU) && !(flags & NTA_CONV_VALID) &&
      fh == 3 && fw == 3 && H == W && ph == pw) {
    int rc2 = launch_stage_fast(in_nngp, in_ntk, n1, n2, H, ph, flags, w_var,
                                b_var, act_k1, act_k0, q1, q2, out_nngp,
                                out_ntk, (hipStream_t)stream);
    if (rc2 != kNotHandled) return rc2;
  }
  float scale = 1.0f / (float)(fh * fw);
  uint64_t T = n2 == 0 ? (uint64_t)n1 : (uint64_t)n1 * n2;
  uint64_t total = T * Ho * Ho * Wo * Wo;
  hipLaunchKernelGGL(k_stage, dim3(grid_for(total, 256)), dim3(256), 0,
                     (hipStream_t)stream, in_nngp, in_ntk, n1, n2, H, W, g,
                     ph, pw, flags, scale, w_var, b_var, act_k1, act_k0, q1,
                     q2, out_nngp, out_ntk);
  return check_launch("nta_stage");
}

extern "C" int nta_head(const float* nngp, const float* ntk, int n1, int n2,
                        int H, int W, int flags, double a, double b, double c,
                        double d, double e, float* out_nngp, float* out_ntk,
                        long long ld, void* stream) {
  if (n1 < 0 || n2 < 0 || H <= 0 || W <= 0)
    return set_error(NTA_EINVAL, "nta_head: bad sizes n1=%d n2=%d H=%d W=%d",
                     n1, n2, H, W);
  if (n1 == 0) return NTA_OK;
  if (!nngp || !out_nngp || ((flags & NTA_NTK) && (!ntk || !out_ntk)))
    return set_error(NTA_EINVAL, "nta_head: null pointer for a requested "
                     "operand");
  if (n2 > 0 && ld < n2)
    return set_error(NTA_EINVAL, "nta_head: ld=%lld < n2=%d", ld, n2);
  uint64_t T = n2 == 0 ? (uint64_t)n1 : (uint64_t)n1 * n2;
  if (T > 0x7fffffffULL)
    return set_error(NTA_EINVAL, "nta_head: too many pairs in one tile");
  hipLaunchKernelGGL(k_head, dim3((unsigned)T), dim3(256), 0,
                     (hipStream_t)stream, nngp, ntk, n1, n2, H, W, flags, a,
                     b, c, d, e, out_nngp, out_ntk, ld);
  return check_launch("nta_head");
}
