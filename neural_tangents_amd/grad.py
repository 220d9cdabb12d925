"""Gradients of the analytic kernels: hand-written adjoint kernels behind a
``torch.autograd.Function``.

The reference has no adjoint code: ``jax.grad`` traces its ``kernel_fn``
(``linear.py:1325-1424``, ``elementwise.py:1283-1310`` for the singular-point
rules).  Here the reverse pass of each fused stage runs on the GPU:

* ``nta_relu_bwd`` (``csrc/nta_adjoint.hip``): the adjoint of the ReLU
  arc-cosine map and of the average pool, per pre-pool element, including
  the cotangents of the normalisers ``q`` and of ``W_std^2``, ``b_std^2``;
* the adjoint of the 3x3 stencil mean is the stencil mean itself (it is
  self-adjoint), run by ``nta_stage`` with only its stencil enabled;
* the adjoint of the first stage's patch Gram is a plain GEMM of the Gram
  cotangent with the other image's patches, folded back to pixels.

Per-stage states are recomputed from the checkpoints the forward engine
keeps at stage (pool) boundaries, one pair tile at a time; the pooled 3x3
stages of the recomputation also return their inputs' stencil means
(``nta_stage_keep_pre``), which the ReLU adjoint consumes.  Cotangents of
the normalisers ``q`` are summed over all pairs of an image and pushed
through that image's self chain (whose spatial diagonal defines ``q``).

Scope: ``serial`` networks of the engine's grammar whose stages all have a
ReLU (Myrtle-style: Conv 2-D SAME stride 1, Relu, optional AvgPool, then
GlobalAvgPool / Flatten + Dense).  ``kernel_vjp`` returns per-stage
cotangents of ``W_std^2`` and ``b_std^2``; ``KernelFunction`` maps them onto
shared ``W_std``, ``b_std`` scalars.
"""
import torch
import torch.nn.functional as F

from . import _engine, _lib


def _stencil(x, n, S, fh, fw, valid, stream):
  """Stencil mean (1/(fh fw)) sum over the diagonal taps, nngp-only mode of
  nta_stage with w_var = 1, b_var = 0; x: [n, S, S, S, S] (self layout)."""
  out = torch.empty_like(x)
  flags = _lib.CONV | (_lib.CONV_VALID if valid else 0)
  _lib.call('nta_stage', _lib.ptr(x), None, n, 0, S, S, fh, fw, 1, 1, flags,
            1.0, 0.0, 1.0, 0.0, None, None, _lib.ptr(out), None, stream)
  return out


def _patches(x, fh, fw):
  """im2col of zero-padded NHWC images: [N, H*W, C*fh*fw] (SAME)."""
  xn = x.permute(0, 3, 1, 2)
  ph, pw = (fh - 1) // 2, (fw - 1) // 2
  xn = F.pad(xn, (pw, fw - 1 - pw, ph, fh - 1 - ph))
  p = F.unfold(xn, (fh, fw))                       # [N, C*fh*fw, H*W]
  return p.transpose(1, 2).contiguous()


def _fold(gp, shape, fh, fw):
  """Adjoint of ``_patches``: [N, H*W, C*fh*fw] -> NHWC gradient."""
  N, H, W, C = shape
  ph, pw = (fh - 1) // 2, (fw - 1) // 2
  g = F.fold(gp.transpose(1, 2), (H + fh - 1, W + fw - 1), (fh, fw))
  g = g[:, :, ph:ph + H, pw:pw + W]
  return g.permute(0, 2, 3, 1).contiguous()


def _to_hw_pairs(g, S):
  """[T, S, S', S, S'] zipped (h, h', w, w') -> [T, (h w), (h' w')]."""
  T = g.shape[0]
  return g.reshape(T, S, S, S, S).permute(0, 1, 3, 2, 4).reshape(
      T, S * S, S * S)


NO_SELF = 2 ** 31 - 1


def _relu_bwd(aN, aT, n1, n2, S, pool, w_var, b_var, q1, q2, gN, gT,
              stream, self_off=NO_SELF, act=(1.0, 0.0), gram_scale=0.0,
              acc=None):
  """Adjoint of one stage's ReLU (+ pool).  Returns (gaN, gaT, gqa, gqb,
  gw, gb): gqa, gqb [T, S, S] the normaliser cotangents summed over the
  partner's positions, gw, gb the d/dw_var, d/db_var contributions as 0-d
  FP64 device tensors (summed from the kernel's per-workgroup partials,
  with no host synchronisation).  gram_scale != 0: gaN in the patch-Gram
  layout [n1, S, S, n2, S, S] times gram_scale (cross tiles).
  acc = (wb_acc [2] FP64, gq1_acc [n1, S, S], gq2_acc [n2, S, S]): the
  library adds gw, gb and the partner-summed gqa, gqb into these instead
  (nta_relu_bwd_accumulate), and gw, gb come back as None."""
  T = aN.shape[0]
  gaN = torch.empty_like(aN)
  gaT = torch.empty_like(aN) if aT is not None else None
  gqa = torch.empty((T, S, S), dtype=torch.float32, device=aN.device)
  gqb = torch.empty_like(gqa)
  parts = _lib.value('nta_relu_bwd_parts', n1, n2, S)
  gwb = torch.empty((max(parts, 1), 2), dtype=torch.float64,
                    device=aN.device)
  nwork = _lib.value('nta_relu_bwd_work_bytes', n1, n2, S)
  work = torch.empty((max(nwork, 1),), dtype=torch.uint8, device=aN.device)
  _lib.call('nta_relu_bwd', _lib.ptr(aN), _lib.ptr(aT), n1, n2, S, pool,
            self_off, w_var, b_var, act[0], act[1], _lib.ptr(q1),
            _lib.ptr(q2), _lib.ptr(gN), _lib.ptr(gT), gram_scale,
            _lib.ptr(gaN), _lib.ptr(gaT), _lib.ptr(gqa), _lib.ptr(gqb),
            _lib.ptr(gwb), _lib.ptr(work), work.numel(), stream)
  if acc is not None:
    wb_acc, gq1_acc, gq2_acc = acc
    _lib.call('nta_relu_bwd_accumulate', _lib.ptr(gqa), _lib.ptr(gqb),
              _lib.ptr(gwb), n1, n2, S, _lib.ptr(wb_acc), _lib.ptr(gq1_acc),
              _lib.ptr(gq2_acc), stream)
    return gaN, gaT, gqa, gqb, None, None
  sums = gwb[:parts].sum(dim=0)
  return gaN, gaT, gqa, gqb, sums[0], sums[1]


def _check_plan(plan):
  if plan.head is None:
    raise NotImplementedError('kernel gradients need a readout head '
                              '(GlobalAvgPool or Flatten).')
  if plan.out_scale != 1.0 or any(s.wscale != 1.0 for s in plan.stages):
    raise NotImplementedError('kernel gradients do not support SumPool.')
  for s in plan.stages:
    if s.act_kind:
      raise NotImplementedError('kernel gradients support the ABRelu family '
                                '(Relu, LeakyRelu, Abs) only.')
    if s.conv is None or not s.relu:
      raise NotImplementedError('kernel gradients support stages of Conv + '
                                'Relu [+ AvgPool] only.')
    if s.hin != s.win:
      raise NotImplementedError('kernel gradients need square images.')
    if s.conv[2] or s.circular:
      raise NotImplementedError('kernel gradients need SAME convolutions.')


def _pair_tile(n1, n2, tile_pairs, side):
  """(rows, columns) of the backward's pair tiles: whole rows of the
  N1 x N2 grid where they fit, at most ``tile_pairs`` pairs and at most
  2^30 elements of a [pairs, side^4] state (32-bit indexing)."""
  tile_pairs = max(1, min(tile_pairs, (1 << 30) // side ** 4))
  tj = max(1, min(n2, tile_pairs))
  ti = max(1, min(n1, tile_pairs // tj))
  return ti, tj


def kernel_vjp(plan, x1, x2, g_nngp, g_ntk, tile_pairs=512):
  """Cotangents of (nngp, ntk) [N1, N2] -> (gx1, gx2, g_wvar[stage],
  g_bvar[stage]).  ``x2=None`` means ``x2 = x1`` (gx2 is then None and the
  x1 cotangent collects both arguments).  Pair tiles hold up to
  ``tile_pairs`` pairs (about 40 MiB each at 32x32), at most 2^30 state
  elements (32-bit indexing in the kernels)."""
  _check_plan(plan)
  dev = x1.device
  stream = _lib.stream_handle(dev)
  L = len(plan.stages)
  sym = x2 is None
  xb_all = x1 if sym else x2
  N1, N2 = x1.shape[0], xb_all.shape[0]
  g_nngp = torch.zeros((N1, N2), device=dev) if g_nngp is None else \
      g_nngp.to(dev, torch.float32).contiguous()
  g_ntk = torch.zeros((N1, N2), device=dev) if g_ntk is None else \
      g_ntk.to(dev, torch.float32).contiguous()
  want_ntk = True
  q1s, _ = _engine._self_chain(plan, x1, False, _engine.DEFAULT_TILE_BYTES,
                               stream)
  q2s = q1s if sym else _engine._self_chain(plan, x2, False,
                                            _engine.DEFAULT_TILE_BYTES,
                                            stream)[0]
  gq1 = [torch.zeros_like(q) for q in q1s]
  gq2 = gq1 if sym else [torch.zeros_like(q) for q in q2s]
  # d/dw_var, d/db_var per stage, accumulated on the device (FP64)
  gwv = [torch.zeros((), dtype=torch.float64, device=dev) for _ in range(L)]
  gbv = [torch.zeros((), dtype=torch.float64, device=dev) for _ in range(L)]
  wb_acc = torch.zeros((L, 2), dtype=torch.float64, device=dev)  # cross
  gx1 = torch.zeros_like(x1)
  gx2 = gx1 if sym else torch.zeros_like(x2)
  first = plan.stages[0]
  fh, fw = first.conv[0], first.conv[1]
  C = x1.shape[-1]
  a, b, c, d, e = plan.head_coef
  last = plan.stages[-1]
  nlast = last.hout ** 2 * last.wout ** 2
  if plan.head == 'flatten':
    raise NotImplementedError('kernel gradients support GlobalAvgPool.')
  ti, tj = _pair_tile(N1, N2, tile_pairs, plan.stages[0].hin)
  P_all1 = _patches(x1, fh, fw)
  P_all2 = P_all1 if sym else _patches(x2, fh, fw)
  # Myrtle-5/7 at 32x32: the first two stages recomputed by one launch
  # (nta_first_pool_stage_keep_pre), from the normalisers' (sqrt q, 1/sqrt q)
  fused = _engine._fused_stages(plan, x1.shape) == 2
  if fused:
    qn1 = [_engine._qnorm(q1s[k], stream) for k in range(2)]
    qn2 = qn1 if sym else [_engine._qnorm(q2s[k], stream) for k in range(2)]
  for i0 in range(0, N1, ti):
    n1 = min(ti, N1 - i0)
    for j0 in range(0, N2, tj):
      n2 = min(tj, N2 - j0)
      T = n1 * n2
      qa = [q[i0:i0 + n1] for q in q1s]
      qb = [q[j0:j0 + n2] for q in q2s]
      keep = []
      pre = {}
      qn = None
      if fused:
        qn = (qn1[0][i0:i0 + n1], qn2[0][j0:j0 + n2], qn1[1][i0:i0 + n1],
              qn2[1][j0:j0 + n2])
      _engine._run_chain(plan, x1[i0:i0 + n1], xb_all[j0:j0 + n2], n1, n2,
                         qa, qb, want_ntk, stream, keep=keep, pre=pre, qn=qn)
      # head: K_nngp = a m + b, K_ntk = c m + d mt + e (m = mean)
      gn = g_nngp[i0:i0 + n1, j0:j0 + n2].reshape(T)
      gt = g_ntk[i0:i0 + n1, j0:j0 + n2].reshape(T)
      shp = (T, last.hout, last.hout, last.wout, last.wout)
      gN = gn.mul(a / nlast).add_(gt, alpha=c / nlast)[:, None].expand(
          T, nlast).reshape(shp).contiguous()
      gT = gt.mul(d / nlast)[:, None].expand(T, nlast).reshape(
          shp).contiguous()
      for si in range(L - 1, -1, -1):
        s = plan.stages[si]
        fh_s, fw_s, valid, w_var, b_var = s.conv
        S = s.hin
        pool = s.pool[0] if s.pool is not None else 1
        if si in pre:                  # kept by the recomputation
          aN, aT = pre.pop(si)
        elif si > 0:
          Nin, Tin = keep[si - 1]
          aN = _stencil(Nin, T, S, fh_s, fw_s, valid, stream)
          aT = _stencil(Tin, T, S, fh_s, fw_s, valid, stream)
        else:
          aN = _first_gram(x1[i0:i0 + n1], xb_all[j0:j0 + n2], n1, n2, S,
                           C, fh, fw, stream)
          aT = None
        first_stage = si == 0
        gaN, gaT, _, _, _, _ = _relu_bwd(
            aN, aT, n1, n2, S, pool, w_var, b_var, qa[si], qb[si], gN, gT,
            stream, self_off=(i0 - j0) if sym else NO_SELF, act=s.act,
            gram_scale=1.0 / (fh * fw * C) if first_stage else 0.0,
            acc=(wb_acc[si], gq1[si][i0:i0 + n1], gq2[si][j0:j0 + n2]))
        if not first_stage:
          gN = _stencil(gaN, T, S, fh_s, fw_s, valid, stream)
          gT = _stencil(gaT, T, S, fh_s, fw_s, valid, stream)
        else:
          # aN = G / (fh fw C), G = P1 P2^T per pair; gaN holds
          # M = dL/dG as [(a, h, w), (b, h', w')]
          HW = S * S
          M = gaN.reshape(n1 * HW, n2 * HW)
          P1 = P_all1[i0:i0 + n1].reshape(n1 * HW, -1)
          P2 = P_all2[j0:j0 + n2].reshape(n2 * HW, -1)
          gP1 = torch.mm(M, P2).reshape(n1, HW, -1)
          gP2 = torch.mm(M.t(), P1).reshape(n2, HW, -1)
          gx1[i0:i0 + n1] += _fold(gP1, (n1,) + tuple(x1.shape[1:]), fh, fw)
          gx2[j0:j0 + n2] += _fold(gP2, (n2,) + tuple(x1.shape[1:]), fh, fw)
  # Self chains: push the normaliser cotangents through each image's chain.
  for x, gq, gx in ((x1, gq1, gx1),) + (() if sym else ((x2, gq2, gx2),)):
    _self_chain_vjp(plan, x, gq, gx, gwv, gbv, stream, C, fh, fw)
  sums = torch.cat([torch.stack(gwv) + wb_acc[:, 0],
                    torch.stack(gbv) + wb_acc[:, 1]]).cpu()   # one sync
  return (gx1, (None if sym else gx2), [float(v) for v in sums[:L]],
          [float(v) for v in sums[L:]])


def _first_gram(xa, xb, n1, n2, S, C, fh, fw, stream):
  """aN of the first stage: G / (fh fw C), the stencil of the input Gram
  (nta_first_stage with w_var = 1, b_var = 0 and no ReLU)."""
  T = n1 if n2 == 0 else n1 * n2
  out = torch.empty((T, S, S, S, S), dtype=torch.float32, device=xa.device)
  H, W = xa.shape[1], xa.shape[2]
  _lib.call('nta_first_stage', _lib.ptr(xa), _lib.ptr(xb), n1, n2, H, W, C,
            fh, fw, 0, 1.0, 0.0, 1.0, 0.0, None, None, _lib.ptr(out), None,
            stream)
  return out


def _self_chain_vjp(plan, x, gq, gx, gwv, gbv, stream, C, fh, fw,
                    chunk=128):
  N = x.shape[0]
  L = len(plan.stages)
  chunk = max(1, min(chunk, (1 << 30) // plan.stages[0].hin ** 4))
  P_all = _patches(x, fh, fw)
  for n0 in range(0, N, chunk):
    n = min(chunk, N - n0)
    xa = x[n0:n0 + n]
    keep = []
    pre = {}
    collect = [None] * L
    _engine._run_chain(plan, xa, None, n, 0, None, None, False, stream,
                       collect_q=collect, keep=keep, pre=pre)
    gR = None                                   # cotangent of stage output
    for si in range(L - 1, -1, -1):
      s = plan.stages[si]
      fh_s, fw_s, valid, w_var, b_var = s.conv
      S = s.hin
      pool = s.pool[0] if s.pool is not None else 1
      q = collect[si]
      if si in pre:
        aN = pre.pop(si)[0]
      elif si > 0:
        aN = _stencil(keep[si - 1][0], n, S, fh_s, fw_s, valid, stream)
      else:
        aN = _first_gram(xa, None, n, 0, S, C, fh, fw, stream)
      gaN, _, gqa, gqb, dw, db = _relu_bwd(aN, None, n, 0, S, pool, w_var,
                                           b_var, q, None, gR, None, stream,
                                           act=s.act)
      gq_tot = gq[si][n0:n0 + n] + gqa + gqb                 # [n, S, S]
      # q = diag(sigma), sigma = w_var aN + b_var
      v = gaN.reshape(n, S, S, S, S)
      an = aN.reshape(n, S, S, S, S)
      idx = torch.arange(S, device=x.device)
      diag_a = an[:, idx[:, None], idx[:, None], idx[None, :], idx[None, :]]
      v[:, idx[:, None], idx[:, None], idx[None, :], idx[None, :]] += \
          w_var * gq_tot
      gwv[si] += dw + (gq_tot * diag_a).sum(dtype=torch.float64)
      gbv[si] += db + gq_tot.sum(dtype=torch.float64)
      if si > 0:
        gR = _stencil(gaN, n, S, fh_s, fw_s, valid, stream)
      else:
        M = _to_hw_pairs(gaN, S) / (fh * fw * C)             # [n, HW, HW]
        P = P_all[n0:n0 + n]
        gP = torch.bmm(M + M.transpose(1, 2), P)
        gx[n0:n0 + n] += _fold(gP, (n,) + tuple(x.shape[1:]), fh, fw)


class KernelFunction(torch.autograd.Function):
  """(nngp, ntk) = kernel(x1, x2; W_std, b_std) of a Myrtle-style network
  whose convolutions share ``W_std``, ``b_std``; differentiable in x1, x2,
  W_std and b_std through the adjoint kernels."""

  @staticmethod
  def forward(ctx, x1, x2, W_std, b_std, builder):
    kernel_fn = builder(float(W_std), float(b_std))
    x2n = None if x2 is None or x2.numel() == 0 else x2
    k = kernel_fn(x1.detach(), None if x2n is None else x2n.detach(),
                  ('nngp', 'ntk'))
    ctx.kernel_fn = kernel_fn
    ctx.sym = x2n is None
    ctx.W = float(W_std)
    ctx.b = float(b_std)
    ctx.save_for_backward(x1.detach(), x1.detach() if x2n is None else
                          x2n.detach())
    return k.nngp, k.ntk

  @staticmethod
  def backward(ctx, g_nngp, g_ntk):
    x1, x2 = ctx.saved_tensors
    x1c = x1.to(torch.float32).contiguous()
    plan = ctx.kernel_fn.plan(x1c.shape)
    gx1, gx2, gwv, gbv = kernel_vjp(
        plan, x1c, None if ctx.sym else x2.to(torch.float32).contiguous(),
        g_nngp, g_ntk)
    dev = x1.device
    need = ctx.needs_input_grad
    # Gradients only for the inputs that require them: W_std / b_std passed
    # as Python floats (fixed hyperparameters) get None.
    gW = torch.tensor(2.0 * ctx.W * sum(gwv), dtype=torch.float64,
                      device=dev) if need[2] else None
    gb = torch.tensor(2.0 * ctx.b * sum(gbv), dtype=torch.float64,
                      device=dev) if need[3] else None
    gx1 = gx1.to(x1.dtype) if need[0] else None
    gx2 = None if ctx.sym or not need[1] else gx2.to(x2.dtype)
    return gx1, gx2, gW, gb, None


def myrtle_builder(depths=(2, 1, 1), readout_W_std=1.0, readout_b_std=0.0):
  """Returns ``builder(W_std, b_std) -> kernel_fn`` of the Myrtle network
  with shared convolution parameters (for ``KernelFunction``)."""
  from . import stax

  def builder(W_std, b_std):
    return stax.myrtle(depths=list(depths), W_std=W_std, b_std=b_std,
                       readout_W_std=readout_W_std,
                       readout_b_std=readout_b_std)[2]
  return builder


def kernel(x1, x2, W_std, b_std, builder):
  """Differentiable (nngp, ntk); pass ``x2=None`` for K(x1, x1)."""
  if x2 is None:
    x2 = torch.empty(0, device=x1.device)
  return KernelFunction.apply(x1, x2, W_std, b_std, builder)


def krr_loss(x_support, y_support, x_target, y_target, W_std, b_std,
             builder, diag_reg=1e-4, get='ntk'):
  """Kernel-regression loss 0.5 * mean((K_ts (K_ss + lambda I)^-1 Y_s -
  Y_t)^2) with lambda = diag_reg * tr(K_ss) / N (predict.py:1224-1232); the
  small dense solve runs in FP64 under torch autograd, the kernels through
  the adjoint kernels."""
  gi = 0 if get == 'nngp' else 1
  K_ss = kernel(x_support, None, W_std, b_std, builder)[gi].double()
  K_ts = kernel(x_target, x_support, W_std, b_std, builder)[gi].double()
  n = K_ss.shape[0]
  lam = diag_reg * torch.trace(K_ss) / n
  A = K_ss + lam * torch.eye(n, dtype=torch.float64, device=K_ss.device)
  Lc = torch.linalg.cholesky(A)
  alpha = torch.cholesky_solve(y_support.double(), Lc)
  f = K_ts @ alpha
  return 0.5 * torch.mean((f - y_target.double()) ** 2)
