"""Adjoint kernels at the differentiation form's image size (needs a GPU).

The differentiation workload is a kernel-regression loss of a support set
against target images at 32x32x3 (Myrtle-5 NTK, learnable support inputs,
W_std and b_std).  Here 2 support and 2 target images at that size (the
fused first-two-stage path runs in the forward) are checked against central
finite differences of the FP64 oracle, Richardson-extrapolated from the
steps h = 1e-3 and h / 2 (each has O(h^2) truncation error; at h = 1e-3 it
is ~4e-6 of dK/dW_std for these eight-layer kernels, K growing like
W_std^8):

* every entry of dK/dW_std and dK/db_std of K_ss and K_ts (the adjoint with
  a one-hot cotangent) at 1e-6 relative, and of the directional derivative
  in the support inputs at 1e-5 (its path adds the patch-Gram adjoint, FP32
  sums over ~10^6 spatial terms) -- the adjoint's own accuracy;
* the loss gradients (W_std, b_std and a direction in the support inputs).
  The KRR loss with a trace-relative ridge is invariant to scaling K, so
  these derivatives are small remainders of cancelling sums of per-entry
  terms dL/dK_ij dK_ij/dtheta (here dL/dW_std = -3.5e-4 while the terms'
  absolute values sum to C_W = 2.8, about 8,000 times more).  FP32 rounding
  of each term (~1e-7 of it) then bounds a component's relative accuracy
  near 1e-3, whatever the adjoint does.  The bar is the first-order
  condition bound of the sum: |g - fd| <= 1e-5 C_theta + 1e-4 |fd|, with
  C_theta = sum |dL/dK_ij dK_ij/dtheta| from the FP64 oracle (dL/dK by
  torch autograd of the FP64 loss, dK/dtheta per entry by the same finite
  differences): every term may carry the per-entry error bounded above.

The oracle's evaluations (about 4 s each at 32x32) run in a thread pool
(numpy releases the GIL in its array loops).
"""
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest
import torch

from neural_tangents_amd import grad
from neural_tangents_amd.data import blurred_images, one_hot_targets
import _fd_oracle

pytestmark = pytest.mark.gpu

DEPTHS = (2, 1, 1)
DIAG_REG = _fd_oracle.DIAG_REG


def _richardson(fp, fm, fp2, fm2, h):
  d1 = (fp - fm) / (2 * h)
  d2 = (fp2 - fm2) / h
  return (4 * d2 - d1) / 3


def test_krr_gradients_32x32_match_finite_differences():
  xs, xt = blurred_images([2, 2], 32, seed=41)
  xs, xt = xs.astype(np.float64), xt.astype(np.float64)
  ys = one_hot_targets(2, seed=42).astype(np.float64)
  yt = one_hot_targets(2, seed=43).astype(np.float64)
  W, b, h = 1.3, 0.2, 1e-3
  v = np.random.default_rng(7).standard_normal(xs.shape)
  v /= np.linalg.norm(v)
  jobs = [(xs, ys, xt, yt, W, b)]
  for s in (h, h / 2):
    jobs += [(xs, ys, xt, yt, W + s, b), (xs, ys, xt, yt, W - s, b),
             (xs, ys, xt, yt, W, b + s), (xs, ys, xt, yt, W, b - s),
             (xs + s * v, ys, xt, yt, W, b), (xs - s * v, ys, xt, yt, W, b)]
  dev = torch.device('cuda', 0)
  with ThreadPoolExecutor(len(jobs)) as pool:
    pending = [pool.submit(_fd_oracle.loss_and_kernels, j) for j in jobs]
    xs_t = torch.tensor(xs, dtype=torch.float32, device=dev,
                        requires_grad=True)
    W_t = torch.tensor(W, dtype=torch.float64, device=dev,
                       requires_grad=True)
    b_t = torch.tensor(b, dtype=torch.float64, device=dev,
                       requires_grad=True)
    loss = grad.krr_loss(xs_t, torch.tensor(ys, device=dev),
                         torch.tensor(xt, dtype=torch.float32, device=dev),
                         torch.tensor(yt, device=dev), W_t, b_t,
                         grad.myrtle_builder(DEPTHS), diag_reg=DIAG_REG)
    loss.backward()
    # per-entry kernel derivatives from the adjoint (one-hot cotangents)
    plan = grad.myrtle_builder(DEPTHS)(W, b).plan(xs.shape)
    adj = {}
    for name, a, c in (('ss', xs, None), ('ts', xt, xs)):
      for i in range(2):
        for j in range(2):
          g = torch.zeros((2, 2), device=dev)
          g[i, j] = 1.0
          gx1, gx2, gwv, gbv = grad.kernel_vjp(
              plan, torch.tensor(a, dtype=torch.float32, device=dev),
              None if c is None else torch.tensor(c, dtype=torch.float32,
                                                  device=dev), None, g)
          gs = (gx1 if c is None else gx2).double().cpu().numpy()
          adj[name, i, j] = (2 * W * sum(gwv), 2 * b * sum(gbv),
                             float(np.sum(gs * v)))
    res = [f.result(timeout=200) for f in pending]
  L = [r[0] for r in res]
  ref = L[0]
  assert abs(float(loss.detach()) - ref) <= 1e-5 * abs(ref)

  def fd(k, idx=None):
    # k = 0: W, 1: b, 2: x; idx selects a kernel entry (name, i, j)
    vals = []
    for s in range(2):
      base = 1 + 6 * s + 2 * k
      for o in (0, 1):
        r = res[base + o]
        if idx is None:
          vals.append(r[0])
        else:
          name, i, j = idx
          vals.append((r[1] if name == 'ss' else r[2])[i, j])
    fp, fm, fp2, fm2 = vals
    return _richardson(fp, fm, fp2, fm2, h)

  for key, (gw, gb, gx) in adj.items():
    dw, db, dx = fd(0, key), fd(1, key), fd(2, key)
    assert abs(gw - dw) <= 1e-6 * abs(dw), (key, gw, dw)
    assert abs(gb - db) <= 1e-6 * abs(db), (key, gb, db)
    assert abs(gx - dx) <= 1e-5 * abs(dx), (key, gx, dx)
  # dL/dK of the FP64 loss at the oracle's kernels
  k_ss = torch.tensor(res[0][1], requires_grad=True)
  k_ts = torch.tensor(res[0][2], requires_grad=True)
  n = k_ss.shape[0]
  lam = DIAG_REG * torch.trace(k_ss) / n
  alpha = torch.linalg.solve(k_ss + lam * torch.eye(n, dtype=torch.float64),
                             torch.tensor(ys))
  loss64 = 0.5 * torch.mean((k_ts @ alpha - torch.tensor(yt)) ** 2)
  loss64.backward()
  g_ss, g_ts = k_ss.grad.numpy(), k_ts.grad.numpy()
  g_x = xs_t.grad.double().cpu().numpy()
  got = [float(W_t.grad), float(b_t.grad), float(np.sum(g_x * v))]
  for k in range(3):
    want = fd(k)
    cond = sum(abs(g_ss[i, j] * fd(k, ('ss', i, j))) +
               abs(g_ts[i, j] * fd(k, ('ts', i, j)))
               for i in range(2) for j in range(2))
    tol = 1e-5 * cond + 1e-4 * abs(want)
    assert abs(got[k] - want) <= tol, (k, got[k], want, cond)


def test_float_hyperparameters_give_input_gradients_only():
  """W_std / b_std passed as Python floats (fixed hyperparameters): the
  backward returns gradients for the inputs and None for the floats."""
  xs, xt = blurred_images([2, 3], 8, seed=44)
  dev = torch.device('cuda', 0)
  xs_t = torch.tensor(xs, device=dev, requires_grad=True)
  loss = grad.krr_loss(xs_t, torch.tensor(one_hot_targets(2, seed=1),
                                          device=dev),
                       torch.tensor(xt, device=dev),
                       torch.tensor(one_hot_targets(3, seed=2), device=dev),
                       1.4, 0.1, grad.myrtle_builder(DEPTHS),
                       diag_reg=DIAG_REG)
  loss.backward()
  assert xs_t.grad is not None and torch.isfinite(xs_t.grad).all()
  assert float(xs_t.grad.abs().max()) > 0
