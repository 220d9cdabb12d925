"""Adjoint kernels at the differentiation form's image size (needs a GPU).

The differentiation workload is a kernel-regression loss of a support set
against target images at 32x32x3 (Myrtle-5 NTK, learnable support inputs,
W_std and b_std).  Its gradients are checked here at that image size (the
fused first-two-stage path runs at 32x32 in the forward) on 2 support and 2
target images against central finite differences (h = 1e-3) of the FP64
oracle's loss, at 1e-4.  The oracle's loss evaluations (about 4 s each at
32x32) run in a thread pool (numpy releases the GIL in its array loops).
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
DIAG_REG = 1e-2


def test_krr_gradients_32x32_match_finite_differences():
  xs, xt = blurred_images([2, 2], 32, seed=41)
  xs, xt = xs.astype(np.float64), xt.astype(np.float64)
  ys = one_hot_targets(2, seed=42).astype(np.float64)
  yt = one_hot_targets(2, seed=43).astype(np.float64)
  W, b, h = 1.3, 0.2, 1e-3
  v = np.random.default_rng(7).standard_normal(xs.shape)
  v /= np.linalg.norm(v)
  base = (xs, ys, xt, yt, W, b)
  jobs = [base,
          (xs, ys, xt, yt, W + h, b), (xs, ys, xt, yt, W - h, b),
          (xs, ys, xt, yt, W, b + h), (xs, ys, xt, yt, W, b - h),
          (xs + h * v, ys, xt, yt, W, b), (xs - h * v, ys, xt, yt, W, b)]
  with ThreadPoolExecutor(len(jobs)) as pool:
    pending = [pool.submit(_fd_oracle.krr_loss, j) for j in jobs]
    dev = torch.device('cuda', 0)
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
    ref, wp, wm, bp, bm, xp, xm = [f.result(timeout=100) for f in pending]
  assert abs(float(loss) - ref) <= 1e-5 * abs(ref)
  fd_w, fd_b = (wp - wm) / (2 * h), (bp - bm) / (2 * h)
  assert abs(float(W_t.grad) - fd_w) <= 1e-4 * abs(fd_w), (W_t.grad, fd_w)
  assert abs(float(b_t.grad) - fd_b) <= 1e-4 * abs(fd_b), (b_t.grad, fd_b)
  g = xs_t.grad.double().cpu().numpy()
  fd = (xp - xm) / (2 * h)
  assert abs(float(np.sum(g * v)) - fd) <= 1e-4 * np.linalg.norm(g), (
      float(np.sum(g * v)), fd)


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
