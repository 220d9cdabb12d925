"""Adjoint kernels: gradients of a kernel-regression loss with respect to
the support inputs, W_std and b_std against central finite differences of
the FP64 oracle (h = 1e-3), at 1e-4 relative error (needs a GPU).

The reference's own gradients (jax.grad) cannot be evaluated here (jax is not
installed), so parity with them is unpinned; the finite differences of the
FP64 restatement are the check.
"""
import numpy as np
import pytest
import torch

from neural_tangents_amd import grad
from neural_tangents_amd.data import blurred_images, one_hot_targets
from oracle import stax_oracle as so

pytestmark = pytest.mark.gpu

DEPTHS = (2, 1, 1)
DIAG_REG = 1e-2


def oracle_loss(xs, ys, xt, yt, W, b, get='ntk'):
  layers = so.myrtle_layers(DEPTHS, W_std=W, b_std=b)
  k_ss = so.kernel_fn(layers, xs, None, get)[get]
  k_ts = so.kernel_fn(layers, xt, xs, get)[get]
  n = k_ss.shape[0]
  lam = DIAG_REG * np.trace(k_ss) / n
  alpha = np.linalg.solve(k_ss + lam * np.eye(n), ys)
  f = k_ts @ alpha
  return 0.5 * np.mean((f - yt) ** 2)


def _setup(side=8, ns=3, nt=4):
  xs, xt = blurred_images([ns, nt], side, seed=21)
  ys = one_hot_targets(ns, seed=22).astype(np.float64)
  yt = one_hot_targets(nt, seed=23).astype(np.float64)
  return xs.astype(np.float64), xt.astype(np.float64), ys, yt


@pytest.mark.parametrize('get', ['ntk', 'nngp'])
def test_krr_gradients_match_finite_differences(get):
  xs, xt, ys, yt = _setup()
  W, b = 1.3, 0.2
  dev = torch.device('cuda', 0)
  xs_t = torch.tensor(xs, dtype=torch.float32, device=dev,
                      requires_grad=True)
  W_t = torch.tensor(W, dtype=torch.float64, device=dev, requires_grad=True)
  b_t = torch.tensor(b, dtype=torch.float64, device=dev, requires_grad=True)
  loss = grad.krr_loss(xs_t, torch.tensor(ys, device=dev),
                       torch.tensor(xt, dtype=torch.float32, device=dev),
                       torch.tensor(yt, device=dev), W_t, b_t,
                       grad.myrtle_builder(DEPTHS), diag_reg=DIAG_REG,
                       get=get)
  loss.backward()
  ref = oracle_loss(xs, ys, xt, yt, W, b, get)
  assert abs(float(loss) - ref) <= 1e-5 * abs(ref)
  h = 1e-3
  fd_w = (oracle_loss(xs, ys, xt, yt, W + h, b, get) -
          oracle_loss(xs, ys, xt, yt, W - h, b, get)) / (2 * h)
  fd_b = (oracle_loss(xs, ys, xt, yt, W, b + h, get) -
          oracle_loss(xs, ys, xt, yt, W, b - h, get)) / (2 * h)
  assert abs(float(W_t.grad) - fd_w) <= 1e-4 * abs(fd_w), (W_t.grad, fd_w)
  assert abs(float(b_t.grad) - fd_b) <= 1e-4 * abs(fd_b), (b_t.grad, fd_b)
  g = xs_t.grad.double().cpu().numpy()
  rng = np.random.default_rng(5)
  for _ in range(3):
    v = rng.standard_normal(xs.shape)
    v /= np.linalg.norm(v)
    fd = (oracle_loss(xs + h * v, ys, xt, yt, W, b, get) -
          oracle_loss(xs - h * v, ys, xt, yt, W, b, get)) / (2 * h)
    got = float(np.sum(g * v))
    assert abs(got - fd) <= 1e-4 * np.linalg.norm(g), (got, fd)


def test_b_std_gradient_is_zero_at_zero_bias():
  """Only b_std^2 enters the kernel (stax/linear.py:1385-1389)."""
  xs, xt, ys, yt = _setup()
  dev = torch.device('cuda', 0)
  b_t = torch.tensor(0.0, dtype=torch.float64, device=dev,
                     requires_grad=True)
  loss = grad.krr_loss(torch.tensor(xs, dtype=torch.float32, device=dev),
                       torch.tensor(ys, device=dev),
                       torch.tensor(xt, dtype=torch.float32, device=dev),
                       torch.tensor(yt, device=dev),
                       torch.tensor(1.4, dtype=torch.float64, device=dev),
                       b_t, grad.myrtle_builder(DEPTHS), diag_reg=DIAG_REG)
  loss.backward()
  assert float(b_t.grad) == 0.0


def test_kernel_vjp_cross_equals_finite_difference_of_kernel():
  """Directional derivative of sum(G * K(x1, x2)) in x2 (second argument,
  cross kernel) against the oracle."""
  x1, x2 = blurred_images([2, 3], 8, seed=31)
  x1 = x1.astype(np.float64)
  x2 = x2.astype(np.float64)
  W, b = 1.1, 0.3
  G = np.random.default_rng(1).standard_normal((2, 3))
  dev = torch.device('cuda', 0)
  x2_t = torch.tensor(x2, dtype=torch.float32, device=dev,
                      requires_grad=True)
  nngp, ntk = grad.kernel(torch.tensor(x1, dtype=torch.float32, device=dev),
                          x2_t, torch.tensor(W, dtype=torch.float64,
                                             device=dev),
                          torch.tensor(b, dtype=torch.float64, device=dev),
                          grad.myrtle_builder(DEPTHS))
  (ntk.double() * torch.tensor(G, device=dev)).sum().backward()
  g = x2_t.grad.double().cpu().numpy()
  layers = so.myrtle_layers(DEPTHS, W_std=W, b_std=b)
  f = lambda z: float(np.sum(G * so.kernel_fn(layers, x1, z, 'ntk')['ntk']))
  v = np.random.default_rng(2).standard_normal(x2.shape)
  v /= np.linalg.norm(v)
  h = 1e-3
  fd = (f(x2 + h * v) - f(x2 - h * v)) / (2 * h)
  assert abs(float(np.sum(g * v)) - fd) <= 1e-4 * np.linalg.norm(g)


@pytest.mark.parametrize('sym', [False, True])
def test_kernel_vjp_is_independent_of_the_pair_tiling(sym):
  """Tiles of 1 x 2 pairs (ragged at the edges, and off the diagonal of
  K(x, x) with a self offset) give the cotangents of one tile."""
  xs, xt, _, _ = _setup(ns=5, nt=3)
  dev = torch.device('cuda', 0)
  plan = grad.myrtle_builder(DEPTHS)(1.3, 0.2).plan(xs.shape)
  x1 = torch.tensor(xs, dtype=torch.float32, device=dev)
  x2 = None if sym else torch.tensor(xt, dtype=torch.float32, device=dev)
  n2 = 5 if sym else 3
  rng = np.random.default_rng(8)
  g_n = torch.tensor(rng.standard_normal((5, n2)), dtype=torch.float32,
                     device=dev)
  g_t = torch.tensor(rng.standard_normal((5, n2)), dtype=torch.float32,
                     device=dev)
  one = grad.kernel_vjp(plan, x1, x2, g_n, g_t)
  tiled = grad.kernel_vjp(plan, x1, x2, g_n, g_t, tile_pairs=2)
  torch.testing.assert_close(tiled[0], one[0], rtol=1e-5,
                             atol=1e-5 * float(one[0].abs().max()))
  if not sym:
    torch.testing.assert_close(tiled[1], one[1], rtol=1e-5,
                               atol=1e-5 * float(one[1].abs().max()))
  for k in (2, 3):                       # d/dw_var, d/db_var per stage
    np.testing.assert_allclose(tiled[k], one[k], rtol=1e-5,
                               atol=1e-6 * max(abs(v) for v in one[k]))
