"""FP64 oracle loss and kernels for the finite-difference checks of the
adjoint kernels (numpy only)."""
import numpy as np

from oracle import stax_oracle as so

DEPTHS = (2, 1, 1)
DIAG_REG = 1e-2


def kernels(xs, xt, W, b):
  """(K_ss, K_ts) NTKs of the Myrtle-5 network with shared conv W_std,
  b_std."""
  layers = so.myrtle_layers(DEPTHS, W_std=W, b_std=b)
  k_ss = so.kernel_fn(layers, xs, None, 'ntk')['ntk']
  k_ts = so.kernel_fn(layers, xt, xs, 'ntk')['ntk']
  return k_ss, k_ts


def loss_of(k_ss, k_ts, ys, yt):
  """0.5 mean((K_ts (K_ss + lambda I)^-1 Y_s - Y_t)^2), lambda =
  DIAG_REG tr(K_ss) / n."""
  n = k_ss.shape[0]
  lam = DIAG_REG * np.trace(k_ss) / n
  alpha = np.linalg.solve(k_ss + lam * np.eye(n), ys)
  f = k_ts @ alpha
  return 0.5 * np.mean((f - yt) ** 2)


def krr_loss(args):
  xs, ys, xt, yt, W, b = args
  return loss_of(*kernels(xs, xt, W, b), ys, yt)


def loss_and_kernels(args):
  """(loss, K_ss, K_ts) at (xs, ys, xt, yt, W, b)."""
  xs, ys, xt, yt, W, b = args
  k_ss, k_ts = kernels(xs, xt, W, b)
  return loss_of(k_ss, k_ts, ys, yt), k_ss, k_ts
