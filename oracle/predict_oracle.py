"""FP64 restatement of the kernel-regression predictors (test infrastructure).

TEST INFRASTRUCTURE ONLY -- never imported by the product package.

Follows ``neural_tangents/_src/predict.py`` of neural-tangents 0.6.6 for the
flat case (``K`` is ``[N, N]``, ``y`` is ``[N, K]``, ``trace_axes=(-1,)``).
"""
import numpy as np
import scipy.linalg


def add_diagonal_regularizer(A, diag_reg, diag_reg_absolute_scale):
  """``predict.py:1224-1232``."""
  n = A.shape[0]
  if not diag_reg_absolute_scale:
    diag_reg = diag_reg * np.trace(A) / n
  return A + diag_reg * np.eye(n, dtype=A.dtype)


def cho_solve_fn(A, diag_reg, diag_reg_absolute_scale):
  """``_get_cho_solve`` ``predict.py:1235-1258``: upper Cholesky (line 1244)
  and ``cho_solve`` (line 1254)."""
  A = add_diagonal_regularizer(np.asarray(A, np.float64), diag_reg,
                               diag_reg_absolute_scale)
  C = scipy.linalg.cho_factor(A, lower=False)
  return lambda b: scipy.linalg.cho_solve(C, np.asarray(b, np.float64))


def gp_inference_mean(k_train_train, y_train, k_test_train, diag_reg=0.0,
                      diag_reg_absolute_scale=False):
  """Mean of ``gp_inference`` ``predict.py:566-745`` (``k_inv_y`` 633-635,
  test mean 696)."""
  solve = cho_solve_fn(k_train_train, diag_reg, diag_reg_absolute_scale)
  alpha = solve(y_train)
  if k_test_train is None:
    return np.asarray(y_train, np.float64), alpha
  return np.asarray(k_test_train, np.float64) @ alpha, alpha


def gp_inference_cov(k_train_train_nngp, k_train_train_g, k_test_train_nngp,
                     k_test_train_g, k_test_test_nngp, g, diag_reg=0.0,
                     diag_reg_absolute_scale=False):
  """Posterior covariance of ``gp_inference`` ``predict.py:699-739``."""
  solve_g = cho_solve_fn(k_train_train_g, diag_reg, diag_reg_absolute_scale)
  k_td = np.asarray(k_test_train_g, np.float64)
  k_td_nngp = np.asarray(k_test_train_nngp, np.float64)
  k_tt = np.asarray(k_test_test_nngp, np.float64)
  if g == 'nngp':
    return k_tt - k_td @ solve_g(k_td.T)
  term_1 = solve_g(k_td.T)                    # [N, T]
  cov = np.asarray(k_train_train_nngp, np.float64) @ term_1
  cov = term_1.T @ cov
  term_2 = k_td @ solve_g(k_td_nngp.T)
  term_2 = term_2 + term_2.T
  return cov - term_2 + k_tt


def ensemble_mean_finite_t(k_train_train, y_train, k_test_train, t,
                           learning_rate=1.0, diag_reg=0.0,
                           diag_reg_absolute_scale=False):
  """Finite-time mean of ``gradient_descent_mse_ensemble``
  (``predict.py:876-883, 994-1036``) via the eigendecomposition of the
  regularised train kernel; ``expm1``/``inv_expm1`` of 1277-1294."""
  A = add_diagonal_regularizer(np.asarray(k_train_train, np.float64),
                               diag_reg, diag_reg_absolute_scale)
  evals, evecs = np.linalg.eigh(A)
  size = np.asarray(y_train).size
  tt = np.asarray(t, np.float64) * learning_rate
  expm1 = np.expm1(-np.maximum(evals, 0.0) * tt / size)
  y = np.asarray(y_train, np.float64)
  proj = evecs.T @ y
  if k_test_train is None:
    return evecs @ (-expm1[:, None] * proj)
  fac = -expm1 / np.abs(evals)
  return np.asarray(k_test_train, np.float64) @ (evecs @ (fac[:, None] * proj))


def ensemble_cov_finite_t(k_train_train_g, nngp_train_train, k_test_train_g,
                          k_test_train_nngp, nngp_test_test, t, g,
                          size, learning_rate=1.0, diag_reg=0.0,
                          diag_reg_absolute_scale=False):
  """Finite-time posterior covariance of ``gradient_descent_mse_ensemble``
  (``predict.py:1038-1090``) for one time ``t``: ``g`` is the kernel that
  drives training ('nngp' or 'ntk'), ``size`` is ``y_train.size`` (the
  normaliser of ``expm1``, 1277-1294).  ``k_test_train_g is None`` gives the
  training-set covariance (1043-1069), else the test set (1072-1090)."""
  A = add_diagonal_regularizer(np.asarray(k_train_train_g, np.float64),
                               diag_reg, diag_reg_absolute_scale)
  evals, V = np.linalg.eigh(A)
  tt = float(t) * learning_rate
  lam = np.maximum(evals, 0.0)

  def inv_expm1(s):
    return np.expm1(-lam * s / size) / np.abs(evals)

  nngp_dd = np.asarray(nngp_train_train, np.float64)
  if k_test_train_g is None:
    if g == 'nngp':
      return (V * (lam * np.exp(-2 * lam * tt / size))) @ V.T
    ex = (V * np.exp(-lam * tt / size)) @ V.T
    return ex @ nngp_dd @ ex.T
  ktd = np.asarray(k_test_train_g, np.float64)
  ntt = np.asarray(nngp_test_test, np.float64)
  if g == 'nngp':
    return ntt - ktd @ ((V * -inv_expm1(2 * tt)) @ V.T) @ ktd.T
  neg = -inv_expm1(tt)
  term_1 = (V * neg) @ V.T @ ktd.T                     # [N, T]
  term_2 = ktd @ ((V * neg) @ V.T) @ np.asarray(k_test_train_nngp,
                                                np.float64).T
  term_2 = term_2 + term_2.T
  return term_1.T @ nngp_dd @ term_1 - term_2 + ntt
