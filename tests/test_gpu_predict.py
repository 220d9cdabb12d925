"""Kernel regression on the GPU against the FP64 oracle (needs a GPU).

The mixed solver (FP32 Cholesky + FP64-residual refinement) must match the
FP64 oracle on the same kernel matrices within 1e-6 relative error; the
plain FP64 solver within 1e-8 (two FP64 solvers of a kernel with condition
number ~1e6 differ at ~1e-10).
"""
import numpy as np
import pytest
import torch

from conftest import rel_err
from neural_tangents_amd import predict, stax
from neural_tangents_amd.data import blurred_images, one_hot_targets
from oracle import predict_oracle as po

pytestmark = pytest.mark.gpu


def _spd(n, rank, seed, scale_mean=0.5):
  g = torch.Generator().manual_seed(seed)
  X = torch.randn(n, rank, generator=g, dtype=torch.float64)
  K = X @ X.T / rank + scale_mean * torch.ones(n, n, dtype=torch.float64)
  return K.float().double()     # an fp32-representable kernel


@pytest.mark.parametrize('n', [200, 512])
@pytest.mark.parametrize('solver,tol', [('fp64', 1e-8), ('mixed', 1e-6)])
def test_gp_inference_mean_matches_oracle(n, solver, tol):
  K = _spd(n + 40, 64, seed=n)
  k_dd, k_td = K[:n, :n], K[n:, :n]
  y = torch.from_numpy(one_hot_targets(n, seed=1)).double()
  fn = predict.gp_inference(k_dd.cuda(), y.cuda(), diag_reg=1e-4,
                            solver=solver)
  got = fn('ntk', k_td.cuda())
  want, _ = po.gp_inference_mean(k_dd.numpy(), y.numpy(), k_td.numpy(),
                                 diag_reg=1e-4)
  assert rel_err(got.cpu().numpy(), want) <= tol


def test_mixed_solver_refines_to_fp64_on_ill_conditioned_kernel():
  """diag_reg = 1e-6 on a rank-48 kernel: the FP32 factorisation of K +
  lambda I breaks down, so the preconditioner is factored with a larger
  shift and the FP64-residual iteration still reaches the FP64 solution."""
  from neural_tangents_amd import _solve
  n = 640
  K = _spd(n, 48, seed=7, scale_mean=2.0).cuda()
  y = torch.randn(n, 10, dtype=torch.float64,
                  generator=torch.Generator().manual_seed(3)).cuda()
  s = _solve.CholeskySolver(K, 1e-6, False, 'mixed')
  x = s.solve(y)
  want = _solve.CholeskySolver(K, 1e-6, False, 'fp64').solve(y)
  assert rel_err(x.cpu().numpy(), want.cpu().numpy()) <= 1e-6
  assert s.impl.shift > s.lam


def test_ensemble_on_myrtle_kernel():
  """gradient_descent_mse_ensemble(t=None) == gp_inference on the same
  kernels == oracle solve; t -> inf matches t=None; train-set outputs."""
  xtr, xte = blurred_images([96, 12], 8, seed=2)
  ytr = one_hot_targets(96, seed=4)
  kernel_fn = stax.myrtle(5)[2]
  pred = predict.gradient_descent_mse_ensemble(kernel_fn, xtr, ytr,
                                               diag_reg=1e-4)
  out = pred(x_test=xte, get=('nngp', 'ntk'))
  k_dd = kernel_fn(xtr, None, ('nngp', 'ntk'))
  k_td = kernel_fn(xte, xtr, ('nngp', 'ntk'))
  for g in ('nngp', 'ntk'):
    want, _ = po.gp_inference_mean(getattr(k_dd, g).double().cpu().numpy(),
                                   ytr, getattr(k_td, g).double().cpu()
                                   .numpy(), diag_reg=1e-4)
    assert rel_err(getattr(out, g).cpu().numpy(), want) <= 1e-6, g
  inf = pred(t=1e12, x_test=xte, get='ntk')
  assert rel_err(inf.cpu().numpy(), out.ntk.cpu().numpy()) <= 1e-5
  train = pred(x_test=None, get='nngp')
  np.testing.assert_allclose(train.cpu().numpy(), ytr, rtol=0, atol=0)
  fin = pred(t=3.0, x_test=xte, get='ntk')
  want_fin = po.ensemble_mean_finite_t(k_dd.ntk.double().cpu().numpy(), ytr,
                                       k_td.ntk.double().cpu().numpy(), 3.0,
                                       diag_reg=1e-4)
  assert rel_err(fin.cpu().numpy(), want_fin) <= 1e-8


def test_posterior_covariance_matches_oracle():
  n, t = 128, 20
  K = _spd(n + t, 32, seed=11)
  Kn = _spd(n + t, 32, seed=12)
  y = torch.from_numpy(one_hot_targets(n, seed=5)).double()
  tt = (Kn[n:, n:] + Kn[n:, n:].T) / 2
  kdd = predict._Kernel(nngp=Kn[:n, :n].cuda(), ntk=K[:n, :n].cuda())
  ktd = predict._Kernel(nngp=Kn[n:, :n].cuda(), ntk=K[n:, :n].cuda())
  fn = predict.gp_inference(kdd, y.cuda(), diag_reg=1e-3, solver='fp64')
  out = fn(('nngp', 'ntk'), ktd, tt.cuda())
  for g in ('nngp', 'ntk'):
    want = po.gp_inference_cov(Kn[:n, :n].numpy(),
                               (Kn if g == 'nngp' else K)[:n, :n].numpy(),
                               Kn[n:, :n].numpy(),
                               (Kn if g == 'nngp' else K)[n:, :n].numpy(),
                               tt.numpy(), g, diag_reg=1e-3)
    assert rel_err(getattr(out, g).covariance.cpu().numpy(), want) <= 1e-9, g


def test_non_spd_raises():
  from neural_tangents_amd import _solve
  K = -torch.eye(128, dtype=torch.float64).cuda()
  with pytest.raises(ValueError):
    _solve.CholeskySolver(K, 0.0, True, 'mixed')
  with pytest.raises(ValueError):
    _solve.CholeskySolver(K, 0.0, True, 'fp64')
