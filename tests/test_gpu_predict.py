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
  """The mixed solver checks the FP32 pivots at the first solve (the
  factorisation is not waited for in the constructor); the FP64 solver at
  construction."""
  from neural_tangents_amd import _solve
  K = -torch.eye(128, dtype=torch.float64).cuda()
  with pytest.raises(ValueError):
    _solve.CholeskySolver(K, 0.0, True, 'mixed').solve(
        torch.ones((128, 2), dtype=torch.float64, device='cuda'))
  with pytest.raises(ValueError):
    _solve.CholeskySolver(K, 0.0, True, 'fp64')


def _spectrum_kernel(n, cond, seed):
  """FP64 SPD kernel with eigenvalues log-spaced over [1/cond, 1] and a
  random orthogonal basis: its entries are not fp32-representable."""
  g = torch.Generator(device='cuda').manual_seed(seed)
  Q, _ = torch.linalg.qr(torch.randn(n, n, generator=g, device='cuda',
                                     dtype=torch.float64))
  ev = torch.logspace(0, -np.log10(cond), n, dtype=torch.float64,
                      device='cuda')
  K = (Q * ev) @ Q.T
  return 0.5 * (K + K.T)


def test_mixed_solver_on_fp64_kernel_solves_the_fp64_system():
  """A float64 kernel with cond ~ 1e6 whose entries are not
  fp32-representable: the mixed solve reads the FP64 entries in its residual
  operator (nta_pcg64_k64), so it matches the FP64 solve of the same system
  (the solution of fp32(K) differs from it by ~cond * 2^-24)."""
  from neural_tangents_amd import _solve
  n = 1024
  K = _spectrum_kernel(n, 1e6, seed=5)
  assert float((K.float().double() - K).abs().max()) > 0
  y = torch.randn(n, 10, dtype=torch.float64, device='cuda',
                  generator=torch.Generator(device='cuda').manual_seed(2))
  s = _solve.CholeskySolver(K, 1e-7, True, 'mixed')
  x = s.solve(y)
  assert s.impl.k64 and not s.impl.fell_back
  want = _solve.CholeskySolver(K, 1e-7, True, 'fp64').solve(y)
  err = rel_err(x.cpu().numpy(), want.cpu().numpy())
  assert err <= 1e-8, err
  # solving the fp32-rounded system instead is visibly different
  x32 = _solve.CholeskySolver(K.float(), 1e-7, True, 'fp64').solve(y)
  assert rel_err(x32.cpu().numpy(), want.cpu().numpy()) > 100 * max(err,
                                                                    1e-12)


def test_mixed_solver_falls_back_when_refinement_stops_short():
  """max_iter = 1 cannot reach the tolerance: the solve warns and returns
  the FP64 Cholesky solution instead of the unconverged iterate."""
  from neural_tangents_amd import _solve
  n = 512
  K = _spectrum_kernel(n, 1e5, seed=6).float()
  y = torch.randn(n, 3, dtype=torch.float64, device='cuda',
                  generator=torch.Generator(device='cuda').manual_seed(4))
  s = _solve.CholeskySolver(K, 1e-6, True, 'mixed', max_iter=1)
  with pytest.warns(RuntimeWarning, match='FP64 Cholesky'):
    x = s.solve(y)
  assert s.impl.fell_back
  assert s.last_residual > s.impl.accept_residual()
  want = _solve.CholeskySolver(K, 1e-6, True, 'fp64').solve(y)
  assert rel_err(x.cpu().numpy(), want.cpu().numpy()) <= 1e-12


def test_mixed_solve_n8192_at_myrtle_conditioning():
  """n = 8192 (the kernel-regression workload's train size) with the
  condition number of its Myrtle-5 NTK at diag_reg = 1e-4 (~2.6e6): the
  mixed solve matches the FP64 Cholesky solve at <= 1e-6 (alpha and a test
  prediction K_td alpha)."""
  from neural_tangents_amd import _solve
  n = 8192
  K = _spectrum_kernel(n, 2.6e6, seed=7).float()
  y = torch.from_numpy(one_hot_targets(n, seed=8)).double().cuda()
  s = _solve.CholeskySolver(K, 0.0, True, 'mixed')
  a = s.solve(y)
  want = _solve.CholeskySolver(K, 0.0, True, 'fp64').solve(y)
  err = rel_err(a.cpu().numpy(), want.cpu().numpy())
  print(f'n=8192 mixed vs fp64: rel err {err:.2e}, {s.iterations} iters, '
        f'residual {s.last_residual:.2e}', flush=True)
  assert err <= 1e-6
  ktd = K[:1024].double()
  assert rel_err((ktd @ a).cpu().numpy(), (ktd @ want).cpu().numpy()) <= 1e-6
