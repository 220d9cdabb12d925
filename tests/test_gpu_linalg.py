"""FP32 blocked Cholesky, its block preconditioner, the FP64 PCG solve and
the FP32 x FP64 test-set product of libnta.so against torch FP64 linear
algebra (needs a GPU).

The factor is FP32, so the preconditioner is checked to FP32 accuracy
(relative error of (L L^T)^-1 R against the FP64 solve, scaled by the
condition number); the PCG solution must match the FP64 solve to 1e-9.
"""
import ctypes

import numpy as np
import pytest
import torch

from neural_tangents_amd import _lib

pytestmark = pytest.mark.gpu

DEV = torch.device('cuda', 0)


def _spd(n, seed, rank=None, const=0.5, ridge=1e-2):
  g = torch.Generator().manual_seed(seed)
  r = rank or n
  X = torch.randn(n, r, generator=g, dtype=torch.float64)
  K = X @ X.T / r + const * torch.ones(n, n, dtype=torch.float64)
  K += ridge * torch.eye(n, dtype=torch.float64)
  return K.float()


def _factor(K32):
  n = K32.shape[0]
  A = K32.clone().to(DEV)
  # NaN on entry: every entry a product reads must have been written
  W = torch.full_like(A, float('nan'))
  WT = torch.full_like(A, float('nan'))     # lower part is never read
  info = torch.zeros((1,), dtype=torch.int32, device=DEV)
  _lib.call('nta_chol_f32', _lib.ptr(A), n, n, _lib.ptr(W), n,
            _lib.ptr(WT), _lib.ptr(info), _lib.stream_handle(DEV))
  torch.cuda.synchronize()
  return W, WT, int(info.item())


@pytest.mark.parametrize('n', [64, 128, 192, 256, 320, 640, 1024, 1216])
def test_factor_blocks_match_fp64(n):
  """W holds L11^-1, L22^-1 on the diagonal blocks and L21 below them."""
  K = _spd(n, seed=n)
  W, WT, info = _factor(K)
  assert info == 0
  # WT holds W^T on and above the diagonal (32-row tiles)
  up = torch.ones(n, n, dtype=torch.bool, device=DEV).triu()
  assert torch.equal(WT[up], W.T.contiguous()[up])
  n1 = _lib.value('nta_chol_split', n)
  L = torch.linalg.cholesky(K.double())
  Wd = W.double().cpu()
  Li = torch.linalg.inv(L)
  tol = 2e-4
  # zeros above the diagonal inside every diagonal 64x64 block (triangular
  # products read whole 64-tiles); the rest of the upper triangle is not
  # written (NaN from entry)
  for b0 in range(0, n, 64):
    d = Wd[b0:b0 + 64, b0:b0 + 64]
    assert float(d.triu(1).abs().max()) == 0.0, b0
  if n1 == n:
    assert torch.allclose(Wd.tril(), Li, rtol=0, atol=tol * Li.abs().max())
    return
  for blk, want in ((Wd[:n1, :n1].tril(), Li[:n1, :n1]),
                    (Wd[n1:, n1:].tril(), torch.linalg.inv(L[n1:, n1:])),
                    (Wd[n1:, :n1], L[n1:, :n1])):
    err = (blk - want).abs().max() / want.abs().max()
    assert err <= tol, (n, float(err))


def test_factor_large_matches_fp64():
  """n = 6144: twelve 512-column panels, so every update kind (the next
  panel's head and middle columns, the bulk right of it beside the next
  panel's factorisation) and the batched inversion joins of both halves
  run; checked block by block against torch FP64 on the GPU."""
  n = 6144
  g = torch.Generator(device=DEV).manual_seed(11)
  X = torch.randn(n, n, generator=g, device=DEV, dtype=torch.float64)
  K = X @ X.T / n + 0.5
  K.diagonal().add_(1e-2)
  W, WT, info = _factor(K.float().cpu())
  assert info == 0
  n1 = _lib.value('nta_chol_split', n)
  L = torch.linalg.cholesky(K.float().double())
  Wd = W.double()
  for blk, want in ((Wd[:n1, :n1].tril(), torch.linalg.inv(L[:n1, :n1])),
                    (Wd[n1:, n1:].tril(), torch.linalg.inv(L[n1:, n1:])),
                    (Wd[n1:, :n1], L[n1:, :n1])):
    err = (blk - want).abs().max() / want.abs().max()
    assert err <= 2e-4, (n, float(err))


@pytest.mark.parametrize('n,nrhs', [(64, 3), (640, 10), (1024, 16),
                                    (1216, 5)])
def test_preconditioner_solves_in_fp32(n, nrhs):
  K = _spd(n, seed=3 * n)
  W, WT, _ = _factor(K)
  n1 = _lib.value('nta_chol_split', n)
  R = torch.randn(n, nrhs, generator=torch.Generator().manual_seed(1))
  Rd = R.to(DEV)
  Z = torch.empty_like(Rd)
  ws = torch.empty((_lib.value('nta_pcg64_work_bytes', n, nrhs),),
                   dtype=torch.uint8, device=DEV)
  _lib.call('nta_chol_precond', _lib.ptr(W), _lib.ptr(WT), n, n, n1,
            _lib.ptr(Rd), nrhs, _lib.ptr(Z), _lib.ptr(ws), ws.numel(),
            _lib.stream_handle(DEV))
  want = torch.linalg.solve(K.double(), R.double())
  kappa = float(torch.linalg.cond(K.double()))
  err = float((Z.double().cpu() - want).norm() / want.norm())
  assert err <= 1e-6 * kappa, (err, kappa)


@pytest.mark.parametrize('n,nrhs,lam', [(640, 10, 1e-3), (1024, 4, 1e-3),
                                        (1216, 16, 1e-2)])
def test_pcg_matches_fp64_solve(n, nrhs, lam):
  K = _spd(n, seed=7 * n, rank=n // 2, const=1.0, ridge=0.0)
  W, WT, info = _factor(K + lam * torch.eye(n))
  assert info == 0
  n1 = _lib.value('nta_chol_split', n)
  B = torch.randn(n, nrhs, dtype=torch.float64,
                  generator=torch.Generator().manual_seed(2))
  X = torch.empty((n, nrhs), dtype=torch.float64, device=DEV)
  ws = torch.empty((_lib.value('nta_pcg64_work_bytes', n, nrhs),),
                   dtype=torch.uint8, device=DEV)
  iters, resid = ctypes.c_int(0), ctypes.c_double(0.0)
  K_dev = K.to(DEV)
  _lib.call('nta_pcg64', _lib.ptr(K_dev), n, n, lam, _lib.ptr(W),
            _lib.ptr(WT), n, n1,
            _lib.ptr(B.to(DEV)), nrhs, _lib.ptr(X), 1e-13, 100,
            _lib.ptr(ws), ws.numel(), ctypes.byref(iters),
            ctypes.byref(resid), _lib.stream_handle(DEV))
  A = K.double() + lam * torch.eye(n, dtype=torch.float64)
  want = torch.linalg.solve(A, B)
  err = float((X.cpu() - want).norm() / want.norm())
  assert err <= 1e-9, (err, iters.value, resid.value)
  # the recurrence residual reaches 1e-13; the recomputed one is floored by
  # the FP64 rounding of the products (~ kappa * 1e-16, kappa ~ 1e6 here)
  assert resid.value <= 1e-8, resid.value
  assert 1 <= iters.value < 100


def test_factor_reports_non_positive_pivot():
  n = 256
  K = _spd(n, seed=5)
  K[100, 100] = -1.0
  _, _, info = _factor(K)
  assert info == 101


def test_invalid_arguments_raise():
  W = torch.zeros((64, 64), device=DEV)
  with pytest.raises(ValueError):
    _lib.call('nta_chol_f32', _lib.ptr(W), 0, 64, _lib.ptr(W), 64,
              _lib.ptr(W), None, None)
  with pytest.raises(NotImplementedError):
    _lib.call('nta_chol_f32', _lib.ptr(W), 60, 64, _lib.ptr(W), 64,
              _lib.ptr(W), _lib.ptr(W), None)
  assert _lib.value('nta_pcg64_work_bytes', 64, 17) == -1
  assert _lib.value('nta_chol_split', 100) == 0


@pytest.mark.parametrize('m,n,c', [(1024, 8192, 10), (96, 640, 16),
                                   (50, 300, 3), (64, 256, 20),
                                   (32, 4000, 5), (8192, 1024, 10)])
def test_kernel_times_matches_fp64_product(m, n, c):
  """nta_gemm_f32_f64 (through _solve.kernel_times, padded when m, n are
  not multiples of 32 and 16, in 16-column chunks) against the FP64 product
  of the same fp32 entries: only the summation order differs."""
  from neural_tangents_amd import _solve
  g = torch.Generator(device=DEV).manual_seed(m + n + c)
  K = torch.randn(m, n, generator=g, device=DEV)
  X = torch.randn(n, c, generator=g, device=DEV, dtype=torch.float64)
  Y = _solve.kernel_times(K, X)
  want = K.double() @ X
  assert Y.shape == (m, c) and Y.dtype == torch.float64
  err = float((Y - want).abs().max() / want.abs().max())
  assert err <= 1e-13, err


def test_kernel_times_rejects_unpadded_shapes():
  K = torch.zeros((48, 64), device=DEV)
  X = torch.zeros((64, 16), dtype=torch.float64, device=DEV)
  Y = torch.empty((48, 16), dtype=torch.float64, device=DEV)
  with pytest.raises(NotImplementedError):
    _lib.call('nta_gemm_f32_f64', _lib.ptr(K), 64, 48, 64, _lib.ptr(X), 16,
              16, _lib.ptr(Y), 16, None, 0, _lib.stream_handle(DEV))
  assert _lib.value('nta_gemm_f32_f64_work_bytes', 48, 64) == -1
  # narrower than 16 columns needs the workspace
  K = torch.zeros((64, 64), device=DEV)
  with pytest.raises(ValueError, match='workspace'):
    _lib.call('nta_gemm_f32_f64', _lib.ptr(K), 64, 64, 64, _lib.ptr(X), 16,
              10, _lib.ptr(Y), 16, None, 0, _lib.stream_handle(DEV))


@pytest.mark.parametrize('ncols,ldx,ldy', [(10, 10, 10), (16, 16, 16),
                                           (3, 7, 5), (16, 20, 18)])
def test_gemm_f32_f64_strided_columns(ncols, ldx, ldy):
  """Y[:, :ncols] = K X[:, :ncols] with leading dimensions ldx, ldy: the
  columns past ncols of Y are left untouched; with and without the
  workspace when the layout is dense."""
  m, n = 96, 512
  g = torch.Generator(device=DEV).manual_seed(ncols + ldx)
  K = torch.randn(m, n, generator=g, device=DEV)
  X = torch.randn(n, ldx, generator=g, device=DEV, dtype=torch.float64)
  Y = torch.full((m, ldy), 5.0, dtype=torch.float64, device=DEV)
  wb = _lib.value('nta_gemm_f32_f64_work_bytes', m, n)
  work = torch.empty((wb,), dtype=torch.uint8, device=DEV)
  _lib.call('nta_gemm_f32_f64', _lib.ptr(K), n, m, n, _lib.ptr(X), ldx,
            ncols, _lib.ptr(Y), ldy, _lib.ptr(work), wb,
            _lib.stream_handle(DEV))
  want = K.double() @ X[:, :ncols]
  err = float((Y[:, :ncols] - want).abs().max() / want.abs().max())
  assert err <= 1e-13, err
  assert bool((Y[:, ncols:] == 5.0).all())
  if ncols == ldx == ldy == 16:
    Y2 = torch.empty_like(Y)
    _lib.call('nta_gemm_f32_f64', _lib.ptr(K), n, m, n, _lib.ptr(X), 16,
              16, _lib.ptr(Y2), 16, None, 0, _lib.stream_handle(DEV))
    err = float((Y2 - want).abs().max() / want.abs().max())
    assert err <= 1e-13, err


def test_copy_lower_shift_writes_lower_tiles_only():
  """nta_copy_lower_shift: K + shift I on the 64x64 tiles on and below the
  diagonal, the tiles above untouched; the solver factors this copy."""
  n = 320
  K = torch.randn(n, n, generator=torch.Generator().manual_seed(4)).to(DEV)
  A = torch.full_like(K, 7.0)
  _lib.call('nta_copy_lower_shift', _lib.ptr(K), n, n, 0.25, _lib.ptr(A), n,
            _lib.stream_handle(DEV))
  torch.cuda.synchronize()
  t = torch.arange(n, device=DEV) // 64
  lower = t[:, None] >= t[None, :]
  want = K + 0.25 * torch.eye(n, device=DEV)
  assert torch.equal(A[lower], want[lower])
  assert bool((A[~lower] == 7.0).all())
  with pytest.raises(NotImplementedError):
    _lib.call('nta_copy_lower_shift', _lib.ptr(K), n, 300, 0.0, _lib.ptr(A),
              n, _lib.stream_handle(DEV))
  with pytest.raises(ValueError):
    _lib.call('nta_copy_lower_shift', None, n, n, 0.0, _lib.ptr(A), n,
              _lib.stream_handle(DEV))
