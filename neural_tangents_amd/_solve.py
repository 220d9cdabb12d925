"""Cholesky solves of the regularised train kernel for ``predict``.

``_get_cho_solve`` / ``_add_diagonal_regularizer`` (``predict.py:1224-1258``)
on the GPU.  ``mode='fp64'`` is the plain FP64 Cholesky solve;
``mode='mixed'`` factors in FP32 with the hand-written MFMA Cholesky of
libnta.so and refines the solution with FP64 residuals (``nta_linalg``).
"""
import torch

from . import _lib


def add_diagonal_regularizer(A: torch.Tensor, diag_reg: float,
                             diag_reg_absolute_scale: bool) -> torch.Tensor:
  """``predict.py:1224-1232`` (returns a new tensor)."""
  n = A.shape[0]
  if not diag_reg_absolute_scale:
    diag_reg = diag_reg * torch.diagonal(A).to(torch.float64).sum() / n
  out = A.clone()
  out.diagonal().add_(diag_reg)
  return out


def regularizer_value(A: torch.Tensor, diag_reg: float,
                      diag_reg_absolute_scale: bool) -> float:
  """The FP64 shift lambda added to the diagonal."""
  if diag_reg_absolute_scale:
    return float(diag_reg)
  trace = float(torch.diagonal(A).to(torch.float64).sum())
  return float(diag_reg) * trace / A.shape[0]


def kernel_times(K: torch.Tensor, X: torch.Tensor) -> torch.Tensor:
  """``K @ X`` in FP64: the test-set mean ``tensordot(k_td, k_inv_y)``
  (``predict.py:696``) of a test-train kernel ``K`` [M, N] and the FP64
  solution ``X`` [N, C].

  An fp32 kernel on the GPU goes through ``nta_gemm_f32_f64`` (its entries
  are exact in FP64; products accumulate on fp64 MFMA), padded to the
  kernel's 32-row / 16-column granularity when needed; a float64 kernel is
  multiplied as given.
  """
  if K.ndim != 2 or X.ndim != 2 or K.shape[1] != X.shape[0]:
    raise ValueError(f'shapes {tuple(K.shape)} and {tuple(X.shape)} do not '
                     'multiply')
  if K.dtype != torch.float32 or not K.is_cuda:
    return K.to(X.device, torch.float64) @ X.to(torch.float64)
  m, n = K.shape
  X = X.to(K.device, torch.float64)
  if X.stride(1) != 1 or X.data_ptr() % 8 != 0:
    X = X.contiguous()
  mp, npad = -(-m // 32) * 32, -(-n // 16) * 16
  if (mp == m and npad == n and K.is_contiguous() and n % 4 == 0 and
      K.data_ptr() % 16 == 0):
    Kp = K
  else:
    Kp = torch.zeros((mp, npad), dtype=torch.float32, device=K.device)
    Kp[:m, :n] = K
  if npad != n:                      # X's rows must cover K's columns
    Xp = torch.zeros((npad, X.shape[1]), dtype=torch.float64,
                     device=K.device)
    Xp[:n] = X
    X = Xp
  stream = _lib.stream_handle(K.device)
  wb = _lib.value('nta_gemm_f32_f64_work_bytes', mp, npad)
  work = torch.empty((wb,), dtype=torch.uint8, device=K.device)
  C = X.shape[1]
  Y = torch.empty((mp, C), dtype=torch.float64, device=K.device)
  for c0 in range(0, C, 16):           # 16 columns per call, strided views
    nc = min(16, C - c0)
    _lib.call('nta_gemm_f32_f64', _lib.ptr(Kp), npad, mp, npad,
              X.data_ptr() + 8 * c0, X.stride(0), nc,
              Y.data_ptr() + 8 * c0, Y.stride(0), _lib.ptr(work), wb, stream)
  return Y[:m] if mp != m else Y


class CholeskySolver:
  """Factor once, then ``solve(b)`` for ``b`` of shape ``[N, K]``."""

  def __init__(self, A: torch.Tensor, diag_reg: float,
               diag_reg_absolute_scale: bool, mode: str = 'mixed',
               tol: float = 1e-10, max_iter: int = 200):
    if A.ndim != 2 or A.shape[0] != A.shape[1]:
      raise ValueError(f'k_train_train must be square, got {tuple(A.shape)}')
    if mode not in ('mixed', 'fp64'):
      raise ValueError(f'Unknown solver {mode!r}; use "mixed" or "fp64".')
    self.mode = mode
    self.device = A.device
    self.n = A.shape[0]
    self.lam = regularizer_value(A, diag_reg, diag_reg_absolute_scale)
    self.iterations = 0
    self.last_residual = None
    if mode == 'fp64':
      A64 = A.to(torch.float64).clone()
      A64.diagonal().add_(self.lam)
      L, info = torch.linalg.cholesky_ex(A64)
      if int(info) != 0:
        raise ValueError(
            f'Cholesky factorisation failed (leading minor {int(info)} is '
            'not positive definite); increase diag_reg.')
      self.L64 = L
    else:
      from . import _mixed
      self.impl = _mixed.MixedCholesky(A, self.lam, tol=tol,
                                       max_iter=max_iter)

  def solve(self, b: torch.Tensor) -> torch.Tensor:
    b = b.to(self.device, torch.float64)
    squeeze = b.ndim == 1
    if squeeze:
      b = b[:, None]
    if self.mode == 'fp64':
      x = torch.cholesky_solve(b, self.L64)
    else:
      x = self.impl.solve(b)
      self.iterations = self.impl.iterations
      self.last_residual = self.impl.last_residual
    return x[:, 0] if squeeze else x
