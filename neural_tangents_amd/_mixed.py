"""Mixed-precision Cholesky solve: FP32 factor on the GPU (libnta.so),
FP64 residuals, preconditioned conjugate gradients.

For ``A = K + lambda I`` (``predict.py:1224-1258``):

1. ``nta_chol_f32`` factors ``K32 + lambda I`` in FP32 (blocked, MFMA
   trailing updates) and forms ``W = L^-1``; ``M^-1 = W^T W`` is the
   preconditioner (two dense triangular products, ``nta_trmv``).
2. Conjugate gradients on ``A`` with FP64 vectors; ``A p`` is
   ``nta_symv64`` (the fp32 kernel upcast exactly, FP64 accumulation and the
   FP64 ``lambda`` shift), so the converged solution is the FP64 solution of
   the regularised system.  The FP32 factor only steers convergence.
   Iteration stops when every column's relative residual is below ``tol``
   or stops decreasing.
"""
import torch

from . import _lib


class MixedCholesky:
  NB = 64

  def __init__(self, K: torch.Tensor, lam: float, tol: float = 1e-12,
               max_iter: int = 200):
    if not K.is_cuda:
      raise RuntimeError('the mixed-precision solve runs on the GPU')
    n = K.shape[0]
    self.n = n
    self.device = K.device
    self.lam = float(lam)
    self.tol = tol
    self.max_iter = max_iter
    self.iterations = 0
    self.last_residual = None
    npad = -(-n // self.NB) * self.NB
    self.npad = npad
    stream = _lib.stream_handle(self.device)
    K32 = torch.zeros((npad, npad), dtype=torch.float32, device=self.device)
    K32[:n, :n] = K.to(torch.float32)
    self.K32 = K32                               # kept for FP64 residuals
    # The preconditioner factors K32 + shift I.  When FP32 breaks down (a
    # non-positive pivot: the kernel is too ill-conditioned for single
    # precision at this lambda), the shift grows by 10x; the FP64-residual
    # iteration still converges to the solution of the unshifted system.
    mean_diag = float(torch.diagonal(K32)[:n].double().mean())
    shift = self.lam
    for attempt in range(8):
      A = K32.clone()
      A.diagonal().add_(shift)
      if npad > n:                                # identity padding
        A.diagonal()[n:] = 1.0
      W = torch.zeros_like(A)
      info = torch.zeros((1,), dtype=torch.int32, device=self.device)
      _lib.call('nta_chol_f32', _lib.ptr(A), npad, npad, _lib.ptr(W), npad,
                _lib.ptr(info), stream)
      if int(info.item()) == 0:
        break
      shift = max(10.0 * shift, 1e-7 * mean_diag)
    else:
      raise ValueError(
          f'FP32 Cholesky failed at leading minor {int(info.item())} even '
          f'with shift {shift:.3g}: the train kernel is not positive '
          'definite; increase diag_reg or use solver="fp64".')
    self.shift = shift
    self.L = A
    self.W = W
    WT = torch.empty_like(W)
    _lib.call('nta_transpose_f32', _lib.ptr(W), npad, npad, _lib.ptr(WT),
              npad, stream)
    self.WT = WT

  # --------------------------------------------------------------- ops
  def _pad(self, x):
    if self.npad == self.n:
      return x.contiguous()
    out = torch.zeros((self.npad, x.shape[1]), dtype=x.dtype,
                      device=x.device)
    out[:self.n] = x
    return out

  def matvec(self, X):
    """(K + lambda I) X in FP64 (padded rows are identity)."""
    Y = torch.empty_like(X)
    _lib.call('nta_symv64', _lib.ptr(self.K32), self.npad, self.npad,
              self.lam, _lib.ptr(X), X.shape[1], _lib.ptr(Y),
              _lib.stream_handle(self.device))
    if self.npad > self.n:
      Y[self.n:] = X[self.n:] * (1.0 + self.lam)
    return Y

  def precond(self, R):
    """W^T (W R) with W = L^-1 (fp32 arithmetic)."""
    T = torch.empty_like(R)
    Z = torch.empty_like(R)
    st = _lib.stream_handle(self.device)
    _lib.call('nta_trmv', _lib.ptr(self.W), self.npad, self.npad, 0,
              _lib.ptr(R), R.shape[1], _lib.ptr(T), st)
    _lib.call('nta_trmv', _lib.ptr(self.WT), self.npad, self.npad, 1,
              _lib.ptr(T), R.shape[1], _lib.ptr(Z), st)
    return Z

  # -------------------------------------------------------------- solve
  def solve(self, b: torch.Tensor) -> torch.Tensor:
    b = b.to(self.device, torch.float64)
    out_cols = b.shape[1]
    chunks = []
    for c0 in range(0, out_cols, 16):
      chunks.append(self._solve16(b[:, c0:c0 + 16]))
    return torch.cat(chunks, dim=1) if len(chunks) > 1 else chunks[0]

  def _solve16(self, b):
    B = self._pad(b)
    bnorm = torch.linalg.vector_norm(B, dim=0).clamp_min(1e-300)
    X = self.precond(B)
    R = B - self.matvec(X)
    Z = self.precond(R)
    P = Z.clone()
    rz = (R * Z).sum(0)
    best = float('inf')
    stall = 0
    it = 0
    rel = torch.linalg.vector_norm(R, dim=0) / bnorm
    for it in range(1, self.max_iter + 1):
      Q = self.matvec(P)
      pq = (P * Q).sum(0)
      alpha = torch.where(pq != 0, rz / pq, torch.zeros_like(pq))
      X = X + alpha * P
      if it % 25 == 0:
        R = B - self.matvec(X)                 # true FP64 residual
      else:
        R = R - alpha * Q
      rel = torch.linalg.vector_norm(R, dim=0) / bnorm
      worst = float(rel.max())
      if worst <= self.tol:
        break
      if worst < best:
        best = worst
        stall = 0
      else:
        stall += 1
        if stall >= 10:           # no progress for 10 iterations
          break
      Z = self.precond(R)
      rz_new = (R * Z).sum(0)
      beta = torch.where(rz != 0, rz_new / rz, torch.zeros_like(rz))
      P = Z + beta * P
      rz = rz_new
    R = B - self.matvec(X)
    self.iterations = it
    self.last_residual = float(
        (torch.linalg.vector_norm(R, dim=0) / bnorm).max())
    return X[:self.n]
