"""Mixed-precision Cholesky solve: FP32 factor on the GPU (libnta.so),
FP64 residuals, preconditioned conjugate gradients.

For ``A = K + lambda I`` (``predict.py:1224-1258``):

1. ``nta_chol_f32`` factors ``K32 + shift I`` in FP32 by recursive blocking
   on fp32 MFMA and inverts the factor on the way; the preconditioner
   ``(L L^T)^-1`` is applied from the blocks it leaves (``nta_chol_precond``).
2. ``nta_pcg64`` runs conjugate gradients on ``A`` with FP64 vectors:
   ``A p`` reads the fp32 kernel (exact in FP64), accumulates in FP64 and
   adds the FP64 ``lambda``, so the converged solution is the FP64 solution
   of the regularised system; the FP32 factor only steers convergence.
"""
import ctypes

import torch

from . import _lib


class MixedCholesky:
  NB = 64

  def __init__(self, K: torch.Tensor, lam: float, tol: float = 1e-10,
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
    self.n1 = _lib.value('nta_chol_split', npad)
    stream = _lib.stream_handle(self.device)
    if npad == n and K.dtype == torch.float32 and K.is_contiguous():
      K32 = K
    else:
      K32 = torch.zeros((npad, npad), dtype=torch.float32, device=self.device)
      K32[:n, :n] = K.to(torch.float32)
      if npad > n:                                # identity padding
        K32.diagonal()[n:] = 1.0
    self.K32 = K32                               # kept for FP64 residuals
    # The preconditioner factors K32 + shift I.  When FP32 breaks down (a
    # non-positive pivot: the kernel is too ill-conditioned for single
    # precision at this lambda), the shift grows by 10x; the FP64-residual
    # iteration still converges to the solution of the unshifted system.
    mean_diag = float(torch.diagonal(K32)[:n].double().mean())
    shift = self.lam
    info = torch.zeros((1,), dtype=torch.int32, device=self.device)
    for attempt in range(8):
      A = K32.clone()
      A.diagonal().add_(shift)
      W = torch.zeros_like(A)
      WT = torch.empty_like(A)
      info.zero_()
      _lib.call('nta_chol_f32', _lib.ptr(A), npad, npad, _lib.ptr(W), npad,
                _lib.ptr(WT), _lib.ptr(info), stream)
      if int(info.item()) == 0:
        break
      shift = max(10.0 * shift, 1e-7 * mean_diag)
    else:
      raise ValueError(
          f'FP32 Cholesky failed at leading minor {int(info.item())} even '
          f'with shift {shift:.3g}: the train kernel is not positive '
          'definite; increase diag_reg or use solver="fp64".')
    del A
    self.shift = shift
    self.W = W
    self.WT = WT
    self._work = {}

  def _workspace(self, nrhs):
    ws = self._work.get(nrhs)
    if ws is None:
      nbytes = _lib.value('nta_pcg64_work_bytes', self.npad, nrhs)
      ws = torch.empty((nbytes,), dtype=torch.uint8, device=self.device)
      self._work[nrhs] = ws
    return ws

  # --------------------------------------------------------------- ops
  def _pad(self, x):
    if self.npad == self.n:
      return x.contiguous()
    out = torch.zeros((self.npad, x.shape[1]), dtype=x.dtype,
                      device=x.device)
    out[:self.n] = x
    return out

  def precond(self, R):
    """(L L^T)^-1 R in fp32 (R: [npad, k], k <= 16)."""
    R32 = R.to(torch.float32).contiguous()
    Z = torch.empty_like(R32)
    ws = self._workspace(R.shape[1])
    _lib.call('nta_chol_precond', _lib.ptr(self.W), _lib.ptr(self.WT),
              self.npad, self.npad, self.n1, _lib.ptr(R32), R.shape[1],
              _lib.ptr(Z), _lib.ptr(ws), ws.numel(),
              _lib.stream_handle(self.device))
    return Z

  # -------------------------------------------------------------- solve
  def solve(self, b: torch.Tensor) -> torch.Tensor:
    b = b.to(self.device, torch.float64)
    if b.ndim == 1:
      return self.solve(b[:, None])[:, 0]
    out_cols = b.shape[1]
    chunks = []
    its, res = 0, 0.0
    for c0 in range(0, out_cols, 16):
      x, it, r = self._solve16(b[:, c0:c0 + 16])
      chunks.append(x)
      its, res = max(its, it), max(res, r)
    self.iterations = its
    self.last_residual = res
    return torch.cat(chunks, dim=1) if len(chunks) > 1 else chunks[0]

  def _solve16(self, b):
    B = self._pad(b).contiguous()
    k = B.shape[1]
    X = torch.empty_like(B)
    ws = self._workspace(k)
    iters = ctypes.c_int(0)
    resid = ctypes.c_double(0.0)
    _lib.call('nta_pcg64', _lib.ptr(self.K32), self.npad, self.npad,
              self.lam, _lib.ptr(self.W), _lib.ptr(self.WT), self.npad,
              self.n1, _lib.ptr(B), k,
              _lib.ptr(X), self.tol, self.max_iter, _lib.ptr(ws), ws.numel(),
              ctypes.byref(iters), ctypes.byref(resid),
              _lib.stream_handle(self.device))
    return X[:self.n], iters.value, resid.value
