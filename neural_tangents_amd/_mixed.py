"""Mixed-precision Cholesky solve: FP32 factor on the GPU (libnta.so),
FP64 residuals, preconditioned conjugate gradients.

For ``A = K + lambda I`` (``predict.py:1224-1258``):

1. ``nta_copy_lower_shift`` forms ``fp32(K) + shift I`` (lower tiles) and
   ``nta_chol_f32`` factors it in FP32 (right-looking
   blocked Cholesky on fp32 MFMA, 128-column leaves inverted in LDS) and
   inverts each half of the factor; the preconditioner ``(L L^T)^-1`` is
   applied from the blocks it leaves (``nta_chol_precond``).
2. ``nta_pcg64`` runs conjugate gradients on ``A`` with FP64 vectors:
   ``A p`` reads the kernel as given -- fp32 (exact in FP64) or, for a
   float64 kernel, its FP64 entries (``nta_pcg64_k64``) -- accumulates in
   FP64 and adds the FP64 ``lambda``, so the converged solution is the FP64
   solution of the regularised system; the FP32 factor only steers
   convergence.
3. A solve whose final relative residual (recomputed from scratch) stays
   above ``max(100 tol, 1e-8)`` falls back to the FP64 Cholesky solve with a
   ``RuntimeWarning``; a result is never returned unverified.
"""
import ctypes
import warnings

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
    self.fell_back = False
    self._L64 = None
    npad = -(-n // self.NB) * self.NB
    self.npad = npad
    self.n1 = _lib.value('nta_chol_split', npad)
    self.k64 = K.dtype == torch.float64
    if (npad == n and K.dtype == torch.float32 and K.is_contiguous() and
        K.data_ptr() % 16 == 0):
      K32 = K
    else:
      K32 = torch.zeros((npad, npad), dtype=torch.float32, device=self.device)
      K32[:n, :n] = K.to(torch.float32)
      if npad > n:                                # identity padding
        K32.diagonal()[n:] = 1.0
    if self.k64:
      # The residual operator reads the FP64 entries: fp32(K) would change
      # the system by up to cond(A) 2^-24 in the solution.
      if npad == n and K.is_contiguous():
        Kop = K
      else:
        Kop = torch.zeros((npad, npad), dtype=torch.float64,
                          device=self.device)
        Kop[:n, :n] = K
        if npad > n:
          Kop.diagonal()[n:] = 1.0
    else:
      Kop = K32
    self.Kop = Kop                               # FP64 residual operator
    # The preconditioner factors K32 + shift I.  When FP32 breaks down (a
    # non-positive pivot: the kernel is too ill-conditioned for single
    # precision at this lambda), the shift grows by 10x; the FP64-residual
    # iteration still converges to the solution of the unshifted system.
    # The pivot check waits for the first solve: the factorisation's status
    # goes to pinned host memory by an asynchronous copy behind it, which
    # the refinement's synchronisation completes, so the factorisation and
    # the refinement run back to back and the check costs no round trip.
    self._K32 = K32
    self._info = torch.zeros((1,), dtype=torch.int32, device=self.device)
    self._info_host = torch.zeros((1,), dtype=torch.int32, pin_memory=True)
    self.shift = self.lam
    self._factor(self.shift)
    self._checked = False
    self._work = {}

  def _factor(self, shift):
    # fp32(K) + shift I, lower tiles only (all the factorisation reads)
    A = torch.empty_like(self._K32)
    _lib.call('nta_copy_lower_shift', _lib.ptr(self._K32), self.npad,
              self.npad, shift, _lib.ptr(A), self.npad,
              _lib.stream_handle(self.device))
    self.W = torch.empty_like(A)        # the factorisation writes what is read
    self.WT = torch.empty_like(A)
    self._info.zero_()
    _lib.call('nta_chol_f32', _lib.ptr(A), self.npad, self.npad,
              _lib.ptr(self.W), self.npad, _lib.ptr(self.WT),
              _lib.ptr(self._info), _lib.stream_handle(self.device))
    self._info_host.copy_(self._info, non_blocking=True)

  def _factor_ok(self, synced=False):
    """True once the factor in W has no non-positive pivot; refactors with
    a growing shift (synchronously) after a breakdown and returns False.
    synced: the stream was synchronised after the factorisation (its status
    is then in the pinned host copy)."""
    if self._checked:
      return True
    info = int(self._info_host[0]) if synced else int(self._info.item())
    if info == 0:
      self._checked = True
      return True
    mean_diag = float(torch.diagonal(self._K32)[:self.n].double().mean())
    shift = self.shift
    for attempt in range(8):
      shift = max(10.0 * shift, 1e-7 * mean_diag)
      self._factor(shift)
      if int(self._info.item()) == 0:
        self.shift = shift
        self._checked = True
        return False
    raise ValueError(
        f'FP32 Cholesky failed at leading minor {int(self._info.item())} '
        f'even with shift {shift:.3g}: the train kernel is not positive '
        'definite; increase diag_reg or use solver="fp64".')

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
    self._factor_ok()
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

  def accept_residual(self):
    """Largest final relative residual a refined solve may return."""
    return max(100.0 * self.tol, 1e-8)

  def _solve16(self, b):
    X, it, r = self._refine(b)       # synchronises the stream
    if not self._factor_ok(synced=True):   # replaced: solve again
      X, it, r = self._refine(b)
    if not r <= self.accept_residual():
      warnings.warn(
          f'mixed-precision solve stopped at relative residual {r:.3g} '
          f'after {it} iterations (limit {self.accept_residual():.3g}); '
          'using the FP64 Cholesky solve instead.', RuntimeWarning,
          stacklevel=3)
      self.fell_back = True
      return self._fp64_solve(b), it, r
    return X[:self.n], it, r

  def _refine(self, b):
    """nta_pcg64 on the padded right-hand side: (X, iterations, final
    relative residual recomputed from scratch)."""
    B = self._pad(b).contiguous()
    k = B.shape[1]
    X = torch.empty_like(B)
    ws = self._workspace(k)
    iters = ctypes.c_int(0)
    resid = ctypes.c_double(0.0)
    _lib.call('nta_pcg64_k64' if self.k64 else 'nta_pcg64',
              _lib.ptr(self.Kop), self.npad, self.npad,
              self.lam, _lib.ptr(self.W), _lib.ptr(self.WT), self.npad,
              self.n1, _lib.ptr(B), k,
              _lib.ptr(X), self.tol, self.max_iter, _lib.ptr(ws), ws.numel(),
              ctypes.byref(iters), ctypes.byref(resid),
              _lib.stream_handle(self.device))
    return X, iters.value, resid.value

  def _fp64_solve(self, b):
    """Plain FP64 Cholesky solve of (K + lambda I) x = b (the fallback)."""
    if self._L64 is None:
      A = self.Kop[:self.n, :self.n].to(torch.float64).clone()
      A.diagonal().add_(self.lam)
      L, info = torch.linalg.cholesky_ex(A)
      if int(info) != 0:
        raise ValueError(
            f'Cholesky factorisation failed (leading minor {int(info)} is '
            'not positive definite); increase diag_reg.')
      self._L64 = L
    return torch.cholesky_solve(b.to(torch.float64), self._L64)
