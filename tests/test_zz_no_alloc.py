"""nta_pcg64_noalloc (the refinement solve with the library's static
workspace) against nta_pcg64 with a caller workspace, through the C ABI
(needs a GPU).

Two calls in a row on the same device, the second smaller than the first
(the static array keeps the first call's contents), each compared with the
workspace entry point on the same factor and right-hand sides: identical
iteration counts and solutions equal to 1e-12 (the same kernels in the same
order; the FP64 bar of the solve tests is 1e-9 against torch).
"""
import ctypes

import pytest
import torch

from neural_tangents_amd import _lib

pytestmark = pytest.mark.gpu

DEV = torch.device('cuda', 0)


def _spd(n, seed, lam):
  g = torch.Generator().manual_seed(seed)
  X = torch.randn(n, n, generator=g, dtype=torch.float64)
  K = X @ X.T / n + torch.ones(n, n, dtype=torch.float64)
  return K.float(), lam


def _factor(K32, lam):
  n = K32.shape[0]
  A = (K32 + lam * torch.eye(n)).to(DEV)
  W = torch.full_like(A, float('nan'))    # written where it is read
  WT = torch.empty_like(A)
  info = torch.zeros((1,), dtype=torch.int32, device=DEV)
  _lib.call('nta_chol_f32', _lib.ptr(A), n, n, _lib.ptr(W), n,
            _lib.ptr(WT), _lib.ptr(info), _lib.stream_handle(DEV))
  torch.cuda.synchronize()
  assert int(info.item()) == 0
  return W, WT


def _solve_both(n, nrhs, seed):
  K, lam = _spd(n, seed, 1e-2)
  W, WT = _factor(K, lam)
  n1 = _lib.value('nta_chol_split', n)
  B = torch.randn(n, nrhs, dtype=torch.float64,
                  generator=torch.Generator().manual_seed(seed + 1)).to(DEV)
  Kd = K.to(DEV)
  st = _lib.stream_handle(DEV)
  X_ws = torch.empty((n, nrhs), dtype=torch.float64, device=DEV)
  ws = torch.empty((_lib.value('nta_pcg64_work_bytes', n, nrhs),),
                   dtype=torch.uint8, device=DEV)
  it_ws, r_ws = ctypes.c_int(0), ctypes.c_double(0.0)
  _lib.call('nta_pcg64', _lib.ptr(Kd), n, n, lam, _lib.ptr(W), _lib.ptr(WT),
            n, n1, _lib.ptr(B), nrhs, _lib.ptr(X_ws), 1e-12, 100,
            _lib.ptr(ws), ws.numel(), ctypes.byref(it_ws),
            ctypes.byref(r_ws), st)
  X_na = torch.full((n, nrhs), float('nan'), dtype=torch.float64,
                    device=DEV)
  it_na, r_na = ctypes.c_int(0), ctypes.c_double(0.0)
  _lib.call('nta_pcg64_noalloc', _lib.ptr(Kd), n, n, lam, _lib.ptr(W),
            _lib.ptr(WT), n, n1, _lib.ptr(B), nrhs, _lib.ptr(X_na), 1e-12,
            100, ctypes.byref(it_na), ctypes.byref(r_na), st)
  torch.cuda.synchronize()
  assert it_na.value == it_ws.value, (it_na.value, it_ws.value)
  err = float((X_na - X_ws).norm() / X_ws.norm())
  assert err <= 1e-12, err
  assert abs(r_na.value - r_ws.value) <= 1e-12 + 1e-6 * r_ws.value
  A = K.double() + lam * torch.eye(n, dtype=torch.float64)
  want = torch.linalg.solve(A, B.cpu())
  assert float((X_na.cpu() - want).norm() / want.norm()) <= 1e-9


def test_noalloc_matches_workspace_entry_point_twice():
  _solve_both(1024, 10, seed=3)
  _solve_both(512, 3, seed=4)        # smaller: stale contents must not leak


def test_noalloc_rejects_sizes_beyond_its_workspace():
  n = 16384 + 64
  with pytest.raises(ValueError, match='static workspace'):
    _lib.call('nta_pcg64_noalloc', None, n, n, 0.0, None, None, n,
              _lib.value('nta_chol_split', n), None, 1, None, 1e-10, 10,
              None, None, _lib.stream_handle(DEV))
