"""The fused launch of the first two Myrtle stages (nta_first_pool_stage)
against the same network run stage by stage (nta_first_stage + nta_stage),
on the same inputs (needs a GPU).  Both are FP32; the fused kernel
accumulates the patch Gram in another order, so the bar is 1e-5 relative
(Frobenius) and 1e-4 of the largest entry elementwise."""
import numpy as np
import pytest
import torch

from neural_tangents_amd import _engine, stax
from neural_tangents_amd.data import blurred_images

pytestmark = pytest.mark.gpu

DEV = torch.device('cuda', 0)


def _net(b1, b2, readout):
  layers = [stax.Conv(1, (3, 3), padding='SAME', W_std=1.3, b_std=b1),
            stax.Relu(),
            stax.Conv(1, (3, 3), padding='SAME', W_std=1.5, b_std=b2),
            stax.Relu(), stax.AvgPool((2, 2), (2, 2))]
  if readout:
    layers += [stax.Conv(1, (3, 3), padding='SAME', W_std=1.4), stax.Relu(),
               stax.AvgPool((2, 2), (2, 2)), stax.GlobalAvgPool(),
               stax.Dense(1, 1.0, 0.0)]
  return stax.serial(*layers)[2]


def _both(monkeypatch, kfn, x1, x2, get):
  fused = kfn(x1, x2, get)
  with monkeypatch.context() as m:
    m.setattr(_engine, '_fused_stages', lambda plan, shape: 0)
    plain = kfn(x1, x2, get)
  return fused, plain


def _close(a, b, name):
  a = a.double().cpu().numpy()
  b = b.double().cpu().numpy()
  e = np.linalg.norm(a - b) / np.linalg.norm(b)
  assert e <= 1e-5, (name, e)
  assert np.max(np.abs(a - b)) <= 1e-4 * np.max(np.abs(b)), name


@pytest.mark.parametrize('C,b1,b2', [(3, 0.0, 0.0), (1, 0.1, 0.2),
                                     (4, 0.3, 0.0)])
def test_fused_state_matches_stagewise(monkeypatch, C, b1, b2):
  rng = np.random.default_rng(C)
  x1 = torch.from_numpy(
      rng.standard_normal((3, 32, 32, C)).astype(np.float32)).to(DEV)
  x2 = torch.from_numpy(
      rng.standard_normal((2, 32, 32, C)).astype(np.float32)).to(DEV)
  kfn = _net(b1, b2, readout=False)
  f, p = _both(monkeypatch, kfn, x1, x2, ('nngp', 'ntk'))
  assert f.nngp.shape == (3, 2, 16, 16, 16, 16)
  _close(f.nngp, p.nngp, 'nngp')
  _close(f.ntk, p.ntk, 'ntk')
  fn, pn = _both(monkeypatch, kfn, x1, x2, 'nngp')
  _close(fn, pn, 'nngp only')


def test_fused_kernel_matrix_matches_stagewise(monkeypatch):
  """Myrtle-5-type readout, same-input (symmetric tiles) and cross."""
  x1, x2 = blurred_images([5, 3], 32, seed=3)
  kfn = _net(0.0, 0.0, readout=True)
  for a, b in ((x1, None), (x1, x2)):
    f, p = _both(monkeypatch, kfn, a, b, ('nngp', 'ntk'))
    _close(f.nngp, p.nngp, 'nngp')
    _close(f.ntk, p.ntk, 'ntk')
  # small tiles: several fused launches per kernel matrix
  f = kfn(x1, x2, ('nngp', 'ntk'), tile_bytes=4 * 2 * 2 * 4 * 16 ** 4)
  _, p = _both(monkeypatch, kfn, x1, x2, ('nngp', 'ntk'))
  _close(f.nngp, p.nngp, 'nngp tiled')
  _close(f.ntk, p.ntk, 'ntk tiled')


def test_split_fp16_gram_keeps_fp32_accuracy(monkeypatch):
  """The fused kernel's patch Gram runs on FP16 MFMAs over a hi/lo split of
  power-of-two scaled images: inputs spanning six decades, a large overall
  scale and an all-zero image must match the FP32 stage-by-stage path."""
  rng = np.random.default_rng(7)
  x1 = rng.standard_normal((3, 32, 32, 3)) * 10.0 ** rng.uniform(
      -3, 3, (3, 32, 32, 3))
  x2 = rng.standard_normal((3, 32, 32, 3)) * 3e4
  x2[1] = 0.0
  x1 = torch.from_numpy(x1.astype(np.float32)).to(DEV)
  x2 = torch.from_numpy(x2.astype(np.float32)).to(DEV)
  kfn = _net(0.2, 0.1, readout=False)
  f, p = _both(monkeypatch, kfn, x1, x2, ('nngp', 'ntk'))
  assert torch.isfinite(f.nngp).all() and torch.isfinite(f.ntk).all()
  _close(f.nngp, p.nngp, 'nngp')
  _close(f.ntk, p.ntk, 'ntk')
