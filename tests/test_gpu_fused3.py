"""The fused launch of the first three Myrtle-10 stages
(nta_first3_pool_stage: Conv-Relu three times, then AvgPool, on 32x32
images) against the same network run stage by stage (nta_first_stage +
nta_stage), and the Myrtle-10 kernel against the FP64 oracle, on the same
inputs (needs a GPU).  FP32 both ways; the fused kernel accumulates the
patch Gram in another order, so the bar is 1e-5 relative (Frobenius) and
1e-4 of the largest entry elementwise, and 1e-5 against the oracle."""
import numpy as np
import pytest
import torch

from conftest import rel_err
from neural_tangents_amd import _engine, stax
from neural_tangents_amd.data import blurred_images
from oracle import stax_oracle as so

pytestmark = pytest.mark.gpu

DEV = torch.device('cuda', 0)


def _net(bs, readout):
  W = (1.3, 1.5, 1.2)
  layers = []
  for w, b in zip(W, bs):
    layers += [stax.Conv(1, (3, 3), padding='SAME', W_std=w, b_std=b),
               stax.Relu()]
  layers += [stax.AvgPool((2, 2), (2, 2))]
  if readout:
    layers += [stax.Conv(1, (3, 3), padding='SAME', W_std=1.4), stax.Relu(),
               stax.AvgPool((2, 2), (2, 2)), stax.GlobalAvgPool(),
               stax.Dense(1, 1.0, 0.0)]
  return stax.serial(*layers)[2]


def _both(monkeypatch, kfn, x1, x2, get, **kw):
  fused = kfn(x1, x2, get, **kw)
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


@pytest.mark.parametrize('C,bs', [(3, (0.0, 0.0, 0.0)), (1, (0.1, 0.2, 0.3)),
                                  (4, (0.3, 0.0, 0.1))])
def test_fused3_state_matches_stagewise(monkeypatch, C, bs):
  rng = np.random.default_rng(10 + C)
  x1 = torch.from_numpy(
      rng.standard_normal((3, 32, 32, C)).astype(np.float32)).to(DEV)
  x2 = torch.from_numpy(
      rng.standard_normal((2, 32, 32, C)).astype(np.float32)).to(DEV)
  kfn = _net(bs, readout=False)
  assert _engine._fused_stages(kfn.plan(x1.shape), tuple(x1.shape)) == 3
  f, p = _both(monkeypatch, kfn, x1, x2, ('nngp', 'ntk'))
  assert f.nngp.shape == (3, 2, 16, 16, 16, 16)
  _close(f.nngp, p.nngp, 'nngp')
  _close(f.ntk, p.ntk, 'ntk')
  fn, pn = _both(monkeypatch, kfn, x1, x2, 'nngp')
  _close(fn, pn, 'nngp only')


def test_fused3_kernel_matrix_matches_stagewise(monkeypatch):
  """Readout network, same-input (symmetric tiles) and cross, and small
  tiles (several fused launches per matrix)."""
  x1, x2 = blurred_images([5, 3], 32, seed=13)
  kfn = _net((0.0, 0.0, 0.0), readout=True)
  for a, b in ((x1, None), (x1, x2)):
    f, p = _both(monkeypatch, kfn, a, b, ('nngp', 'ntk'))
    _close(f.nngp, p.nngp, 'nngp')
    _close(f.ntk, p.ntk, 'ntk')
  f, p = _both(monkeypatch, kfn, x1, x2, ('nngp', 'ntk'),
               tile_bytes=4 * 2 * 2 * 4 * 16 ** 4)
  _close(f.nngp, p.nngp, 'nngp tiled')
  _close(f.ntk, p.ntk, 'ntk tiled')


def test_myrtle10_32x32_fused_matches_oracle():
  """Myrtle-10 ([3, 3, 3]) on 32x32 images of the workload's generator
  against the FP64 oracle."""
  x1, x2 = blurred_images([2, 3], 32, seed=0)
  kfn = stax.myrtle(10)[2]
  k = kfn(x1, x2, ('nngp', 'ntk'))
  ref = so.kernel_fn(so.myrtle_layers((3, 3, 3)), x1, x2, ('nngp', 'ntk'))
  for g in ('nngp', 'ntk'):
    e = rel_err(getattr(k, g).double().cpu().numpy(), ref[g])
    assert e <= 1e-5, (g, e)
