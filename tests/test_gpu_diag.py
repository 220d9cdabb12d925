"""Diagonal-spatial path (Flatten readout, no pooling) against the FP64
oracle's diagonal_spatial restatement and against the engine's own
full-spatial path (needs a GPU)."""
import numpy as np
import pytest
import torch

from conftest import rel_err
from neural_tangents_amd import _engine, stax
from neural_tangents_amd.data import blurred_images
from oracle import stax_oracle as so

pytestmark = pytest.mark.gpu


def _nets(depth, W_std=2 ** 0.5, b_std=0.0, padding='SAME', filt=(3, 3),
          readout=(1.5, 0.1)):
  ours, ref = [], []
  for _ in range(depth):
    ours += [stax.Conv(1, filt, padding=padding, W_std=W_std, b_std=b_std),
             stax.Relu()]
    ref += [so.conv_layer(filt, (1, 1), padding, W_std, b_std),
            so.relu_layer()]
  ours += [stax.Flatten(), stax.Dense(1, *readout)]
  ref += [so.flatten_layer(), so.dense_layer(*readout)]
  return stax.serial(*ours)[2], ref


@pytest.mark.parametrize('side,depth,padding,b_std', [
    (8, 3, 'SAME', 0.0), (16, 4, 'SAME', 0.3), (10, 2, 'VALID', 0.2)])
def test_cross_and_self_match_oracle(side, depth, padding, b_std):
  x1, x2 = blurred_images([5, 4], side, seed=side)
  kfn, layers = _nets(depth, b_std=b_std, padding=padding)
  assert kfn.plan(x1.shape).diag
  got = kfn(x1, x2, ('nngp', 'ntk'))
  want = so.kernel_fn(layers, x1.astype(np.float64), x2.astype(np.float64),
                      ('nngp', 'ntk'), diagonal_spatial=True)
  for g in ('nngp', 'ntk'):
    assert rel_err(getattr(got, g).double().cpu().numpy(), want[g]) <= 1e-5
  full = kfn(x1, None, None)
  want_s = so.kernel_fn(layers, x1.astype(np.float64), None, None,
                        diagonal_spatial=True)
  assert rel_err(full.nngp.double().cpu().numpy(), want_s.nngp) <= 1e-5
  assert rel_err(full.ntk.double().cpu().numpy(), want_s.ntk) <= 1e-5
  assert rel_err(full.cov1.double().cpu().numpy(), want_s.cov1) <= 1e-5
  # x2 = None computes the upper triangle and mirrors it
  assert torch.equal(full.nngp, full.nngp.T)


def test_non_square_images():
  x = blurred_images([6], 8, seed=3)[0][:, :, :6, :].copy()
  kfn, layers = _nets(3, b_std=0.1)
  got = kfn(x, None, 'ntk')
  want = so.kernel_fn(layers, x.astype(np.float64), None, 'ntk',
                      diagonal_spatial=True)['ntk']
  assert rel_err(got.double().cpu().numpy(), want) <= 1e-5


def test_matches_full_spatial_engine_path():
  """The same network through the full-spatial kernels (trace readout)."""
  x1, x2 = blurred_images([3, 4], 8, seed=11)
  kfn, _ = _nets(3, b_std=0.2)
  got = kfn(x1, x2, ('nngp', 'ntk'))
  plan = kfn.plan(x1.shape)
  plan_full = _engine.Plan(**{**plan.__dict__, 'diag': False})
  out = _engine.run(plan_full, torch.from_numpy(x1).cuda(),
                    torch.from_numpy(x2).cuda(), True, True, False)
  for g in ('nngp', 'ntk'):
    assert rel_err(getattr(got, g).double().cpu().numpy(),
                   out[g].double().cpu().numpy()) <= 1e-5


def test_readme_21_layer_cnn_without_pooling():
  """README.md:399-416: 21 x [Conv 3x3 SAME, Relu] on 32x32x3 + Flatten."""
  x1, x2 = blurred_images([3, 2], 32, seed=5)
  ours, ref = [], []
  for _ in range(21):
    ours += [stax.Conv(1, (3, 3), (1, 1), 'SAME'), stax.Relu()]
    ref += [so.conv_layer((3, 3), (1, 1), 'SAME'), so.relu_layer()]
  ours += [stax.Flatten()]
  ref += [so.flatten_layer()]
  got = stax.serial(*ours)[2](x1, x2, ('nngp', 'ntk'))
  want = so.kernel_fn(ref, x1.astype(np.float64), x2.astype(np.float64),
                      ('nngp', 'ntk'), diagonal_spatial=True)
  for g in ('nngp', 'ntk'):
    assert rel_err(getattr(got, g).double().cpu().numpy(), want[g]) <= 1e-5


def test_readme_21_layer_cnn_with_global_pooling():
  """README.md:380-396: the same 21 layers with a GlobalAvgPool readout run
  as 21 full-spatial Conv-Relu stages; one cross entry and the self chains
  against the FP64 oracle (about 7 s of numpy)."""
  x1, x2 = blurred_images([1, 1], 32, seed=6)
  ours, ref = [], []
  for _ in range(21):
    ours += [stax.Conv(1, (3, 3), (1, 1), 'SAME'), stax.Relu()]
    ref += [so.conv_layer((3, 3), (1, 1), 'SAME'), so.relu_layer()]
  ours += [stax.GlobalAvgPool()]
  ref += [so.global_avg_pool_layer()]
  got = stax.serial(*ours)[2](x1, x2, ('nngp', 'ntk'))
  want = so.kernel_fn(ref, x1.astype(np.float64), x2.astype(np.float64),
                      ('nngp', 'ntk'))
  for g in ('nngp', 'ntk'):
    assert rel_err(getattr(got, g).double().cpu().numpy(), want[g]) <= 1e-5
