"""Parity of the HIP ``kernel_fn`` with the FP64 oracle (needs a GPU).

Bar: relative Frobenius error <= 1e-5 on every returned array (FP32 kernels
vs the FP64 restatement), plus an elementwise check at 1e-4.
"""
import numpy as np
import pytest
import torch

from conftest import rel_err
from neural_tangents_amd import stax
from neural_tangents_amd.data import blurred_images
from oracle import stax_oracle as so

pytestmark = pytest.mark.gpu

TOL = 1e-5


def _np(t):
  return None if t is None else t.detach().cpu().numpy()


def _check(got, want, tol=TOL, name=''):
  got = np.asarray(got, np.float64)
  want = np.asarray(want, np.float64)
  assert got.shape == want.shape, (name, got.shape, want.shape)
  e = rel_err(got, want)
  assert e <= tol, f'{name}: relative error {e:.3e} > {tol}'
  scale = np.max(np.abs(want)) + 1e-30
  assert np.max(np.abs(got - want)) <= 10 * tol * scale, name


def _myrtle_pair(depths, W_std=2 ** 0.5, b_std=0.0):
  layers = []
  olayers = []
  for d in depths:
    for _ in range(d):
      layers += [stax.Conv(1, (3, 3), padding='SAME', W_std=W_std,
                           b_std=b_std), stax.Relu()]
      olayers += [so.conv_layer((3, 3), (1, 1), 'SAME', W_std, b_std),
                  so.relu_layer()]
    layers += [stax.AvgPool((2, 2), (2, 2))]
    olayers += [so.avg_pool_layer((2, 2), (2, 2), 'VALID')]
  layers += [stax.GlobalAvgPool(), stax.Dense(1, 1.0, 0.0)]
  olayers += [so.global_avg_pool_layer(), so.dense_layer(1.0, 0.0)]
  return stax.serial(*layers)[2], olayers


@pytest.mark.parametrize('side,n1,n2', [(8, 3, 2), (16, 4, 3)])
def test_myrtle5_cross(side, n1, n2):
  x1, x2 = blurred_images([n1, n2], side, seed=side)
  kfn, ol = _myrtle_pair((2, 1, 1))
  k = kfn(x1, x2, ('nngp', 'ntk'))
  ref = so.kernel_fn(ol, x1, x2, ('nngp', 'ntk'))
  _check(_np(k.nngp), ref['nngp'], name='nngp')
  _check(_np(k.ntk), ref['ntk'], name='ntk')


def test_myrtle5_same_inputs_kernel_record():
  x1, = blurred_images([5], 8, seed=3)
  kfn, ol = _myrtle_pair((2, 1, 1))
  k = kfn(x1, None)
  ref = so.kernel_fn(ol, x1, None, None)
  _check(_np(k.nngp), ref.nngp, name='nngp')
  _check(_np(k.ntk), ref.ntk, name='ntk')
  _check(_np(k.cov1), ref.cov1, name='cov1')
  assert k.cov2 is None and ref.cov2 is None
  assert k.is_reversed == ref.is_reversed
  assert k.x1_is_x2
  np.testing.assert_allclose(_np(k.nngp), _np(k.nngp).T, rtol=0, atol=0)


def test_myrtle10_cross():
  x1, x2 = blurred_images([2, 3], 16, seed=10)
  kfn, ol = _myrtle_pair((3, 3, 3))
  k = kfn(x1, x2, ('nngp', 'ntk'))
  ref = so.kernel_fn(ol, x1, x2, ('nngp', 'ntk'))
  _check(_np(k.nngp), ref['nngp'], name='nngp')
  _check(_np(k.ntk), ref['ntk'], name='ntk')


def test_cfg1_block():
  """Myrtle-5 16x16, a 6x5 block of the 64x64 workload (pairs are
  independent, so a block of the full matrix is exact)."""
  x1, x2 = blurred_images([64, 64], 16, seed=0)
  kfn, ol = _myrtle_pair((2, 1, 1))
  a, b = x1[10:16], x2[40:45]
  k = kfn(x1, x2, ('nngp', 'ntk'))
  ref = so.kernel_fn(ol, a, b, ('nngp', 'ntk'))
  _check(_np(k.nngp)[10:16, 40:45], ref['nngp'], name='nngp')
  _check(_np(k.ntk)[10:16, 40:45], ref['ntk'], name='ntk')


@pytest.mark.parametrize('b_std', [None, 0.3])
def test_full_spatial_output_layout(b_std):
  """No readout: the full-spatial state, in the reference axis order
  (reversed after an odd number of Convs), for nngp/ntk/cov1/cov2."""
  rng = np.random.default_rng(1)
  x1 = rng.standard_normal((2, 5, 4, 3)).astype(np.float32)
  x2 = rng.standard_normal((3, 5, 4, 3)).astype(np.float32)
  kfn = stax.serial(stax.Conv(1, (3, 3), padding='SAME', W_std=1.3,
                              b_std=b_std), stax.Relu())[2]
  ol = [so.conv_layer((3, 3), (1, 1), 'SAME', 1.3, b_std), so.relu_layer()]
  k = kfn(x1, x2)
  ref = so.kernel_fn(ol, x1, x2, None)
  assert k.is_reversed and ref.is_reversed
  for f in ('nngp', 'ntk', 'cov1', 'cov2'):
    _check(_np(getattr(k, f)), getattr(ref, f), name=f)


def test_valid_conv_flatten_dense_chain():
  rng = np.random.default_rng(2)
  x1 = rng.standard_normal((3, 6, 6, 2)).astype(np.float32)
  x2 = rng.standard_normal((2, 6, 6, 2)).astype(np.float32)
  kfn = stax.serial(
      stax.Conv(4, (3, 3), padding='VALID', W_std=1.5, b_std=0.2),
      stax.Relu(),
      stax.Conv(4, (2, 2), padding='SAME', W_std=0.9, b_std=0.1),
      stax.Relu(),
      stax.Flatten(),
      stax.Dense(3, 1.2, 0.4), stax.Dense(2, 0.7, 0.3))[2]
  ol = [so.conv_layer((3, 3), (1, 1), 'VALID', 1.5, 0.2), so.relu_layer(),
        so.conv_layer((2, 2), (1, 1), 'SAME', 0.9, 0.1), so.relu_layer(),
        so.flatten_layer(), so.dense_layer(1.2, 0.4),
        so.dense_layer(0.7, 0.3)]
  k = kfn(x1, x2, ('nngp', 'ntk', 'cov1', 'cov2'))
  ref = so.kernel_fn(ol, x1, x2, ('nngp', 'ntk', 'cov1', 'cov2'))
  for f in ('nngp', 'ntk', 'cov1', 'cov2'):
    _check(_np(getattr(k, f)), ref[f], name=f)


def test_pool_relu_order_and_conv_conv():
  """Conv -> AvgPool -> Relu (normaliser from a pooled state) and two Convs
  in a row (stage without ReLU)."""
  x1, x2 = blurred_images([2, 2], 8, seed=4)
  kfn = stax.serial(stax.Conv(1, (3, 3), padding='SAME', W_std=1.1),
                    stax.AvgPool((2, 2), (2, 2)), stax.Relu(),
                    stax.Conv(1, (3, 3), padding='SAME', W_std=1.4,
                              b_std=0.5),
                    stax.Conv(1, (3, 3), padding='SAME', W_std=0.8),
                    stax.Relu(), stax.GlobalAvgPool())[2]
  ol = [so.conv_layer((3, 3), (1, 1), 'SAME', 1.1, None),
        so.avg_pool_layer((2, 2), (2, 2), 'VALID'), so.relu_layer(),
        so.conv_layer((3, 3), (1, 1), 'SAME', 1.4, 0.5),
        so.conv_layer((3, 3), (1, 1), 'SAME', 0.8, None), so.relu_layer(),
        so.global_avg_pool_layer()]
  k = kfn(x1, x2, ('nngp', 'ntk'))
  ref = so.kernel_fn(ol, x1, x2, ('nngp', 'ntk'))
  _check(_np(k.nngp), ref['nngp'], name='nngp')
  _check(_np(k.ntk), ref['ntk'], name='ntk')


def test_tiling_and_symmetry_match_untiled():
  """Small tiles (many launches, mirrored lower triangle) give the same
  matrix as one tile."""
  x1, = blurred_images([7], 8, seed=5)
  kfn, _ = _myrtle_pair((2, 1, 1))
  full = kfn(x1, None, ('nngp', 'ntk'))
  small = kfn(x1, None, ('nngp', 'ntk'), tile_bytes=3 * 2 * 2 * 4 * 8 ** 4)
  np.testing.assert_array_equal(_np(full.nngp), _np(small.nngp))
  np.testing.assert_array_equal(_np(full.ntk), _np(small.ntk))
  x2, = blurred_images([3], 8, seed=6)
  cross = kfn(x1, x2, 'nngp', tile_bytes=2 * 2 * 2 * 4 * 8 ** 4)
  cross1 = kfn(x1, x2, 'nngp')
  np.testing.assert_array_equal(_np(cross), _np(cross1))


def test_errors_match_reference_types():
  x, = blurred_images([2], 8)
  with pytest.raises(ValueError):
    stax.serial(stax.Relu(), stax.GlobalAvgPool())[2](x, None, 'nngp')
  with pytest.raises(NotImplementedError):
    stax.serial(stax.Conv(1, (3, 3), (2, 2), 'SAME'), stax.Relu(),
                stax.GlobalAvgPool())[2](x, None, 'nngp')
  with pytest.raises(ValueError):
    stax.serial(stax.Conv(1, (3, 3), padding='SAME'),
                stax.GlobalAvgPool())[2](x, x[:, :4], 'nngp')


def test_cfg2_block_32x32():
  """Myrtle-5 32x32 (the benchmark workload's shape): a 2x3 block against
  the FP64 oracle, plus the same-input kernel of 2 images."""
  x1, x2 = blurred_images([256, 256], 32, seed=0)
  kfn, ol = _myrtle_pair((2, 1, 1))
  a, b = x1[100:102], x2[7:10]
  k = kfn(a, b, ('nngp', 'ntk'))
  ref = so.kernel_fn(ol, a, b, ('nngp', 'ntk'))
  _check(_np(k.nngp), ref['nngp'], name='nngp')
  _check(_np(k.ntk), ref['ntk'], name='ntk')
  ks = kfn(a[:2], None, ('nngp', 'ntk'))
  refs = so.kernel_fn(ol, a[:2], None, ('nngp', 'ntk'))
  _check(_np(ks.nngp), refs['nngp'], name='nngp self')
  _check(_np(ks.ntk), refs['ntk'], name='ntk self')


def test_myrtle10_32x32_block():
  x1, x2 = blurred_images([2, 2], 32, seed=11)
  kfn, ol = _myrtle_pair((3, 3, 3))
  k = kfn(x1[:1], x2, ('nngp', 'ntk'))
  ref = so.kernel_fn(ol, x1[:1], x2, ('nngp', 'ntk'))
  _check(_np(k.nngp), ref['nngp'], name='nngp')
  _check(_np(k.ntk), ref['ntk'], name='ntk')


@pytest.mark.parametrize('side', [8, 16, 32])
def test_single_stage_full_spatial(side):
  """One Conv+Relu+Conv+Relu(+pool) network without readout: the
  full-spatial state of every specialised stage kernel against the oracle,
  with a bias and an odd number of pairs (tail workgroups)."""
  x1, x2 = blurred_images([3, 1], side, seed=side + 1)
  for pool in (False, True):
    layers = [stax.Conv(1, (3, 3), padding='SAME', W_std=1.2, b_std=0.1),
              stax.Relu(),
              stax.Conv(1, (3, 3), padding='SAME', W_std=1.4, b_std=0.2),
              stax.Relu()]
    ol = [so.conv_layer((3, 3), (1, 1), 'SAME', 1.2, 0.1), so.relu_layer(),
          so.conv_layer((3, 3), (1, 1), 'SAME', 1.4, 0.2), so.relu_layer()]
    if pool:
      layers.append(stax.AvgPool((2, 2), (2, 2)))
      ol.append(so.avg_pool_layer((2, 2), (2, 2), 'VALID'))
    k = stax.serial(*layers)[2](x1, x2, ('nngp', 'ntk', 'cov1', 'cov2'))
    ref = so.kernel_fn(ol, x1, x2, ('nngp', 'ntk', 'cov1', 'cov2'))
    for f in ('nngp', 'ntk', 'cov1', 'cov2'):
      _check(_np(getattr(k, f)), ref[f], name=f'{f} pool={pool}')
