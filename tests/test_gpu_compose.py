"""``kernel_fn`` applied to a ``Kernel``, ``Kernel.reverse``/``transpose``,
fully-connected networks on flat inputs, ``batch`` and the notebook readout
on the HIP engine (needs a GPU).

The composition identities are the reference's own tests
(``tests/stax/stax_test.py:636-687``): ``serial(A, B)(x) == B(A(x))`` for a
Dense block on flat inputs and for a Conv block followed by a GlobalAvgPool or
Flatten readout.  Each composed result is also compared with the FP64 oracle
at the 1e-5 bar of the other kernel tests.
"""
import numpy as np
import pytest
import torch

from conftest import rel_err
from neural_tangents_amd import batching, stax
from neural_tangents_amd.data import blurred_images
from oracle import stax_oracle as so

pytestmark = pytest.mark.gpu

TOL = 1e-5


def _np(t):
  return t.detach().double().cpu().numpy()


def _close(got, want, tol=TOL):
  got, want = np.asarray(got, np.float64), np.asarray(want, np.float64)
  assert got.shape == want.shape, (got.shape, want.shape)
  e = rel_err(got, want)
  assert e <= tol, e


@pytest.mark.parametrize('same_inputs', [True, False])
def test_composition_dense(same_inputs):
  rng = np.random.default_rng(0)
  x1 = rng.standard_normal((2, 3)).astype(np.float32)
  x2 = None if same_inputs else rng.standard_normal((4, 3)).astype(
      np.float32)
  block = stax.serial(stax.Dense(256), stax.Relu())
  ker_fn = block[2]
  composed = stax.serial(block, block)[2]
  got = ker_fn(ker_fn(x1, x2))
  want = composed(x1, x2)
  for f in ('nngp', 'ntk', 'cov1'):
    _close(_np(getattr(got, f)), _np(getattr(want, f)), 1e-6)
  ol = [so.dense_layer(1.0), so.relu_layer()] * 2
  ref = so.kernel_fn(ol, x1, x2, None)
  for f in ('nngp', 'ntk', 'cov1'):
    _close(_np(getattr(want, f)), getattr(ref, f))
  assert want.nngp.shape == (2, 2 if same_inputs else 4)


@pytest.mark.parametrize('avg_pool', [True, False])
@pytest.mark.parametrize('same_inputs', [True, False])
def test_composition_conv(avg_pool, same_inputs):
  x1, x2 = blurred_images([3, 4], 5, seed=3)
  x2 = None if same_inputs else x2
  block = stax.serial(stax.Conv(256, (3, 3)), stax.Relu())
  if avg_pool:
    readout = stax.serial(stax.Conv(256, (3, 3)), stax.GlobalAvgPool(),
                          stax.Dense(10))
    oread = [so.conv_layer((3, 3), (1, 1), 'VALID', 1.0, None),
             so.global_avg_pool_layer(), so.dense_layer(1.0)]
  else:
    readout = stax.serial(stax.Flatten(), stax.Dense(10))
    oread = [so.flatten_layer(), so.dense_layer(1.0)]
  composed = stax.serial(block, readout)[2]
  want = composed(x1, x2)
  inner = block[2](x1, x2, diagonal_spatial=False)
  assert inner.is_reversed and inner.nngp.ndim == 6
  got = readout[2](inner)
  for f in ('nngp', 'ntk', 'cov1'):
    _close(_np(getattr(got, f)), _np(getattr(want, f)), 1e-6)
  oblock = [so.conv_layer((3, 3), (1, 1), 'VALID', 1.0, None),
            so.relu_layer()]
  ref = so.kernel_fn(oblock + oread, x1, x2, None)
  for f in ('nngp', 'ntk', 'cov1'):
    _close(_np(getattr(want, f)), getattr(ref, f))
  if not same_inputs:
    _close(_np(got.cov2), ref.cov2)


def test_kernel_input_keeps_reference_layout_and_reverse():
  """A block's full-spatial output, its ``reverse()`` and ``transpose``
  carry the same kernel: the readout of each agrees, and the state matches
  the oracle in the reference's (reversed) axis order."""
  x1, x2 = blurred_images([2, 3], 6, seed=4)
  block = stax.serial(stax.Conv(1, (3, 3), padding='SAME', W_std=1.3,
                                b_std=0.2), stax.Relu(),
                      stax.Conv(1, (3, 1), padding='SAME'))
  k = block[2](x1, x2)
  ref = so.kernel_fn([so.conv_layer((3, 3), (1, 1), 'SAME', 1.3, 0.2),
                      so.relu_layer(),
                      so.conv_layer((3, 1), (1, 1), 'SAME', 1.0, None)],
                     x1, x2, None)
  assert k.is_reversed == ref.is_reversed
  _close(_np(k.nngp), ref.nngp)
  _close(_np(k.ntk), ref.ntk)
  r = k.reverse()
  assert r.is_reversed != k.is_reversed
  assert torch.equal(r.reverse().nngp, k.nngp)
  t = k.transpose((1, 0))
  assert torch.equal(t.nngp, r.nngp)
  readout = stax.serial(stax.Relu(), stax.AvgPool((2, 2), (2, 2)),
                        stax.GlobalAvgPool(), stax.Dense(1, 2.0, 0.1))[2]
  a = readout(k, get=('nngp', 'ntk'))
  b = readout(r, get=('nngp', 'ntk'))
  _close(_np(b.nngp), _np(a.nngp), 1e-6)
  _close(_np(b.ntk), _np(a.ntk), 1e-6)
  oref = so.serial(so.relu_layer(), so.avg_pool_layer((2, 2), (2, 2),
                                                      'VALID'),
                   so.global_avg_pool_layer(),
                   so.dense_layer(2.0, 0.1))(ref)
  _close(_np(a.nngp), oref.nngp)
  _close(_np(a.ntk), oref.ntk)


@pytest.mark.parametrize('same_inputs', [True, False])
def test_mlp_on_flat_inputs(same_inputs):
  """The README's fully-connected network on [N, D] inputs."""
  rng = np.random.default_rng(1)
  x1 = rng.standard_normal((5, 20)).astype(np.float32)
  x2 = None if same_inputs else rng.standard_normal((7, 20)).astype(
      np.float32)
  kfn = stax.serial(stax.Dense(512, 1.5, 0.05), stax.Relu(),
                    stax.Dense(512, 1.5, 0.05), stax.Relu(),
                    stax.Dense(1, 1.5, 0.05))[2]
  k = kfn(x1, x2, ('nngp', 'ntk'))
  ol = [so.dense_layer(1.5, 0.05), so.relu_layer()] * 2 + [
      so.dense_layer(1.5, 0.05)]
  ref = so.kernel_fn(ol, x1, x2, ('nngp', 'ntk'))
  _close(_np(k.nngp), ref['nngp'])
  _close(_np(k.ntk), ref['ntk'])


def test_conv_on_flat_inputs_raises():
  kfn = stax.serial(stax.Conv(1, (3, 3)), stax.Relu())[2]
  with pytest.raises(ValueError):
    kfn(np.ones((2, 5), np.float32))


@pytest.mark.parametrize('store_on_device', [True, False])
@pytest.mark.parametrize('same_inputs', [True, False])
def test_batch_on_hip_kernel_fn(store_on_device, same_inputs):
  """``nt.batch`` (``batching.py:76-133, 314-503``) around the HIP
  ``kernel_fn``, against the FP64 oracle of the whole matrix."""
  x1, x2 = blurred_images([6, 4], 8, seed=9)
  x2 = None if same_inputs else x2
  kfn = stax.myrtle(5)[2]
  bfn = batching.batch(kfn, batch_size=2, store_on_device=store_on_device)
  got = bfn(x1, x2, ('nngp', 'ntk'))
  assert got.nngp.is_cuda == store_on_device
  ref = so.kernel_fn(so.myrtle_layers((2, 1, 1)), x1, x2, ('nngp', 'ntk'))
  _close(_np(got.nngp), ref['nngp'])
  _close(_np(got.ntk), ref['ntk'])
  full = bfn(x1, x2, None)
  _close(_np(full.cov1), so.kernel_fn(so.myrtle_layers((2, 1, 1)), x1, x2,
                                      None).cov1)


def test_notebook_readout_matches_oracle():
  """The notebook's Myrtle readout: AvgPool x3 -> Flatten -> Dense(sqrt 2)
  (``notebooks/myrtle_kernel_with_neural_tangents.ipynb``, cells 114-128;
  ``linear.py:1868-1899``) at 32x32, against the FP64 oracle."""
  x1, x2 = blurred_images([2, 3], 32, seed=12)
  W = 2 ** 0.5
  layers, ol = [], []
  for d in (2, 1, 1):
    for _ in range(d):
      layers += [stax.Conv(1, (3, 3), padding='SAME', W_std=W), stax.Relu()]
      ol += [so.conv_layer((3, 3), (1, 1), 'SAME', W, None),
             so.relu_layer()]
    layers += [stax.AvgPool((2, 2), (2, 2))]
    ol += [so.avg_pool_layer((2, 2), (2, 2), 'VALID')]
  layers += [stax.AvgPool((2, 2), (2, 2))] * 2 + [stax.Flatten(),
                                                 stax.Dense(10, W)]
  ol += [so.avg_pool_layer((2, 2), (2, 2), 'VALID')] * 2 + [
      so.flatten_layer(), so.dense_layer(W)]
  kfn = stax.serial(*layers)[2]
  k = kfn(x1, x2, ('nngp', 'ntk'))
  ref = so.kernel_fn(ol, x1, x2, ('nngp', 'ntk'))
  _close(_np(k.nngp), ref['nngp'])
  _close(_np(k.ntk), ref['ntk'])
