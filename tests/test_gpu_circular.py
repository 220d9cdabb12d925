"""CIRCULAR (periodic) Conv padding on the HIP engine against the FP64 oracle
(linear.py:1191-1199, 2977-3021, 3158-3164, 3422-3426): full-spatial stages
with pooling and readout, a head-less state on non-square images, and the
diagonal-spatial chain (needs a GPU)."""
import numpy as np
import pytest

from conftest import rel_err
from neural_tangents_amd import stax
from neural_tangents_amd.data import blurred_images
from oracle import stax_oracle as so

pytestmark = pytest.mark.gpu


def _check(got, want, tol=1e-5):
  got = got.double().cpu().numpy()
  assert got.shape == want.shape, (got.shape, want.shape)
  assert rel_err(got, want) <= tol, rel_err(got, want)


def test_full_spatial_with_pool_and_readout():
  x1, x2 = blurred_images([3, 2], 8, seed=21)
  ours = [stax.Conv(1, (3, 3), padding='CIRCULAR', W_std=1.3, b_std=0.1),
          stax.Relu(), stax.AvgPool((2, 2), (2, 2)),
          stax.Conv(1, (3, 2), padding='CIRCULAR', W_std=1.2, b_std=0.2),
          stax.Erf(), stax.GlobalAvgPool(), stax.Dense(1, 1.0, 0.0)]
  ref = [so.conv_layer((3, 3), (1, 1), 'CIRCULAR', 1.3, 0.1), so.relu_layer(),
         so.avg_pool_layer((2, 2), (2, 2)),
         so.conv_layer((3, 2), (1, 1), 'CIRCULAR', 1.2, 0.2), so.erf_layer(),
         so.global_avg_pool_layer(), so.dense_layer(1.0, 0.0)]
  kfn = stax.serial(*ours)[2]
  got = kfn(x1, x2, ('nngp', 'ntk', 'cov1', 'cov2'))
  want = so.kernel_fn(ref, x1.astype(np.float64), x2.astype(np.float64),
                      ('nngp', 'ntk', 'cov1', 'cov2'))
  for g in ('nngp', 'ntk', 'cov1', 'cov2'):
    _check(getattr(got, g), want[g])
  sym = kfn(x1, None, 'ntk')
  _check(sym, so.kernel_fn(ref, x1.astype(np.float64), None, 'ntk')['ntk'])


@pytest.mark.parametrize('side', [16, 32])
def test_myrtle_shaped_circular(side):
  """Myrtle-5 with periodic convs (the per-layer kernels, not the fused
  SAME ones)."""
  x1, x2 = blurred_images([2, 2], side, seed=side)
  ours, ref = [], []
  for d in (2, 1, 1):
    for _ in range(d):
      ours += [stax.Conv(1, (3, 3), padding='CIRCULAR', W_std=2 ** 0.5),
               stax.Relu()]
      ref += [so.conv_layer((3, 3), (1, 1), 'CIRCULAR', 2 ** 0.5),
              so.relu_layer()]
    ours.append(stax.AvgPool((2, 2), (2, 2)))
    ref.append(so.avg_pool_layer((2, 2), (2, 2)))
  ours += [stax.GlobalAvgPool(), stax.Dense(1, 1.0, 0.0)]
  ref += [so.global_avg_pool_layer(), so.dense_layer(1.0, 0.0)]
  got = stax.serial(*ours)[2](x1, x2, ('nngp', 'ntk'))
  want = so.kernel_fn(ref, x1.astype(np.float64), x2.astype(np.float64),
                      ('nngp', 'ntk'))
  _check(got.nngp, want['nngp'])
  _check(got.ntk, want['ntk'])


def test_head_less_state_non_square():
  rng = np.random.default_rng(8)
  x1 = rng.standard_normal((2, 6, 5, 2)).astype(np.float32)
  x2 = rng.standard_normal((3, 6, 5, 2)).astype(np.float32)
  ours = [stax.Conv(1, (3, 3), padding='CIRCULAR', W_std=1.1, b_std=0.3),
          stax.Relu(), stax.Conv(1, (2, 3), padding='CIRCULAR', W_std=1.4),
          stax.Gelu()]
  ref = [so.conv_layer((3, 3), (1, 1), 'CIRCULAR', 1.1, 0.3),
         so.relu_layer(), so.conv_layer((2, 3), (1, 1), 'CIRCULAR', 1.4),
         so.gelu_layer()]
  got = stax.serial(*ours)[2](x1, x2, ('nngp', 'ntk'))
  want = so.kernel_fn(ref, x1.astype(np.float64), x2.astype(np.float64),
                      ('nngp', 'ntk'))
  _check(got.nngp, want['nngp'])
  _check(got.ntk, want['ntk'])


def test_diagonal_spatial_chain():
  x1, x2 = blurred_images([4, 3], 10, seed=12)
  ours = [stax.Conv(1, (3, 3), padding='CIRCULAR', b_std=0.2), stax.Relu(),
          stax.Conv(1, (3, 3), padding='CIRCULAR', W_std=1.5), stax.Erf(),
          stax.Flatten(), stax.Dense(1, 1.2, 0.1)]
  ref = [so.conv_layer((3, 3), (1, 1), 'CIRCULAR', 1.0, 0.2),
         so.relu_layer(), so.conv_layer((3, 3), (1, 1), 'CIRCULAR', 1.5),
         so.erf_layer(), so.flatten_layer(), so.dense_layer(1.2, 0.1)]
  got = stax.serial(*ours)[2](x1, x2, ('nngp', 'ntk'))
  want = so.kernel_fn(ref, x1.astype(np.float64), x2.astype(np.float64),
                      ('nngp', 'ntk'), diagonal_spatial=True)
  _check(got.nngp, want['nngp'])
  _check(got.ntk, want['ntk'])
