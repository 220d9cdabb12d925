"""Erf and Gelu (elementwise.py:49-108, 197-277) on the HIP engine against
the FP64 oracle: full-spatial stages with pooling (first stage and later
stages), the diagonal-spatial chain, self kernels with cov1/cov2 (needs a
GPU)."""
import numpy as np
import pytest

from conftest import rel_err
from neural_tangents_amd import stax
from neural_tangents_amd.data import blurred_images
from oracle import stax_oracle as so

pytestmark = pytest.mark.gpu

ACTS = [('erf', lambda: stax.Erf(), lambda: so.erf_layer()),
        ('erf_abc', lambda: stax.Erf(0.8, 1.7, 0.3),
         lambda: so.erf_layer(0.8, 1.7, 0.3)),
        ('gelu', lambda: stax.Gelu(), lambda: so.gelu_layer())]


def _check(got, want, tol=1e-5):
  got = got.double().cpu().numpy()
  assert got.shape == want.shape, (got.shape, want.shape)
  assert rel_err(got, want) <= tol, rel_err(got, want)


@pytest.mark.parametrize('name,act,oact', ACTS, ids=[a[0] for a in ACTS])
@pytest.mark.parametrize('side', [8, 16])
def test_full_spatial_matches_oracle(name, act, oact, side):
  x1, x2 = blurred_images([3, 2], side, seed=side + 3)
  ours = [stax.Conv(1, (3, 3), padding='SAME', W_std=1.3, b_std=0.1), act(),
          stax.AvgPool((2, 2), (2, 2)),
          stax.Conv(1, (3, 3), padding='SAME', W_std=1.2, b_std=0.2), act(),
          stax.GlobalAvgPool(), stax.Dense(1, 1.0, 0.0)]
  ref = [so.conv_layer((3, 3), (1, 1), 'SAME', 1.3, 0.1), oact(),
         so.avg_pool_layer((2, 2), (2, 2)),
         so.conv_layer((3, 3), (1, 1), 'SAME', 1.2, 0.2), oact(),
         so.global_avg_pool_layer(), so.dense_layer(1.0, 0.0)]
  kfn = stax.serial(*ours)[2]
  got = kfn(x1, x2, ('nngp', 'ntk', 'cov1', 'cov2'))
  want = so.kernel_fn(ref, x1.astype(np.float64), x2.astype(np.float64),
                      ('nngp', 'ntk', 'cov1', 'cov2'))
  for g in ('nngp', 'ntk', 'cov1', 'cov2'):
    _check(getattr(got, g), want[g])
  sym = kfn(x1, None, 'ntk')
  _check(sym, so.kernel_fn(ref, x1.astype(np.float64), None, 'ntk')['ntk'])


@pytest.mark.parametrize('name,act,oact', ACTS, ids=[a[0] for a in ACTS])
def test_full_spatial_state_without_readout(name, act, oact):
  """Head-less network: the [N1, N2, H, H, W, W] state after the
  activation, including a stage whose activation follows a pool."""
  x1, x2 = blurred_images([2, 3], 6, seed=4)
  ours = [stax.Conv(1, (3, 3), padding='SAME', W_std=1.1, b_std=0.3), act(),
          stax.Conv(1, (3, 3), padding='SAME', W_std=1.4), act()]
  ref = [so.conv_layer((3, 3), (1, 1), 'SAME', 1.1, 0.3), oact(),
         so.conv_layer((3, 3), (1, 1), 'SAME', 1.4), oact()]
  got = stax.serial(*ours)[2](x1, x2, ('nngp', 'ntk'))
  want = so.kernel_fn(ref, x1.astype(np.float64), x2.astype(np.float64),
                      ('nngp', 'ntk'))
  _check(got.nngp, want['nngp'])
  _check(got.ntk, want['ntk'])


@pytest.mark.parametrize('name,act,oact', ACTS, ids=[a[0] for a in ACTS])
def test_diagonal_spatial_matches_oracle(name, act, oact):
  x1, x2 = blurred_images([4, 3], 12, seed=9)
  ours = [stax.Conv(1, (3, 3), padding='SAME', b_std=0.2), act(),
          stax.Conv(1, (3, 3), padding='SAME', W_std=1.5), act(),
          stax.Flatten(), stax.Dense(1, 1.2, 0.1)]
  ref = [so.conv_layer((3, 3), (1, 1), 'SAME', 1.0, 0.2), oact(),
         so.conv_layer((3, 3), (1, 1), 'SAME', 1.5), oact(),
         so.flatten_layer(), so.dense_layer(1.2, 0.1)]
  kfn = stax.serial(*ours)[2]
  got = kfn(x1, x2, ('nngp', 'ntk'))
  want = so.kernel_fn(ref, x1.astype(np.float64), x2.astype(np.float64),
                      ('nngp', 'ntk'), diagonal_spatial=True)
  _check(got.nngp, want['nngp'])
  _check(got.ntk, want['ntk'])
