"""ABRelu / LeakyRelu / Abs on the HIP engine against the FP64 oracle
(ABRelu.kernel_fn, elementwise.py:405-520), full-spatial and
diagonal-spatial, plus their kernel gradients against finite differences
(needs a GPU)."""
import numpy as np
import pytest
import torch

from conftest import rel_err
from neural_tangents_amd import grad, stax
from neural_tangents_amd.data import blurred_images
from oracle import stax_oracle as so

pytestmark = pytest.mark.gpu

ACTS = [('leaky', stax.LeakyRelu(0.1), (0.1, 1.0)),
        ('abs', stax.Abs(), (-1.0, 1.0)),
        ('ab', stax.ABRelu(-0.5, 2.0), (-0.5, 2.0))]


def _pooled_net(act, ab, side_stages=2, b_std=0.1):
  ours, ref = [], []
  for _ in range(side_stages):
    ours += [stax.Conv(1, (3, 3), padding='SAME', W_std=1.3, b_std=b_std),
             act, stax.AvgPool((2, 2), (2, 2))]
    ref += [so.conv_layer((3, 3), (1, 1), 'SAME', 1.3, b_std),
            so.abrelu_layer(*ab), so.avg_pool_layer((2, 2), (2, 2))]
  ours += [stax.GlobalAvgPool(), stax.Dense(1, 1.0, 0.0)]
  ref += [so.global_avg_pool_layer(), so.dense_layer(1.0, 0.0)]
  return stax.serial(*ours)[2], ref


@pytest.mark.parametrize('name,act,ab', ACTS, ids=[a[0] for a in ACTS])
@pytest.mark.parametrize('side', [8, 16])
def test_full_spatial_matches_oracle(name, act, ab, side):
  x1, x2 = blurred_images([3, 4], side, seed=side + 1)
  kfn, layers = _pooled_net(act, ab)
  got = kfn(x1, x2, ('nngp', 'ntk'))
  want = so.kernel_fn(layers, x1.astype(np.float64), x2.astype(np.float64),
                      ('nngp', 'ntk'))
  for g in ('nngp', 'ntk'):
    assert rel_err(getattr(got, g).double().cpu().numpy(), want[g]) <= 1e-5
  sym = kfn(x1, None, 'ntk')
  want_s = so.kernel_fn(layers, x1.astype(np.float64), None, 'ntk')['ntk']
  assert rel_err(sym.double().cpu().numpy(), want_s) <= 1e-5


@pytest.mark.parametrize('name,act,ab', ACTS, ids=[a[0] for a in ACTS])
def test_diagonal_spatial_matches_oracle(name, act, ab):
  x1, x2 = blurred_images([4, 3], 12, seed=7)
  ours = [stax.Conv(1, (3, 3), padding='SAME', b_std=0.2), act,
          stax.Conv(1, (3, 3), padding='SAME'), act, stax.Flatten()]
  ref = [so.conv_layer((3, 3), (1, 1), 'SAME', 1.0, 0.2), so.abrelu_layer(*ab),
         so.conv_layer((3, 3), (1, 1), 'SAME'), so.abrelu_layer(*ab),
         so.flatten_layer()]
  got = stax.serial(*ours)[2](x1, x2, ('nngp', 'ntk'))
  want = so.kernel_fn(ref, x1.astype(np.float64), x2.astype(np.float64),
                      ('nngp', 'ntk'), diagonal_spatial=True)
  for g in ('nngp', 'ntk'):
    assert rel_err(getattr(got, g).double().cpu().numpy(), want[g]) <= 1e-5


def test_leaky_relu_kernel_gradient_matches_finite_differences():
  """d/dx2 of sum(G * ntk(x1, x2)) through LeakyRelu stages."""
  x1, x2 = blurred_images([2, 3], 8, seed=41)
  x1 = x1.astype(np.float64)
  x2 = x2.astype(np.float64)
  alpha = 0.2
  conv = lambda: stax.Conv(1, (3, 3), padding='SAME', W_std=1.2, b_std=0.3)
  oconv = lambda: so.conv_layer((3, 3), (1, 1), 'SAME', 1.2, 0.3)
  ours = [conv(), stax.LeakyRelu(alpha)]
  ref = [oconv(), so.abrelu_layer(alpha, 1.0)]
  for _ in range(2):
    ours += [conv(), stax.LeakyRelu(alpha), stax.AvgPool((2, 2), (2, 2))]
    ref += [oconv(), so.abrelu_layer(alpha, 1.0),
            so.avg_pool_layer((2, 2), (2, 2))]
  ours += [stax.GlobalAvgPool()]
  ref += [so.global_avg_pool_layer()]
  kfn = stax.serial(*ours)[2]
  plan = kfn.plan(x1.shape)
  G = np.random.default_rng(3).standard_normal((2, 3))
  dev = torch.device('cuda', 0)
  gx1, gx2, _, _ = grad.kernel_vjp(
      plan, torch.tensor(x1, dtype=torch.float32, device=dev),
      torch.tensor(x2, dtype=torch.float32, device=dev), None,
      torch.tensor(G, dtype=torch.float32, device=dev))
  g = gx2.double().cpu().numpy()
  f = lambda z: float(np.sum(G * so.kernel_fn(ref, x1, z, 'ntk')['ntk']))
  v = np.random.default_rng(4).standard_normal(x2.shape)
  v /= np.linalg.norm(v)
  h = 1e-3
  fd = (f(x2 + h * v) - f(x2 - h * v)) / (2 * h)
  assert abs(float(np.sum(g * v)) - fd) <= 1e-4 * np.linalg.norm(g)


def test_sum_pools_match_oracle():
  """SumPool / GlobalSumPool through the AvgPool kernels with the window
  factors folded into the next Conv, the readout and the outputs."""
  x1, x2 = blurred_images([3, 2], 16, seed=9)
  conv = lambda: stax.Conv(1, (3, 3), padding='SAME', W_std=1.1, b_std=0.2)
  oconv = lambda: so.conv_layer((3, 3), (1, 1), 'SAME', 1.1, 0.2)
  ours = [conv(), stax.Relu(), conv(), stax.Relu(),
          stax.SumPool((2, 2), (2, 2)), conv(), stax.Relu(),
          stax.AvgPool((2, 2), (2, 2)), stax.GlobalSumPool(),
          stax.Dense(1, 0.5, 0.1)]
  ref = [oconv(), so.relu_layer(), oconv(), so.relu_layer(),
         so.sum_pool_layer((2, 2), (2, 2)), oconv(), so.relu_layer(),
         so.avg_pool_layer((2, 2), (2, 2)),
         so.global_avg_pool_layer(pool_sum=True), so.dense_layer(0.5, 0.1)]
  got = stax.serial(*ours)[2](x1, x2, ('nngp', 'ntk'))
  want = so.kernel_fn(ref, x1.astype(np.float64), x2.astype(np.float64),
                      ('nngp', 'ntk'))
  for g in ('nngp', 'ntk'):
    assert rel_err(getattr(got, g).double().cpu().numpy(), want[g]) <= 1e-5
  # a head-less network ending in SumPool returns the scaled state
  ours = [conv(), stax.Relu(), stax.SumPool((2, 2), (2, 2))]
  ref = [oconv(), so.relu_layer(), so.sum_pool_layer((2, 2), (2, 2))]
  k = stax.serial(*ours)[2](x1[:2], x2[:1], None)
  want = so.kernel_fn(ref, x1[:2].astype(np.float64),
                      x2[:1].astype(np.float64), None)
  assert rel_err(k.nngp.double().cpu().numpy(), want.nngp) <= 1e-5
  assert rel_err(k.ntk.double().cpu().numpy(), want.ntk) <= 1e-5
  assert rel_err(k.cov1.double().cpu().numpy(), want.cov1) <= 1e-5
