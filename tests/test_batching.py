"""Host logic of ``batch`` (pair tiling) with a CPU stand-in kernel_fn."""
import collections
import warnings

import numpy as np
import pytest
import torch

from neural_tangents_amd.batching import batch, _n_batches

AK = collections.namedtuple('AnalyticKernel', ('nngp', 'ntk'))


def fake_kernel_fn(x1, x2=None, get=None, **kw):
  """K(x1, x2) = x1 x2^T (+1 for ntk); not symmetric in its spatial part
  when full-spatial (tests the block transpose)."""
  x2 = x1 if x2 is None else x2
  a = torch.as_tensor(x1).reshape(x1.shape[0], -1).double()
  b = torch.as_tensor(x2).reshape(x2.shape[0], -1).double()
  k = a @ b.T
  if isinstance(get, str):
    return k
  return AK(k, k + 1.0)


def test_batched_equals_unbatched_cross_and_symmetric():
  rng = np.random.default_rng(0)
  x1 = rng.standard_normal((12, 5))
  x2 = rng.standard_normal((8, 5))
  full = fake_kernel_fn(x1, x2, ('nngp', 'ntk'))
  got = batch(fake_kernel_fn, batch_size=4)(x1, x2, ('nngp', 'ntk'))
  torch.testing.assert_close(got.nngp, full.nngp)
  torch.testing.assert_close(got.ntk, full.ntk)
  sym = batch(fake_kernel_fn, batch_size=4)(x1, None, 'nngp')
  torch.testing.assert_close(sym, fake_kernel_fn(x1, None, 'nngp'))
  host = batch(fake_kernel_fn, batch_size=4, store_on_device=False)(
      x1, x2, 'nngp')
  torch.testing.assert_close(host, full.nngp)


def test_batch_size_rules():
  """batching.py:647-679: reduced to gcd with a warning, else must divide."""
  with warnings.catch_warnings(record=True) as w:
    warnings.simplefilter('always')
    assert _n_batches(12, 8, 16) == (3, 2, 4)
    assert any('reduced' in str(m.message) for m in w)
  assert _n_batches(12, 8, 4) == (3, 2, 4)
  with pytest.raises(ValueError):
    _n_batches(12, 8, 3)


def test_multi_device_rejected_and_zero_batch_passthrough():
  with pytest.raises(NotImplementedError):
    batch(fake_kernel_fn, 4, device_count=2)
  assert batch(fake_kernel_fn, 0) is fake_kernel_fn
