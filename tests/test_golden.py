"""The oracle reproduces its committed golden fixtures, and the fixture
inputs are the pinned synthetic workload (CPU only)."""
import numpy as np
import pytest

from golden_util import NAMES, load
from neural_tangents_amd.data import blurred_images, sha256
from oracle import stax_oracle as so


def _layers(spec):
  import importlib.util
  import os
  path = os.path.join(os.path.dirname(__file__), 'golden', 'make_golden.py')
  s = importlib.util.spec_from_file_location('make_golden', path)
  m = importlib.util.module_from_spec(s)
  s.loader.exec_module(m)
  return m.layers_from_spec(spec)


@pytest.mark.parametrize('name', [n for n in NAMES
                                  if n != 'myrtle5_32x32_block'])
def test_oracle_reproduces_fixture(name):
  d = load(name)
  meta = d['meta']
  x2 = d.get('x2')
  res = so.kernel_fn(_layers(meta['spec']), d['x1'], x2, tuple(meta['get']))
  for g in meta['get']:
    np.testing.assert_allclose(res[g], d[g], rtol=1e-12, atol=1e-14)


def test_workload_generator_is_pinned():
  d = load('myrtle5_16x16_block')
  x1, x2 = blurred_images([64, 64], 16, seed=0)
  assert sha256(x1) == d['meta']['workload_x1_sha256']
  assert sha256(x2) == d['meta']['workload_x2_sha256']
  np.testing.assert_array_equal(x1[:4], d['x1'])
