"""The HIP engine against the committed golden fixtures (needs a GPU)."""
import numpy as np
import pytest

from conftest import rel_err
from golden_util import NAMES, load
from neural_tangents_amd import stax

pytestmark = pytest.mark.gpu


def _net(spec):
  layers = []
  for kind, kw in spec:
    if kind == 'conv':
      layers.append(stax.Conv(1, tuple(kw['filter']), padding=kw['padding'],
                              W_std=kw['W_std'], b_std=kw['b_std']))
    elif kind == 'relu':
      layers.append(stax.Relu())
    elif kind == 'avgpool':
      layers.append(stax.AvgPool(tuple(kw['window']), tuple(kw['window'])))
    elif kind == 'gap':
      layers.append(stax.GlobalAvgPool())
    elif kind == 'flatten':
      layers.append(stax.Flatten())
    elif kind == 'dense':
      layers.append(stax.Dense(1, kw['W_std'], kw['b_std']))
  return stax.serial(*layers)[2]


@pytest.mark.parametrize('name', NAMES)
def test_engine_matches_fixture(name):
  d = load(name)
  meta = d['meta']
  k = _net(meta['spec'])(d['x1'], d.get('x2'), tuple(meta['get']))
  for g in meta['get']:
    got = getattr(k, g).double().cpu().numpy()
    assert got.shape == d[g].shape, g
    assert rel_err(got, d[g]) <= 1e-5, (g, rel_err(got, d[g]))
