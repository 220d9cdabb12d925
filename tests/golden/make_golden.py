"""Writes the golden fixtures of this directory from the FP64 oracle.

Run from the repository root: ``python tests/golden/make_golden.py``.
Each fixture holds the inputs, the layer description and the oracle's
outputs; ``tests/test_golden.py`` checks the oracle against them (CPU) and
``tests/test_gpu_golden.py`` checks the HIP engine against them.
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from neural_tangents_amd.data import blurred_images, sha256  # noqa: E402
from oracle import stax_oracle as so  # noqa: E402


def layers_from_spec(spec):
  out = []
  for kind, kw in spec:
    if kind == 'conv':
      out.append(so.conv_layer(kw['filter'], (1, 1), kw['padding'],
                               kw['W_std'], kw['b_std']))
    elif kind == 'relu':
      out.append(so.relu_layer())
    elif kind == 'avgpool':
      out.append(so.avg_pool_layer(kw['window'], kw['window'], 'VALID'))
    elif kind == 'gap':
      out.append(so.global_avg_pool_layer())
    elif kind == 'flatten':
      out.append(so.flatten_layer())
    elif kind == 'dense':
      out.append(so.dense_layer(kw['W_std'], kw['b_std']))
    else:
      raise ValueError(kind)
  return out


def myrtle_spec(depths):
  spec = []
  for d in depths:
    for _ in range(d):
      spec += [('conv', dict(filter=[3, 3], padding='SAME', W_std=2 ** 0.5,
                             b_std=0.0)), ('relu', {})]
    spec += [('avgpool', dict(window=[2, 2]))]
  spec += [('gap', {}), ('dense', dict(W_std=1.0, b_std=0.0))]
  return spec


def write(name, spec, x1, x2, get, extra=None):
  res = so.kernel_fn(layers_from_spec(spec), x1, x2, get)
  arrays = {'x1': x1}
  if x2 is not None:
    arrays['x2'] = x2
  for g in get:
    arrays[g] = np.asarray(res[g], np.float64)
  meta = {'spec': spec, 'get': list(get), 'x1_sha256': sha256(x1),
          'x2_sha256': None if x2 is None else sha256(x2)}
  if extra:
    meta.update(extra)
  arrays['meta'] = np.array(json.dumps(meta))
  np.savez_compressed(os.path.join(HERE, name + '.npz'), **arrays)
  print(name, {k: v.shape for k, v in arrays.items()})


def main():
  x1, x2 = blurred_images([3, 2], 8, seed=8)
  write('myrtle5_8x8', myrtle_spec((2, 1, 1)), x1, x2, ('nngp', 'ntk'))
  x1, x2 = blurred_images([2, 2], 16, seed=16)
  write('myrtle10_16x16', myrtle_spec((3, 3, 3)), x1, x2, ('nngp', 'ntk'))
  x1, = blurred_images([4], 8, seed=9)
  write('myrtle5_8x8_same_inputs', myrtle_spec((2, 1, 1)), x1, None,
        ('nngp', 'ntk', 'cov1'))
  rng = np.random.default_rng(4)
  a = rng.standard_normal((2, 4, 5, 3)).astype(np.float32)
  b = rng.standard_normal((3, 4, 5, 3)).astype(np.float32)
  write('conv_relu_full_spatial_4x5',
        [('conv', dict(filter=[3, 3], padding='SAME', W_std=1.3, b_std=0.2)),
         ('relu', {})], a, b, ('nngp', 'ntk', 'cov1', 'cov2'))
  x1, x2 = blurred_images([64, 64], 16, seed=0)
  write('myrtle5_16x16_block', myrtle_spec((2, 1, 1)), x1[:4], x2[:4],
        ('nngp', 'ntk'),
        extra={'workload': 'myrtle5 16x16 64x64 seed 0, rows 0:4 cols 0:4',
               'workload_x1_sha256': sha256(x1),
               'workload_x2_sha256': sha256(x2)})
  x1, x2 = blurred_images([256, 256], 32, seed=0)
  write('myrtle5_32x32_block', myrtle_spec((2, 1, 1)), x1[:2], x2[:2],
        ('nngp', 'ntk'),
        extra={'workload': 'myrtle5 32x32 256x256 seed 0, rows 0:2 cols 0:2',
               'workload_x1_sha256': sha256(x1),
               'workload_x2_sha256': sha256(x2)})


if __name__ == '__main__':
  main()
