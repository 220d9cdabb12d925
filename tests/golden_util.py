"""Loads the golden fixtures written by tests/golden/make_golden.py."""
import json
import os

import numpy as np

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'golden')
NAMES = ('myrtle5_8x8', 'myrtle10_16x16', 'myrtle5_8x8_same_inputs',
         'conv_relu_full_spatial_4x5', 'myrtle5_16x16_block',
         'myrtle5_32x32_block')


def load(name):
  with np.load(os.path.join(HERE, name + '.npz'), allow_pickle=False) as z:
    d = {k: z[k] for k in z.files}
  d['meta'] = json.loads(str(d['meta']))
  return d
