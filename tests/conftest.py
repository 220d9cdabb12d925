import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
  sys.path.insert(0, ROOT)


def pytest_configure(config):
  config.addinivalue_line('markers',
                          'gpu: needs an MI355X (runs the HIP kernels)')
  config.addinivalue_line('markers', 'slow: long CPU test')


def rel_err(a, b):
  """Relative Frobenius error ||a - b|| / max(||b||, 1e-12), the matrix
  metric of tests/test_utils.py:390-415 in neural-tangents."""
  import numpy as np
  a = np.asarray(a, np.float64)
  b = np.asarray(b, np.float64)
  return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-12))
