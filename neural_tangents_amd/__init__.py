"""MI355X-native engine for the analytic infinite-width CNN kernels of
neural-tangents: ``stax`` (``serial`` -> ``kernel_fn``), ``predict``
(kernel regression) and ``batch`` (pair tiling) over hand-written HIP
kernels in libnta.so.
"""
from . import stax
from .kernel import Kernel

__version__ = '0.1.0'

__all__ = ['stax', 'Kernel', 'predict', 'batch']


def __getattr__(name):
  import importlib
  if name == 'predict':
    return importlib.import_module(__name__ + '.predict')
  if name == 'batch':
    return importlib.import_module(__name__ + '.batching').batch
  raise AttributeError(name)
