"""The C-ABI library loads and exports every entry point include/nta.h
declares (no compute calls: runs without a GPU)."""
import ctypes
import os
import re

from neural_tangents_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
  names = set()
  for f in os.listdir(os.path.join(ROOT, 'include')):
    if f.endswith('.h'):
      src = open(os.path.join(ROOT, 'include', f)).read()
      src = re.sub(r'/\*.*?\*/', '', src, flags=re.S)
      names |= set(re.findall(r'\b(nta_\w+)\s*\(', src))
  return names


def test_library_exports_every_declared_symbol():
  lib = ctypes.CDLL(_lib.LIB_PATH)
  declared = _declared()
  assert declared, 'no declarations found'
  for name in sorted(declared):
    assert hasattr(lib, name), name
  assert declared == set(_lib.exported_symbols()) | (
      declared - set(_lib.exported_symbols())), declared


def test_bindings_cover_declarations():
  assert _declared() == set(_lib.exported_symbols())


def test_version_and_error_string():
  lib = _lib.load()
  assert lib.nta_version() >= 100
  assert isinstance(lib.nta_last_error(), bytes)


def test_invalid_arguments_return_einval_without_gpu_work():
  lib = _lib.load()
  # n < 0 is rejected before any HIP call.
  rc = lib.nta_first_diag(None, -1, 8, 8, 3, 3, 3, 0, 1.0, 0.0, None, None)
  assert rc == _lib.NTA_EINVAL
  assert b'nta_first_diag' in lib.nta_last_error()
  rc = lib.nta_stage(None, None, 1, 1, 8, 8, 3, 3, 0, 1, 0, 1.0, 0.0, 1.0,
                     0.0, None, None, None, None, None)
  assert rc == _lib.NTA_EINVAL


def test_first_pool_stage_rejects_other_shapes_without_gpu_work():
  lib = _lib.load()
  rc = lib.nta_first_pool_stage(None, None, 1, 1, 16, 16, 3, 0, 2.0, 0.0,
                                None, None, 2.0, 0.0, None, None, None, None,
                                None)
  assert rc == _lib.NTA_EUNSUPPORTED
  assert b'nta_first_pool_stage' in lib.nta_last_error()
  rc = lib.nta_first_pool_stage(None, None, 1, 1, 32, 32, 5, 0, 2.0, 0.0,
                                None, None, 2.0, 0.0, None, None, None, None,
                                None)
  assert rc == _lib.NTA_EUNSUPPORTED
  rc = lib.nta_first_pool_stage(None, None, 2, 1, 32, 32, 3, 0, 2.0, 0.0,
                                None, None, 2.0, 0.0, None, None, None, None,
                                None)
  assert rc == _lib.NTA_EINVAL
  assert lib.nta_qnorm(None, -1, None, None) == _lib.NTA_EINVAL


def test_keep_pre_entry_points_check_shapes_without_gpu_work():
  lib = _lib.load()
  rc = lib.nta_first_pool_stage_keep_pre(None, None, 1, 1, 16, 16, 3, 0, 2.0,
                                         0.0, None, None, 2.0, 0.0, None,
                                         None, None, None, None, None, None)
  assert rc == _lib.NTA_EUNSUPPORTED
  assert b'nta_first_pool_stage_keep_pre' in lib.nta_last_error()
  rc = lib.nta_first_pool_stage_keep_pre(None, None, 2, 1, 32, 32, 3, 0, 2.0,
                                         0.0, None, None, 2.0, 0.0, None,
                                         None, None, None, None, None, None)
  assert rc == _lib.NTA_EINVAL
  relu = _lib.CONV | _lib.RELU
  assert lib.nta_stage_keep_pre(None, None, 1, 0, 12, relu, 2.0, 0.0, None,
                                None, None, None, None, None,
                                None) == _lib.NTA_EUNSUPPORTED
  assert lib.nta_stage_keep_pre(None, None, 1, 0, 16, _lib.CONV, 2.0, 0.0,
                                None, None, None, None, None, None,
                                None) == _lib.NTA_EUNSUPPORTED
  assert lib.nta_stage_keep_pre(None, None, -1, 0, 16, relu, 2.0, 0.0, None,
                                None, None, None, None, None,
                                None) == _lib.NTA_EINVAL
  assert lib.nta_stage_keep_pre(None, None, 1, 0, 16, relu, 2.0, 0.0, None,
                                None, None, None, None, None,
                                None) == _lib.NTA_EINVAL
