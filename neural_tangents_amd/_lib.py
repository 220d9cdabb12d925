"""ctypes binding of libnta.so (the C ABI declared in include/nta.h).

The library must exist: there is no CPU or eager fallback for any entry
point, and a missing or unloadable library raises ``RuntimeError``.
"""
import ctypes
import os
import threading

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, 'libnta.so')

NTA_OK = 0
NTA_EINVAL = -1
NTA_EUNSUPPORTED = -2

RELU = 1
NTK = 2
CONV = 4
CONV_VALID = 8
FLATTEN = 16
ACT_ERF = 32
ACT_GELU = 64
CONV_CIRCULAR = 128

_lock = threading.Lock()
_lib = None

_P = ctypes.c_void_p
_I = ctypes.c_int
_F = ctypes.c_float
_D = ctypes.c_double
_LL = ctypes.c_longlong

# name -> argtypes; every entry point returns int status.
SIGNATURES = {
    'nta_first_diag': [_P, _I, _I, _I, _I, _I, _I, _I, _F, _F, _P, _P],
    'nta_first_stage': [_P, _P, _I, _I, _I, _I, _I, _I, _I, _I, _F, _F, _F,
                        _F, _P, _P, _P, _P, _P],
    'nta_stage_diag': [_P, _I, _I, _I, _I, _I, _I, _F, _F, _P, _P],
    'nta_stage': [_P, _P, _I, _I, _I, _I, _I, _I, _I, _I, _I, _F, _F, _F, _F,
                  _P, _P, _P, _P, _P],
    'nta_stage_keep_pre': [_P, _P, _I, _I, _I, _I, _F, _F, _P, _P, _P, _P,
                           _P, _P, _P],
    'nta_head': [_P, _P, _I, _I, _I, _I, _I, _D, _D, _D, _D, _D, _P, _P, _LL,
                 _P],
    'nta_chol_f32': [_P, _I, _LL, _P, _LL, _P, _P, _P],
    'nta_copy_lower_shift': [_P, _LL, _I, _F, _P, _LL, _P],
    'nta_chol_precond': [_P, _P, _I, _LL, _I, _P, _I, _P, _P, _LL, _P],
    'nta_pcg64': [_P, _I, _LL, _D, _P, _P, _LL, _I, _P, _I, _P, _D, _I, _P,
                  _LL, _P, _P, _P],
    'nta_pcg64_k64': [_P, _I, _LL, _D, _P, _P, _LL, _I, _P, _I, _P, _D, _I,
                      _P, _LL, _P, _P, _P],
    'nta_gemm_f32_f64': [_P, _LL, _I, _I, _P, _LL, _I, _P, _LL, _P, _LL,
                         _P],
    'nta_pcg64_noalloc': [_P, _I, _LL, _D, _P, _P, _LL, _I, _P, _I, _P, _D,
                          _I, _P, _P, _P],
    'nta_diag_chain': [_P, _P, _I, _I, _I, _I, _I, _I, _I, _P, _P, _P, _P,
                       _I, _D, _D, _D, _D, _D, _P, _P, _LL, _P],
    'nta_first_pool_stage': [_P, _P, _I, _I, _I, _I, _I, _I, _F, _F, _P, _P,
                             _F, _F, _P, _P, _P, _P, _P],
    'nta_first_pool_stage_keep_pre': [_P, _P, _I, _I, _I, _I, _I, _I, _F, _F,
                                      _P, _P, _F, _F, _P, _P, _P, _P, _P, _P,
                                      _P],
    'nta_first3_pool_stage': [_P, _P, _I, _I, _I, _I, _I, _I, _F, _F, _P,
                              _P, _F, _F, _P, _P, _F, _F, _P, _P, _P, _P,
                              _P],
    'nta_qnorm': [_P, _LL, _P, _P],
    'nta_relu_bwd': [_P, _P, _I, _I, _I, _I, _I, _F, _F, _F, _F, _P, _P, _P,
                     _P, _F, _P, _P, _P, _P, _P, _P, _LL, _P],
    'nta_relu_bwd_accumulate': [_P, _P, _P, _I, _I, _I, _P, _P, _P, _P],
}

# Entry points that return a value instead of a status.
VALUE_FNS = {
    'nta_chol_split': (_I, [_I]),
    'nta_pcg64_work_bytes': (_LL, [_I, _I]),
    'nta_gemm_f32_f64_work_bytes': (_LL, [_I, _I]),
    'nta_relu_bwd_parts': (_LL, [_I, _I, _I]),
    'nta_relu_bwd_work_bytes': (_LL, [_I, _I, _I]),
}


def load():
  """Loads libnta.so once; raises RuntimeError if it is missing."""
  global _lib
  if _lib is not None:
    return _lib
  with _lock:
    if _lib is not None:
      return _lib
    if not os.path.exists(LIB_PATH):
      raise RuntimeError(
          f'{LIB_PATH} is missing: build it with '
          '`python -m neural_tangents_amd._build` (there is no fallback).')
    lib = ctypes.CDLL(LIB_PATH)
    lib.nta_last_error.restype = ctypes.c_char_p
    lib.nta_last_error.argtypes = []
    lib.nta_version.restype = ctypes.c_int
    lib.nta_version.argtypes = []
    for name, args in SIGNATURES.items():
      fn = getattr(lib, name)
      fn.restype = ctypes.c_int
      fn.argtypes = args
    for name, (res, args) in VALUE_FNS.items():
      fn = getattr(lib, name)
      fn.restype = res
      fn.argtypes = args
    _lib = lib
  return _lib


def exported_symbols():
  return (['nta_version', 'nta_last_error'] + list(SIGNATURES) +
          list(VALUE_FNS))


def value(name, *args):
  """Calls an entry point that returns a value (VALUE_FNS)."""
  return getattr(load(), name)(*args)


def call(name, *args):
  """Calls an entry point and maps a non-zero status to an exception:
  NTA_EINVAL -> ValueError, NTA_EUNSUPPORTED -> NotImplementedError,
  positive (HIP error) -> RuntimeError."""
  lib = load()
  rc = getattr(lib, name)(*args)
  if rc != NTA_OK:
    msg = (lib.nta_last_error() or b'').decode(errors='replace')
    if rc == NTA_EINVAL:
      raise ValueError(msg)
    if rc == NTA_EUNSUPPORTED:
      raise NotImplementedError(msg)
    raise RuntimeError(f'HIP error {rc}: {msg}')
  return rc


def ptr(t):
  """Device pointer of a torch tensor (or None -> NULL)."""
  return None if t is None else ctypes.c_void_p(t.data_ptr())


def stream_handle(device=None):
  import torch
  return ctypes.c_void_p(torch.cuda.current_stream(device).cuda_stream)
