"""Kernel-regression predictors over GPU kernels: ``gp_inference`` and
``gradient_descent_mse_ensemble`` with the reference's call surface
(``predict.py:566-1100`` of neural-tangents).

``t=None`` (infinite time) solves ``(K + lambda I) alpha = y`` with a
Cholesky factorisation, ``lambda = diag_reg * tr(K) / N`` unless
``diag_reg_absolute_scale`` (``_add_diagonal_regularizer``,
``predict.py:1224-1232``).  Two solvers sit behind the same interface
(``solver=``):

* ``'mixed'`` (default): FP32 Cholesky factor, FP64 residuals, refined to
  FP64 accuracy (``neural_tangents_amd._solve``);
* ``'fp64'``: the plain FP64 Cholesky solve (rocSOLVER through torch), the
  reference's ``cho_factor``/``cho_solve`` on the GPU.

Finite ``t`` uses the FP64 eigendecomposition of the regularised train
kernel (``predict.py:876-883, 994-1098``).  Kernels are ``[N, N]`` matrices
(``trace_axes=(-1,)``); targets are ``[N]`` or ``[N, K]``.
"""
import collections
import functools
from typing import NamedTuple, Optional

import torch

from . import _solve


class Gaussian(NamedTuple):
  """``(mean, covariance)`` (``predict.py:552-563``)."""
  mean: torch.Tensor
  covariance: torch.Tensor


_NT_CACHE = {}


def _namedtuple(name, fields):
  key = (name, fields)
  if key not in _NT_CACHE:
    _NT_CACHE[key] = collections.namedtuple(name, fields)
  return _NT_CACHE[key]


def _canonicalize_get(get):
  """``utils.py:139-156``."""
  if get is None:
    return True, None
  if not get:
    raise ValueError('"get" must be non-empty.')
  single = isinstance(get, str)
  if single:
    get = (get,)
  get = tuple(s.lower() for s in get)
  if len(set(get)) < len(get):
    raise ValueError(f'All entries in "get" must be unique. Got {get}')
  return single, get


def _format(out: dict, get, single):
  """``utils.get_namedtuple('Gaussians')`` output convention."""
  if get is None:
    return _namedtuple('Gaussians', tuple(out))(*out.values())
  if single:
    return out[get[0]]
  return _namedtuple('Gaussians', get)(*[out[g] for g in get])


def _as_2d(k, name):
  if isinstance(k, torch.Tensor):
    t = k
  else:
    t = torch.as_tensor(k)
  if t.ndim != 2:
    raise NotImplementedError(
        f'{name} must be a [N1, N2] kernel matrix (trace_axes=(-1,)); got '
        f'shape {tuple(t.shape)}.')
  return t


def _get_attr(k, g):
  """``predict.py:1343-1346``."""
  if isinstance(k, torch.Tensor) or not hasattr(k, g):
    return k
  return getattr(k, g)


def _check_trace_axes(trace_axes, y):
  axes = tuple(a % y.ndim for a in trace_axes)
  if y.ndim == 2 and axes == (1,):
    return
  if y.ndim == 1 and axes == ():
    return
  raise NotImplementedError(
      'Only trace_axes=(-1,) with targets [N, K] (or trace_axes=() with '
      f'[N]) is supported; got trace_axes={trace_axes}, y shape '
      f'{tuple(y.shape)}.')


def _device_of(*ts):
  for t in ts:
    if isinstance(t, torch.Tensor) and t.is_cuda:
      return t.device
  if not torch.cuda.is_available():
    raise RuntimeError('predict runs on the GPU and no GPU is visible.')
  return torch.device('cuda', 0)


def gp_inference(k_train_train, y_train, diag_reg: float = 0.,
                 diag_reg_absolute_scale: bool = False, trace_axes=(-1,),
                 solver: str = 'mixed'):
  """``gp_inference`` (``predict.py:566-745``): returns
  ``predict_fn(get=None, k_test_train=None, k_test_test=None)``."""
  y_t = y_train if isinstance(y_train, torch.Tensor) else torch.as_tensor(
      y_train)
  if y_t.ndim == 1:
    trace_axes = ()
  _check_trace_axes(trace_axes, y_t)

  @functools.lru_cache(2)
  def solve(g):
    k_dd = _as_2d(_get_attr(k_train_train, g), 'k_train_train')
    dev = _device_of(k_dd, y_t)
    return _solve.CholeskySolver(k_dd.to(dev), diag_reg,
                                 diag_reg_absolute_scale, solver)

  @functools.lru_cache(2)
  def k_inv_y(g):
    s = solve(g)
    y = y_t.to(s.device)
    return s.solve(y if y.ndim == 2 else y[:, None])

  def predict_fn(get=None, k_test_train=None, k_test_test=None):
    single, get_c = _canonicalize_get(get)
    gets = ('nngp', 'ntk') if get_c is None else get_c
    out = {}
    for g in gets:
      if g not in ('nngp', 'ntk', 'ntkgp'):
        raise ValueError(g)
      k = g if g != 'ntkgp' else 'ntk'
      k_dd = _as_2d(_get_attr(k_train_train, k), 'k_train_train')
      k_td = None if k_test_train is None else _as_2d(
          _get_attr(k_test_train, k), 'k_test_train')
      if k_td is None:
        mean = y_t.to(dtype=k_dd.dtype)
      else:
        alpha = k_inv_y(k)
        mean = _solve.kernel_times(k_td.to(alpha.device), alpha)
        if y_t.ndim == 1:
          mean = mean[:, 0]
      if k_test_test is not None:
        if k_td is None:
          out[g] = Gaussian(mean, torch.zeros_like(k_dd))
        else:
          if g == 'ntk' and (not hasattr(k_train_train, 'nngp') or
                             not hasattr(k_test_train, 'nngp')):
            raise ValueError(
                'If `"ntk" in get`, and `k_test_test is not None`, and '
                '`k_test_train is not None`, i.e. you request the NTK '
                'posterior covariance on the test set, you need both NTK '
                'and NNGP train-train and test-train matrices contained in '
                '`k_test_train` and `k_train_train`. Hence they must be '
                '`namedtuple`s with `nngp` and `ntk` attributes.')
          g_init = 'nngp' if g != 'ntkgp' else 'ntk'
          s = solve(k)
          dev = s.device
          k_td64 = k_td.to(dev, torch.float64)
          k_td_g = _as_2d(_get_attr(k_test_train, g_init),
                          'k_test_train').to(dev, torch.float64)
          k_tt = _as_2d(_get_attr(k_test_test, g_init),
                        'k_test_test').to(dev, torch.float64)
          k_td_g_inv_y = s.solve(k_td_g.T)                # [N, T]
          if g in ('nngp', 'ntkgp'):
            cov = k_tt - k_td64 @ k_td_g_inv_y
          else:
            term_1 = s.solve(k_td64.T)                    # [N, T]
            nngp_dd = _as_2d(_get_attr(k_train_train, 'nngp'),
                             'k_train_train').to(dev, torch.float64)
            cov = term_1.T @ (nngp_dd @ term_1)
            term_2 = k_td64 @ k_td_g_inv_y
            term_2 = term_2 + term_2.T
            cov = cov - term_2 + k_tt
          out[g] = Gaussian(mean, cov)
      else:
        out[g] = mean
    return _format(out, get_c, single)

  return predict_fn


def _get_dependency(get, compute_cov):
  """``predict.py:1165-1177``."""
  _, get = _canonicalize_get(get)
  for g in get:
    if g not in ('nngp', 'ntk'):
      raise NotImplementedError(
          'Can only get either "nngp" or "ntk" predictions, got %s.' % g)
  dep = ()
  if 'nngp' in get or ('ntk' in get and compute_cov):
    dep += ('nngp',)
  if 'ntk' in get:
    dep += ('ntk',)
  return dep


_Kernel = collections.namedtuple('Kernel', 'nngp ntk')
_Kernel.__new__.__defaults__ = (None,) * len(_Kernel._fields)


def gradient_descent_mse_ensemble(kernel_fn, x_train, y_train,
                                  learning_rate: float = 1.,
                                  diag_reg: float = 0.0,
                                  diag_reg_absolute_scale: bool = False,
                                  trace_axes=(-1,), solver: str = 'mixed',
                                  **kernel_fn_train_train_kwargs):
  """``gradient_descent_mse_ensemble`` (``predict.py:753-1100``): returns
  ``predict_fn(t=None, x_test=None, get=None, compute_cov=False)``."""
  y_t = y_train if isinstance(y_train, torch.Tensor) else torch.as_tensor(
      y_train)
  if y_t.ndim == 1:
    trace_axes = ()
  _check_trace_axes(trace_axes, y_t)
  k_dd_cache = {}

  def get_k_train_train(get):
    """``predict.py:854-874``."""
    missing = tuple(g for g in get if g not in k_dd_cache)
    if missing:
      k = kernel_fn(x_train, None, missing, **kernel_fn_train_train_kwargs)
      for g in missing:
        k_dd_cache[g] = getattr(k, g)
    return _Kernel(**k_dd_cache)

  @functools.lru_cache(4)
  def predict_inf(get):
    k_dd = get_k_train_train(get)
    return gp_inference(k_dd, y_t, diag_reg, diag_reg_absolute_scale,
                        trace_axes, solver)

  @functools.lru_cache(2)
  def eigenspace(g):
    """``predict.py:876-883``."""
    k_dd = _as_2d(getattr(get_k_train_train((g,)), g), 'k_train_train')
    A = _solve.add_diagonal_regularizer(k_dd.to(torch.float64), diag_reg,
                                        diag_reg_absolute_scale)
    evals, evecs = torch.linalg.eigh(A)
    return evals, evecs

  def get_kernels(get, x_test, compute_cov, **kw):
    """``predict.py:892-942``."""
    get = _get_dependency(get, compute_cov)
    k_dd = get_k_train_train(get)
    if x_test is None:
      k_td = None
      nngp_tt = compute_cov or None
    else:
      kw_all = dict(kernel_fn_train_train_kwargs)
      for k, v in kw.items():
        if k in kw_all and kw_all[k] != v:
          raise ValueError(
              f'Same keyword argument {k} of `kernel_fn` is set to different '
              f'values {kw_all[k]} != {v} when computing the train-train and '
              'test-train/test-test kernels.')
        kw_all[k] = v
      k_td = kernel_fn(x_test, x_train, get, **kw_all)
      nngp_tt = kernel_fn(x_test, None, 'nngp', **kw_all) if compute_cov \
          else None
    return k_dd, k_td, nngp_tt

  def predict_fn(t=None, x_test=None, get=None, compute_cov=False,
                 **kernel_fn_test_test_kwargs):
    single, get_c = _canonicalize_get(get)
    gets = ('nngp', 'ntk') if get_c is None else get_c
    k_dd, k_td, nngp_tt = get_kernels(gets, x_test, compute_cov,
                                      **kernel_fn_test_test_kwargs)
    if t is None:
      res = predict_inf(gets)(get=gets, k_test_train=k_td,
                              k_test_test=nngp_tt)
      out = dict(zip(gets, res))
      return _format(out, get_c, single)
    return _format(_finite_t(t, gets, k_dd, k_td, nngp_tt), get_c, single)

  def _finite_t(t, gets, k_dd, k_td, nngp_tt):
    """``predict.py:1000-1098`` (``expm1``/``inv_expm1`` 1277-1294)."""
    dev = _device_of(*(v for v in k_dd if v is not None))
    tt = torch.as_tensor(t, dtype=torch.float64, device=dev) * learning_rate
    t_shape = tt.shape
    tt = tt.reshape(-1, 1)
    size = y_t.numel()
    y = y_t.to(dev, torch.float64)
    y2 = y if y.ndim == 2 else y[:, None]
    out = {}
    for g in gets:
      evals, evecs = eigenspace(g)
      ev = evals[None, :]
      expm1 = torch.expm1(-torch.clamp(ev, min=0.) * tt / size)   # [t, N]
      proj = evecs.T @ y2                                          # [N, K]
      if k_td is None:
        mean = torch.einsum('ji,ti,ik->tjk', evecs, -expm1, proj)
      else:
        ktd = _as_2d(getattr(k_td, g), 'k_test_train').to(dev,
                                                             torch.float64)
        inv = -expm1 / torch.abs(ev)
        mean = torch.einsum('lj,ji,ti,ik->tlk', ktd, evecs, inv, proj)
      if y.ndim == 1:
        mean = mean[..., 0]
      mean = mean.reshape(t_shape + mean.shape[1:])
      if nngp_tt is not None:
        nngp_dd = _as_2d(k_dd.nngp, 'k_train_train').to(dev, torch.float64)
        cp = torch.clamp(ev, min=0.)
        if k_td is None:
          if g == 'nngp':
            cov = torch.einsum('ji,ti,ki->tjk', evecs,
                               cp * torch.exp(-2 * cp * tt / size), evecs)
          else:
            ex = torch.einsum('mi,ti,ki->tmk', evecs,
                              torch.exp(-cp * tt / size), evecs)
            cov = torch.einsum('tmk,kl,tnl->tmn', ex, nngp_dd, ex)
        else:
          ntt = _as_2d(nngp_tt, 'k_test_test').to(dev, torch.float64)[None]
          ktd = _as_2d(getattr(k_td, g), 'k_test_train').to(dev,
                                                               torch.float64)
          if g == 'nngp':
            inv2 = torch.expm1(-cp * 2 * tt / size) / torch.abs(ev)
            cov = ntt - torch.einsum('mj,ji,ti,ki,lk->tml', ktd, evecs,
                                     -inv2, evecs, ktd)
          else:
            neg_inv = -expm1 / torch.abs(ev)
            term_1 = torch.einsum('mi,ti,ki,lk->tml', evecs, neg_inv, evecs,
                                  ktd)
            ktd_n = _as_2d(k_td.nngp, 'k_test_train').to(dev, torch.float64)
            term_2 = torch.einsum('mj,ji,ti,ki,lk->tml', ktd, evecs, neg_inv,
                                  evecs, ktd_n)
            term_2 = term_2 + term_2.transpose(1, 2)
            cov = torch.einsum('tji,jk,tkl->til', term_1, nngp_dd, term_1)
            cov = cov - term_2 + ntt
        cov = cov.reshape(t_shape + cov.shape[1:])
        out[g] = Gaussian(mean, cov)
      else:
        out[g] = mean
    return out

  return predict_fn
