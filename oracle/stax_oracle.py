"""FP64 CPU restatement of the analytic-kernel layers (test infrastructure).

TEST INFRASTRUCTURE ONLY -- never imported by the product package.

Every array follows the reference's layouts: ``nngp``/``ntk`` are
``[N1, N2, X, X', Y, Y']`` with spatial axes zipped (``utils.py:324-342``),
``cov1``/``cov2`` are ``[N, X, X', Y, Y']`` (``diagonal_batch=True``), and
the spatial order flips after every Conv exactly as ``is_reversed`` records
(``linear.py:1343-1349, 3366-3376``).  Arithmetic runs in the dtype of the
inputs after the cast in ``inputs_to_kernel`` (float64 by default, float32 for
the timed CPU baseline).

Reference: neural-tangents 0.6.6, paths relative to
``neural_tangents/_src/``.
"""
import dataclasses
import math
from typing import Optional, Sequence

import numpy as np


@dataclasses.dataclass
class OKernel:
  """Mirror of ``utils/kernel.py:27-149`` (fields 124-145) for the oracle."""
  nngp: np.ndarray
  ntk: Optional[np.ndarray]
  cov1: np.ndarray
  cov2: Optional[np.ndarray]
  x1_is_x2: bool
  is_gaussian: bool
  is_reversed: bool
  diagonal_batch: bool
  diagonal_spatial: bool
  shape1: tuple
  shape2: tuple

  def replace(self, **kw) -> 'OKernel':
    return dataclasses.replace(self, **kw)


# ---------------------------------------------------------------- axis utils

def zip_axes(x, start_axis=0):
  """``utils.py:324-342, 380-402``: [.., X, Y, X', Y'] -> [.., X, X', Y, Y']."""
  half = (x.ndim - start_axis) // 2
  odd_axes = list(range(start_axis + 1, x.ndim, 2))
  last_axes = list(range(x.ndim - half, x.ndim))
  return np.moveaxis(x, last_axes, odd_axes)


def unzip_axes(x, start_axis=0):
  """``utils.py:345-361``: [.., X, X', Y, Y'] -> [.., X, Y, X', Y']."""
  half = (x.ndim - start_axis) // 2
  odd_axes = list(range(start_axis + 1, x.ndim, 2))
  last_axes = list(range(x.ndim - half, x.ndim))
  return np.moveaxis(x, odd_axes, last_axes)


def diagonal_between(x, start_axis):
  """``utils.py:405-429`` with ``end_axis = x.ndim``."""
  half = (x.ndim - start_axis) // 2
  if half == 0:
    return x
  side = x.shape[start_axis:start_axis + half]
  size = int(np.prod(side))
  x2 = x.reshape(x.shape[:start_axis] + (size, size))
  d = np.diagonal(x2, axis1=start_axis, axis2=start_axis + 1)
  return d.reshape(x.shape[:start_axis] + side)


def get_diagonal(cov, diagonal_spatial):
  """``stax/requirements.py:1057-1074`` for ``diagonal_batch=True``."""
  if cov is None:
    return None
  if diagonal_spatial:
    return cov
  return diagonal_between(unzip_axes(cov, 1), 1)


def outer_prod(x, y, start_axis):
  """``utils.py:432-452`` with ``op.mul``: interleaves the axes of x and y."""
  if y is None:
    y = x
  xs = x.shape[:start_axis]
  xa = x.shape[start_axis:]
  ya = y.shape[start_axis:]
  xr = x.reshape(xs + tuple(v for a in xa for v in (a, 1)))
  yr = y.reshape(y.shape[:start_axis] + tuple(v for a in ya for v in (1, a)))
  return xr * yr


def reverse_zipped(x, start_axis):
  """``utils.py:465-480``: reverse the order of zipped spatial axis pairs."""
  if x is None or np.ndim(x) == 0:
    return x
  src = [j for i in range(x.ndim - 2, start_axis - 1, -2) for j in (i, i + 1)]
  return np.moveaxis(x, src, list(range(start_axis, x.ndim)))


# ------------------------------------------------------------- input kernel

def inputs_to_kernel(x1, x2, compute_ntk=True, dtype=np.float64,
                     diagonal_spatial=False):
  """``stax/requirements.py:641-830`` (NHWC, ``diagonal_batch=True``).

  ``cov`` via ``_cov_diag_batch`` (572-582) -> ``_cov_diag_batch_full_spatial``
  (529-539); ``nngp`` via ``_cov`` (585-638) -> ``_cov_full_batch_full_spatial``
  (542-553); the cast is line 794; ``ntk = zeros(())`` is line 807.
  """
  x1 = np.asarray(x1).astype(dtype)
  x2a = None if x2 is None else np.asarray(x2).astype(dtype)
  c = x1.shape[-1]
  xx2 = x1 if x2a is None else x2a

  def cov_self(x):
    if x is None:
      return None
    if diagonal_spatial:
      return np.sum(x * x, axis=-1) / c
    n = x.ndim - 2
    if n == 0:
      return np.einsum('ic,ic->i', x, x) / c
    # [N, H, W, H', W'] -> zipped [N, H, H', W, W']
    letters = 'hwdefg'[:n]
    primes = 'HWDEFG'[:n]
    g = np.einsum(f'n{letters}c,n{primes}c->n{letters}{primes}', x, x)
    return zip_axes(g, 1) / c

  def cov_cross(a, b):
    n = a.ndim - 2
    if diagonal_spatial:
      letters = 'hwdefg'[:n]
      return np.einsum(f'i{letters}c,j{letters}c->ij{letters}', a, b) / c
    if n == 0:
      return a @ b.T / c
    letters = 'hwdefg'[:n]
    primes = 'HWDEFG'[:n]
    g = np.einsum(f'i{letters}c,j{primes}c->ij{letters}{primes}', a, b)
    return zip_axes(g, 2) / c

  cov1 = cov_self(x1)
  cov2 = cov_self(x2a)
  nngp = cov_cross(x1, xx2)
  ntk = np.zeros((), dtype) if compute_ntk else None
  x1_is_x2 = x2a is None or (x1.shape == x2a.shape and
                             bool(np.all(np.abs(x1 - x2a) < 1e-12)))
  return OKernel(nngp=nngp, ntk=ntk, cov1=cov1, cov2=cov2, x1_is_x2=x1_is_x2,
                 is_gaussian=False, is_reversed=False, diagonal_batch=True,
                 diagonal_spatial=diagonal_spatial, shape1=x1.shape,
                 shape2=xx2.shape)


# ---------------------------------------------------------------- padding

def same_pads(n, window, stride):
  """XLA ``SAME`` padding (lo, hi) for one axis (``lax.padtype_to_pads``)."""
  out = -(-n // stride)
  total = max((out - 1) * stride + window - n, 0)
  return total // 2, total - total // 2


# --------------------------------------------------------------------- Conv

def _diag_conv_last2(x, f, stride, padding):
  """One depthwise pass of ``linear.py:3341-3378``: filter ``diag(1/f)``
  (line 3353) over the last two axes, i.e.
  ``out[.., p, p'] = (1/f) sum_k xpad[.., p*s + k, p'*s + k]``."""
  n = x.shape[-1]
  assert x.shape[-2] == n
  if padding in ('SAME', 'CIRCULAR'):
    lo, hi = same_pads(n, f, stride)
  elif padding == 'VALID':
    lo = hi = 0
  else:
    raise NotImplementedError(padding)
  # CIRCULAR: the SAME pads filled by wrapping (``_same_pad_for_filter_shape``
  # with mode 'wrap', ``linear.py:2977-3021``, applied at 3158-3164), then a
  # VALID pass.
  xp = np.pad(x, [(0, 0)] * (x.ndim - 2) + [(lo, hi), (lo, hi)],
              mode='wrap' if padding == 'CIRCULAR' else 'constant')
  np_ = n + lo + hi
  out_n = (np_ - f) // stride + 1
  acc = np.zeros(x.shape[:-2] + (out_n, out_n), x.dtype)
  w = x.dtype.type(1.0 / f)
  for k in range(f):
    sl = slice(k, k + stride * (out_n - 1) + 1, stride)
    acc += w * xp[..., sl, sl]
  return acc


def conv_kernel_full_spatial_shared(lhs, filter_shape, strides, padding,
                                    batch_ndim):
  """``linear.py:3115-3207`` + ``_conv_kernel_full_spatial_loop`` 3341-3378.

  Loops over spatial axis pairs from last to first, moving each pair to the
  end before its pass (3366-3376), which reverses the spatial order."""
  if lhs is None or np.ndim(lhs) == 0:
    return lhs
  for i in range(lhs.ndim - 1, batch_ndim, -2):
    spatial_i = (i - batch_ndim) // 2
    lhs = np.moveaxis(lhs, (i - 1, i), (-2, -1))
    lhs = _diag_conv_last2(lhs, filter_shape[spatial_i], strides[spatial_i],
                           padding)
  return lhs


def conv_kernel_diagonal_spatial(lhs, filter_shape, strides, padding,
                                 batch_ndim):
  """``linear.py:3381-3437``: window sum over the spatial axes divided by the
  filter size (diagonal-spatial layout ``[.., X, Y]``)."""
  if lhs is None or np.ndim(lhs) == 0:
    return lhs
  window = (1,) * batch_ndim + tuple(filter_shape)
  st = (1,) * batch_ndim + tuple(strides)
  if padding == 'CIRCULAR':                 # 3422-3426
    pads = [(0, 0)] * batch_ndim + [
        same_pads(lhs.shape[batch_ndim + a], filter_shape[a], strides[a])
        for a in range(len(filter_shape))]
    lhs = np.pad(lhs, pads, mode='wrap')
    padding = 'VALID'
  out = _reduce_window_sum(lhs, window, st, padding)
  return out / lhs.dtype.type(np.prod(filter_shape))


def conv_layer(filter_shape=(3, 3), strides=None, padding='SAME', W_std=1.0,
               b_std=None):
  """``Conv`` ``linear.py:936-991`` -> ``_Conv.kernel_fn`` 1325-1424 (ntk
  parameterization, shared weights)."""
  filter_shape = tuple(filter_shape)
  strides = tuple(strides) if strides is not None else (1,) * len(
      filter_shape)

  def kernel_fn(k: OKernel) -> OKernel:
    fs, st = filter_shape, strides
    if k.diagonal_spatial:
      conv_kernel = conv_kernel_diagonal_spatial
      is_rev = k.is_reversed
    else:
      if k.is_reversed:                       # 1343-1349
        fs, st = fs[::-1], st[::-1]
      is_rev = not k.is_reversed
      conv_kernel = conv_kernel_full_spatial_shared

    def conv_unscaled(lhs, bnd):
      return conv_kernel(lhs, fs, st, padding, bnd)

    w_var = W_std ** 2                        # Python float, as 1385
    b_var = None if b_std is None else b_std ** 2

    def affine(out):
      if out is not None:
        out = out * out.dtype.type(w_var)
        if b_var is not None:
          out = out + out.dtype.type(b_var)
      return out

    cov1 = affine(conv_unscaled(k.cov1, 1))   # 1392
    cov2 = affine(conv_unscaled(k.cov2, 1))   # 1393
    nngp = affine(conv_unscaled(k.nngp, 2))   # 1396
    ntk = k.ntk
    if ntk is not None:                       # 1397-1398
      ntk = nngp.dtype.type(w_var) * conv_unscaled(ntk, 2) + nngp
    return k.replace(cov1=cov1, nngp=nngp, cov2=cov2, ntk=ntk,
                     is_gaussian=True, is_reversed=is_rev)
  return kernel_fn


# --------------------------------------------------------------------- ReLU

def _sqrt(x):
  """``elementwise.py:1278-1280``: sqrt(max(x, 0))."""
  return np.sqrt(np.maximum(x, 0))


def _arctan2(x, y, fill_zero):
  """``elementwise.py:1293-1298``: atan2 with ``fill_zero`` where x=y=0."""
  return np.where((x == 0) & (y == 0), fill_zero, np.arctan2(x, y))


def relu_nngp_ntk(nngp, prod, ntk, a=0.0, b=1.0):
  """``ABRelu.nngp_ntk_fn`` ``elementwise.py:444-455``."""
  dt = nngp.dtype.type
  square_root = _sqrt(prod - nngp ** 2)
  angles = _arctan2(square_root, nngp, dt(math.pi / 2))
  factor = dt((a - b) ** 2 / (2 * math.pi))
  dot_sigma = dt((a ** 2 + b ** 2) / 2) - factor * angles
  nngp = factor * square_root + dot_sigma * nngp
  if ntk is not None:
    ntk = ntk * dot_sigma
  return nngp, ntk


def get_diagonal_outer_prods(cov1, cov2, diagonal_spatial):
  """``stax/requirements.py:1077-1117`` (``op.mul``, no masks)."""
  d1 = get_diagonal(cov1, diagonal_spatial)
  d2 = get_diagonal(cov2, diagonal_spatial)
  if diagonal_spatial:
    # end_axis = 1: plain batch outer products, spatial stays diagonal.
    prod12 = d1[:, None] * (d1 if d2 is None else d2)[None, :]
    return d1 * d1, prod12, (d2 * d2 if d2 is not None else d1 * d1)
  prod12 = outer_prod(d1, d2, 0)
  prod11 = outer_prod(d1, d1, 1)
  prod22 = outer_prod(d2, d2, 1) if d2 is not None else prod11
  return prod11, prod12, prod22


def abrelu_layer(a=0.0, b=1.0):
  """``ABRelu`` ``elementwise.py:405-479`` wrapped by ``_elementwise``
  1250-1275 (Gaussian check 1267-1271).  ``Relu`` is ``ABRelu(0, 1)``
  (482-491)."""
  def kernel_fn(k: OKernel) -> OKernel:
    if not k.is_gaussian:
      raise ValueError('The input to the activation function must be '
                       'Gaussian, i.e. a random affine transform is required '
                       'before the activation function.')
    prod11, prod12, prod22 = get_diagonal_outer_prods(k.cov1, k.cov2,
                                                      k.diagonal_spatial)
    nngp, ntk = relu_nngp_ntk(k.nngp, prod12, k.ntk, a, b)
    if k.diagonal_spatial:                    # 462-464
      s = k.cov1.dtype.type((a ** 2 + b ** 2) / 2)
      cov1 = s * k.cov1
      cov2 = None if k.cov2 is None else s * k.cov2
    else:                                     # 465-469
      cov1, _ = relu_nngp_ntk(k.cov1, prod11, None, a, b)
      cov2 = None if k.cov2 is None else relu_nngp_ntk(k.cov2, prod22, None,
                                                       a, b)[0]
    return k.replace(cov1=cov1, nngp=nngp, cov2=cov2, ntk=ntk,
                     is_gaussian=False)
  return kernel_fn


def relu_layer():
  return abrelu_layer(0.0, 1.0)


def _scale_kernel(k: OKernel, a: float) -> OKernel:
  """``Kernel.__mul__`` ``utils/kernel.py:426-431``: every covariance times
  a^2."""
  var = k.nngp.dtype.type(a ** 2)
  return k.replace(nngp=var * k.nngp, cov1=var * k.cov1,
                   cov2=None if k.cov2 is None else var * k.cov2,
                   ntk=None if k.ntk is None else var * k.ntk)


def _check_gaussian(k: OKernel):
  if not k.is_gaussian:
    raise ValueError('The input to the activation function must be '
                     'Gaussian, i.e. a random affine transform is required '
                     'before the activation function.')


def erf_nngp_ntk(nngp, prod, ntk):
  """``Erf.nngp_ntk_fn`` ``elementwise.py:80-95`` (prod = outer product of
  the diagonals of 1 + 2 cov)."""
  dt = nngp.dtype.type
  factor = dt(2 / math.pi)
  square_root = _sqrt(prod - 4 * nngp ** 2)
  out = factor * np.arctan2(2 * nngp, square_root)
  if ntk is not None:
    ntk = ntk * (2 * factor / square_root)
  return out, ntk


def erf_layer(a=1.0, b=1.0, c=0.0):
  """``Erf(a, b, c)`` ``elementwise.py:49-108``: ``k *= b``, the arcsine
  map, then ``a * k + c`` (``utils/kernel.py:426-439``: a^2 on every
  covariance, c^2 added to nngp and cov)."""
  def kernel_fn(k: OKernel) -> OKernel:
    _check_gaussian(k)
    k = _scale_kernel(k, b)
    dt = k.nngp.dtype.type
    d1 = 1 + 2 * k.cov1
    d2 = None if k.cov2 is None else 1 + 2 * k.cov2
    prod11, prod12, prod22 = get_diagonal_outer_prods(d1, d2,
                                                      k.diagonal_spatial)
    nngp, ntk = erf_nngp_ntk(k.nngp, prod12, k.ntk)
    if k.diagonal_spatial:                    # 97-98, 102-105
      diag = lambda x: dt(2 / math.pi) * np.arctan2(x, np.sqrt(x + 0.25))
      cov1 = diag(k.cov1)
      cov2 = None if k.cov2 is None else diag(k.cov2)
    else:
      cov1, _ = erf_nngp_ntk(k.cov1, prod11, None)
      cov2 = None if k.cov2 is None else erf_nngp_ntk(k.cov2, prod22,
                                                      None)[0]
    k = k.replace(cov1=cov1, nngp=nngp, cov2=cov2, ntk=ntk,
                  is_gaussian=False)
    k = _scale_kernel(k, a)
    cc = dt(c ** 2)
    return k.replace(nngp=k.nngp + cc, cov1=k.cov1 + cc,
                     cov2=None if k.cov2 is None else k.cov2 + cc)
  return kernel_fn


def gelu_nngp_ntk(nngp, prod, prod_plus_1, ntk):
  """``Gelu.nngp_ntk_fn`` ``elementwise.py:228-248``."""
  dt = nngp.dtype.type
  delta_squared = prod_plus_1 - nngp ** 2
  delta = _sqrt(delta_squared)
  angles = np.arctan2(nngp, delta)
  new = (nngp ** 2 + prod * delta_squared) / (prod_plus_1 * delta)
  new = new + nngp * angles
  new = new / dt(2 * math.pi) + dt(0.25) * nngp
  if ntk is not None:
    second = dt(0.25) + angles / dt(2 * math.pi)
    first = 1 / delta_squared + (1 - prod) / prod_plus_1 + 1
    first = first * nngp / delta / dt(2 * math.pi)
    ntk = ntk * (first + second)
  return new, ntk


def gelu_layer():
  """``Gelu`` ``elementwise.py:197-277``."""
  def kernel_fn(k: OKernel) -> OKernel:
    _check_gaussian(k)
    dt = k.nngp.dtype.type
    c1p = k.cov1 + 1
    c2p = None if k.cov2 is None else k.cov2 + 1
    p11p, p12p, p22p = get_diagonal_outer_prods(c1p, c2p, k.diagonal_spatial)
    p11, p12, p22 = get_diagonal_outer_prods(k.cov1, k.cov2,
                                             k.diagonal_spatial)
    nngp, ntk = gelu_nngp_ntk(k.nngp, p12, p12p, k.ntk)
    if k.diagonal_spatial:                    # 250-257, 261-264
      def diag(x):
        sq = np.sqrt(1 + 2 * x)
        out = x / ((x + 1) * sq) + np.arctan2(x, sq) / 2
        return (out / dt(math.pi) + dt(0.25)) * x
      cov1 = diag(k.cov1)
      cov2 = None if k.cov2 is None else diag(k.cov2)
    else:
      cov1, _ = gelu_nngp_ntk(k.cov1, p11, p11p, None)
      cov2 = None if k.cov2 is None else gelu_nngp_ntk(k.cov2, p22, p22p,
                                                       None)[0]
    return k.replace(cov1=cov1, nngp=nngp, cov2=cov2, ntk=ntk,
                     is_gaussian=False)
  return kernel_fn


# --------------------------------------------------------------------- Pool

def _window_sum_axis(x, axis, window, stride, padding):
  """One axis of ``lax.reduce_window(add)`` with XLA padding semantics."""
  n = x.shape[axis]
  if padding == 'SAME':
    lo, hi = same_pads(n, window, stride)
  elif padding == 'VALID':
    lo = hi = 0
  else:
    raise NotImplementedError(padding)
  pad = [(0, 0)] * x.ndim
  pad[axis] = (lo, hi)
  xp = np.pad(x, pad)
  out_n = (n + lo + hi - window) // stride + 1
  acc = None
  for k in range(window):
    idx = [slice(None)] * x.ndim
    idx[axis] = slice(k, k + stride * (out_n - 1) + 1, stride)
    v = xp[tuple(idx)]
    acc = v.copy() if acc is None else acc + v
  return acc


def _reduce_window_sum(x, window, strides, padding):
  """``lax.reduce_window(x, 0., lax.add, window, strides, padding)``; the
  box sum is separable, so it runs one axis at a time."""
  for ax, (w, s) in enumerate(zip(window, strides)):
    if w != 1 or s != 1:
      x = _window_sum_axis(x, ax, w, s, padding)
  return x


def pool_kernel(lhs, window_shape, strides, padding, normalize_edges,
                batch_ndim, pool_sum=False):
  """``_pool_kernel`` ``linear.py:3499-3559`` + ``_normalize`` 3562-3572
  (average pooling; ``pool_sum``: the SUM pooling type, no ``_normalize``,
  3551-3556)."""
  window = (1,) * batch_ndim + tuple(window_shape)
  st = (1,) * batch_ndim + tuple(strides)
  out = _reduce_window_sum(lhs, window, st, padding)
  if pool_sum:
    return out
  if padding == 'SAME' and normalize_edges:
    ones = np.ones_like(lhs)
    out = out / _reduce_window_sum(ones, window, st, padding)
  else:
    out = out / lhs.dtype.type(np.prod(window))
  return out


def avg_pool_layer(window_shape=(2, 2), strides=None, padding='VALID',
                   normalize_edges=False, pool_sum=False):
  """``AvgPool`` ``linear.py:1459-1500`` -> ``_Pool.kernel_fn`` 1634-1664
  (``pool_sum``: ``SumPool`` 1505-1539, the same kernel_fn with the SUM
  pooling type)."""
  window_shape = tuple(window_shape)
  strides = tuple(strides) if strides is not None else (1,) * len(
      window_shape)

  def kernel_fn(k: OKernel) -> OKernel:
    if k.diagonal_spatial:
      wk, sk = window_shape, strides
    else:                                     # 1642-1644
      r = -1 if k.is_reversed else 1
      wk = tuple(v for a in window_shape[::r] for v in (a, a))
      sk = tuple(v for a in strides[::r] for v in (a, a))

    def pool(mat, bnd):
      if mat is None or np.ndim(mat) == 0:
        return mat
      out = pool_kernel(mat, wk, sk, padding, normalize_edges, bnd, pool_sum)
      if k.diagonal_spatial and not pool_sum:  # 1651-1655
        out = pool_kernel_diag_normalize(mat, out, normalize_edges, padding,
                                         (1,) * bnd + tuple(strides),
                                         (1,) * bnd + tuple(window_shape))
      return out

    return k.replace(nngp=pool(k.nngp, 2), ntk=pool(k.ntk, 2),
                     cov1=pool(k.cov1, 1), cov2=pool(k.cov2, 1))
  return kernel_fn


def pool_kernel_diag_normalize(lhs, out, normalize_edges, padding, strides,
                               window):
  """Second ``_normalize`` applied on the diagonal-spatial path (1651-1655)."""
  if padding == 'SAME' and normalize_edges:
    ones = np.ones_like(lhs)
    return out / _reduce_window_sum(ones, window, strides, padding)
  return out / lhs.dtype.type(np.prod(window))


def sum_pool_layer(window_shape=(2, 2), strides=None, padding='VALID'):
  """``SumPool`` ``linear.py:1505-1539``."""
  return avg_pool_layer(window_shape, strides, padding, False, pool_sum=True)


def global_avg_pool_layer(pool_sum=False):
  """``GlobalAvgPool`` ``linear.py:1701-1721`` -> ``_GlobalPool.kernel_fn``
  1774-1801 (``mean_and_var`` without masks; ``pool_sum``:
  ``GlobalSumPool`` 1676-1696, ``jnp.sum`` at 1751)."""
  def kernel_fn(k: OKernel) -> OKernel:
    def _pool(mat, bnd):
      if mat is None or np.ndim(mat) == 0:
        return mat
      axes = tuple(range(bnd, mat.ndim))
      out = np.sum(mat, axis=axes) if pool_sum else np.mean(mat, axis=axes)
      if k.diagonal_spatial and not pool_sum:
        out = out / int(np.prod([mat.shape[a] for a in axes]))
      return out
    return k.replace(cov1=_pool(k.cov1, 1), cov2=_pool(k.cov2, 1),
                     nngp=_pool(k.nngp, 2), ntk=_pool(k.ntk, 2),
                     is_reversed=False)
  return kernel_fn


# ------------------------------------------------------------ Dense/Flatten

def _affine(mat, W_std, b_std):
  """``linear.py:2938-2971``."""
  if mat is None:
    return None
  mat = mat * mat.dtype.type(W_std ** 2)
  if b_std is not None:
    mat = mat + mat.dtype.type(b_std ** 2)
  return mat


def dense_layer(W_std=1.0, b_std=None, parameterization='ntk',
                input_width=None):
  """``Dense`` ``linear.py:762-931``, ``kernel_fn`` 902-926.  The
  ``'standard'`` parameterization (913-918) needs ``input_width`` =
  ``shape1[channel_axis] / s[0]`` of the layer's input."""
  def kernel_fn(k: OKernel) -> OKernel:
    ntk = k.ntk
    if parameterization == 'standard':
      if ntk is not None:
        ntk = k.nngp.dtype.type(input_width) * k.nngp + k.nngp.dtype.type(
            W_std ** 2) * ntk
        if b_std is not None:
          ntk = ntk + 1.0
    cov1 = _affine(k.cov1, W_std, b_std)
    nngp = _affine(k.nngp, W_std, b_std)
    cov2 = _affine(k.cov2, W_std, b_std)
    if parameterization == 'ntk' and ntk is not None:
      ntk = nngp + nngp.dtype.type(W_std ** 2) * ntk
    return k.replace(cov1=cov1, nngp=nngp, cov2=cov2, ntk=ntk,
                     is_gaussian=True)
  return kernel_fn


def flatten_layer():
  """``Flatten`` ``linear.py:1815-1906``, ``kernel_fn`` 1868-1899."""
  def kernel_fn(k: OKernel) -> OKernel:
    def trace(x, bnd):
      if x is None or np.ndim(x) == 0:
        return x
      if k.diagonal_spatial:
        return np.mean(x, axis=tuple(range(bnd, x.ndim)))
      while x.ndim > bnd:                     # 1881-1882
        x = np.trace(x, axis1=-2, axis2=-1) / x.shape[-1]
      return x
    return k.replace(cov1=trace(k.cov1, 1), cov2=trace(k.cov2, 1),
                     nngp=trace(k.nngp, 2), ntk=trace(k.ntk, 2),
                     is_gaussian=False, is_reversed=False,
                     diagonal_spatial=False)
  return kernel_fn


# ------------------------------------------------------------------- serial

def serial(*layers):
  """``combinators.py:40-68``: apply layer kernel transforms in order."""
  def kernel_fn(k: OKernel) -> OKernel:
    for f in layers:
      k = f(k)
    return k
  return kernel_fn


def kernel_fn(layers, x1, x2=None, get=('nngp', 'ntk'), dtype=np.float64,
              diagonal_spatial=False):
  """``kernel_fn(x1, x2, get)`` of ``stax/requirements.py:939-1051``.

  Returns a dict with the requested entries.  ``layers`` is a list of
  kernel transforms built by the ``*_layer`` factories above."""
  if isinstance(get, str):
    get = (get,)
  compute_ntk = get is None or 'ntk' in get
  k = inputs_to_kernel(x1, x2, compute_ntk=compute_ntk, dtype=dtype,
                       diagonal_spatial=diagonal_spatial)
  k = serial(*layers)(k)
  if get is None:
    return k
  return {g: getattr(k, g) for g in get}


def myrtle_layers(depths=(2, 1, 1), W_std=2 ** 0.5, b_std=0.0,
                  readout_W_std=1.0, readout_b_std=0.0):
  """Myrtle CNN: per stage ``depth x [Conv 3x3 SAME, Relu]`` then AvgPool
  2x2/2; GlobalAvgPool and a Dense readout (the architecture of
  ``notebooks/myrtle_kernel_with_neural_tangents.ipynb`` with GAP)."""
  layers = []
  for d in depths:
    for _ in range(d):
      layers += [conv_layer((3, 3), (1, 1), 'SAME', W_std, b_std),
                 relu_layer()]
    layers += [avg_pool_layer((2, 2), (2, 2), 'VALID')]
  layers += [global_avg_pool_layer(), dense_layer(readout_W_std,
                                                  readout_b_std)]
  return layers
