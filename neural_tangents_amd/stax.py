"""Drop-in ``stax`` surface for the analytic CNN-kernel path.

Every layer factory returns the reference's triple ``(init_fn, apply_fn,
kernel_fn)`` (``utils/typing.py:225-229``):

* ``init_fn(rng, input_shape) -> (output_shape, params)`` and
  ``apply_fn(params, inputs)`` define the finite-width network (torch, NHWC,
  ntk parameterization);
* ``kernel_fn(x1, x2=None, get=None, **kw)`` computes the infinite-width
  NNGP / NTK on the GPU through libnta.so (``_engine``).  Arguments, return
  types and errors mirror ``AnalyticKernelFn`` of ``stax/requirements.py``
  (``kernel_fn_any`` 955-1051): ``get=None`` returns a ``Kernel``, a string
  returns that tensor, a tuple returns an ``AnalyticKernel`` namedtuple.

Layers: ``Conv``, ``Relu``, ``LeakyRelu``, ``Abs``, ``ABRelu``,
``AvgPool``, ``SumPool``, ``GlobalAvgPool``, ``GlobalSumPool``, ``Dense``,
``Flatten``, ``Identity``, combined with ``serial``.  Networks outside what
the engine fuses raise ``NotImplementedError``; there is no CPU path.
"""
import collections
import dataclasses
import functools
import math
import operator
from typing import Optional, Sequence

import numpy as np
import torch
import torch.nn.functional as F

from . import _engine
from .kernel import Kernel

# ------------------------------------------------------------------ specs


@dataclasses.dataclass(frozen=True)
class _ConvSpec:
  out_chan: int
  filter_shape: tuple
  strides: tuple
  padding: str
  W_std: float
  b_std: Optional[float]
  parameterization: str = 'ntk'
  name: str = 'Conv'


@dataclasses.dataclass(frozen=True)
class _ReluSpec:
  """ABRelu(a, b) = a min(x, 0) + b max(x, 0); Relu is (0, 1)."""
  a: float = 0.0
  b: float = 1.0
  name: str = 'Relu'


@dataclasses.dataclass(frozen=True)
class _ActSpec:
  """Erf: a erf(b x) + c; Gelu (``approximate`` only changes apply_fn)."""
  kind: str
  a: float = 1.0
  b: float = 1.0
  c: float = 0.0
  approximate: bool = False
  name: str = ''


@dataclasses.dataclass(frozen=True)
class _AvgPoolSpec:
  window_shape: tuple
  strides: tuple
  padding: str
  normalize_edges: bool
  name: str = 'AvgPool'
  sum: bool = False                # SumPool: no 1 / window normalisation


@dataclasses.dataclass(frozen=True)
class _GlobalAvgPoolSpec:
  name: str = 'GlobalAvgPool'
  sum: bool = False                # GlobalSumPool


@dataclasses.dataclass(frozen=True)
class _FlattenSpec:
  name: str = 'Flatten'


@dataclasses.dataclass(frozen=True)
class _DenseSpec:
  out_dim: int
  W_std: float
  b_std: Optional[float]
  parameterization: str = 'ntk'
  s: tuple = (1, 1)
  name: str = 'Dense'


@dataclasses.dataclass(frozen=True)
class _IdentitySpec:
  name: str = 'Identity'


# ------------------------------------------------------- finite networks

def _as_generator(rng):
  if isinstance(rng, torch.Generator):
    return rng
  g = torch.Generator()
  g.manual_seed(int(rng) if rng is not None else 0)
  return g


def _same_pads(n, window, stride):
  out = -(-n // stride)
  total = max((out - 1) * stride + window - n, 0)
  return total // 2, total - total // 2


def _conv_layer_fns(spec: _ConvSpec):
  fh, fw = spec.filter_shape

  def init_fn(rng, input_shape):
    g = _as_generator(rng)
    n, h, w, c = input_shape
    W = torch.randn((fh, fw, c, spec.out_chan), generator=g,
                    dtype=torch.float64)
    b = None if spec.b_std is None else torch.randn((spec.out_chan,),
                                                    generator=g,
                                                    dtype=torch.float64)
    if spec.padding in ('SAME', 'CIRCULAR'):
      ho, wo = -(-h // spec.strides[0]), -(-w // spec.strides[1])
    else:
      ho = (h - fh) // spec.strides[0] + 1
      wo = (w - fw) // spec.strides[1] + 1
    return (n, ho, wo, spec.out_chan), (W, b)

  def apply_fn(params, inputs, **kwargs):
    W, b = params
    x = inputs.permute(0, 3, 1, 2)
    W = W.to(x.device, x.dtype)
    if spec.padding in ('SAME', 'CIRCULAR'):
      ph = _same_pads(x.shape[2], fh, spec.strides[0])
      pw = _same_pads(x.shape[3], fw, spec.strides[1])
      x = F.pad(x, (pw[0], pw[1], ph[0], ph[1]),
                mode='circular' if spec.padding == 'CIRCULAR' else 'constant')
    fan_in = fh * fw * W.shape[2]
    y = F.conv2d(x, W.permute(3, 2, 0, 1), stride=tuple(spec.strides))
    y = y * (spec.W_std / math.sqrt(fan_in))
    if b is not None:
      y = y + spec.b_std * b.to(y.device, y.dtype)[None, :, None, None]
    return y.permute(0, 2, 3, 1)
  return init_fn, apply_fn


def _dense_layer_fns(spec: _DenseSpec):
  """``Dense`` init / apply (``linear.py:846-895``): ntk parameterization
  draws N(0, 1) weights and scales the output by ``W_std / sqrt(fan_in)``;
  standard parameterization draws ``out_dim * s[1]`` outputs with weights
  N(0, W_std^2 / (fan_in / s[0])), biases N(0, b_std^2), and divides the
  product by ``sqrt(s[0])``."""
  standard = spec.parameterization == 'standard'
  s0, s1 = spec.s

  def init_fn(rng, input_shape):
    g = _as_generator(rng)
    c = input_shape[-1]
    out = spec.out_dim * s1 if standard else spec.out_dim
    W = torch.randn((c, out), generator=g, dtype=torch.float64)
    b = None if spec.b_std is None else torch.randn((out,), generator=g,
                                                    dtype=torch.float64)
    if standard:
      W = W * (spec.W_std / math.sqrt(c / s0))
      b = None if b is None else b * spec.b_std
    return tuple(input_shape[:-1]) + (out,), (W, b)

  def apply_fn(params, inputs, **kwargs):
    W, b = params
    W = W.to(inputs.device, inputs.dtype)
    if standard:
      y = inputs @ W / math.sqrt(s0)
      if b is not None:
        y = y + b.to(y.device, y.dtype)
      return y
    y = inputs @ W * (spec.W_std / math.sqrt(W.shape[0]))
    if b is not None:
      y = y + spec.b_std * b.to(y.device, y.dtype)
    return y
  return init_fn, apply_fn


def _relu_layer_fns(spec):
  a, b = spec.a, spec.b

  def apply_fn(p, x, **kw):
    return a * torch.clamp(x, max=0) + b * torch.clamp(x, min=0)
  return (lambda rng, s: (tuple(s), ())), apply_fn


def _act_layer_fns(spec):
  if spec.kind == 'erf':
    def apply_fn(p, x, **kw):
      return spec.a * torch.erf(spec.b * x) + spec.c
  else:
    def apply_fn(p, x, **kw):
      return F.gelu(x, approximate='tanh' if spec.approximate else 'none')
  return (lambda rng, s: (tuple(s), ())), apply_fn


def _avg_pool_layer_fns(spec: _AvgPoolSpec):
  wh, ww = spec.window_shape
  sh, sw = spec.strides

  def out_hw(h, w):
    if spec.padding == 'SAME':
      return -(-h // sh), -(-w // sw)
    return (h - wh) // sh + 1, (w - ww) // sw + 1

  def init_fn(rng, input_shape):
    n, h, w, c = input_shape
    return (n,) + out_hw(h, w) + (c,), ()

  def apply_fn(params, inputs, **kwargs):
    x = inputs.permute(0, 3, 1, 2)
    if spec.padding == 'SAME':
      ph = _same_pads(x.shape[2], wh, sh)
      pw = _same_pads(x.shape[3], ww, sw)
      ones = torch.ones_like(x[:1, :1])
      x = F.pad(x, (pw[0], pw[1], ph[0], ph[1]))
      y = F.avg_pool2d(x, (wh, ww), (sh, sw))
      if spec.normalize_edges:
        cnt = F.avg_pool2d(F.pad(ones, (pw[0], pw[1], ph[0], ph[1])),
                           (wh, ww), (sh, sw))
        y = y / cnt
    else:
      y = F.avg_pool2d(x, (wh, ww), (sh, sw))
    if spec.sum:
      y = y * (wh * ww)
    return y.permute(0, 2, 3, 1)
  return init_fn, apply_fn


def _gap_layer_fns(spec):
  def init_fn(rng, input_shape):
    return (input_shape[0], input_shape[-1]), ()

  def apply_fn(p, x, **kw):
    axes = tuple(range(1, x.ndim - 1))
    return x.sum(dim=axes) if spec.sum else x.mean(dim=axes)
  return init_fn, apply_fn


def _flatten_layer_fns(spec):
  def init_fn(rng, input_shape):
    return (input_shape[0], int(np.prod(input_shape[1:]))), ()
  return init_fn, (lambda p, x, **kw: x.reshape(x.shape[0], -1))


def _identity_layer_fns(spec):
  return (lambda rng, s: (tuple(s), ())), (lambda p, x, **kw: x)


_FNS = {
    _ConvSpec: _conv_layer_fns,
    _DenseSpec: _dense_layer_fns,
    _ReluSpec: _relu_layer_fns,
    _ActSpec: _act_layer_fns,
    _AvgPoolSpec: _avg_pool_layer_fns,
    _GlobalAvgPoolSpec: _gap_layer_fns,
    _FlattenSpec: _flatten_layer_fns,
    _IdentitySpec: _identity_layer_fns,
}


# ----------------------------------------------------------- kernel_fn

_NT_CACHE = {}


def _namedtuple(name, fields):
  """Cached namedtuple types (``utils.py:161-167``)."""
  key = (name, fields)
  if key not in _NT_CACHE:
    _NT_CACHE[key] = collections.namedtuple(name, fields)
  return _NT_CACHE[key]


def _canonicalize_get(get):
  """``utils.py:139-156`` (``canonicalize_get``): a string returns the bare
  value, a tuple (of any length) a namedtuple; entries are lower-cased and
  must be unique and non-empty."""
  if get is None:
    return True, ('nngp', 'ntk', 'cov1', 'cov2')
  if not get:
    raise ValueError('"get" must be non-empty.')
  if isinstance(get, str):
    return False, (get.lower(),)
  get = tuple(g.lower() for g in get)
  if len(set(get)) < len(get):
    raise ValueError(f'All entries in "get" must be unique. Got {get}')
  return False, get


def _to_device_input(x, name):
  if x is None:
    return None
  if isinstance(x, np.ndarray):
    x = torch.from_numpy(np.ascontiguousarray(x))
  if not isinstance(x, torch.Tensor):
    raise TypeError(f'Wrong input types given. Found `{name}` of type '
                    f'{type(x)}, need a torch.Tensor or numpy.ndarray.')
  if x.ndim < 2:
    raise ValueError('Inputs must be at least 2D (a batch dimension and a '
                     f'channel/feature dimension), got {x.ndim}.')
  if not torch.cuda.is_available():
    raise RuntimeError('kernel_fn runs on the GPU through libnta.so and no '
                       'GPU is visible.')
  dev = x.device if x.is_cuda else torch.device('cuda', 0)
  return x.to(device=dev, dtype=torch.float32).contiguous()


class AnalyticKernelFn:
  """``kernel_fn(x1, x2=None, get=None, **kw)`` of a ``serial`` network."""

  def __init__(self, specs, init_fn=None):
    self.specs = tuple(specs)
    self._init_fn = init_fn
    self._plans = {}
    self.__name__ = 'kernel_fn'

  def plan(self, in_shape, from_state=False, in_gaussian=False):
    if len(in_shape) == 2:
      self._check_flat()
      in_shape = (in_shape[0], 1, 1, in_shape[1])
    key = tuple(in_shape[1:]) + (bool(from_state), bool(in_gaussian))
    if key not in self._plans:
      self._plans[key] = _engine.compile_plan(
          self.specs, (1,) + tuple(in_shape[1:]), from_state=from_state,
          in_gaussian=in_gaussian)
    return self._plans[key]

  def _check_flat(self):
    for spec in self.specs:
      if isinstance(spec, (_ConvSpec, _AvgPoolSpec)):
        raise ValueError(f'{spec.name} needs spatial inputs [N, H, W, C]; '
                         'got flat inputs [N, D].')

  def __call__(self, x1_or_kernel, x2=None, get=None, *, pattern=None,
               mask_constant=None, diagonal_batch=None,
               diagonal_spatial=None, tile_bytes=None, **kwargs):
    if isinstance(x1_or_kernel, Kernel):
      if x2 is not None:
        raise ValueError('`x2` must be None when `kernel_fn` is applied to '
                         'a `Kernel`.')
      return self._apply_to_kernel(x1_or_kernel, get, tile_bytes)
    if pattern is not None or mask_constant is not None:
      raise NotImplementedError('pattern / mask_constant are not supported '
                                'by the HIP engine.')
    if diagonal_batch is False:
      raise NotImplementedError('diagonal_batch=False is not supported by '
                                'the HIP engine.')
    return_kernel, get_t = _canonicalize_get(get)
    for g in get_t:
      if g not in Kernel.__dataclass_fields__:
        raise ValueError(f'Unknown `get` entry {g!r}.')
    x1 = _to_device_input(x1_or_kernel, 'x1')
    x2t = _to_device_input(x2, 'x2')
    if x2t is not None and tuple(x2t.shape[1:]) != tuple(x1.shape[1:]):
      raise ValueError(f'x1 and x2 must agree apart from the batch axis, got '
                       f'{tuple(x1.shape)} and {tuple(x2t.shape)}.')
    flat = x1.ndim == 2
    shape1_in = tuple(x1.shape)
    shape2_in = tuple((x1 if x2t is None else x2t).shape)
    if flat:
      # [N, D] -> [N, 1, 1, D]: Dense layers run as 1x1 stages.
      x1 = x1.reshape(x1.shape[0], 1, 1, x1.shape[1])
      if x2t is not None:
        x2t = x2t.reshape(x2t.shape[0], 1, 1, x2t.shape[1])
    plan = self.plan(shape1_in)
    if diagonal_spatial and not plan.diag:
      # The diagonal-spatial path runs when the network allows it (a Flatten
      # readout and no pooling); other networks need the full spatial
      # kernel.
      raise NotImplementedError('diagonal_spatial=True needs a Flatten '
                                'readout and no pooling layers.')
    want_nngp = True
    want_ntk = 'ntk' in get_t
    want_cov = return_kernel or 'cov1' in get_t or 'cov2' in get_t
    res = _engine.run(plan, x1, x2t, want_nngp, want_ntk, want_cov,
                      tile_bytes=tile_bytes or _engine.DEFAULT_TILE_BYTES)
    is_rev = plan.head is None and plan.n_convs % 2 == 1

    def order(t, bnd):
      if t is None:
        return t
      if flat and plan.head is None:
        return t.reshape(t.shape[:bnd])
      if not is_rev:
        return t
      return _swap_spatial_pairs(t, bnd)

    if x2t is None:
      x1_is_x2 = True
    elif x1.shape != x2t.shape:
      x1_is_x2 = False
    else:
      # a 0-d device tensor, like the reference's jnp bool (utils.py:255-282);
      # evaluated without a host synchronisation.
      x1_is_x2 = torch.all(torch.abs(x1 - x2t) < 1e-12)
    shape1 = self._out_shape(shape1_in)
    shape2 = self._out_shape(shape2_in)
    k = Kernel(nngp=order(res['nngp'], 2), ntk=order(res['ntk'], 2),
               cov1=order(res['cov1'], 1), cov2=order(res['cov2'], 1),
               x1_is_x2=x1_is_x2, is_gaussian=plan.is_gaussian,
               is_reversed=is_rev, is_input=False, diagonal_batch=True,
               diagonal_spatial=False, shape1=shape1, shape2=shape2,
               batch_axis=0, channel_axis=len(shape1) - 1)
    return self._format(k, get, return_kernel, get_t)

  @staticmethod
  def _format(k, get, return_kernel, get_t):
    if return_kernel:
      return k
    if isinstance(get, str):
      return getattr(k, get_t[0])
    return _namedtuple('AnalyticKernel', get_t)(
        *[getattr(k, g) for g in get_t])

  def _apply_to_kernel(self, k: Kernel, get, tile_bytes):
    """``kernel_fn(kernel)``: the network applied to a given ``Kernel``
    (``kernel_fn_kernel``, ``stax/requirements.py:955-975``), e.g. the output
    of another ``kernel_fn`` with spatial axes.  The state is taken in the
    reference's layout (``[N1, N2, X, X', Y, Y']`` or ``[N1, N2]``, reversed
    when ``is_reversed``) and the result is laid out the same way."""
    return_kernel, get_t = _canonicalize_get(get)
    for g in get_t:
      if g not in Kernel.__dataclass_fields__:
        raise ValueError(f'Unknown `get` entry {g!r}.')
    if k.diagonal_spatial:
      raise NotImplementedError('Kernels with diagonal_spatial=True are not '
                                'supported as kernel_fn inputs.')
    if not k.diagonal_batch:
      raise NotImplementedError('Kernels with diagonal_batch=False are not '
                                'supported as kernel_fn inputs.')

    def dev(t):
      if t is None:
        return None
      if isinstance(t, np.ndarray):
        t = torch.from_numpy(np.ascontiguousarray(t))
      if not torch.cuda.is_available():
        raise RuntimeError('kernel_fn runs on the GPU through libnta.so and '
                           'no GPU is visible.')
      d = t.device if t.is_cuda else torch.device('cuda', 0)
      return t.to(device=d, dtype=torch.float32)

    nngp = dev(k.nngp)
    if nngp.ndim == 2:
      flat = True
      N1, N2 = nngp.shape
      H = W = 1
    elif nngp.ndim == 6:
      flat = False
      N1, N2 = nngp.shape[:2]
    else:
      raise NotImplementedError('kernel_fn(Kernel) takes kernels with two '
                                f'spatial axes or none; got nngp of shape '
                                f'{tuple(nngp.shape)}.')
    if flat:
      self._check_flat()

    def natural(t, bnd):
      if t is None:
        return None
      if flat:
        return t.reshape(tuple(t.shape[:bnd]) + (1, 1, 1, 1)).contiguous()
      t = _swap_spatial_pairs(t, bnd) if k.is_reversed else t
      return t.contiguous()

    nngp = natural(nngp, 2)
    if not flat:
      H, W = nngp.shape[2], nngp.shape[4]
    ntk_in = dev(k.ntk) if isinstance(k.ntk, torch.Tensor) or isinstance(
        k.ntk, np.ndarray) else k.ntk
    want_ntk = 'ntk' in get_t and ntk_in is not None
    if isinstance(ntk_in, (int, float)) or (
        isinstance(ntk_in, torch.Tensor) and ntk_in.ndim == 0):
      if float(ntk_in) != 0.0:
        raise NotImplementedError('a scalar ntk other than 0 is not '
                                  'supported as kernel_fn input.')
      ntk_state = None                         # zero
    else:
      ntk_state = natural(ntk_in, 2) if want_ntk else None
    cov1 = natural(dev(k.cov1), 1)
    cov2 = natural(dev(k.cov2), 1)
    C = k.shape1[-1] if k.shape1 else 1
    plan = self.plan((1, H, W, C), from_state=True,
                     in_gaussian=bool(k.is_gaussian))
    want_cov = return_kernel or 'cov1' in get_t or 'cov2' in get_t
    res = _engine.run_state(
        plan, nngp, ntk_state, cov1, cov2, want_ntk, want_cov,
        tile_bytes=tile_bytes or _engine.DEFAULT_TILE_BYTES)
    is_rev = plan.head is None and (bool(k.is_reversed) !=
                                    (plan.n_convs % 2 == 1))

    def order(t, bnd):
      if t is None:
        return t
      if flat and plan.head is None:
        return t.reshape(t.shape[:bnd])
      return _swap_spatial_pairs(t, bnd) if is_rev else t

    out = Kernel(nngp=order(res['nngp'], 2),
                 ntk=order(res['ntk'], 2) if want_ntk else None,
                 cov1=order(res['cov1'], 1), cov2=order(res['cov2'], 1),
                 x1_is_x2=k.x1_is_x2, is_gaussian=plan.is_gaussian,
                 is_reversed=is_rev, is_input=False, diagonal_batch=True,
                 diagonal_spatial=False,
                 shape1=self._out_shape(tuple(k.shape1)),
                 shape2=self._out_shape(tuple(k.shape2)),
                 batch_axis=0, channel_axis=len(self._out_shape(
                     tuple(k.shape1))) - 1)
    return self._format(out, get, return_kernel, get_t)

  def _out_shape(self, in_shape):
    if self._init_fn is None:
      return in_shape
    out_shape, _ = self._init_fn(0, in_shape)
    return tuple(out_shape)


def _swap_spatial_pairs(t, bnd):
  """[.., X, X', Y, Y'] <-> [.., Y, Y', X, X'] after ``bnd`` batch axes: the
  reference's axis reversal of a 2-D zipped kernel (``utils.py:465-480``),
  its own inverse."""
  return t.permute(tuple(range(bnd)) + (bnd + 2, bnd + 3, bnd, bnd + 1))


def _layer(spec):
  init_fn, apply_fn = _FNS[type(spec)](spec)
  kernel_fn = AnalyticKernelFn([spec], init_fn)
  init_fn.__name__ = apply_fn.__name__ = spec.name
  return init_fn, apply_fn, kernel_fn


# ---------------------------------------------------------------- layers

def Conv(out_chan: int, filter_shape: Sequence[int],
         strides: Optional[Sequence[int]] = None, padding: str = 'VALID',
         W_std: float = 1.0, b_std: Optional[float] = None,
         dimension_numbers=None, parameterization: str = 'ntk',
         s=(1, 1)):
  """``Conv`` (``linear.py:936-991``)."""
  if dimension_numbers not in (None, ('NHWC', 'HWIO', 'NHWC')):
    raise NotImplementedError('Only NHWC/HWIO/NHWC dimension numbers.')
  filter_shape = tuple(int(f) for f in filter_shape)
  strides = tuple(strides) if strides is not None else (1,) * len(
      filter_shape)
  padding = padding.upper()
  if padding not in ('SAME', 'VALID', 'CIRCULAR'):
    raise ValueError(f'Unknown padding {padding!r}.')
  parameterization = parameterization.lower()
  if parameterization not in ('ntk', 'standard'):
    raise ValueError(f'Parameterization not supported: {parameterization}')
  return _layer(_ConvSpec(out_chan, filter_shape, strides, padding,
                          float(W_std), None if b_std is None else
                          float(b_std), parameterization))


def Relu(do_stabilize: bool = False):
  """``Relu`` (``elementwise.py:482-491``).  ``do_stabilize`` rescales the
  reference's intermediate kernels only; the map is positively homogeneous,
  so the result is the same and the flag is accepted and ignored."""
  return _layer(_ReluSpec())


def ABRelu(a: float, b: float, do_stabilize: bool = False):
  """``ABRelu`` (``elementwise.py:405-479``): ``a min(x, 0) + b max(x, 0)``;
  on kernels ``(a - b)^2 Relu + a b Id``."""
  return _layer(_ReluSpec(float(a), float(b), name=f'ABReLU({a}, {b})'))


def LeakyRelu(alpha: float, do_stabilize: bool = False):
  """``LeakyRelu`` (``elementwise.py:494-506``) = ``ABRelu(alpha, 1)``."""
  return _layer(_ReluSpec(float(alpha), 1.0, name=f'LeakyRelu({alpha})'))


def Abs(do_stabilize: bool = False):
  """``Abs`` (``elementwise.py:509-520``) = ``ABRelu(-1, 1)``."""
  return _layer(_ReluSpec(-1.0, 1.0, name='Abs'))


def Erf(a: float = 1., b: float = 1., c: float = 0.):
  """``Erf`` (``elementwise.py:49-108``): ``a erf(b x) + c``; on kernels the
  arcsine map of ``b^2 k`` scaled by ``a^2`` plus ``c^2``."""
  return _layer(_ActSpec('erf', float(a), float(b), float(c),
                         name=f'Erf({a}, {b}, {c})'))


def Gelu(approximate: bool = False):
  """``Gelu`` (``elementwise.py:197-277``); ``approximate`` only affects the
  finite-width ``apply_fn``, as in the reference."""
  return _layer(_ActSpec('gelu', approximate=bool(approximate), name='Gelu'))


def AvgPool(window_shape: Sequence[int],
            strides: Optional[Sequence[int]] = None, padding: str = 'VALID',
            normalize_edges: bool = False, batch_axis: int = 0,
            channel_axis: int = -1):
  """``AvgPool`` (``linear.py:1459-1500``)."""
  if batch_axis != 0 or channel_axis not in (-1, 3):
    raise NotImplementedError('Only NHWC pooling is supported.')
  window_shape = tuple(int(w) for w in window_shape)
  strides = tuple(strides) if strides is not None else (1,) * len(
      window_shape)
  return _layer(_AvgPoolSpec(window_shape, strides, padding.upper(),
                             bool(normalize_edges)))


def SumPool(window_shape: Sequence[int],
            strides: Optional[Sequence[int]] = None, padding: str = 'VALID',
            batch_axis: int = 0, channel_axis: int = -1):
  """``SumPool`` (``linear.py:1505-1539``): on kernels, AvgPool times the
  square of the window size (``_pool_kernel`` without ``_normalize``,
  3499-3572)."""
  if batch_axis != 0 or channel_axis not in (-1, 3):
    raise NotImplementedError('Only NHWC pooling is supported.')
  window_shape = tuple(int(w) for w in window_shape)
  strides = tuple(strides) if strides is not None else (1,) * len(
      window_shape)
  return _layer(_AvgPoolSpec(window_shape, strides, padding.upper(), False,
                             name='SumPool', sum=True))


def GlobalAvgPool(batch_axis: int = 0, channel_axis: int = -1):
  """``GlobalAvgPool`` (``linear.py:1701-1721``)."""
  if batch_axis != 0 or channel_axis not in (-1, 3):
    raise NotImplementedError('Only NHWC global pooling is supported.')
  return _layer(_GlobalAvgPoolSpec())


def GlobalSumPool(batch_axis: int = 0, channel_axis: int = -1):
  """``GlobalSumPool`` (``linear.py:1676-1696``): the spatial sum, i.e.
  GlobalAvgPool times the number of (h, h', w, w') entries."""
  if batch_axis != 0 or channel_axis not in (-1, 3):
    raise NotImplementedError('Only NHWC global pooling is supported.')
  return _layer(_GlobalAvgPoolSpec(name='GlobalSumPool', sum=True))


def Dense(out_dim: int, W_std: float = 1., b_std: Optional[float] = None,
          batch_axis: int = 0, channel_axis: int = -1,
          parameterization: str = 'ntk', s=(1, 1)):
  """``Dense`` (``linear.py:762-931``); ``parameterization`` is ``'ntk'``
  or ``'standard'`` (``s``: the width scales of the standard form)."""
  parameterization = parameterization.lower()
  if parameterization not in ('ntk', 'standard'):
    raise ValueError(f'Parameterization not supported: {parameterization}')
  return _layer(_DenseSpec(int(out_dim), float(W_std),
                           None if b_std is None else float(b_std),
                           parameterization, tuple(int(v) for v in s)))


def Flatten(batch_axis: int = 0, batch_axis_out: int = 0):
  """``Flatten`` (``linear.py:1815-1906``)."""
  if batch_axis != 0 or batch_axis_out != 0:
    raise NotImplementedError('Only batch_axis=0 Flatten is supported.')
  return _layer(_FlattenSpec())


def Identity():
  """``Identity`` (``linear.py``)."""
  return _layer(_IdentitySpec())


def serial(*layers):
  """``serial`` (``combinators.py:40-68``): composes layer triples."""
  init_fns, apply_fns, kernel_fns = zip(*layers)
  specs = []
  for kf in kernel_fns:
    if not isinstance(kf, AnalyticKernelFn):
      raise NotImplementedError(f'{kf!r} is not a layer of this package.')
    specs.extend(kf.specs)

  def init_fn(rng, input_shape):
    g = _as_generator(rng)
    params = []
    shape = tuple(input_shape)
    for f in init_fns:
      shape, p = f(g, shape)
      params.append(p)
    return shape, tuple(params)

  def apply_fn(params, inputs, **kwargs):
    for f, p in zip(apply_fns, params):
      inputs = f(p, inputs, **kwargs)
    return inputs

  return init_fn, apply_fn, AnalyticKernelFn(specs, init_fn)


def myrtle(depth: int = 5, W_std: float = 2 ** 0.5, b_std: float = 0.,
           readout_W_std: float = 1., readout_b_std: float = 0.,
           depths: Optional[Sequence[int]] = None):
  """Myrtle CNN with GlobalAvgPool readout: per stage ``d x [Conv 3x3 SAME,
  Relu]`` then ``AvgPool 2x2/2``; depths 5: [2,1,1], 7: [2,2,2],
  10: [3,3,3] (``notebooks/myrtle_kernel_with_neural_tangents.ipynb``)."""
  if depths is None:
    depths = {5: [2, 1, 1], 7: [2, 2, 2], 10: [3, 3, 3]}[depth]
  layers = []
  for d in depths:
    layers += [Conv(1, (3, 3), padding='SAME', W_std=W_std, b_std=b_std),
               Relu()] * d
    layers += [AvgPool((2, 2), strides=(2, 2))]
  layers += [GlobalAvgPool(), Dense(1, readout_W_std, readout_b_std)]
  return serial(*layers)
