"""Schedule compiler and executor for ``kernel_fn`` on the HIP kernels.

A ``serial`` network is compiled into fused *stages*, each one launch per
tile of pairs:

* the first stage fuses the input covariance with the first Conv (the
  stencil of the input Gram is the Gram of zero-padded patches) and, when a
  Relu follows, the arc-cosine map;
* every later stage is ``[Conv] -> [Relu] -> [AvgPool]`` in one kernel;
* the head is ``GlobalAvgPool`` or ``Flatten`` followed by the trailing Dense
  layers folded into one affine map.

Per-image self chains (``x1`` and ``x2`` against themselves, nngp only) run
first and leave the ReLU normalisers ``q[n, h, w]`` of every stage; pair
tiles then run the cross chain.  All arithmetic happens in libnta.so; this
module only plans, allocates (torch caching allocator) and launches.

Reference call path replaced: ``_preprocess_kernel_fn`` /
``_inputs_to_kernel`` (``stax/requirements.py:641-830, 912-1054``), the
``serial`` loop (``stax/combinators.py:40-68``) and the per-layer
``kernel_fn``s listed in each stage; pair tiling follows ``nt.batch``
(``batching.py:314-503``).
"""
import dataclasses
import math
from typing import List, Optional, Tuple

import torch

from . import _lib


@dataclasses.dataclass
class Stage:
  conv: Optional[tuple] = None     # (fh, fw, valid, w_var, b_var)
  relu: bool = False
  act: Tuple[float, float] = (1.0, 0.0)   # ABRelu (k1, k0) = ((a-b)^2, ab)
  act_kind: int = 0                # 0 ABRelu, _lib.ACT_ERF, _lib.ACT_GELU
  circular: bool = False           # CIRCULAR (periodic) conv padding
  wscale: float = 1.0              # SumPool factor folded into conv W_std^2
  pool: Optional[tuple] = None     # (ph, pw)
  first: bool = False
  # sizes, filled by the planner
  hin: int = 0
  win: int = 0
  hc: int = 0
  wc: int = 0
  hout: int = 0
  wout: int = 0

  def flags(self, ntk: bool) -> int:
    f = 0
    if self.conv is not None:
      f |= _lib.CONV
      if self.conv[2]:
        f |= _lib.CONV_VALID
      if self.circular:
        f |= _lib.CONV_CIRCULAR
    if self.relu:
      f |= _lib.RELU | self.act_kind
    if ntk:
      f |= _lib.NTK
    return f


@dataclasses.dataclass
class Plan:
  stages: List[Stage]
  head: Optional[str]                   # None | 'gap' | 'flatten'
  head_coef: Tuple[float, ...]          # (a, b, c, d, e)
  n_convs: int
  is_gaussian: bool
  out_channels: Optional[int]
  out_spatial: Tuple[int, int]
  # Flatten readout and no pooling: only the spatial diagonal is needed
  # (the reference's diagonal_spatial requirement, requirements.py:912-1054).
  diag: bool = False
  # Factor applied to head-less outputs (SumPool scales left after the last
  # layer that could absorb them).
  out_scale: float = 1.0


class _Unsupported(NotImplementedError):
  pass


def compile_plan(specs, in_shape, from_state: bool = False,
                 in_gaussian: bool = False) -> Plan:
  """Groups layer specs into fused stages; validates shapes and options.

  ``in_shape`` is NHWC ``[N, H, W, C]`` (a flat ``[N, D]`` input is passed as
  ``[N, 1, 1, D]``).  ``from_state``: the network is applied to a given
  kernel (``kernel_fn(Kernel)``, ``stax/requirements.py:955-975``) whose
  full-spatial state has spatial size ``H x W``, instead of to inputs; then
  no stage computes an input covariance and ``in_gaussian`` is the input
  kernel's ``is_gaussian``.

  ``Dense`` before a readout acts on the channel axis at every spatial
  position, so on kernels it is a 1x1 Conv without the axis reversal
  (``linear.py:902-926`` is the same affine map as ``linear.py:1387-1398``
  with a 1x1 filter)."""
  from . import stax as S
  if len(in_shape) != 4:
    raise NotImplementedError(
        'The HIP engine runs NHWC image inputs [N, H, W, C] or flat inputs '
        f'[N, D]; got input shape {tuple(in_shape)}.')
  _, H, W, C = in_shape
  stages: List[Stage] = []
  cur: Optional[Stage] = None
  head = None
  dense = []
  is_gaussian = bool(in_gaussian) if from_state else False
  n_convs = 0
  n_affine = 0
  chans = C
  # SumPool = AvgPool x window^2 on kernels.  Every layer here is linear or
  # positively homogeneous of degree 1 in (nngp, ntk, cov) jointly (the
  # ABRelu map is), so the factor is carried forward and folded into the
  # next Conv's W_std^2 (which scales before b_std^2 is added), into the
  # readout coefficients, or into the outputs.
  pend = 1.0
  gsum = False

  def close():
    nonlocal cur
    if cur is not None:
      stages.append(cur)
      cur = None

  for spec in specs:
    if head is not None:
      if isinstance(spec, S._DenseSpec):
        dense.append((spec, chans))
        is_gaussian = True
        chans = spec.out_dim * (spec.s[1] if spec.parameterization ==
                                'standard' else 1)
        continue
      if isinstance(spec, S._IdentitySpec):
        continue
      raise NotImplementedError(
          f'{spec.name} after a readout is not supported by the HIP engine; '
          'only Dense layers may follow GlobalAvgPool/Flatten.')
    if isinstance(spec, S._IdentitySpec):
      continue
    if isinstance(spec, S._ConvSpec):
      if any(s != 1 for s in spec.strides):
        raise NotImplementedError('Conv strides other than 1 are not '
                                  'supported by the HIP engine.')
      if spec.padding not in ('SAME', 'VALID', 'CIRCULAR'):
        raise NotImplementedError(f'Conv padding {spec.padding!r} is not '
                                  'supported by the HIP engine.')
      if len(spec.filter_shape) != 2:
        raise NotImplementedError('Only 2-D convolutions are supported.')
      if spec.parameterization != 'ntk':
        raise NotImplementedError('Only the ntk parameterization is '
                                  'supported by the HIP engine.')
      close()
      fh, fw = spec.filter_shape
      w_var = float(spec.W_std) ** 2 * pend
      pend = 1.0
      b_var = 0.0 if spec.b_std is None else float(spec.b_std) ** 2
      cur = Stage(conv=(fh, fw, spec.padding == 'VALID', w_var, b_var),
                  circular=spec.padding == 'CIRCULAR',
                  first=not from_state and not stages and n_affine == 0,
                  wscale=w_var / float(spec.W_std) ** 2
                  if spec.W_std != 0 else 1.0)
      if not from_state and not cur.first and not stages:
        raise NotImplementedError('The network must start with a Conv.')
      n_convs += 1
      n_affine += 1
      is_gaussian = True
      chans = spec.out_chan
      continue
    if isinstance(spec, S._DenseSpec):
      if spec.parameterization != 'ntk':
        raise NotImplementedError('Dense before a readout runs with the ntk '
                                  'parameterization only.')
      close()
      w_var = float(spec.W_std) ** 2 * pend
      pend = 1.0
      b_var = 0.0 if spec.b_std is None else float(spec.b_std) ** 2
      cur = Stage(conv=(1, 1, False, w_var, b_var),
                  first=not from_state and not stages and n_affine == 0,
                  wscale=w_var / float(spec.W_std) ** 2
                  if spec.W_std != 0 else 1.0)
      n_affine += 1
      is_gaussian = True
      chans = spec.out_dim
      continue
    if isinstance(spec, (S._ReluSpec, S._ActSpec)) and not is_gaussian:
      raise ValueError('The input to the activation function must be '
                       'Gaussian, i.e. a random affine transform is '
                       'required before the activation function.')
    if not from_state and not stages and cur is None:
      raise NotImplementedError(
          'The HIP engine needs a Conv or Dense as the first layer, got '
          f'{spec.name}.')
    if isinstance(spec, S._ReluSpec):
      if cur is None or cur.relu or cur.pool is not None:
        close()
        cur = Stage()
      cur.relu = True
      cur.act = ((spec.a - spec.b) ** 2, spec.a * spec.b)
      is_gaussian = False
      continue
    if isinstance(spec, S._ActSpec):
      if cur is None or cur.relu or cur.pool is not None:
        close()
        cur = Stage()
      cur.relu = True
      if spec.kind == 'erf':
        # a Erf(b x) + c: the input scale b multiplies the pre-activation
        # (nngp, ntk, cov) by b^2 (elementwise.py:67-68, kernel.py:426-431),
        # i.e. the stage's conv W_std^2 and b_std^2.
        if spec.b != 1.0:
          if cur.conv is None:
            raise NotImplementedError('Erf with an input scale b != 1 needs '
                                      'a Conv right before it.')
          fh, fw, valid, wv, bv = cur.conv
          b2 = float(spec.b) ** 2
          cur.conv = (fh, fw, valid, wv * b2, bv * b2)
          cur.wscale *= b2
        cur.act_kind = _lib.ACT_ERF
        cur.act = (float(spec.a) ** 2, float(spec.c) ** 2)
      else:
        cur.act_kind = _lib.ACT_GELU
        cur.act = (0.0, 0.0)
      is_gaussian = False
      continue
    if isinstance(spec, S._AvgPoolSpec):
      win = tuple(spec.window_shape)
      st = tuple(spec.strides)
      if len(win) != 2 or win != st:
        raise NotImplementedError('AvgPool is supported with a 2-D window '
                                  'equal to its strides.')
      if spec.padding not in ('VALID', 'SAME'):
        raise NotImplementedError(f'AvgPool padding {spec.padding!r}.')
      if cur is None or cur.pool is not None:
        close()
        cur = Stage()
      cur.pool = (spec.padding,) + win
      if spec.sum:
        pend *= float(win[0] * win[1]) ** 2
      continue
    if isinstance(spec, (S._GlobalAvgPoolSpec, S._FlattenSpec)):
      close()
      head = 'gap' if isinstance(spec, S._GlobalAvgPoolSpec) else 'flatten'
      gsum = head == 'gap' and spec.sum
      if isinstance(spec, S._FlattenSpec):
        is_gaussian = False
      continue
    raise NotImplementedError(f'Layer {spec.name} is not supported by the '
                              'HIP engine.')
  close()
  if not stages and not from_state:
    raise NotImplementedError('The network has no Conv or Dense layer.')

  # The first stage carries only Conv (+ Relu); its pool becomes a stage.
  if stages and stages[0].first and stages[0].pool is not None:
    p = stages[0].pool
    stages[0].pool = None
    stages.insert(1, Stage(pool=p))

  # Sizes.
  h, w = H, W
  for s in stages:
    s.hin, s.win = h, w
    if s.conv is not None:
      fh, fw, valid = s.conv[:3]
      if valid:
        if fh > h or fw > w:
          raise ValueError(f'VALID {fh}x{fw} filter on a {h}x{w} kernel.')
        h, w = h - fh + 1, w - fw + 1
    s.hc, s.wc = h, w
    if s.pool is not None:
      pad, ph, pw = s.pool
      if pad == 'SAME' and (h % ph or w % pw):
        raise NotImplementedError('AvgPool SAME with a window that does not '
                                  'divide the spatial size.')
      if ph > h or pw > w:
        raise ValueError(f'AvgPool window {ph}x{pw} larger than {h}x{w}.')
      h, w = h // ph, w // pw
      s.pool = (ph, pw)
    s.hout, s.wout = h, w

  # Dense layers after the readout fold into nngp = a m + b and
  # ntk = c m + d mt + e (linear.py:902-926 applied in sequence).
  a, b, c, d, e = 1.0, 0.0, 1.0, 0.0, 0.0   # nngp = m, ntk = mt
  ntk_m, ntk_t, ntk_1 = 0.0, 1.0, 0.0
  if head is not None:
    # m is the mean over the readout's entries; GlobalSumPool sums them.
    a = pend * (float(h * h * w * w) if gsum else 1.0)
    ntk_t = a
    pend = 1.0
  for spec, fan_in in dense:
    wv = float(spec.W_std) ** 2
    bv = 0.0 if spec.b_std is None else float(spec.b_std) ** 2
    if spec.parameterization == 'standard':
      # ntk' = (fan_in / s0) nngp + W^2 ntk (+ 1 with a bias), from the
      # layer's input nngp (linear.py:913-918)
      iw = fan_in / spec.s[0]
      ntk_m, ntk_t = iw * a + wv * ntk_m, wv * ntk_t
      ntk_1 = iw * b + wv * ntk_1 + (1.0 if spec.b_std is not None else 0.0)
      a, b = wv * a, wv * b + bv
    else:
      a, b = wv * a, wv * b + bv
      ntk_m, ntk_t, ntk_1 = a + wv * ntk_m, wv * ntk_t, b + wv * ntk_1
  coef = (a, b, ntk_m, ntk_t, ntk_1)
  diag = (not from_state and head == 'flatten' and
          all(s.pool is None for s in stages))
  return Plan(stages=stages, head=head, head_coef=coef, n_convs=n_convs,
              is_gaussian=is_gaussian, out_channels=chans,
              out_spatial=(h, w), diag=diag, out_scale=pend)


# ------------------------------------------------------------------ runner

DEFAULT_TILE_BYTES = 6 << 30

# When a list, every launch appends (entry point, start event, end event,
# algorithmic HBM bytes of the launch, pairs in the launch).  Used by the
# benchmark to time each kernel with HIP events on the launch stream.
TRACE = None


def _launch(name, alg_bytes, pairs, *args):
  """Calls entry point ``name`` (a trace label may follow it in
  brackets)."""
  entry = name.split('[')[0]
  if TRACE is None:
    _lib.call(entry, *args)
    return
  s = torch.cuda.Event(enable_timing=True)
  e = torch.cuda.Event(enable_timing=True)
  s.record()
  _lib.call(entry, *args)
  e.record()
  TRACE.append((name, s, e, alg_bytes, pairs))


def _state_elems(h, w):
  return h * h * w * w


def _keeps_pre(s) -> bool:
  """Conv(3x3 SAME) -> Relu -> AvgPool(2x2) on an 8x8, 16x16 or 32x32
  state: nta_stage_keep_pre can also return the stage's stencil means."""
  return (not s.first and s.conv is not None and s.conv[:3] == (3, 3, False)
          and not s.circular and s.relu and s.act == (1.0, 0.0)
          and s.act_kind == 0 and s.pool == (2, 2) and s.hin == s.win
          and s.hin in (8, 16, 32))


def _run_chain(plan: Plan, xa, xb, n1, n2, qa, qb, want_ntk, stream,
               collect_q=None, head_out=None, ld=0, keep=None, qn=None,
               init=None, pre=None):
  """Runs all stages for one tile.  Self tile when ``n2 == 0`` (then
  ``collect_q`` receives the normaliser of each ReLU stage, computed from the
  self state on the way).  Returns the final (nngp, ntk) state when the plan
  has no head, else writes the head output into ``head_out``.  ``keep``
  (a list) receives the output state of every stage (the checkpoints the
  adjoint recomputes from; ``(None, None)`` for a stage inside a fused
  launch, whose state is never formed).  ``qn`` = (first-ReLU normaliser tables of the
  tile's x1 and x2 images, then the second ReLU's), from ``nta_qnorm``,
  selects the fused launch of the first two stages (cross tiles only).
  ``init`` = (nngp, ntk) input state of the tile ([T, H, H, W, W], ntk may
  be None for zero) for plans compiled ``from_state``; then ``xa`` and
  ``xb`` are unused.  ``pre`` (a dict) receives, for every stage index
  whose stage ``_keeps_pre`` accepts, the stencil means of its inputs
  (nngp, ntk) at the conv-output resolution (the adjoint's ``aN``,
  ``aT``)."""
  T = n1 if n2 == 0 else n1 * n2
  nngp = ntk = None
  fl_ntk = want_ntk and n2 != 0
  if init is not None:
    nngp, ntk = init
    dev = nngp.device
    if fl_ntk and ntk is None:
      ntk = torch.zeros_like(nngp)
  else:
    dev = xa.device
    _, H, W, C = xa.shape
  skip = 0
  if qn is not None and n2 != 0 and len(qn) == 6:
    s0, s1, s2 = plan.stages[0], plan.stages[1], plan.stages[2]
    out = torch.empty((T, s2.hout, s2.hout, s2.wout, s2.wout),
                      dtype=torch.float32, device=dev)
    out_t = torch.empty_like(out) if fl_ntk else None
    # Algorithmic bytes: those of the three per-layer kernels it replaces
    # (two 32^4 states written and read back, the pooled state written).
    nt = 2 if fl_ntk else 1
    alg = T * nt * 4 * (4 * _state_elems(s0.hc, s0.wc) + out[0].numel())
    _launch('nta_first3_pool_stage', alg, T, _lib.ptr(xa), _lib.ptr(xb), n1,
            n2, H, W, C, _lib.NTK if fl_ntk else 0, s0.conv[3], s0.conv[4],
            _lib.ptr(qn[0]), _lib.ptr(qn[1]), s1.conv[3], s1.conv[4],
            _lib.ptr(qn[2]), _lib.ptr(qn[3]), s2.conv[3], s2.conv[4],
            _lib.ptr(qn[4]), _lib.ptr(qn[5]), _lib.ptr(out), _lib.ptr(out_t),
            stream)
    if keep is not None:         # the first two stages' states never formed
      keep.extend([(None, None), (None, None), (out, out_t)])
    nngp, ntk = out, out_t
    skip = 3
  elif qn is not None and n2 != 0:
    s0, s1 = plan.stages[0], plan.stages[1]
    out = torch.empty((T, s1.hout, s1.hout, s1.wout, s1.wout),
                      dtype=torch.float32, device=dev)
    out_t = torch.empty_like(out) if fl_ntk else None
    # Algorithmic bytes: those of the two per-layer kernels it replaces
    # (the first stage's state written and read back, the pooled state
    # written), the accounting the per-layer schedule is measured against.
    nt = 2 if fl_ntk else 1
    alg = T * nt * 4 * (2 * _state_elems(s0.hc, s0.wc) + out[0].numel())
    args = (_lib.ptr(xa), _lib.ptr(xb), n1, n2, H, W, C,
            _lib.NTK if fl_ntk else 0, s0.conv[3], s0.conv[4],
            _lib.ptr(qn[0]), _lib.ptr(qn[1]), s1.conv[3], s1.conv[4],
            _lib.ptr(qn[2]), _lib.ptr(qn[3]), _lib.ptr(out), _lib.ptr(out_t))
    if pre is not None:
      # the second stage's stencil means, [T, 32, 32, 32, 32] per tensor
      shp = (T, s1.hin, s1.hin, s1.win, s1.win)
      pN = torch.empty(shp, dtype=torch.float32, device=dev)
      pT = torch.empty_like(pN) if fl_ntk else None
      _launch('nta_first_pool_stage_keep_pre', alg + T * nt * 4 * pN[0].numel(),
              T, *args, _lib.ptr(pN), _lib.ptr(pT), stream)
      pre[1] = (pN, pT)
    else:
      _launch('nta_first_pool_stage', alg, T, *args, stream)
    if keep is not None:         # the first stage's state is never formed
      keep.extend([(None, None), (out, out_t)])
    nngp, ntk = out, out_t
    skip = 2
  for si, s in enumerate(plan.stages):
    if si < skip:
      continue
    if s.first:
      fh, fw, valid, w_var, b_var = s.conv
      flags = s.flags(fl_ntk)
      q1 = q2 = None
      if s.relu:
        if n2 == 0:
          q = torch.empty((n1, s.hc, s.wc), dtype=torch.float32, device=dev)
          _launch('nta_first_diag', 4 * n1 * s.hc * s.wc, 0, _lib.ptr(xa),
                  n1, H, W, C, fh, fw, flags, w_var, b_var, _lib.ptr(q),
                  stream)
          collect_q[si] = q
          q1 = q
        else:
          q1, q2 = qa[si], qb[si]
      out = torch.empty((T,) + (s.hc, s.hc, s.wc, s.wc), dtype=torch.float32,
                        device=dev)
      out_t = torch.empty_like(out) if fl_ntk else None
      nt = 2 if fl_ntk else 1
      _launch('nta_first_stage', T * nt * 4 * out[0].numel(), T,
              _lib.ptr(xa), _lib.ptr(xb), n1, n2, H, W, C, fh, fw, flags,
              w_var, b_var, s.act[0], s.act[1], _lib.ptr(q1), _lib.ptr(q2),
              _lib.ptr(out), _lib.ptr(out_t), stream)
      nngp, ntk = out, out_t
      if keep is not None:
        keep.append((nngp, ntk))
      continue
    conv = s.conv if s.conv is not None else (1, 1, False, 1.0, 0.0)
    fh, fw, valid, w_var, b_var = conv
    flags = s.flags(fl_ntk)
    q1 = q2 = None
    if s.relu:
      if n2 == 0:
        q = torch.empty((n1, s.hc, s.wc), dtype=torch.float32, device=dev)
        _launch('nta_stage_diag', 4 * n1 * s.hc * s.wc, 0, _lib.ptr(nngp),
                n1, s.hin, s.win, fh, fw, flags, w_var, b_var, _lib.ptr(q),
                stream)
        collect_q[si] = q
        q1 = q
      else:
        q1, q2 = qa[si], qb[si]
    ph, pw = s.pool if s.pool is not None else (1, 1)
    out = torch.empty((T, s.hout, s.hout, s.wout, s.wout),
                      dtype=torch.float32, device=dev)
    out_t = torch.empty_like(out) if fl_ntk else None
    nt = 2 if fl_ntk else 1
    if pre is not None and _keeps_pre(s):
      pN = torch.empty_like(nngp)
      pT = torch.empty_like(nngp) if fl_ntk else None
      _launch(f'nta_stage_keep_pre[{s.hin}->{s.hout}]',
              T * nt * 4 * (2 * nngp[0].numel() + out[0].numel()), T,
              _lib.ptr(nngp), _lib.ptr(ntk), n1, n2, s.hin, flags, w_var,
              b_var, _lib.ptr(q1), _lib.ptr(q2), _lib.ptr(out),
              _lib.ptr(out_t), _lib.ptr(pN), _lib.ptr(pT), stream)
      pre[si] = (pN, pT)
    else:
      _launch(f'nta_stage[{s.hin}->{s.hout}]',
              T * nt * 4 * (nngp[0].numel() + out[0].numel()), T,
              _lib.ptr(nngp), _lib.ptr(ntk), n1, n2, s.hin, s.win, fh, fw,
              ph, pw, flags, w_var, b_var, s.act[0], s.act[1], _lib.ptr(q1),
              _lib.ptr(q2), _lib.ptr(out), _lib.ptr(out_t), stream)
    nngp, ntk = out, out_t
    if keep is not None:
      keep.append((nngp, ntk))
  if plan.head is None or head_out is None:
    return nngp, ntk
  a, b, c, d, e = plan.head_coef
  hflags = (_lib.FLATTEN if plan.head == 'flatten' else 0) | (
      _lib.NTK if fl_ntk else 0)
  ho, wo = plan.out_spatial
  out_n, out_t = head_out
  nt = 2 if fl_ntk else 1
  _launch('nta_head', T * nt * 4 * (nngp[0].numel() + 1), T, _lib.ptr(nngp),
          _lib.ptr(ntk), n1, n2, ho, wo, hflags, a, b, c, d, e,
          _lib.ptr(out_n), _lib.ptr(out_t), ld, stream)
  return None


def _fuses_first_two(plan: Plan, in_shape) -> bool:
  """Conv(3x3 SAME) -> Relu -> Conv(3x3 SAME) -> Relu -> AvgPool(2x2) at the
  start of the network on 32x32 images with <= 4 channels: the first two
  stages run as one launch (nta_first_pool_stage)."""
  if len(plan.stages) < 2:
    return False
  _, H, W, C = in_shape
  s0, s1 = plan.stages[0], plan.stages[1]
  def conv33(s):
    return (s.conv is not None and s.conv[:3] == (3, 3, False)
            and not s.circular)
  return (H == 32 and W == 32 and 1 <= C <= 4 and s0.first and conv33(s0)
          and s0.relu and s0.act == (1.0, 0.0) and s0.act_kind == 0
          and s0.pool is None
          and conv33(s1) and s1.relu and s1.act == (1.0, 0.0)
          and s1.act_kind == 0
          and s1.pool == (2, 2))


def _fuses_first_three(plan: Plan, in_shape) -> bool:
  """Conv(3x3 SAME) -> Relu, three times, -> AvgPool(2x2) at the start of
  the network (Myrtle-10) on 32x32 images with <= 4 channels: the first
  three stages run as one launch (nta_first3_pool_stage)."""
  if len(plan.stages) < 3:
    return False
  _, H, W, C = in_shape
  s0, s1, s2 = plan.stages[0], plan.stages[1], plan.stages[2]

  def conv_relu(s):
    return (s.conv is not None and s.conv[:3] == (3, 3, False)
            and not s.circular and s.relu and s.act == (1.0, 0.0)
            and s.act_kind == 0)
  return (H == 32 and W == 32 and 1 <= C <= 4 and s0.first
          and conv_relu(s0) and s0.pool is None and conv_relu(s1)
          and s1.pool is None and conv_relu(s2) and s2.pool == (2, 2))


def _fused_stages(plan: Plan, in_shape) -> int:
  """Number of leading stages one fused launch runs (0, 2 or 3)."""
  if _fuses_first_three(plan, in_shape):
    return 3
  if _fuses_first_two(plan, in_shape):
    return 2
  return 0


def _per_pair_bytes(plan: Plan, ntk: bool, fused=False,
                    in_hw=None) -> int:
  m = 0 if in_hw is None else _state_elems(*in_hw)
  nf = 2 if fused is True else int(fused)
  for si, s in enumerate(plan.stages):
    if si < nf:
      # the fused stages' states are never materialised, except the last
      # one's (pooled) output
      m = max(m, _state_elems(s.hout, s.wout)) if si == nf - 1 else m
      continue
    m = max(m, _state_elems(s.hc, s.wc), _state_elems(s.hout, s.wout))
  # ping-pong input/output state, nngp (+ ntk)
  return 4 * m * 2 * (2 if ntk else 1)


def _self_chain(plan: Plan, x, want_cov, tile_bytes, stream):
  """Returns (q list per stage [N, hc, wc], cov) for the images of x."""
  N = x.shape[0]
  per = _per_pair_bytes(plan, False)
  chunk = max(1, min(N, tile_bytes // per))
  qs = [None] * len(plan.stages)
  q_parts = [[] for _ in plan.stages]
  cov = None
  if plan.head is not None:
    cov = torch.empty((N,), dtype=torch.float32, device=x.device)
  full = []
  for i0 in range(0, N, chunk):
    n = min(chunk, N - i0)
    xa = x[i0:i0 + n]
    collect = [None] * len(plan.stages)
    if plan.head is not None:
      _run_chain(plan, xa, None, n, 0, None, None, False, stream,
                 collect_q=collect, head_out=(cov[i0:i0 + n], None), ld=0)
    else:
      st, _ = _run_chain(plan, xa, None, n, 0, None, None, False, stream,
                         collect_q=collect)
      if want_cov:
        full.append(st)
    for si, q in enumerate(collect):
      if q is not None:
        q_parts[si].append(q)
  for si in range(len(plan.stages)):
    if q_parts[si]:
      qs[si] = torch.cat(q_parts[si]) if len(q_parts[si]) > 1 else \
          q_parts[si][0]
  if plan.head is None and want_cov:
    cov = torch.cat(full) if len(full) > 1 else full[0]
  return qs, cov


def _self_chain_state(plan: Plan, cov, want_cov, tile_bytes, stream):
  """``_self_chain`` for a plan compiled ``from_state``: the self chains
  start from the per-image states ``cov`` [N, H, H, W, W]."""
  N, H = cov.shape[0], cov.shape[1]
  W = cov.shape[3]
  per = _per_pair_bytes(plan, False, in_hw=(H, W))
  chunk = max(1, min(N, tile_bytes // per))
  q_parts = [[] for _ in plan.stages]
  out = None
  if plan.head is not None:
    out = torch.empty((N,), dtype=torch.float32, device=cov.device)
  full = []
  for i0 in range(0, N, chunk):
    n = min(chunk, N - i0)
    init = (cov[i0:i0 + n].contiguous(), None)
    collect = [None] * len(plan.stages)
    if plan.head is not None:
      _run_chain(plan, None, None, n, 0, None, None, False, stream,
                 collect_q=collect, head_out=(out[i0:i0 + n], None), ld=0,
                 init=init)
    else:
      st, _ = _run_chain(plan, None, None, n, 0, None, None, False, stream,
                         collect_q=collect, init=init)
      if want_cov:
        full.append(st if st is not init[0] else st.clone())
    for si, q in enumerate(collect):
      if q is not None:
        q_parts[si].append(q)
  qs = [torch.cat(p) if len(p) > 1 else (p[0] if p else None)
        for p in q_parts]
  if plan.head is None and want_cov:
    out = torch.cat(full) if len(full) > 1 else full[0]
  return qs, out


def run_state(plan: Plan, nngp, ntk, cov1, cov2, want_ntk, want_cov,
              tile_bytes=DEFAULT_TILE_BYTES):
  """Applies a plan compiled ``from_state`` to a kernel: ``nngp``, ``ntk``
  [N1, N2, H, H, W, W] (natural axis order; ``ntk`` None for zero), per-image
  states ``cov1`` [N1, H, H, W, W] and ``cov2`` [N2, ...] or None (x2 is x1).
  Returns the dict of ``run``."""
  dev = nngp.device
  stream = _lib.stream_handle(dev)
  N1, N2, H = nngp.shape[0], nngp.shape[1], nngp.shape[2]
  W = nngp.shape[4]
  sym = cov2 is None
  q1s, c1 = _self_chain_state(plan, cov1, want_cov, tile_bytes, stream)
  if sym:
    q2s, c2 = q1s, None
  else:
    q2s, c2 = _self_chain_state(plan, cov2, want_cov, tile_bytes, stream)
  per = _per_pair_bytes(plan, want_ntk, in_hw=(H, W)) + 4 * _state_elems(
      H, W) * (2 if want_ntk else 1)
  T = max(1, tile_bytes // per)
  if sym and plan.head is not None and T < N1 * N2:
    ti = tj = max(1, min(N1, math.isqrt(T)))
  else:
    tj = min(N2, T)
    ti = max(1, min(N1, T // tj))
  ho, wo = plan.out_spatial
  if plan.head is not None:
    out_n = torch.empty((N1, N2), dtype=torch.float32, device=dev)
    out_t = torch.empty((N1, N2), dtype=torch.float32, device=dev) \
        if want_ntk else None
  else:
    shp = (N1, N2, ho, ho, wo, wo)
    out_n = torch.empty(shp, dtype=torch.float32, device=dev)
    out_t = torch.empty(shp, dtype=torch.float32, device=dev) \
        if want_ntk else None
  tri = sym and plan.head is not None and ti == tj
  for i0 in range(0, N1, ti):
    n1 = min(ti, N1 - i0)
    for j0 in range(0, N2, tj):
      n2 = min(tj, N2 - j0)
      if tri and j0 + n2 <= i0:
        continue
      qa = [None if q is None else q[i0:i0 + n1] for q in q1s]
      qb = [None if q is None else q[j0:j0 + n2] for q in q2s]
      sl = (slice(i0, i0 + n1), slice(j0, j0 + n2))
      init_n = nngp[sl].reshape((n1 * n2,) + tuple(nngp.shape[2:]))
      init_t = None
      if want_ntk and ntk is not None:
        init_t = ntk[sl].reshape((n1 * n2,) + tuple(ntk.shape[2:]))
      init = (init_n.contiguous(), None if init_t is None else
              init_t.contiguous())
      if plan.head is not None:
        _run_chain(plan, None, None, n1, n2, qa, qb, want_ntk, stream,
                   head_out=(out_n[i0:, j0:],
                             out_t[i0:, j0:] if want_ntk else None),
                   ld=N2, init=init)
      else:
        st_n, st_t = _run_chain(plan, None, None, n1, n2, qa, qb, want_ntk,
                                stream, init=init)
        sh = st_n.shape[1:]
        out_n[sl] = st_n.view((n1, n2) + sh)
        if want_ntk:
          out_t[sl] = st_t.view((n1, n2) + sh)
  if tri:
    _symmetrize_from_upper(out_n)
    if want_ntk:
      _symmetrize_from_upper(out_t)
  out = {'cov1': c1, 'cov2': c2, 'nngp': out_n, 'ntk': out_t}
  if plan.out_scale != 1.0:
    for k in ('nngp', 'ntk', 'cov1', 'cov2'):
      if out[k] is not None:
        out[k] = out[k] * plan.out_scale
  return out


def _diag_tables(plan: Plan):
  """Host stage tables of nta_diag_chain."""
  import ctypes
  n = len(plan.stages)
  desc = (ctypes.c_int * (3 * n))()
  wb = (ctypes.c_float * (4 * n))()
  for k, s in enumerate(plan.stages):
    f = 0
    fh = fw = 1
    wv, bv = 1.0, 0.0
    if s.conv is not None:
      fh, fw, valid, wv, bv = s.conv
      f |= _lib.CONV | (_lib.CONV_VALID if valid else 0)
      f |= _lib.CONV_CIRCULAR if s.circular else 0
    if s.relu:
      f |= _lib.RELU | s.act_kind
    desc[3 * k:3 * k + 3] = [f, fh, fw]
    wb[4 * k:4 * k + 4] = [wv, bv, s.act[0], s.act[1]]
  return desc, wb


def _run_diag(plan: Plan, x1, x2, want_ntk, stream):
  """Diagonal-spatial plans: one nta_diag_chain launch per self chain and
  one for all cross pairs (states never leave the chip, so no tiling)."""
  desc, wb = _diag_tables(plan)
  n = len(plan.stages)
  _, H, W, C = x1.shape
  a, b, c, d, e = plan.head_coef
  dev = x1.device
  has_relu = any(s.relu for s in plan.stages)

  def self_chain(x):
    N = x.shape[0]
    q = torch.empty((n, N, H, W), dtype=torch.float32, device=dev) \
        if has_relu else None
    cov = torch.empty((N,), dtype=torch.float32, device=dev)
    _launch('nta_diag_chain[self]', 4 * N * H * W * C, N, _lib.ptr(x),
            None, N, 0, 0, H, W, C, n, desc, wb, _lib.ptr(q), None, 0, a, b,
            c, d, e, _lib.ptr(cov), None, 0, stream)
    return q, cov

  q1, cov1 = self_chain(x1)
  if x2 is not None:
    q2, cov2 = self_chain(x2)
    xb = x2
  else:
    q2, cov2, xb = q1, None, x1
  N1, N2 = x1.shape[0], xb.shape[0]
  nngp = torch.empty((N1, N2), dtype=torch.float32, device=dev)
  ntk = torch.empty((N1, N2), dtype=torch.float32, device=dev) \
      if want_ntk else None
  sym = 1 if x2 is None else 0
  pairs = N1 * (N1 + 1) // 2 if sym else N1 * N2
  _launch('nta_diag_chain[cross]', 4 * pairs * (2 * H * W * C + 2), pairs,
          _lib.ptr(x1), _lib.ptr(xb), N1, N2, sym, H, W, C, n, desc, wb,
          _lib.ptr(q1), _lib.ptr(q2), _lib.NTK if want_ntk else 0, a, b, c,
          d, e, _lib.ptr(nngp), _lib.ptr(ntk), N2, stream)
  return {'cov1': cov1, 'cov2': cov2, 'nngp': nngp, 'ntk': ntk}


def run(plan: Plan, x1, x2, want_nngp, want_ntk, want_cov,
        tile_bytes=DEFAULT_TILE_BYTES, symmetric=True):
  """Computes nngp/ntk (and cov1/cov2) for inputs x1 [N1,H,W,C], x2 or None.

  Returns dict with 'nngp', 'ntk', 'cov1', 'cov2' tensors (or None)."""
  stream = _lib.stream_handle(x1.device)
  if plan.diag:
    return _run_diag(plan, x1, x2, want_ntk, stream)
  q1s, cov1 = _self_chain(plan, x1, want_cov, tile_bytes, stream)
  if x2 is not None:
    q2s, cov2 = _self_chain(plan, x2, want_cov, tile_bytes, stream)
    xb_all = x2
  else:
    q2s, cov2 = q1s, None
    xb_all = x1
  N1, N2 = x1.shape[0], xb_all.shape[0]
  dev = x1.device
  fused = _fused_stages(plan, x1.shape)
  qn1 = qn2 = None
  if fused:
    qn1 = [_qnorm(q1s[k], stream) for k in range(fused)]
    qn2 = qn1 if x2 is None else [_qnorm(q2s[k], stream)
                                  for k in range(fused)]
  per = _per_pair_bytes(plan, want_ntk, fused)
  T = max(1, tile_bytes // per)
  if x2 is None and symmetric and plan.head is not None and T < N1 * N2:
    ti = tj = max(1, min(N1, math.isqrt(T)))
  else:
    tj = min(N2, T)
    ti = max(1, min(N1, T // tj))
  out = {'cov1': cov1, 'cov2': cov2, 'nngp': None, 'ntk': None}
  if plan.head is not None:
    nngp = torch.empty((N1, N2), dtype=torch.float32, device=dev)
    ntk = torch.empty((N1, N2), dtype=torch.float32, device=dev) \
        if want_ntk else None
  else:
    last = plan.stages[-1]
    shp = (N1, N2, last.hout, last.hout, last.wout, last.wout)
    nngp = torch.empty(shp, dtype=torch.float32, device=dev)
    ntk = torch.empty(shp, dtype=torch.float32, device=dev) \
        if want_ntk else None
  sym = symmetric and x2 is None and plan.head is not None and ti == tj
  for i0 in range(0, N1, ti):
    n1 = min(ti, N1 - i0)
    for j0 in range(0, N2, tj):
      n2 = min(tj, N2 - j0)
      if sym and j0 + n2 <= i0:
        continue           # strictly below the diagonal: mirrored later
      qa = [None if q is None else q[i0:i0 + n1] for q in q1s]
      qb = [None if q is None else q[j0:j0 + n2] for q in q2s]
      xa = x1[i0:i0 + n1]
      xb = xb_all[j0:j0 + n2]
      qn = None
      if fused:
        qn = tuple(t for k in range(fused)
                   for t in (qn1[k][i0:i0 + n1], qn2[k][j0:j0 + n2]))
      if plan.head is not None:
        on = nngp[i0:, j0:]
        ot = ntk[i0:, j0:] if want_ntk else None
        _run_chain(plan, xa, xb, n1, n2, qa, qb, want_ntk, stream,
                   head_out=(on, ot), ld=N2, qn=qn)
      else:
        st_n, st_t = _run_chain(plan, xa, xb, n1, n2, qa, qb, want_ntk,
                                stream, qn=qn)
        sh = st_n.shape[1:]
        nngp[i0:i0 + n1, j0:j0 + n2] = st_n.view((n1, n2) + sh)
        if want_ntk:
          ntk[i0:i0 + n1, j0:j0 + n2] = st_t.view((n1, n2) + sh)
  if x2 is None and plan.head is not None:
    _symmetrize_from_upper(nngp)
    if want_ntk:
      _symmetrize_from_upper(ntk)
  out['nngp'] = nngp
  out['ntk'] = ntk
  if plan.out_scale != 1.0:
    for k in ('nngp', 'ntk', 'cov1', 'cov2'):
      if out[k] is not None:
        out[k].mul_(plan.out_scale)
  return out


def _qnorm(q, stream):
  """(sqrt(q), 1/sqrt(q)) tables [N, H, W, 2] of a normaliser [N, H, W]."""
  q = q.contiguous()
  out = torch.empty(q.shape + (2,), dtype=torch.float32, device=q.device)
  _launch('nta_qnorm', 12 * q.numel(), 0, _lib.ptr(q), q.numel(),
          _lib.ptr(out), stream)
  return out


def _symmetrize_from_upper(m):
  """K(x, x) is symmetric: keep the upper triangle (which every tile above
  the block diagonal computed) and mirror it, so the result is exactly
  symmetric."""
  up = torch.triu(m)
  m.copy_(up + torch.triu(m, 1).T)
