"""``batch``: the reference's pair-tiling wrapper (``batching.py:76-133``).

``batch(kernel_fn, batch_size, device_count, store_on_device)`` returns a
kernel function that evaluates ``kernel_fn`` on ``batch_size x batch_size``
blocks of pairs and assembles the result, as ``_serial``
(``batching.py:314-503``) does.  The block sizes follow
``_get_n_batches_and_batch_sizes`` (``batching.py:647-679``): the batch size is
reduced to ``gcd(n1, n2)`` with a warning, and must divide both batch sizes
otherwise (``ValueError``).  With ``x2=None`` only blocks on or above the
diagonal are computed and the lower ones are mirrored (the reference's TODO
at ``batching.py:370``).  ``store_on_device=False`` moves each block to host
memory as it is produced.  This engine runs on one GPU: ``device_count``
other than -1, 0 or 1 raises ``NotImplementedError``.

The HIP engine already tiles pairs internally (``kernel_fn(..., tile_bytes=)``);
``batch`` bounds the size of each call and the placement of the output.
"""
import collections
import math
import warnings

import numpy as np
import torch

from .kernel import Kernel


def _n_batches(n1, n2, batch_size):
  """``batching.py:647-679`` with ``device_count = 1``."""
  max_bs = math.gcd(n1, n2)
  bs = min(batch_size, max_bs)
  if bs != batch_size:
    warnings.warn('Batch size is reduced from requested %d to effective %d '
                  'to fit the dataset.' % (batch_size, bs))
  if n1 % bs:
    raise ValueError('Number of rows of kernel must divide batch size. '
                     f'Found n1 = {n1} and batch size = {bs}.')
  if n2 % bs:
    raise ValueError('Number of columns of kernel must divide batch size. '
                     f'Found n2 = {n2} and batch size = {bs}')
  return n1 // bs, n2 // bs, bs


def batch(kernel_fn, batch_size: int = 0, device_count: int = -1,
          store_on_device: bool = True):
  if device_count not in (-1, 0, 1):
    raise NotImplementedError('this engine runs on one GPU; device_count '
                              f'must be -1, 0 or 1, got {device_count}.')
  if not batch_size:
    return kernel_fn

  def batched(x1, x2=None, get=None, **kwargs):
    n1 = x1.shape[0]
    sym = x2 is None
    xb = x1 if sym else x2
    n2 = xb.shape[0]
    nb1, nb2, bs = _n_batches(n1, n2, batch_size)
    blocks = {}
    for i in range(nb1):
      for j in range(nb2):
        if sym and j < i:
          continue
        a = x1[i * bs:(i + 1) * bs]
        b = xb[j * bs:(j + 1) * bs]
        out = kernel_fn(a, None if (sym and i == j) else b, get, **kwargs)
        if not store_on_device:
          out = _to_host(out)
        blocks[i, j] = out
    return _assemble(blocks, nb1, nb2, sym, get)

  return batched


def _to_host(out):
  if isinstance(out, torch.Tensor):
    return out.cpu()
  if isinstance(out, Kernel):
    return out.replace(**{f: (v.cpu() if isinstance(v, torch.Tensor) else v)
                          for f, v in out.asdict().items()})
  if hasattr(out, '_fields'):
    return type(out)(*[_to_host(v) for v in out])
  return out


def _cat_pairs(blocks, nb1, nb2, sym, key):
  rows = []
  for i in range(nb1):
    row = []
    for j in range(nb2):
      if sym and j < i:
        t = key(blocks[j, i])
        t = t.transpose(0, 1) if t.ndim == 2 else _swap_pair_axes(t)
      else:
        t = key(blocks[i, j])
      row.append(t)
    rows.append(torch.cat(row, dim=1))
  return torch.cat(rows, dim=0)


def _swap_pair_axes(t):
  """K(x2, x1) from K(x1, x2) for a full-spatial [N1, N2, X, X', Y, Y']
  kernel: swap the batch axes and each primed/unprimed spatial pair."""
  perm = [1, 0]
  for a in range(2, t.ndim, 2):
    perm += [a + 1, a]
  return t.permute(perm)


def _assemble(blocks, nb1, nb2, sym, get):
  first = blocks[0, 0]
  if isinstance(first, torch.Tensor):
    return _cat_pairs(blocks, nb1, nb2, sym, lambda b: b)
  if hasattr(first, '_fields'):
    vals = []
    for f in first._fields:
      if f in ('nngp', 'ntk'):
        vals.append(_cat_pairs(blocks, nb1, nb2, sym,
                               lambda b, f=f: getattr(b, f)))
      elif f in ('cov1', 'cov2'):
        vals.append(_cat_diag(blocks, nb1, nb2, sym, f))
      else:
        vals.append(getattr(first, f))
    return type(first)(*vals)
  if isinstance(first, Kernel):
    nngp = _cat_pairs(blocks, nb1, nb2, sym, lambda b: b.nngp)
    ntk = None if first.ntk is None else _cat_pairs(
        blocks, nb1, nb2, sym, lambda b: b.ntk)
    cov1 = torch.cat([blocks[i, i if sym else 0].cov1 for i in range(nb1)])
    cov2 = None if sym else torch.cat([blocks[0, j].cov2
                                       for j in range(nb2)])
    return first.replace(nngp=nngp, ntk=ntk, cov1=cov1, cov2=cov2,
                         shape1=(nngp.shape[0],) + tuple(first.shape1[1:]),
                         shape2=(nngp.shape[1],) + tuple(first.shape2[1:]))
  raise TypeError(type(first))


def _cat_diag(blocks, nb1, nb2, sym, f):
  if f == 'cov1':
    return torch.cat([getattr(blocks[i, i if sym else 0], f)
                      for i in range(nb1)])
  if sym:
    return None
  return torch.cat([getattr(blocks[0, j], f) for j in range(nb2)])
