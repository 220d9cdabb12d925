"""The ``Kernel`` record returned by ``kernel_fn(..., get=None)``.

Field names, meanings and layouts follow ``utils/kernel.py:27-149`` of
neural-tangents (fields at 124-145): ``nngp``/``ntk`` are
``[N1, N2, X, X', Y, Y']`` (zipped spatial axes, reversed after each Conv as
``is_reversed`` records) or ``[N1, N2]`` after a readout; ``cov1``/``cov2``
are ``[N, X, X', Y, Y']`` or ``[N]`` (``diagonal_batch=True``).  Arrays are
torch tensors on the device the computation ran on.
"""
import dataclasses
from typing import Any, Optional, Sequence


def _reverse_zipped(x, start_axis):
  """``utils.py:465-480``: the zipped spatial axis pairs after
  ``start_axis`` in reverse order."""
  if x is None or not hasattr(x, 'ndim') or x.ndim <= start_axis:
    return x
  ndim = x.ndim
  src = tuple(j for i in range(ndim - 2, start_axis - 1, -2)
              for j in (i, i + 1))
  return x.permute(tuple(range(start_axis)) + src)


@dataclasses.dataclass
class Kernel:
  nngp: Any
  ntk: Optional[Any]
  cov1: Any
  cov2: Optional[Any]
  x1_is_x2: bool
  is_gaussian: bool
  is_reversed: bool
  is_input: bool
  diagonal_batch: bool
  diagonal_spatial: bool
  shape1: tuple
  shape2: tuple
  batch_axis: int
  channel_axis: int
  mask1: Optional[Any] = None
  mask2: Optional[Any] = None

  def replace(self, **kwargs) -> 'Kernel':
    return dataclasses.replace(self, **kwargs)

  def asdict(self):
    return dataclasses.asdict(self)

  def astuple(self):
    return dataclasses.astuple(self)

  def reverse(self) -> 'Kernel':
    """``utils/kernel.py:169-189``: the order of the spatial axis pairs of
    every covariance is reversed (``[.., X, X', Y, Y']`` ->
    ``[.., Y, Y', X, X']``) and ``is_reversed`` flips."""
    bnd = 1 if self.diagonal_batch else 2
    return self.replace(cov1=_reverse_zipped(self.cov1, bnd),
                        nngp=_reverse_zipped(self.nngp, 2),
                        cov2=_reverse_zipped(self.cov2, bnd),
                        ntk=_reverse_zipped(self.ntk, 2),
                        is_reversed=not self.is_reversed)

  def transpose(self, axes: Optional[Sequence[int]] = None) -> 'Kernel':
    """``utils/kernel.py:191-222``: permutes the spatial dimensions by
    ``axes`` (batch axes stay leading; each spatial dimension's pair of
    zipped axes moves together unless ``diagonal_spatial``)."""
    if axes is None:
      axes = tuple(range(len(self.shape1) - 2))

    def permute(mat, bnd):
      if mat is None or not hasattr(mat, 'ndim') or mat.ndim == 0:
        return mat
      ax = tuple(bnd + a for a in axes)
      if not self.diagonal_spatial:
        ax = tuple(j for a in ax for j in (2 * a - bnd, 2 * a - bnd + 1))
      return mat.permute(tuple(range(bnd)) + ax)

    bnd = 1 if self.diagonal_batch else 2
    return self.replace(cov1=permute(self.cov1, bnd),
                        nngp=permute(self.nngp, 2),
                        cov2=permute(self.cov2, bnd),
                        ntk=permute(self.ntk, 2))

  def slice(self, n1_slice: slice, n2_slice: slice) -> 'Kernel':
    """``utils/kernel.py:151-167``."""
    cov1 = self.cov1[n1_slice]
    cov2 = self.cov1[n2_slice] if self.cov2 is None else self.cov2[n2_slice]
    ntk = self.ntk
    return self.replace(
        cov1=cov1, nngp=self.nngp[n1_slice, n2_slice], cov2=cov2,
        ntk=ntk if ntk is None or ntk.ndim == 0 else ntk[n1_slice, n2_slice],
        shape1=(cov1.shape[0],) + tuple(self.shape1[1:]),
        shape2=(cov2.shape[0],) + tuple(self.shape2[1:]))
