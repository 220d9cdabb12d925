"""The ``Kernel`` record returned by ``kernel_fn(..., get=None)``.

Field names, meanings and layouts follow ``utils/kernel.py:27-149`` of
neural-tangents (fields at 124-145): ``nngp``/``ntk`` are
``[N1, N2, X, X', Y, Y']`` (zipped spatial axes, reversed after each Conv as
``is_reversed`` records) or ``[N1, N2]`` after a readout; ``cov1``/``cov2``
are ``[N, X, X', Y, Y']`` or ``[N]`` (``diagonal_batch=True``).  Arrays are
torch tensors on the device the computation ran on.
"""
import dataclasses
from typing import Any, Optional


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
