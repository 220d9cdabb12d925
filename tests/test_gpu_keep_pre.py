"""nta_stage_keep_pre and nta_first_pool_stage_keep_pre: the pooled 3x3
stages that also return their inputs' stencil means (the pre-activations the
stage's adjoint consumes).  Their pooled outputs must match the launches
without the means, and their means the stencil kernel's, on the states of a
Myrtle-5 network (needs a GPU)."""
import pytest
import torch

from neural_tangents_amd import _engine, _lib, grad, stax
from neural_tangents_amd.data import blurred_images

pytestmark = pytest.mark.gpu


def _plan_and_images(side, n, depths):
  kfn = stax.myrtle(depths=depths, W_std=1.4, b_std=0.3)[2]
  x = torch.tensor(blurred_images([n], side, seed=side)[0],
                   dtype=torch.float32, device='cuda:0')
  return kfn.plan(tuple(x.shape)), x


@pytest.mark.parametrize('side,n1,n2,depths', [
    (32, 1, 3, [2, 1, 1]), (32, 2, 0, [2, 1, 1]), (16, 2, 3, [2, 1, 1]),
    (16, 5, 0, [2, 1, 1]), (8, 3, 7, [2, 1, 1]), (8, 3, 0, [2, 1, 1]),
    (16, 2, 3, [1, 1, 1])])
def test_keep_pre_matches_stage_and_stencil(side, n1, n2, depths):
  """Tiles of 3, 6 and 21 pairs leave tail pairs in the 16x16 and 8x8
  stages' groups of 4 and 16 pairs; self tiles run without the ntk.
  Myrtle-5 (depths 2, 1, 1) keeps every pooled stage at 32, 16 and 8; with
  depths 1, 1, 1 the first pool has no conv and runs nta_stage."""
  plan, x = _plan_and_images(side, n1 + n2, depths)
  dev = x.device
  stream = _lib.stream_handle(dev)
  xa, xb = x[:n1], (x[n1:] if n2 else None)
  kept = [i for i, s in enumerate(plan.stages) if _engine._keeps_pre(s)]
  assert kept, 'the network has pooled 3x3 stages'
  if n2:
    qa = _engine._self_chain(plan, xa, False, _engine.DEFAULT_TILE_BYTES,
                             stream)[0]
    qb = _engine._self_chain(plan, xb, False, _engine.DEFAULT_TILE_BYTES,
                             stream)[0]
    args = (plan, xa, xb, n1, n2, qa, qb, True, stream)
  else:
    args = (plan, xa, None, n1, 0, None, None, False, stream)
  T = n1 * n2 if n2 else n1
  keep0, keep1, pre = [], [], {}
  L = len(plan.stages)
  _engine._run_chain(*args, keep=keep0, collect_q=[None] * L)
  _engine._run_chain(*args, keep=keep1, pre=pre, collect_q=[None] * L)
  assert sorted(pre) == kept
  for si in range(L):
    for a, b in zip(keep0[si], keep1[si]):
      if a is None:
        assert b is None
        continue
      torch.testing.assert_close(b, a, rtol=2e-6,
                                 atol=2e-6 * float(a.abs().max()))
  for si in kept:
    s = plan.stages[si]
    pN, pT = pre[si]
    Nin, Tin = keep0[si - 1]
    want = grad._stencil(Nin, T, s.hin, 3, 3, False, stream)
    torch.testing.assert_close(pN, want, rtol=2e-6,
                               atol=2e-6 * float(want.abs().max()))
    if n2:
      want_t = grad._stencil(Tin, T, s.hin, 3, 3, False, stream)
      torch.testing.assert_close(pT, want_t, rtol=2e-6,
                                 atol=2e-6 * float(want_t.abs().max()))
    else:
      assert pT is None


def test_fused_first_two_stages_keep_the_second_stage_means():
  """Myrtle-5 32x32 cross tiles (3 and 4 x 5 pairs): the fused launch's
  second-stage means and every later stage match the per-layer path's (the
  fused first stage's split-FP16 Gram is FP32-accurate)."""
  plan, x = _plan_and_images(32, 9, [2, 1, 1])
  assert _engine._fused_stages(plan, x.shape) == 2
  stream = _lib.stream_handle(x.device)
  L = len(plan.stages)
  for n1, n2 in ((1, 3), (4, 5)):
    xa, xb = x[:n1], x[n1:n1 + n2]
    qa = _engine._self_chain(plan, xa, False, _engine.DEFAULT_TILE_BYTES,
                             stream)[0]
    qb = _engine._self_chain(plan, xb, False, _engine.DEFAULT_TILE_BYTES,
                             stream)[0]
    qn = tuple(t for k in range(2) for t in (_engine._qnorm(qa[k], stream),
                                             _engine._qnorm(qb[k], stream)))
    args = (plan, xa, xb, n1, n2, qa, qb, True, stream)
    keep0, pre0, keep1, pre1 = [], {}, [], {}
    _engine._run_chain(*args, keep=keep0, pre=pre0)
    _engine._run_chain(*args, keep=keep1, pre=pre1, qn=qn)
    assert sorted(pre1) == sorted(pre0) == [1, 2, 3]
    assert keep1[0] == (None, None)
    for si in range(1, L):
      for a, b in zip(keep0[si], keep1[si]):
        assert float((a - b).abs().max()) <= 1e-5 * float(a.abs().max())
    for si in (1, 2, 3):
      for a, b in zip(pre0[si], pre1[si]):
        assert float((a - b).abs().max()) <= 1e-5 * float(a.abs().max())


def test_keep_pre_rejects_other_stages():
  dev = torch.device('cuda', 0)
  stream = _lib.stream_handle(dev)
  buf = torch.zeros((12 ** 4,), device=dev)
  q = torch.ones((144,), device=dev)
  p = _lib.ptr(buf)
  with pytest.raises(NotImplementedError, match='keep_pre'):
    _lib.call('nta_stage_keep_pre', p, None, 1, 0, 12,
              _lib.CONV | _lib.RELU, 2.0, 0.0, _lib.ptr(q), None, p, None,
              p, None, stream)
  with pytest.raises(NotImplementedError, match='keep_pre'):
    _lib.call('nta_stage_keep_pre', p, None, 1, 0, 8,
              _lib.CONV | _lib.RELU | _lib.CONV_VALID, 2.0, 0.0, _lib.ptr(q),
              None, p, None, p, None, stream)
  with pytest.raises(ValueError, match='null pointer'):
    _lib.call('nta_stage_keep_pre', p, None, 1, 0, 8,
              _lib.CONV | _lib.RELU, 2.0, 0.0, _lib.ptr(q), None, p, None,
              None, None, stream)
