"""nta_relu_bwd (the ReLU + pool adjoint of one stage) against a float64
torch evaluation of the same per-element derivative rules, on random
inputs, for the row kernel (sides 8, 16, 32), the per-element kernel (other
sides), self and cross tiles, tiles with pairs of identical images
(self_off) and the patch-Gram output layout (needs a GPU).

These compare the kernel's indexing and its on-chip sums (gqa, gqb, the
w_var/b_var partials) with a direct evaluation; that evaluation is itself
checked against torch autograd of the forward map (test_relu_bwd_reference),
and the gradients end to end against finite differences of the oracle.
"""
import pytest
import torch

from neural_tangents_amd import _lib
from _relu_bwd_ref import _expand_q, _inputs, _reference

pytestmark = pytest.mark.gpu

NO_SELF = 2 ** 31 - 1


def _run(aN, aT, n1, n2, S, pool, self_off, w, b, k1, k0, q1, q2, gN, gT,
         gram_scale=0.0):
  dev = torch.device('cuda', 0)
  f = lambda x: None if x is None else x.to(dev, torch.float32).contiguous()
  aN_d, aT_d, q1_d, gN_d, gT_d = f(aN), f(aT), f(q1), f(gN), f(gT)
  q2_d = None if n2 == 0 else f(q2)
  T = aN.shape[0]
  gaN = torch.empty_like(aN_d)
  gaT = torch.empty_like(aN_d) if aT is not None else None
  gqa = torch.empty((T, S, S), dtype=torch.float32, device=dev)
  gqb = torch.empty_like(gqa)
  parts = _lib.value('nta_relu_bwd_parts', n1, n2, S)
  gwb = torch.empty((parts, 2), dtype=torch.float64, device=dev)
  nwork = _lib.value('nta_relu_bwd_work_bytes', n1, n2, S)
  work = torch.empty((nwork,), dtype=torch.uint8, device=dev)
  _lib.call('nta_relu_bwd', _lib.ptr(aN_d), _lib.ptr(aT_d), n1, n2, S, pool,
            self_off, w, b, k1, k0, _lib.ptr(q1_d), _lib.ptr(q2_d),
            _lib.ptr(gN_d), _lib.ptr(gT_d), gram_scale, _lib.ptr(gaN),
            _lib.ptr(gaT), _lib.ptr(gqa), _lib.ptr(gqb), _lib.ptr(gwb),
            _lib.ptr(work), nwork, _lib.stream_handle(dev))
  sums = gwb.sum(dim=0).cpu()
  cpu = lambda x: None if x is None else x.double().cpu()
  return cpu(gaN), cpu(gaT), cpu(gqa), cpu(gqb), float(sums[0]), \
      float(sums[1])


def _close(got, want, rtol=2e-5):
  scale = want.abs().max().item()
  err = (got - want).abs().max().item()
  assert err <= rtol * scale, (err, scale)


CASES = [
    # n1, n2, S, pool, has_t, self_off
    (2, 3, 32, 2, True, NO_SELF),      # row kernel, side 32, cross
    (2, 2, 32, 1, False, 0),           # first stage, pairs (a, a)
    (3, 0, 16, 2, True, NO_SELF),      # side 16, self tile
    (2, 3, 16, 1, True, -1),           # pairs (a, a - 1)
    (3, 2, 8, 2, True, 1),             # side 8 (four h' planes at once)
    (2, 0, 8, 1, False, NO_SELF),
    (2, 3, 4, 2, True, NO_SELF),       # per-element kernel
    (3, 0, 6, 2, True, NO_SELF),
    (2, 2, 12, 1, False, 0),
]


@pytest.mark.parametrize('n1,n2,S,pool,has_t,self_off', CASES)
def test_relu_bwd_matches_float64_evaluation(n1, n2, S, pool, has_t,
                                             self_off):
  w, b, k1, k0 = 1.7, 0.2, 1.0, 0.0
  q1, q2, rho, gN, gT, aT = _inputs(n1, n2, S, pool, has_t, seed=S + n1)
  qa, qb = _expand_q(q1, q2, n1, n2, S)
  aN = (rho * torch.sqrt(qa * qb) - b) / w
  # float32 inputs, evaluated in float64
  aN = aN.float().double()
  aT = None if aT is None else aT.float().double()
  q1, q2 = q1.float().double(), q2.float().double()
  qa, qb = _expand_q(q1, q2, n1, n2, S)
  gN, gT = gN.float().double(), gT.float().double()
  T = aN.shape[0]
  ii = torch.arange(T) if n2 == 0 else torch.arange(T) // n2
  jj = torch.arange(T) if n2 == 0 else torch.arange(T) % n2
  pair_same = (n2 == 0) | (ii + self_off == jj)
  h = torch.arange(S)
  diag = (h[:, None, None, None] == h[None, :, None, None]) & \
      (h[None, None, :, None] == h[None, None, None, :])
  same = pair_same[:, None, None, None, None] & diag[None]
  want = _reference(aN, aT, qa, qb, gN, gT, pool, same, w, b, k1, k0)
  got = _run(aN, aT, n1, n2, S, pool, self_off, w, b, k1, k0, q1, q2, gN,
             gT)
  for k in range(4):
    if want[k] is None or k == 1 and aT is None:
      continue
    _close(got[k], want[k])
  assert abs(got[4] - want[4]) <= 1e-5 * (abs(want[4]) + 1.0)
  assert abs(got[5] - want[5]) <= 1e-5 * (abs(want[5]) + 1.0)


@pytest.mark.parametrize('S', [8, 32, 6])
def test_relu_bwd_patch_gram_layout(S):
  n1, n2, pool = 2, 3, 1
  q1, q2, rho, gN, gT, _ = _inputs(n1, n2, S, pool, False, seed=5)
  qa, qb = _expand_q(q1, q2, n1, n2, S)
  aN = ((rho * torch.sqrt(qa * qb) - 0.1) / 1.3).float().double()
  args = (aN, None, n1, n2, S, pool, NO_SELF, 1.3, 0.1, 1.0, 0.0, q1, q2,
          gN, gT)
  base = _run(*args)
  gram = _run(*args, gram_scale=0.25)
  # stage layout [a, b, h, h', w, w'] -> [a, h, w, b, h', w']
  want = 0.25 * base[0].reshape(n1, n2, S, S, S, S).permute(0, 2, 4, 1, 3, 5)
  assert torch.equal(gram[0].reshape(n1, S, S, n2, S, S),
                     want.float().double())
  for k in (2, 3):
    assert torch.equal(gram[k], base[k])


def test_relu_bwd_rejects_bad_arguments():
  dev = torch.device('cuda', 0)
  x = torch.zeros((1, 8, 8, 8, 8), device=dev)
  q = torch.ones((1, 8, 8), device=dev)
  out = torch.empty_like(x)
  small = torch.empty((16,), dtype=torch.uint8, device=dev)
  gwb = torch.empty((64, 2), dtype=torch.float64, device=dev)
  qq = torch.empty((1, 8, 8), device=dev)
  base = [_lib.ptr(x), None, 1, 0, 8, 2, NO_SELF, 1.0, 0.0, 1.0, 0.0,
          _lib.ptr(q), None, None, None]
  tail = [_lib.ptr(out), None, _lib.ptr(qq), _lib.ptr(qq), _lib.ptr(gwb)]
  with pytest.raises(ValueError, match='workspace'):
    _lib.call('nta_relu_bwd', *base, 0.0, *tail, _lib.ptr(small), 16,
              _lib.stream_handle(dev))
  need = _lib.value('nta_relu_bwd_work_bytes', 1, 0, 8)
  work = torch.empty((need,), dtype=torch.uint8, device=dev)
  with pytest.raises(ValueError, match='cross tile'):
    _lib.call('nta_relu_bwd', *base, 0.5, *tail, _lib.ptr(work), need,
              _lib.stream_handle(dev))
  assert _lib.value('nta_relu_bwd_work_bytes', 2, 3, 32) == 6 * 32 ** 3 * 4
  assert _lib.value('nta_relu_bwd_work_bytes', 2, 3, 6) == 2 * 6 * 6 ** 4 * 4
  assert _lib.value('nta_relu_bwd_parts', 2, 3, 16) == 6 * 16


@pytest.mark.parametrize('n1,n2,S', [(2, 3, 8), (3, 0, 16), (1, 5, 32),
                                     (1, 64, 32), (37, 9, 8)])
def test_relu_bwd_accumulate_matches_torch_sums(n1, n2, S):
  """nta_relu_bwd_accumulate: the partner sums of gqa / gqb and the slot
  sums of gwb added into the accumulators (self tiles: gqa + gqb into
  q1's), and the two accumulators as rows of one array (K(x, x) tiles)."""
  dev = torch.device('cuda', 0)
  g = torch.Generator(device=dev).manual_seed(n1 + n2 + S)
  T = n1 if n2 == 0 else n1 * n2
  gqa = torch.randn((T, S, S), generator=g, device=dev)
  gqb = torch.randn((T, S, S), generator=g, device=dev)
  parts = _lib.value('nta_relu_bwd_parts', n1, n2, S)
  gwb = torch.randn((parts, 2), generator=g, device=dev, dtype=torch.float64)
  wb = torch.randn((2,), generator=g, device=dev, dtype=torch.float64)
  q1 = torch.randn((n1, S, S), generator=g, device=dev)
  q2 = torch.randn((max(n2, 1), S, S), generator=g, device=dev)
  want_wb = wb + gwb.sum(0)
  if n2 == 0:
    want1 = q1 + gqa + gqb
  else:
    want1 = q1 + gqa.reshape(n1, n2, S, S).sum(1)
    want2 = q2 + gqb.reshape(n1, n2, S, S).sum(0)
  _lib.call('nta_relu_bwd_accumulate', _lib.ptr(gqa), _lib.ptr(gqb),
            _lib.ptr(gwb), n1, n2, S, _lib.ptr(wb), _lib.ptr(q1),
            _lib.ptr(q2) if n2 else None, _lib.stream_handle(dev))
  torch.testing.assert_close(wb, want_wb, rtol=1e-12, atol=1e-12)
  torch.testing.assert_close(q1, want1, rtol=1e-5, atol=1e-5)
  if n2:
    torch.testing.assert_close(q2, want2, rtol=1e-5, atol=1e-5)
    # one array, overlapping rows: q1 rows [0, n1), q2 rows [1, 1 + n2)
    Q = torch.randn((1 + max(n1, n2), S, S), generator=g, device=dev)
    want = Q.clone()
    want[:n1] += gqa.reshape(n1, n2, S, S).sum(1)
    want[1:1 + n2] += gqb.reshape(n1, n2, S, S).sum(0)
    _lib.call('nta_relu_bwd_accumulate', _lib.ptr(gqa), _lib.ptr(gqb),
              _lib.ptr(gwb), n1, n2, S, None, _lib.ptr(Q[:n1]),
              _lib.ptr(Q[1:1 + n2]), _lib.stream_handle(dev))
    torch.testing.assert_close(Q, want, rtol=1e-5, atol=1e-5)
