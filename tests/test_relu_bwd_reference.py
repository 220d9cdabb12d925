"""The float64 adjoint evaluation used by the nta_relu_bwd tests against
torch autograd of the forward stage map (ABRelu arc-cosine map of
stax/elementwise.py:444-455, then the p x p x p x p average pool): its
cotangents of aN, aT, q1, q2, w_var and b_var."""
import math

import torch

from _relu_bwd_ref import _expand_q, _inputs, _reference


def _forward_loss(aN, aT, q1, q2, w, b, gN, gT, n1, n2, S, pool, k1, k0):
  qa, qb = _expand_q(q1, q2, n1, n2, S)
  c = w * aN + b
  t = w * aT + c
  s = torch.sqrt(torch.clamp(qa * qb - c * c, min=0.0))
  th = torch.atan2(s, c)
  J = (s + (math.pi - th) * c) / (2 * math.pi)
  ds = (math.pi - th) / (2 * math.pi)
  So = S // pool

  def pooled(x):
    return x.reshape(x.shape[0], So, pool, So, pool, So, pool, So,
                     pool).mean(dim=(2, 4, 6, 8))
  return (pooled(k1 * J + k0 * c) * gN).sum() + \
      (pooled(t * (k1 * ds + k0)) * gT).sum()


def test_reference_matches_autograd_of_the_forward_map():
  n1, n2, S, pool = 2, 3, 4, 2
  w, b, k1, k0 = 1.7, 0.2, 0.7, 0.3        # an ABRelu with k0 != 0
  q1, q2, rho, gN, gT, aT = _inputs(n1, n2, S, pool, True, seed=3)
  qa, qb = _expand_q(q1, q2, n1, n2, S)
  aN = (rho * torch.sqrt(qa * qb) - b) / w
  same = torch.zeros(aN.shape, dtype=torch.bool)
  ref = _reference(aN, aT, qa, qb, gN, gT, pool, same, w, b, k1, k0)
  x = [aN.clone().requires_grad_(), aT.clone().requires_grad_(),
       q1.clone().requires_grad_(), q2.clone().requires_grad_(),
       torch.tensor(w, dtype=torch.float64, requires_grad=True),
       torch.tensor(b, dtype=torch.float64, requires_grad=True)]
  _forward_loss(*x, gN, gT, n1, n2, S, pool, k1, k0).backward()
  torch.testing.assert_close(ref[0], x[0].grad, rtol=1e-12, atol=1e-14)
  torch.testing.assert_close(ref[1], x[1].grad, rtol=1e-12, atol=1e-14)
  # per-pair normaliser terms summed over the pairs of each image
  torch.testing.assert_close(ref[2].reshape(n1, n2, S, S).sum(1), x[2].grad,
                             rtol=1e-12, atol=1e-14)
  torch.testing.assert_close(ref[3].reshape(n1, n2, S, S).sum(0), x[3].grad,
                             rtol=1e-12, atol=1e-14)
  assert abs(ref[4] - float(x[4].grad)) <= 1e-12 * abs(ref[4])
  assert abs(ref[5] - float(x[5].grad)) <= 1e-12 * abs(ref[5])
