"""Float64 evaluation of the per-element ReLU + pool adjoint that
nta_relu_bwd computes (test helper; stax/elementwise.py:444-455 with the
_sqrt/_arctan2 JVP rules of 1278-1310, stax/linear.py:3551-3571)."""
import math

import torch


def _inputs(n1, n2, S, pool, has_t, seed):
  g = torch.Generator().manual_seed(seed)
  T = n1 if n2 == 0 else n1 * n2
  q1 = 0.5 + torch.rand((n1, S, S), generator=g, dtype=torch.float64)
  q2 = q1 if n2 == 0 else 0.5 + torch.rand((n2, S, S), generator=g,
                                           dtype=torch.float64)
  rho = 1.8 * torch.rand((T, S, S, S, S), generator=g,
                         dtype=torch.float64) - 0.9
  So = S // pool
  gN = torch.randn((T, So, So, So, So), generator=g, dtype=torch.float64)
  gT = torch.randn((T, So, So, So, So), generator=g, dtype=torch.float64)
  aT = torch.randn((T, S, S, S, S), generator=g,
                   dtype=torch.float64) if has_t else None
  return q1, q2, rho, gN, gT, aT


def _expand_q(q1, q2, n1, n2, S):
  """qa, qb per element [T, h, h', w, w']."""
  if n2 == 0:
    qa = q1[:, :, None, :, None]
    qb = q1[:, None, :, None, :]
  else:
    qa = q1.repeat_interleave(n2, 0)[:, :, None, :, None]
    qb = q2.repeat(n1, 1, 1)[:, None, :, None, :]
  shp = (qa.shape[0], S, S, S, S)
  return qa.expand(shp), qb.expand(shp)


def _reference(aN, aT, qa, qb, gN, gT, pool, same, w, b, k1, k0):
  c = w * aN + b
  t = c if aT is None else w * aT + c
  s = torch.sqrt(torch.clamp(qa * qb - c * c, min=0.0))
  s = torch.where(same, torch.zeros_like(s), s)
  theta = torch.where((s == 0) & (c == 0), torch.full_like(s, math.pi / 2),
                      torch.atan2(s, c))
  ds = k1 * (math.pi - theta) / (2 * math.pi) + k0
  up = lambda g: g.repeat_interleave(pool, 1).repeat_interleave(
      pool, 2).repeat_interleave(pool, 3).repeat_interleave(
          pool, 4) / pool ** 4
  gj, gtd = up(gN), up(gT)
  rs = torch.where(s > 0, 1.0 / torch.where(s > 0, s, 1.0), 0.0)
  gt = gtd * ds
  gc = gj * ds + gtd * t * k1 / (2 * math.pi) * rs
  gct = gc + gt
  v = gj * s - gtd * t * c * rs
  rq = lambda q: torch.where(q > 0, k1 / (4 * math.pi * q), 0.0)
  gqa = (rq(qa) * v).sum(dim=(2, 4))                 # over (h', w')
  gqb = (rq(qb) * v).sum(dim=(1, 3))                 # over (h, w)
  at = torch.zeros_like(aN) if aT is None else aT
  gw = float((gct * aN + gt * at).sum())
  gb = float(gct.sum())
  return w * gct, w * gt, gqa, gqb, gw, gb


