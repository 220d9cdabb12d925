"""Pins of the FP64 oracle against independent facts (CPU only).

Each test checks an oracle function against something that does not reuse
its formula: brute-force loops, the patch-Gram identity, textbook Gaussian
integrals, the reference's own known answer (AvgPool), relational identities
of the reference's tests, and library routines (scipy / numpy).
"""
import math

import numpy as np
import pytest
import scipy.integrate
import scipy.special
import scipy.stats

from conftest import rel_err
from oracle import predict_oracle as po
from oracle import stax_oracle as so


# ------------------------------------------------------------ Conv stencil

def _brute_conv(nngp, fh, fw, padding):
  """Direct 4-D loop: out[h,h',w,w'] = mean over taps (k, l) of
  in[h+k+o, h'+k+o, w+l+o, w'+l+o], zero outside; natural axis order."""
  n1, n2, H, _, W, _ = nngp.shape
  circ = padding == 'CIRCULAR'               # taps wrap around
  if padding in ('SAME', 'CIRCULAR'):
    oh, ow, Hc, Wc = -((fh - 1) // 2), -((fw - 1) // 2), H, W
  else:
    oh, ow, Hc, Wc = 0, 0, H - fh + 1, W - fw + 1
  out = np.zeros((n1, n2, Hc, Hc, Wc, Wc))
  for h in range(Hc):
    for hp in range(Hc):
      for w in range(Wc):
        for wp in range(Wc):
          acc = 0.0
          for k in range(fh):
            a, ap = h + k + oh, hp + k + oh
            if circ:
              a, ap = a % H, ap % H
            if not (0 <= a < H and 0 <= ap < H):
              continue
            for l in range(fw):
              b, bp = w + l + ow, wp + l + ow
              if circ:
                b, bp = b % W, bp % W
              if not (0 <= b < W and 0 <= bp < W):
                continue
              acc = acc + nngp[:, :, a, ap, b, bp]
          out[:, :, h, hp, w, wp] = acc / (fh * fw)
  return out


@pytest.mark.parametrize('fh,fw,padding', [(3, 3, 'SAME'), (2, 3, 'SAME'),
                                           (3, 2, 'VALID'), (1, 3, 'SAME'),
                                           (3, 3, 'CIRCULAR'),
                                           (2, 3, 'CIRCULAR'),
                                           (5, 3, 'CIRCULAR')])
def test_conv_stencil_bruteforce(fh, fw, padding):
  rng = np.random.default_rng(0)
  x = rng.standard_normal((2, 3, 5, 5, 4, 4))
  got = so.conv_kernel_full_spatial_shared(x, (fh, fw), (1, 1), padding, 2)
  want = _brute_conv(x, fh, fw, padding)
  # the reference leaves the spatial order reversed: [.., w, w', h, h']
  np.testing.assert_allclose(got, np.transpose(want, (0, 1, 4, 5, 2, 3)),
                             rtol=1e-13, atol=1e-13)


def test_patch_gram_identity():
  """stencil(input Gram) = Gram of zero-padded 3x3xC patches / (9C)."""
  rng = np.random.default_rng(1)
  x1 = rng.standard_normal((2, 5, 6, 3))
  x2 = rng.standard_normal((3, 5, 6, 3))
  k = so.inputs_to_kernel(x1, x2)
  got = so.conv_kernel_full_spatial_shared(k.nngp, (3, 3), (1, 1), 'SAME', 2)

  def patches(x):
    xp = np.pad(x, ((0, 0), (1, 1), (1, 1), (0, 0)))
    p = np.stack([xp[:, a:a + 5, e:e + 6, :] for a in range(3)
                  for e in range(3)], axis=3)
    return p.reshape(x.shape[0], 5, 6, -1)

  p1, p2 = patches(x1), patches(x2)
  want = np.einsum('ihwk,jgvk->ijhgwv', p1, p2) / 27.0
  np.testing.assert_allclose(got, np.transpose(want, (0, 1, 4, 5, 2, 3)),
                             rtol=1e-12, atol=1e-12)


def test_input_kernel_einsum_and_shapes():
  """requirements.py:668-695 docstring shapes and the closed-form einsum."""
  x = np.ones((10, 32, 16, 3))
  k = so.inputs_to_kernel(x, None)
  assert k.cov1.shape == (10, 32, 32, 16, 16)
  assert k.nngp.shape == (10, 10, 32, 32, 16, 16)
  rng = np.random.default_rng(2)
  x1 = rng.standard_normal((2, 3, 4, 5))
  x2 = rng.standard_normal((3, 3, 4, 5))
  k = so.inputs_to_kernel(x1, x2)
  for i in range(2):
    for j in range(3):
      for h in range(3):
        for hp in range(3):
          for w in range(4):
            for wp in range(4):
              v = np.dot(x1[i, h, w], x2[j, hp, wp]) / 5
              assert abs(k.nngp[i, j, h, hp, w, wp] - v) < 1e-12
  np.testing.assert_allclose(k.cov1[1, 2, 0, 3, 1],
                             np.dot(x1[1, 2, 3], x1[1, 0, 1]) / 5)


# ------------------------------------------------------------------ ReLU

def _gauss_phi_expectations(q1, q2, c, a, b):
  """E[phi(u) phi(v)] and E[phi'(u) phi'(v)] for phi = a min(x,0) +
  b max(x,0), (u, v) ~ N(0, [[q1, c], [c, q2]]), by 1-D quadrature of the
  textbook conditional moments of v | u."""
  beta = c / q1
  sig = math.sqrt(max(q2 - c * c / q1, 0.0))

  def e_relu(mu):
    if sig == 0:
      return max(mu, 0.0)
    z = mu / sig
    return mu * scipy.stats.norm.cdf(z) + sig * scipy.stats.norm.pdf(z)

  def e_step(mu):
    if sig == 0:
      return 1.0 if mu > 0 else 0.0
    return scipy.stats.norm.cdf(mu / sig)

  def phi(u):
    return a * min(u, 0.0) + b * max(u, 0.0)

  su = math.sqrt(q1)

  def dens(u):
    return scipy.stats.norm.pdf(u / su) / su

  def f_nngp(u):
    mu = beta * u
    return phi(u) * ((b - a) * e_relu(mu) + a * mu) * dens(u)

  def f_ntk(u):
    mu = beta * u
    du = a if u < 0 else b
    return du * (a + (b - a) * e_step(mu)) * dens(u)

  lim = 12 * su
  kw = dict(limit=400, epsabs=1e-13, epsrel=1e-12)
  en = sum(scipy.integrate.quad(f_nngp, lo, hi, **kw)[0]
           for lo, hi in ((-lim, 0.0), (0.0, lim)))
  et = sum(scipy.integrate.quad(f_ntk, lo, hi, **kw)[0]
           for lo, hi in ((-lim, 0.0), (0.0, lim)))
  return en, et


@pytest.mark.parametrize('a,b', [(0.0, 1.0), (-1.0, 1.0), (0.2, 1.0),
                                 (0.5, -0.7)])
@pytest.mark.parametrize('rho', [-0.95, -0.4, 0.0, 0.3, 0.8, 0.999])
def test_abrelu_matches_gaussian_integrals(a, b, rho):
  q1, q2 = 1.7, 0.6
  c = rho * math.sqrt(q1 * q2)
  nngp = np.array([c])
  prod = np.array([q1 * q2])
  got_n, got_d = so.relu_nngp_ntk(nngp, prod, np.array([1.0]), a, b)
  want_n, want_d = _gauss_phi_expectations(q1, q2, c, a, b)
  assert abs(got_n[0] - want_n) <= 1e-8 * max(1.0, abs(want_n))
  assert abs(got_d[0] - want_d) <= 1e-8


def test_relu_closed_forms_and_special_points():
  """Arc-cosine kernel of degree 1 (Cho & Saul 2009) at rho = 1, 0, -1 and
  the fill_zero=pi/2 convention (elementwise.py:446, 1293-1298)."""
  P = 2.5
  sq = math.sqrt(P)
  c = np.array([sq, 0.0, -sq, 0.0])
  prod = np.array([P, P, P, 0.0])
  n, d = so.relu_nngp_ntk(c, prod, np.ones(4))
  np.testing.assert_allclose(n[:3], [sq / 2, sq / (2 * math.pi), 0.0],
                             atol=1e-15)
  np.testing.assert_allclose(d[:3], [0.5, 0.25, 0.0], atol=1e-15)
  assert n[3] == 0.0 and d[3] == 0.25          # theta = pi/2 fill
  # prod < nngp^2 (clamped by max(., 0)): s = 0, theta = 0 or pi.
  n, d = so.relu_nngp_ntk(np.array([2.0, -2.0]), np.array([3.0, 3.0]),
                          np.ones(2))
  np.testing.assert_allclose(n, [1.0, 0.0])
  np.testing.assert_allclose(d, [0.5, 0.0])


def test_abrelu_identities():
  """elementwise_test.py:537-637: ABRelu(a, a) == a^2 * Identity on the
  nngp and ABRelu(0, 1) == Relu, on a Conv network."""
  rng = np.random.default_rng(3)
  x1 = rng.standard_normal((2, 4, 4, 3))
  x2 = rng.standard_normal((3, 4, 4, 3))
  conv = so.conv_layer((3, 3), (1, 1), 'SAME', 1.3, 0.2)
  base = so.kernel_fn([conv], x1, x2, ('nngp', 'ntk', 'cov1'))
  lin = so.kernel_fn([conv, so.abrelu_layer(0.7, 0.7)], x1, x2,
                     ('nngp', 'ntk', 'cov1'))
  np.testing.assert_allclose(lin['nngp'], 0.49 * base['nngp'], rtol=1e-12)
  np.testing.assert_allclose(lin['cov1'], 0.49 * base['cov1'], rtol=1e-12)
  np.testing.assert_allclose(lin['ntk'], 0.49 * base['ntk'], rtol=1e-12)
  r1 = so.kernel_fn([conv, so.relu_layer()], x1, x2, 'nngp')
  r2 = so.kernel_fn([conv, so.abrelu_layer(0.0, 1.0)], x1, x2, 'nngp')
  np.testing.assert_array_equal(r1['nngp'], r2['nngp'])


def test_relu_requires_gaussian_input():
  x = np.ones((1, 3, 3, 2))
  with pytest.raises(ValueError):
    so.kernel_fn([so.relu_layer()], x, None)


# ------------------------------------------------------------------ pools

def test_avg_pool_known_answer():
  """stax_test.py:494-539 (AvgPool SAME 2x2/1 on all-ones inputs)."""
  X1 = np.ones((4, 2, 3, 2))
  X2 = np.ones((3, 2, 3, 2))
  ker = so.kernel_fn([so.avg_pool_layer((2, 2), (1, 1), 'SAME', False)], X1,
                     X2, None)
  ker_norm = so.kernel_fn([so.avg_pool_layer((2, 2), (1, 1), 'SAME', True)],
                          X1, X2, None)
  np.testing.assert_allclose(ker_norm.nngp, np.ones_like(ker_norm.nngp))
  np.testing.assert_allclose(ker_norm.cov1, np.ones_like(ker_norm.cov1))
  np.testing.assert_allclose(ker_norm.cov2, np.ones_like(ker_norm.cov2))
  out_unnorm = np.array([[1., 1., 0.5], [0.5, 0.5, 0.25]])
  ker_unnorm = np.outer(out_unnorm, out_unnorm).reshape((2, 3, 2, 3))
  ker_unnorm = np.transpose(ker_unnorm, axes=(0, 2, 1, 3))
  np.testing.assert_allclose(
      ker.nngp, np.broadcast_to(ker_unnorm[None, None], ker.nngp.shape))
  np.testing.assert_allclose(
      ker.cov1, np.broadcast_to(ker_unnorm[None], ker.cov1.shape))
  np.testing.assert_allclose(
      ker.cov2, np.broadcast_to(ker_unnorm[None], ker.cov2.shape))


def test_valid_avg_pool_is_block_mean():
  rng = np.random.default_rng(4)
  x = rng.standard_normal((2, 3, 4, 4, 6, 6))
  got = so.pool_kernel(x, (2, 2, 2, 2), (2, 2, 2, 2), 'VALID', False, 2)
  want = x.reshape(2, 3, 2, 2, 2, 2, 3, 2, 3, 2).mean(axis=(3, 5, 7, 9))
  np.testing.assert_allclose(got, want, rtol=1e-13)


def test_sum_pools_are_window_sums():
  """SumPool is _pool_kernel without _normalize (linear.py:3551-3556):
  brute-force window sums of a random kernel, SAME padding included;
  GlobalSumPool is the plain sum (linear.py:1751)."""
  rng = np.random.default_rng(5)
  x = rng.standard_normal((2, 3, 4, 4, 6, 6))
  got = so.pool_kernel(x, (2, 2, 2, 2), (2, 2, 2, 2), 'VALID', False, 2,
                       pool_sum=True)
  want = np.zeros((2, 3, 2, 2, 3, 3))
  for a in range(2):
    for b in range(2):
      for c in range(3):
        for d in range(3):
          want[:, :, a, b, c, d] = x[:, :, 2 * a:2 * a + 2, 2 * b:2 * b + 2,
                                     2 * c:2 * c + 2, 2 * d:2 * d + 2].sum(
                                         axis=(2, 3, 4, 5))
  np.testing.assert_allclose(got, want, rtol=1e-13)
  # SAME 3x3 / 1: zero padded sums
  y = rng.standard_normal((1, 1, 3, 3))
  got = so.pool_kernel(y, (3, 3), (1, 1), 'SAME', False, 2, pool_sum=True)
  yp = np.pad(y, ((0, 0), (0, 0), (1, 1), (1, 1)))
  want = np.array([[yp[0, 0, i:i + 3, j:j + 3].sum() for j in range(3)]
                   for i in range(3)])
  np.testing.assert_allclose(got[0, 0], want, rtol=1e-13)
  k = so.OKernel(nngp=x, ntk=None, cov1=x[:, 0], cov2=None, x1_is_x2=False,
                 is_gaussian=True, is_reversed=False, diagonal_batch=True,
                 diagonal_spatial=False, shape1=None, shape2=None)
  out = so.global_avg_pool_layer(pool_sum=True)(k)
  np.testing.assert_allclose(out.nngp, x.sum(axis=(2, 3, 4, 5)), rtol=1e-13)


def test_gap_flatten_readout_equivalence():
  """Notebook readout [AvgPool x3 -> 1x1, Flatten, Dense(sqrt 2)] equals
  [AvgPool -> 4x4, GlobalAvgPool, Dense(1)] times W^2 = 2 (both are the
  mean over the 8x8 stage-3 grid)."""
  rng = np.random.default_rng(5)
  x1 = rng.standard_normal((2, 8, 8, 3))
  x2 = rng.standard_normal((2, 8, 8, 3))
  body = [so.conv_layer((3, 3), (1, 1), 'SAME', 2 ** 0.5, 0.0),
          so.relu_layer()]
  nb = body + [so.avg_pool_layer((2, 2), (2, 2), 'VALID')] * 3 + [
      so.flatten_layer(), so.dense_layer(2 ** 0.5, 0.0)]
  cf = body + [so.avg_pool_layer((2, 2), (2, 2), 'VALID'),
               so.global_avg_pool_layer(), so.dense_layer(1.0, 0.0)]
  a = so.kernel_fn(nb, x1, x2)
  b = so.kernel_fn(cf, x1, x2)
  np.testing.assert_allclose(a['nngp'], 2 * b['nngp'], rtol=1e-12)
  np.testing.assert_allclose(a['ntk'], 2 * b['ntk'], rtol=1e-12)


def test_full_vs_diagonal_spatial():
  """requirements_test.py:87-112: the diagonal-spatial path equals the
  spatial diagonal ('...iijj->...ij') of the full-spatial nngp."""
  rng = np.random.default_rng(6)
  x1 = rng.standard_normal((2, 5, 6, 3))
  x2 = rng.standard_normal((3, 5, 6, 3))
  layers = [so.conv_layer((2, 3), (1, 1), 'SAME', 1.1, 0.3), so.relu_layer(),
            so.conv_layer((2, 3), (1, 1), 'SAME', 0.9, 0.1)]
  full = so.kernel_fn(layers, x1, x2, ('nngp', 'ntk'))
  diag = so.kernel_fn(layers, x1, x2, ('nngp', 'ntk'), diagonal_spatial=True)
  for g in ('nngp', 'ntk'):
    # two convs: spatial order back to (h, h', w, w')
    np.testing.assert_allclose(diag[g],
                               np.einsum('...iijj->...ij', full[g]),
                               rtol=1e-12, atol=1e-14)


# ----------------------------------------------------------- predict

def test_cho_solve_matches_library_solve():
  rng = np.random.default_rng(7)
  A = rng.standard_normal((30, 30))
  K = A @ A.T / 30
  y = rng.standard_normal((30, 4))
  Kt = rng.standard_normal((5, 30))
  mean, alpha = po.gp_inference_mean(K, y, Kt, diag_reg=1e-3)
  lam = 1e-3 * np.trace(K) / 30
  want = np.linalg.solve(K + lam * np.eye(30), y)
  np.testing.assert_allclose(alpha, want, rtol=1e-10)
  np.testing.assert_allclose(mean, Kt @ want, rtol=1e-10)
  mean_abs, _ = po.gp_inference_mean(K, y, Kt, diag_reg=0.05,
                                     diag_reg_absolute_scale=True)
  np.testing.assert_allclose(
      mean_abs, Kt @ np.linalg.solve(K + 0.05 * np.eye(30), y), rtol=1e-10)


def test_finite_t_limits():
  """predict_test.py:388-484: t -> inf equals t=None; t=0 gives zero
  test mean (zero-mean prior) and zero train change."""
  rng = np.random.default_rng(8)
  A = rng.standard_normal((20, 20))
  K = A @ A.T / 20 + np.eye(20)
  y = rng.standard_normal((20, 3))
  Kt = rng.standard_normal((4, 20))
  inf, _ = po.gp_inference_mean(K, y, Kt, diag_reg=1e-2)
  big = po.ensemble_mean_finite_t(K, y, Kt, t=1e9, diag_reg=1e-2)
  np.testing.assert_allclose(big, inf, rtol=1e-7, atol=1e-9)
  zero = po.ensemble_mean_finite_t(K, y, Kt, t=0.0, diag_reg=1e-2)
  np.testing.assert_allclose(zero, 0.0, atol=1e-14)
  train_big = po.ensemble_mean_finite_t(K, y, None, t=1e9)
  np.testing.assert_allclose(train_big, y, rtol=1e-7, atol=1e-9)


def test_fast_arccos_polynomial_bound():
  """acos_pos5 (csrc/nta_common.h), evaluated in FP32 with fused
  multiply-adds, is within 9e-7 rad of arccos on [0, 1] (its coefficients are
  read from the source)."""
  import os
  import re
  src = open(os.path.join(os.path.dirname(__file__), '..',
                          'neural_tangents_amd', 'csrc',
                          'nta_common.h')).read()
  body = src[src.index('float acos_pos5('):]
  body = body[:body.index('return r * p;')]
  coef = [float(x) for x in re.findall(r'(-?\d\.\d+e[-+]\d+)f', body)]
  assert len(coef) == 6                    # highest degree first
  a = np.linspace(0.0, 1.0, 400001).astype(np.float32)
  p = np.full_like(a, np.float32(coef[0]))
  for c in coef[1:]:                       # fma: one rounding per step
    p = (p.astype(np.float64) * a + np.float32(c)).astype(np.float32)
  u = (np.float32(1.0) - a).astype(np.float32)
  r = np.sqrt(u.astype(np.float64)).astype(np.float32)
  th = (r.astype(np.float64) * p).astype(np.float32).astype(np.float64)
  err = np.abs(th - np.arccos(a.astype(np.float64)))
  assert err.max() <= 9e-7, err.max()
  assert th[-1] == 0.0                     # arccos(1) = 0 exactly


def _gauss2(q1, q2, c, g, n=96):
  """E[g(u, v)] for (u, v) ~ N(0, [[q1, c], [c, q2]]) by tensor
  Gauss-Hermite quadrature (probabilists' weights)."""
  z, w = np.polynomial.hermite_e.hermegauss(n)
  w = w / np.sqrt(2 * np.pi)
  z1, z2 = np.meshgrid(z, z, indexing='ij')
  rho = c / np.sqrt(q1 * q2)
  u = np.sqrt(q1) * z1
  v = np.sqrt(q2) * (rho * z1 + np.sqrt(max(1 - rho ** 2, 0.0)) * z2)
  return float(np.sum(w[:, None] * w[None, :] * g(u, v)))


_ACT_POINTS = [(1.0, 1.0, 0.0), (0.7, 1.3, 0.5), (2.0, 0.5, -0.9),
               (0.3, 0.4, 0.34), (1.5, 1.5, 1.499)]


@pytest.mark.parametrize('q1,q2,c', _ACT_POINTS)
def test_erf_map_matches_gaussian_integrals(q1, q2, c):
  """Williams' arcsine kernel: E[erf(u) erf(v)] and E[erf'(u) erf'(v)]."""
  nngp, ntk = so.erf_nngp_ntk(np.array([c]), np.array([(1 + 2 * q1) *
                                                       (1 + 2 * q2)]),
                              np.array([1.0]))
  e = _gauss2(q1, q2, c, lambda u, v: scipy.special.erf(u) *
              scipy.special.erf(v))
  d = _gauss2(q1, q2, c, lambda u, v: 4 / np.pi * np.exp(-u ** 2 - v ** 2))
  assert abs(nngp[0] - e) <= 1e-10, (nngp[0], e)
  assert abs(ntk[0] - d) <= 1e-10, (ntk[0], d)


@pytest.mark.parametrize('q1,q2,c', _ACT_POINTS)
def test_gelu_map_matches_gaussian_integrals(q1, q2, c):
  """E[gelu(u) gelu(v)] and E[gelu'(u) gelu'(v)], gelu(x) = x Phi(x)."""
  Phi = scipy.stats.norm.cdf
  phi = scipy.stats.norm.pdf
  nngp, ntk = so.gelu_nngp_ntk(np.array([c]), np.array([q1 * q2]),
                               np.array([(q1 + 1) * (q2 + 1)]),
                               np.array([1.0]))
  e = _gauss2(q1, q2, c, lambda u, v: u * Phi(u) * v * Phi(v))
  d = _gauss2(q1, q2, c, lambda u, v: (Phi(u) + u * phi(u)) *
              (Phi(v) + v * phi(v)))
  assert abs(nngp[0] - e) <= 1e-9, (nngp[0], e)
  assert abs(ntk[0] - d) <= 1e-9, (ntk[0], d)


def test_erf_gelu_layers_scales_and_diagonal_forms():
  """Erf(a, b, c) = a^2 map(b^2 k) + c^2 on nngp/cov (ntk: a^2 b^2 dsigma);
  the diagonal_spatial self forms equal the general map at c = q1 = q2."""
  rng = np.random.default_rng(3)
  x1 = rng.standard_normal((2, 3, 3, 2))
  x2 = rng.standard_normal((2, 3, 3, 2))
  conv = so.conv_layer((3, 3), (1, 1), 'SAME', 1.3, 0.2)
  base = so.kernel_fn([conv], x1, x2, ('nngp', 'ntk', 'cov1', 'cov2'))
  a, b, c = 0.7, 1.6, 0.3
  got = so.kernel_fn([conv, so.erf_layer(a, b, c)], x1, x2,
                     ('nngp', 'ntk', 'cov1'))
  q1 = np.einsum('ihhww->ihw', base['cov1'])
  q2 = np.einsum('ihhww->ihw', base['cov2'])
  # entry (i, j, h, h', w, w'): q1[i, h, w], q2[j, h', w']
  P = (1 + 2 * b * b * q1[:, None, :, None, :, None]) * (
      1 + 2 * b * b * q2[None, :, None, :, None, :])
  n_ref, t_ref = so.erf_nngp_ntk(b * b * base['nngp'], P,
                                 b * b * base['ntk'])
  np.testing.assert_allclose(got['nngp'], a * a * n_ref + c * c, rtol=1e-12)
  np.testing.assert_allclose(got['ntk'], a * a * t_ref, rtol=1e-12)
  # the two layouts differ by the conv's spatial-axis reversal: take the
  # orientation from the conv alone
  full0 = so.kernel_fn([conv], x1, None, ('cov1',))['cov1']
  diag0 = so.kernel_fn([conv], x1, None, ('cov1',),
                       diagonal_spatial=True)['cov1']
  ext = np.einsum('ihhww->ihw', full0)
  flip = not np.allclose(ext, diag0, rtol=1e-12)
  if flip:
    np.testing.assert_allclose(ext.transpose(0, 2, 1), diag0, rtol=1e-12)
  for layer, fn in ((so.erf_layer(1.0, 1.0, 0.0), 'erf'),
                    (so.gelu_layer(), 'gelu')):
    full = so.kernel_fn([conv, layer], x1, None, ('cov1',))['cov1']
    diag = so.kernel_fn([conv, layer], x1, None, ('cov1',),
                        diagonal_spatial=True)['cov1']
    ext = np.einsum('ihhww->ihw', full)
    np.testing.assert_allclose(ext.transpose(0, 2, 1) if flip else ext,
                               diag, rtol=1e-12, err_msg=fn)


def test_circular_full_vs_diagonal_spatial():
  """A Flatten-readout network with CIRCULAR convs: the spatial diagonal of
  the full-spatial kernel equals the diagonal-spatial computation (a check
  of the wrap-padded window sums against the wrap-padded 4-D stencil)."""
  rng = np.random.default_rng(5)
  x1 = rng.standard_normal((2, 5, 4, 2))
  x2 = rng.standard_normal((3, 5, 4, 2))
  layers = [so.conv_layer((3, 3), (1, 1), 'CIRCULAR', 1.2, 0.1),
            so.relu_layer(), so.conv_layer((3, 2), (1, 1), 'CIRCULAR'),
            so.relu_layer(), so.flatten_layer()]
  full = so.kernel_fn(layers, x1, x2, ('nngp', 'ntk'))
  diag = so.kernel_fn(layers, x1, x2, ('nngp', 'ntk'), diagonal_spatial=True)
  for g in ('nngp', 'ntk'):
    np.testing.assert_allclose(full[g], diag[g], rtol=1e-12)


# ------------------------------------------ posterior covariance, finite t

def _spd(rng, n, shift=0.5):
  A = rng.standard_normal((n, n))
  return A @ A.T / n + shift * np.eye(n)


def _joint(rng, nd, nt):
  """A joint (test + train) NNGP covariance and an NTK of the same shape."""
  Kj = _spd(rng, nd + nt, 0.3)
  Tj = _spd(rng, nd + nt, 0.6) + 0.5 * Kj
  sl_t, sl_d = slice(0, nt), slice(nt, nt + nd)
  return Kj, Tj, sl_t, sl_d


def test_gp_inference_cov_nngp_is_schur_complement():
  """NNGP posterior covariance = inverse of the test block of the inverse
  joint covariance (the Schur complement), with the diagonal regulariser on
  the train block; numpy block inverses, not a solve."""
  rng = np.random.default_rng(11)
  nd, nt = 12, 5
  Kj, _, st, sd = _joint(rng, nd, nt)
  reg = 1e-2
  lam = reg * np.trace(Kj[sd, sd]) / nd
  J = Kj.copy()
  J[sd, sd] += lam * np.eye(nd)
  want = np.linalg.inv(np.linalg.inv(J)[st, st])
  got = po.gp_inference_cov(Kj[sd, sd], Kj[sd, sd], Kj[st, sd], Kj[st, sd],
                            Kj[st, st], 'nngp', diag_reg=reg)
  np.testing.assert_allclose(got, want, rtol=1e-9, atol=1e-11)


def test_gp_inference_cov_ntk_is_linear_map_covariance():
  """NTK posterior covariance = Cov[f_inf] for the infinite-time
  gradient-descent output f_inf = f0_test - T_td T_dd^-1 (f0_train - y) with
  f0 ~ N(0, joint NNGP): M J M^T for M = [I, -T_td T_dd^-1].  With the NTK
  equal to the NNGP (and no regulariser) it reduces to the NNGP branch."""
  rng = np.random.default_rng(12)
  nd, nt = 10, 4
  Kj, Tj, st, sd = _joint(rng, nd, nt)
  reg = 3e-3
  lam = reg * np.trace(Tj[sd, sd]) / nd
  Treg = Tj[sd, sd] + lam * np.eye(nd)
  M = np.concatenate([np.eye(nt), -np.linalg.solve(Treg, Tj[sd, st]).T],
                     axis=1)
  Jn = np.block([[Kj[st, st], Kj[st, sd]], [Kj[sd, st], Kj[sd, sd]]])
  want = M @ Jn @ M.T
  got = po.gp_inference_cov(Kj[sd, sd], Tj[sd, sd], Kj[st, sd], Tj[st, sd],
                            Kj[st, st], 'ntk', diag_reg=reg)
  np.testing.assert_allclose(got, want, rtol=1e-9, atol=1e-11)
  same_n = po.gp_inference_cov(Kj[sd, sd], Kj[sd, sd], Kj[st, sd],
                               Kj[st, sd], Kj[st, st], 'nngp')
  same_t = po.gp_inference_cov(Kj[sd, sd], Kj[sd, sd], Kj[st, sd],
                               Kj[st, sd], Kj[st, st], 'ntk')
  np.testing.assert_allclose(same_t, same_n, rtol=1e-9, atol=1e-11)


def test_finite_t_mean_solves_gradient_flow():
  """Finite-t mean = solution of the gradient-flow ODE of the MSE loss,
  df_train/dt = -(lr / y.size) A (f_train - y), df_test/dt = -(lr / y.size)
  K_td (f_train - y), f(0) = 0 (scipy solve_ivp, no eigendecomposition), at
  three times, with the regulariser and a learning rate."""
  import scipy.integrate as si
  rng = np.random.default_rng(13)
  nd, nt, k = 9, 4, 3
  Kd = _spd(rng, nd)
  Ktd = rng.standard_normal((nt, nd)) * 0.5
  y = rng.standard_normal((nd, k))
  reg, lr = 1e-2, 0.7
  A = Kd + reg * np.trace(Kd) / nd * np.eye(nd)
  size = y.size

  def rhs(_, z):
    f = z[:nd * k].reshape(nd, k)
    d = -(lr / size) * A @ (f - y)
    dt = -(lr / size) * Ktd @ (f - y)
    return np.concatenate([d.ravel(), dt.ravel()])

  ts = [3.0, 40.0, 400.0]
  sol = si.solve_ivp(rhs, (0, ts[-1]), np.zeros(nd * k + nt * k),
                     t_eval=ts, rtol=1e-11, atol=1e-13, method='DOP853')
  for i, t in enumerate(ts):
    f_tr = sol.y[:nd * k, i].reshape(nd, k)
    f_te = sol.y[nd * k:, i].reshape(nt, k)
    got_te = po.ensemble_mean_finite_t(Kd, y, Ktd, t, learning_rate=lr,
                                       diag_reg=reg)
    got_tr = po.ensemble_mean_finite_t(Kd, y, None, t, learning_rate=lr,
                                       diag_reg=reg)
    np.testing.assert_allclose(got_te, f_te, rtol=1e-7, atol=1e-9)
    np.testing.assert_allclose(got_tr, f_tr, rtol=1e-7, atol=1e-9)


@pytest.mark.parametrize('g', ['nngp', 'ntk'])
def test_finite_t_cov_is_linear_map_covariance(g):
  """Finite-t covariance = Cov[f(t)] of the gradient-flow output, a linear
  map of the prior sample f0 ~ N(0, joint NNGP) (train block regularised
  when the NNGP drives training): f_train(t) = E f0_d + (I - E) y,
  f_test(t) = f0_t - T_td T^-1 (I - E)(f0_d - y), E = expm(-T t lr / size)
  (scipy.linalg.expm, no eigendecomposition)."""
  import scipy.linalg as sl
  rng = np.random.default_rng(14 if g == 'nngp' else 15)
  nd, nt = 8, 3
  Kj, Tj, st, sd = _joint(rng, nd, nt)
  if g == 'nngp':
    Tj = Kj
  reg, lr, size = 2e-2, 1.3, nd * 2
  lam = reg * np.trace(Tj[sd, sd]) / nd
  T = Tj[sd, sd] + lam * np.eye(nd)
  Kdd = Kj[sd, sd] + (lam * np.eye(nd) if g == 'nngp' else 0.0)
  for t in (0.5, 6.0, 60.0):
    E = sl.expm(-T * t * lr / size)
    got_tr = po.ensemble_cov_finite_t(Tj[sd, sd], Kj[sd, sd], None, None,
                                      None, t, g, size, lr, reg)
    np.testing.assert_allclose(got_tr, E @ Kdd @ E.T, rtol=1e-8, atol=1e-11)
    M = np.concatenate(
        [np.eye(nt), -Tj[st, sd] @ np.linalg.solve(T, np.eye(nd) - E)],
        axis=1)
    J = np.block([[Kj[st, st], Kj[st, sd]], [Kj[sd, st], Kdd]])
    got_te = po.ensemble_cov_finite_t(Tj[sd, sd], Kj[sd, sd], Tj[st, sd],
                                      Kj[st, sd], Kj[st, st], t, g, size, lr,
                                      reg)
    np.testing.assert_allclose(got_te, M @ J @ M.T, rtol=1e-8, atol=1e-11)
  # t -> inf: the infinite-time (gp_inference) covariance
  big = po.ensemble_cov_finite_t(Tj[sd, sd], Kj[sd, sd], Tj[st, sd],
                                 Kj[st, sd], Kj[st, st], 1e9, g, size, lr,
                                 reg)
  inf = po.gp_inference_cov(Kj[sd, sd], Tj[sd, sd], Kj[st, sd], Tj[st, sd],
                            Kj[st, st], g, diag_reg=reg)
  np.testing.assert_allclose(big, inf, rtol=1e-7, atol=1e-9)
