"""Schedule compilation on the host (no GPU)."""
import pytest

from neural_tangents_amd import _engine, stax


def _plan(*layers, shape=(1, 32, 32, 3)):
  return _engine.compile_plan(stax.serial(*layers)[2].specs, shape)


def test_myrtle5_stages():
  p = _engine.compile_plan(stax.myrtle(5)[2].specs, (1, 32, 32, 3))
  assert [(s.first, s.conv is not None, s.relu, s.pool, s.hin, s.hout)
          for s in p.stages] == [
              (True, True, True, None, 32, 32),
              (False, True, True, (2, 2), 32, 16),
              (False, True, True, (2, 2), 16, 8),
              (False, True, True, (2, 2), 8, 4)]
  assert p.head == 'gap'
  assert p.head_coef == (1.0, 0.0, 1.0, 1.0, 0.0)


def test_myrtle10_stages():
  p = _engine.compile_plan(stax.myrtle(10)[2].specs, (1, 32, 32, 3))
  assert len(p.stages) == 9
  assert [s.hout for s in p.stages] == [32, 32, 16, 16, 16, 8, 8, 8, 4]


def test_first_stage_pool_split():
  p = _plan(stax.Conv(1, (3, 3), padding='SAME'), stax.Relu(),
            stax.AvgPool((2, 2), (2, 2)), stax.GlobalAvgPool())
  assert len(p.stages) == 2 and p.stages[1].conv is None
  assert p.stages[1].pool == (2, 2) and not p.stages[1].relu


def test_dense_chain_coefficients():
  p = _plan(stax.Conv(1, (3, 3), padding='SAME'), stax.GlobalAvgPool(),
            stax.Dense(1, 2.0, 1.0), stax.Dense(1, 3.0, 0.5))
  a, b, c, d, e = p.head_coef
  # nngp1 = 4m + 1; ntk1 = nngp1 + 4mt
  # nngp2 = 9 nngp1 + .25 = 36m + 9.25; ntk2 = nngp2 + 9 ntk1
  assert (a, b) == (36.0, 9.25)
  assert (c, d, e) == (36.0 + 36.0, 36.0, 9.25 + 9.0)


@pytest.mark.parametrize('layers,exc', [
    ((stax.Relu(),), ValueError),
    ((stax.AvgPool((2, 2), (2, 2)),), NotImplementedError),
    ((stax.Conv(1, (3, 3), (2, 2), 'SAME'),), NotImplementedError),
    ((stax.Conv(1, (3, 3), (1, 2), padding='CIRCULAR'),),
     NotImplementedError),
    ((stax.Conv(1, (3, 3), padding='SAME'), stax.AvgPool((3, 3), (2, 2))),
     NotImplementedError),
    ((stax.Conv(1, (3, 3), padding='SAME'), stax.GlobalAvgPool(),
      stax.Relu()), NotImplementedError),
])
def test_rejections(layers, exc):
  with pytest.raises(exc):
    _plan(*layers)


def test_diagonal_spatial_plan_selection():
  """Flatten readout without pooling -> diagonal-spatial plan (the
  reference's diagonal_spatial requirement); any pooling or a GAP readout
  needs the full spatial kernel."""
  conv = lambda: stax.Conv(1, (3, 3), padding='SAME')
  assert _plan(conv(), stax.Relu(), conv(), stax.Relu(), stax.Flatten(),
               stax.Dense(1)).diag
  assert not _plan(conv(), stax.Relu(), stax.AvgPool((2, 2), (2, 2)),
                   stax.Flatten()).diag
  assert not _plan(conv(), stax.Relu(), stax.GlobalAvgPool()).diag
  assert not _plan(conv(), stax.Relu()).diag


def test_diag_stage_tables():
  p = _plan(stax.Conv(1, (3, 3), padding='SAME', W_std=1.5, b_std=0.5),
            stax.LeakyRelu(0.25), stax.Conv(1, (2, 2), padding='VALID'),
            stax.Flatten())
  desc, wb = _engine._diag_tables(p)
  from neural_tangents_amd import _lib
  assert list(desc) == [_lib.CONV | _lib.RELU, 3, 3,
                        _lib.CONV | _lib.CONV_VALID, 2, 2]
  # LeakyRelu(alpha) = ABRelu(alpha, 1): k1 = (1 - alpha)^2, k0 = alpha
  assert list(wb) == [2.25, 0.25, 0.5625, 0.25, 1.0, 0.0, 1.0, 0.0]


@pytest.mark.parametrize('layer,act', [(stax.Relu(), (1.0, 0.0)),
                                       (stax.Abs(), (4.0, -1.0)),
                                       (stax.ABRelu(-0.5, 2.0), (6.25, -1.0)),
                                       (stax.LeakyRelu(0.1), (0.81, 0.1))])
def test_activation_constants(layer, act):
  p = _plan(stax.Conv(1, (3, 3), padding='SAME'), layer, stax.GlobalAvgPool())
  assert p.stages[0].act == pytest.approx(act)


def test_sum_pool_scale_folding():
  """SumPool = AvgPool x window^2 on kernels: folded into the next Conv's
  W_std^2, the readout, or the outputs; GlobalSumPool multiplies the mean
  by the number of (h, h', w, w') entries."""
  conv = lambda: stax.Conv(1, (3, 3), padding='SAME', W_std=2.0)
  p = _plan(conv(), stax.Relu(), stax.SumPool((2, 2), (2, 2)), conv(),
            stax.Relu(), stax.GlobalAvgPool())
  assert [s.conv[3] if s.conv else None for s in p.stages] == [
      4.0, None, 64.0]
  assert p.head_coef == (1.0, 0.0, 0.0, 1.0, 0.0)
  p = _plan(conv(), stax.Relu(), stax.SumPool((2, 2), (2, 2)),
            stax.GlobalSumPool(), stax.Dense(1, 2.0))
  # m = mean over 16^4 entries; GlobalSumPool -> 16^4 m; SumPool -> x16
  a = 16.0 * 16 ** 4
  assert p.head_coef == (4 * a, 0.0, 4 * a, 4 * a, 0.0)
  p = _plan(conv(), stax.Relu(), stax.SumPool((2, 2), (2, 2)))
  assert p.out_scale == 16.0


def test_first_two_stage_fusion_selection():
  """The fused first-two-stage launch applies to Conv-Relu-Conv-Relu-Pool
  heads on 32x32 images with <= 4 channels and nowhere else."""
  for depth in (5, 7):
    p = _engine.compile_plan(stax.myrtle(depth)[2].specs, (1, 32, 32, 3))
    assert _engine._fuses_first_two(p, (1, 32, 32, 3))
    assert not _engine._fuses_first_two(p, (1, 32, 32, 5))
  # Myrtle-10 has a third Conv before the first pool: the three-stage
  # launch takes it
  p = _engine.compile_plan(stax.myrtle(10)[2].specs, (1, 32, 32, 3))
  assert not _engine._fuses_first_two(p, (1, 32, 32, 3))
  assert _engine._fuses_first_three(p, (1, 32, 32, 3))
  assert _engine._fused_stages(p, (1, 32, 32, 3)) == 3
  assert _engine._fused_stages(p, (1, 32, 32, 5)) == 0
  assert _engine._per_pair_bytes(p, True, 3) == 4 * 16 ** 4 * 2 * 2
  for depth in (5, 7):
    p = _engine.compile_plan(stax.myrtle(depth)[2].specs, (1, 32, 32, 3))
    assert not _engine._fuses_first_three(p, (1, 32, 32, 3))
    assert _engine._fused_stages(p, (1, 32, 32, 3)) == 2
  p = _engine.compile_plan(stax.myrtle(5)[2].specs, (1, 16, 16, 3))
  assert not _engine._fuses_first_two(p, (1, 16, 16, 3))
  conv = lambda: stax.Conv(1, (3, 3), padding='SAME')
  others = [
      [conv(), stax.Relu(), conv(), stax.Relu(), stax.GlobalAvgPool()],
      [conv(), stax.LeakyRelu(0.1), conv(), stax.Relu(),
       stax.AvgPool((2, 2), (2, 2)), stax.GlobalAvgPool()],
      [conv(), stax.Relu(), stax.Conv(1, (5, 5), padding='SAME'),
       stax.Relu(), stax.AvgPool((2, 2), (2, 2)), stax.GlobalAvgPool()],
      [conv(), stax.Relu(), stax.AvgPool((2, 2), (2, 2)), conv(),
       stax.Relu(), stax.GlobalAvgPool()],
  ]
  for layers in others:
    p = _plan(*layers)
    assert not _engine._fuses_first_two(p, (1, 32, 32, 3))
  # the fused tile budget counts only the pooled second-stage state
  p = _engine.compile_plan(stax.myrtle(5)[2].specs, (1, 32, 32, 3))
  assert _engine._per_pair_bytes(p, True, True) == 4 * 16 ** 4 * 2 * 2
  assert _engine._per_pair_bytes(p, True) == 4 * 32 ** 4 * 2 * 2


def test_erf_gelu_stage_constants():
  """Erf(a, b, c): flags carry the activation, (k1, k0) = (a^2, c^2) and
  b^2 multiplies the preceding conv's W_std^2 and b_std^2; Gelu has no
  constants; neither is fused; b != 1 without a Conv right before raises."""
  from neural_tangents_amd import _lib
  p = _plan(stax.Conv(1, (3, 3), padding='SAME', W_std=1.5, b_std=0.5),
            stax.Erf(0.8, 2.0, 0.3), stax.AvgPool((2, 2), (2, 2)),
            stax.Conv(1, (3, 3), padding='SAME'), stax.Gelu(),
            stax.GlobalAvgPool())
  s0, s1 = p.stages[0], p.stages[2]        # stages[1]: the first pool
  assert s0.act_kind == _lib.ACT_ERF and s1.act_kind == _lib.ACT_GELU
  assert s0.act == pytest.approx((0.64, 0.09))
  assert s0.conv[3] == pytest.approx(1.5 ** 2 * 4.0)
  assert s0.conv[4] == pytest.approx(0.5 ** 2 * 4.0)
  assert s0.flags(True) & _lib.ACT_ERF and s1.flags(False) & _lib.ACT_GELU
  assert not _engine._fuses_first_two(p, (1, 32, 32, 3))
  with pytest.raises(NotImplementedError):
    _plan(stax.Conv(1, (3, 3), padding='SAME'), stax.AvgPool((2, 2), (2, 2)),
          stax.Erf(1.0, 2.0), stax.GlobalAvgPool())
  with pytest.raises(ValueError):
    _plan(stax.Conv(1, (3, 3), padding='SAME'), stax.Gelu(), stax.Erf(),
          stax.GlobalAvgPool())


def test_circular_padding_flags_and_no_fusion():
  from neural_tangents_amd import _lib
  conv = lambda: stax.Conv(1, (3, 3), padding='CIRCULAR')
  p = _plan(conv(), stax.Relu(), conv(), stax.Relu(),
            stax.AvgPool((2, 2), (2, 2)), stax.GlobalAvgPool())
  assert all(s.circular for s in p.stages if s.conv is not None)
  assert p.stages[0].flags(True) & _lib.CONV_CIRCULAR
  assert not p.stages[0].flags(True) & _lib.CONV_VALID
  assert not _engine._fuses_first_two(p, (1, 32, 32, 3))


def test_standard_dense_head_coefficients_match_oracle():
  """Dense(parameterization='standard') in the readout fold equals the
  oracle's dense layers applied to a scalar kernel (nngp = m, ntk = mt),
  with the fan-in taken from the last Conv's out_chan and the s scales."""
  import numpy as np
  from oracle import stax_oracle as so
  p = _plan(stax.Conv(6, (3, 3), padding='SAME'), stax.GlobalAvgPool(),
            stax.Dense(5, 1.3, 0.4, parameterization='standard', s=(2, 3)),
            stax.Dense(2, 0.7, 0.2),
            stax.Dense(1, 0.9, None, parameterization='STANDARD'))
  a, b, c, d, e = p.head_coef
  m, mt = 0.37, 1.91
  k = so.OKernel(nngp=np.array(m), ntk=np.array(mt), cov1=np.array(m),
                 cov2=None, x1_is_x2=False, is_gaussian=False,
                 is_reversed=False, diagonal_batch=True,
                 diagonal_spatial=False, shape1=None, shape2=None)
  k = so.dense_layer(1.3, 0.4, 'standard', input_width=6 / 2)(k)
  k = so.dense_layer(0.7, 0.2)(k)
  k = so.dense_layer(0.9, None, 'standard', input_width=2)(k)
  assert abs(a * m + b - k.nngp) < 1e-12
  assert abs(c * m + d * mt + e - k.ntk) < 1e-12


def test_unknown_parameterization_raises_value_error():
  with pytest.raises(ValueError):
    stax.Dense(1, parameterization='mup')
  with pytest.raises(ValueError):
    stax.Conv(1, (3, 3), parameterization='xyz')
  with pytest.raises(NotImplementedError):
    _plan(stax.Conv(1, (3, 3), padding='SAME', parameterization='standard'),
          stax.GlobalAvgPool())


def test_from_state_plans():
  """Plans for kernel_fn(Kernel): no input-covariance stage, layers may
  start with Relu / pools, and a readout alone is a plan with no stages."""
  kfn = stax.serial(stax.Relu(), stax.AvgPool((2, 2), (2, 2)),
                    stax.GlobalAvgPool(), stax.Dense(1, 2.0))[2]
  p = kfn.plan((1, 8, 8, 1), from_state=True, in_gaussian=True)
  assert not any(s.first for s in p.stages)
  assert p.stages[0].relu and p.stages[0].pool == (2, 2)
  assert p.head == 'gap' and p.out_spatial == (4, 4)
  with pytest.raises(ValueError):
    kfn.plan((1, 8, 8, 1), from_state=True, in_gaussian=False)
  head_only = stax.serial(stax.Flatten(), stax.Dense(10))[2]
  p = head_only.plan((1, 3, 3, 1), from_state=True)
  assert p.stages == [] and p.head == 'flatten' and not p.diag
  with pytest.raises(NotImplementedError):
    _plan(stax.Flatten(), stax.Dense(10))


def test_dense_before_readout_is_a_1x1_stage():
  """Dense on spatial (or flat) kernels is the affine map of a 1x1 Conv
  without the axis reversal (linear.py:902-926, 1387-1398)."""
  p = _plan(stax.Dense(64, 1.5, 0.1), stax.Relu(), stax.Dense(8, 2.0),
            shape=(1, 1, 1, 20))
  assert [(s.first, s.conv, s.relu) for s in p.stages] == [
      (True, (1, 1, False, 2.25, 0.010000000000000002), True),
      (False, (1, 1, False, 4.0, 0.0), False)]
  assert p.n_convs == 0 and p.head is None
  with pytest.raises(NotImplementedError):
    _plan(stax.Dense(8, parameterization='standard'), stax.Relu())
  kfn = stax.serial(stax.Conv(1, (3, 3)), stax.Relu())[2]
  with pytest.raises(ValueError):
    kfn.plan((4, 20))


def test_kernel_reverse_transpose_follow_reference_axes():
  """Kernel.reverse / transpose (utils/kernel.py:169-222) against the
  oracle's reverse_zipped (utils.py:465-480) and numpy transposes."""
  import numpy as np
  import torch
  from neural_tangents_amd.kernel import Kernel
  from oracle import stax_oracle as so
  rng = np.random.default_rng(0)
  nngp = rng.standard_normal((2, 3, 4, 4, 5, 5))
  cov1 = rng.standard_normal((2, 4, 4, 5, 5))
  k = Kernel(nngp=torch.from_numpy(nngp), ntk=torch.zeros(()),
             cov1=torch.from_numpy(cov1), cov2=None, x1_is_x2=False,
             is_gaussian=True, is_reversed=False, is_input=False,
             diagonal_batch=True, diagonal_spatial=False,
             shape1=(2, 4, 5, 7), shape2=(3, 4, 5, 7), batch_axis=0,
             channel_axis=3)
  r = k.reverse()
  assert r.is_reversed and r.ntk.ndim == 0
  np.testing.assert_array_equal(r.nngp.numpy(), so.reverse_zipped(nngp, 2))
  np.testing.assert_array_equal(r.cov1.numpy(), so.reverse_zipped(cov1, 1))
  t = k.transpose((1, 0))
  np.testing.assert_array_equal(t.nngp.numpy(),
                                nngp.transpose(0, 1, 4, 5, 2, 3))
  np.testing.assert_array_equal(t.cov1.numpy(), cov1.transpose(0, 3, 4, 1, 2))
  assert torch.equal(k.transpose().nngp, k.nngp)
