"""Kernel regression on the Myrtle-5 32x32 workload (needs a GPU).

The precision form's workload is 8,192 train / 1,024 test images
(``gradient_descent_mse_ensemble(t=None)``, ``diag_reg = 1e-4``,
``predict.py:753-1100, 1224-1258``).  Here the first 1,536 train and 256
test images of that workload (same generator, same order) run end to end:

* the train and test kernels come from the HIP ``kernel_fn``; a block of
  them (entries on both sides of the block diagonal and the diagonal) is
  checked against the FP64 oracle at 1e-5;
* ``gradient_descent_mse_ensemble`` with the mixed solver (FP32 MFMA
  Cholesky + FP64 refinement) is checked against scipy's FP64
  ``cho_factor`` / ``cho_solve`` of the same kernels at 1e-6 (alpha and the
  test predictions), the accuracy bar of the precision form.
"""
import numpy as np
import pytest
import torch

from conftest import rel_err
from neural_tangents_amd import predict, stax
from neural_tangents_amd.data import krr_workload
from oracle import predict_oracle as po
from oracle import stax_oracle as so

pytestmark = pytest.mark.gpu

N_TRAIN, N_TEST = 1536, 256


@pytest.fixture(scope='module')
def workload():
  xtr, xte, ytr, _ = krr_workload(8192, 1024, seed=0)
  return xtr[:N_TRAIN], xte[:N_TEST], ytr[:N_TRAIN]


def test_krr_subset_matches_fp64_solve(workload):
  xtr, xte, ytr = workload
  kernel_fn = stax.myrtle(5)[2]
  dev = torch.device('cuda', 0)
  xtr_d = torch.from_numpy(xtr).to(dev)
  xte_d = torch.from_numpy(xte).to(dev)
  k_dd = kernel_fn(xtr_d, None, 'ntk')
  k_td = kernel_fn(xte_d, xtr_d, 'ntk')
  torch.cuda.synchronize()
  # kernel entries against the FP64 oracle: rows {0, 1, 700}, columns
  # {1, 699, 700, 1535} (both triangles and the diagonal of K_dd)
  rows, cols = [0, 1, 700], [1, 699, 700, 1535]
  ref = so.kernel_fn(so.myrtle_layers((2, 1, 1)), xtr[rows], xtr[cols],
                     ('ntk',))['ntk']
  got = k_dd.double().cpu().numpy()[np.ix_(rows, cols)]
  assert rel_err(got, ref) <= 1e-5, rel_err(got, ref)
  # the solve
  pred = predict.gradient_descent_mse_ensemble(kernel_fn, xtr_d, ytr,
                                               diag_reg=1e-4)
  f_test = pred(x_test=xte_d, get='ntk')
  want, alpha64 = po.gp_inference_mean(k_dd.double().cpu().numpy(), ytr,
                                       k_td.double().cpu().numpy(),
                                       diag_reg=1e-4)
  err_f = rel_err(f_test.cpu().numpy(), want)
  from neural_tangents_amd import _solve
  s = _solve.CholeskySolver(k_dd, 1e-4, False, 'mixed')
  alpha = s.solve(torch.from_numpy(ytr).to(dev))
  err_a = rel_err(alpha.cpu().numpy(), alpha64)
  print(f'KRR {N_TRAIN}/{N_TEST}: alpha rel err {err_a:.2e}, f_test rel err '
        f'{err_f:.2e}, {s.iterations} PCG iterations', flush=True)
  assert err_a <= 1e-6 and err_f <= 1e-6
  assert not s.impl.fell_back
