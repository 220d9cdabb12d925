"""CPU restatement of the neural-tangents analytic-kernel path (the oracle).

TEST INFRASTRUCTURE ONLY.  Nothing in the product package imports this
package: only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline``
leg of ``bench.py`` use it, and only as the checker (or as the timed CPU
baseline), never as the thing measured or shipped.

* ``stax_oracle``    -- Conv / Relu / AvgPool / GlobalAvgPool / Dense /
  Flatten kernel transforms and the input-covariance builder, restated in
  numpy (float64 by default, any float dtype accepted) from
  ``neural_tangents/_src/stax/{requirements,linear,elementwise,combinators}.py``
  of neural-tangents 0.6.6.
* ``predict_oracle`` -- ``gp_inference`` / ``gradient_descent_mse_ensemble``
  restated with scipy's Cholesky (``neural_tangents/_src/predict.py``).
* ``data``           -- the pinned synthetic-input generator (blurred,
  standardised Gaussian images; one-hot targets).

Each function cites the reference file:line it follows.  What pins each
function lives in ``tests/test_oracle_pins.py`` (closed forms, brute force,
the reference's own known answers, Gaussian quadrature, library routines,
the gradient-flow ODE and matrix-exponential forms of the predictors) and
``tests/test_oracle_mc.py`` (Monte-Carlo agreement of the NNGP / NTK
recursions with empirical kernels of finite networks, the reference's own
validation method, with planted mistakes shown to fail).
"""
