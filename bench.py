"""Benchmark: Myrtle-5 full-spatial NNGP+NTK on one MI355X, plus the
kernel-regression, Myrtle-10 and gradient forms.

One step = one ``kernel_fn(x1, x2, ('nngp', 'ntk'))`` over the whole
workload (Myrtle-5, 32x32x3, x1 and x2 each [256, 32, 32, 3] fp32 blurred
Gaussian images, seed 0): 65,536 cross pairs plus the 512 per-image self
chains.  Inputs are resident in HBM before the timed region.

Prints ONE JSON line.  ``value`` is pairs/s; ``roofline`` describes the
dominant kernel, timed with HIP events on its launch stream over the timed
steps.  Extra keys, all measured after the timed steps:

* ``krr``: the kernel-regression form (Myrtle-5 32x32 NTK of 8,192 train and
  1,024 test images, ``diag_reg = 1e-4``): mixed solve (FP32 MFMA Cholesky +
  FP64 refinement) against the plain FP64 solve on the same kernel, both
  times, the speed-up and the relative errors of alpha and the test
  predictions against FP64;
* ``myrtle10``: the Myrtle-10 workload (x1 [128], x2 [512]) pairs/s and
  per-kernel GB/s;
* ``grad``: wall time of the differentiation form (loss and gradients with
  respect to 100 support images, W_std and b_std against 2,048 targets);
* ``cpu_baseline``: the oracle (the reference algorithm restated in numpy,
  FP32) on a bounded block of the workload, one thread per host core; the
  same leg evaluates the FP64 oracle on a block of the benchmarked output
  that spans pair-tile boundaries and reports ``rel_err_vs_oracle``.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]
       [--no-krr] [--no-m10] [--no-grad]
       python bench.py --krr [--krr-reps R]      (the solve alone)
       python bench.py --grad [--grad-targets N] (the gradients alone)
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0

WORKLOADS = {
    # name: (depths, side, n1, n2)
    'myrtle5_32x32_256x256': ((2, 1, 1), 32, 256, 256),
    'myrtle10_32x32_128x512': ((3, 3, 3), 32, 128, 512),
    'myrtle5_16x16_64x64': ((2, 1, 1), 16, 64, 64),
    # one tile of 256 pairs: the launch the counter summaries are taken on
    'myrtle5_32x32_16x16': ((2, 1, 1), 32, 16, 16),
}


def log(msg):
  print(msg, file=sys.stderr, flush=True)


def canonical_bytes_per_pair(depths, side):
  """Algorithmic HBM bytes per cross pair of the per-layer fused schedule:
  every fused Conv+ReLU(+Pool) kernel reads its (nngp, ntk) fp32 input state
  once and writes its output state once; the first reads only images, the
  last writes one (nngp, ntk) value."""
  sizes = []
  s = side
  for d in depths:
    for k in range(d):
      s_in = s
      s_out = s // 2 if k == d - 1 else s
      sizes.append((s_in, s_out))
    s //= 2
  total = 0
  for idx, (s_in, s_out) in enumerate(sizes):
    if idx > 0:
      total += 2 * 4 * s_in ** 4
    if idx < len(sizes) - 1:
      total += 2 * 4 * s_out ** 4
  return total + 8


# HBM traffic per launch measured with rocprofv3 FETCH_SIZE / WRITE_SIZE
# (scaled by a plain copy of known size), per trace label of the engine.
PMC_FILE = os.path.join(ROOT, 'profiles', 'myrtle5_32x32_hbm_traffic_pmc.json')
# Vector-ALU counters of the fused first kernel (rocprofv3 --pmc).
VALU_FILE = os.path.join(ROOT, 'profiles', 'myrtle5_32x32_valu_pmc.json')


def _load_json(path):
  try:
    with open(path) as f:
      return json.load(f)
  except (OSError, ValueError):
    return None


def measured_traffic(label, pairs_per_launch):
  """HBM bytes per launch of ``label`` from the committed counter summary
  (scaled to the launch's pair count), or None."""
  d = _load_json(PMC_FILE)
  try:
    k = d['kernels'][label]
  except (TypeError, KeyError):
    return None
  return k['traffic_bytes_per_launch'] * pairs_per_launch / k[
      'pairs_per_launch']


def host_threads():
  """Host threads for the CPU leg: the process's CPU affinity, capped by
  OMP_NUM_THREADS when set (the GPU box's CPU share)."""
  n = len(os.sched_getaffinity(0)) if hasattr(os, 'sched_getaffinity') \
      else (os.cpu_count() or 1)
  omp = os.environ.get('OMP_NUM_THREADS')
  if omp and omp.isdigit() and int(omp) > 0:
    n = min(n, int(omp))
  return max(1, n)


def cpu_leg(depths, x1, x2, got, rows, cols, err_rows, err_cols):
  """The CPU leg: times the oracle (FP32, numpy, one thread per host core,
  each thread an independent row block of the sample against the sample's
  columns) and evaluates the FP64 oracle on the (err_rows x err_cols) block
  of the benchmarked output ``got`` = (nngp, ntk) host arrays."""
  import numpy as np
  from concurrent.futures import ThreadPoolExecutor
  from threadpoolctl import threadpool_limits
  from oracle import stax_oracle as so
  layers = so.myrtle_layers(tuple(depths))
  nthr = host_threads()
  xb = x2[:cols]

  def block(i0, i1, dtype=np.float32, xa=x1, xc=xb):
    return so.kernel_fn(layers, xa[i0:i1], xc, ('nngp', 'ntk'), dtype=dtype)

  per = max(1, rows // nthr)
  chunks = [(i, min(rows, i + per)) for i in range(0, rows, per)]
  with threadpool_limits(1):
    with ThreadPoolExecutor(nthr) as ex:
      t0 = time.perf_counter()
      list(ex.map(lambda c: block(*c), chunks))
      dt = time.perf_counter() - t0
      # FP64 oracle on the error block (one row per task)
      ea, eb = x1[err_rows], x2[err_cols]
      refs = list(ex.map(lambda i: block(i, i + 1, np.float64, ea, eb),
                         range(len(err_rows))))
  pairs = rows * cols
  base = {'value': pairs / dt, 'unit': 'pairs/s', 'cores': nthr,
          'kind': 'port',
          'sample': f'{rows}x{cols} block of the workload ({pairs} pairs + '
                    f'{rows + cols} self chains; {len(chunks)} row blocks '
                    f'on {nthr} threads, each with its own x2 self chains), '
                    f'FP32 numpy, {dt:.1f} s'}
  err = {}
  for gi, g in enumerate(('nngp', 'ntk')):
    ref = np.concatenate([r[g] for r in refs], axis=0)
    blk = got[gi][np.ix_(err_rows, err_cols)].astype(np.float64)
    err[g] = float(np.linalg.norm(blk - ref) / np.linalg.norm(ref))
  return base, err


def tile_rows(kernel_fn, x1, x2):
  """Row-tile size of the engine's pair tiling for this call."""
  from neural_tangents_amd import _engine
  plan = kernel_fn.plan(x1.shape)
  fused = _engine._fused_stages(plan, x1.shape)
  per = _engine._per_pair_bytes(plan, True, fused)
  T = max(1, _engine.DEFAULT_TILE_BYTES // per)
  tj = min(x2.shape[0], T)
  return max(1, min(x1.shape[0], T // tj))


def timed_kernel_fn(kernel_fn, x1, x2, steps, warmup):
  """Runs kernel_fn (nngp, ntk) warmup + steps times; returns (seconds of
  the timed steps, per-entry-point launch statistics, last output)."""
  import torch
  from neural_tangents_amd import _engine
  for _ in range(warmup):
    kernel_fn(x1, x2, ('nngp', 'ntk'))
  torch.cuda.synchronize()
  _engine.TRACE = []
  torch.cuda.synchronize()
  t0 = time.perf_counter()
  for _ in range(steps):
    k = kernel_fn(x1, x2, ('nngp', 'ntk'))
  torch.cuda.synchronize()
  elapsed = time.perf_counter() - t0
  trace, _engine.TRACE = _engine.TRACE, None
  per = {}
  for name, s, e, nbytes, pairs in trace:
    d = per.setdefault(name, {'ms': 0.0, 'launches': 0, 'bytes': 0,
                              'pairs': 0})
    d['ms'] += s.elapsed_time(e)
    d['launches'] += 1
    d['bytes'] += nbytes
    d['pairs'] += pairs
  return elapsed, per, k


def kernel_table(per, steps):
  return {n: {'ms_per_step': v['ms'] / steps,
              'launches_per_step': v['launches'] / steps,
              'avg_launch_ms': v['ms'] / v['launches'],
              'gbs': v['bytes'] / (v['ms'] * 1e-3) / 1e9
              if v['ms'] > 0 else None}
          for n, v in per.items()}


def krr_result(ntr, nte, reps, tol=1e-7):
  """Kernel-regression solve (infinite-time gradient_descent_mse_ensemble):
  Myrtle-5 32x32 NTK of ntr synthetic images, one-hot targets (-0.1),
  diag_reg = 1e-4 (relative to the mean trace).  Times the plain FP64
  Cholesky solve (rocSOLVER through torch) and the mixed solve (FP32 MFMA
  Cholesky + FP64-residual PCG), both on the same kernel, and reports the
  relative error of alpha and of the test predictions against FP64.  The
  refinement stops when its estimated relative error (or residual) is at
  most ``tol``, a tenth of the 1e-6 bar by default."""
  import torch
  from neural_tangents_amd import _solve, stax
  from neural_tangents_amd.data import krr_workload
  xtr, xte, ytr, _ = krr_workload(ntr, nte, seed=0)
  dev = torch.device('cuda', 0)
  kernel_fn = stax.myrtle(5)[2]
  t0 = time.perf_counter()
  xtr_d = torch.from_numpy(xtr).to(dev)
  log(f'krr: train kernel {ntr}x{ntr} (upper triangle) ...')
  k_dd = kernel_fn(xtr_d, None, 'ntk')
  torch.cuda.synchronize()
  log(f'krr: train kernel {time.perf_counter() - t0:.1f} s')
  k_td = kernel_fn(torch.from_numpy(xte).to(dev), xtr_d, 'ntk')
  torch.cuda.synchronize()
  t_kernel = time.perf_counter() - t0
  log(f'krr: kernels {t_kernel:.1f} s')
  y = torch.from_numpy(ytr).to(dev, torch.float64)

  def solve_once(mode):
    torch.cuda.synchronize()
    t = time.perf_counter()
    s = _solve.CholeskySolver(k_dd, 1e-4, False, mode, tol=tol)
    alpha = s.solve(y)
    f = _solve.kernel_times(k_td, alpha)
    torch.cuda.synchronize()
    return time.perf_counter() - t, alpha, f, s

  # One untimed solve per form first (library streams and events, code
  # loading, the allocator's first blocks), then reps runs of each form in
  # a row (each form's steady state when solving repeatedly); the median of
  # each is the reported time.  A second, interleaved set (each run right
  # after the other form's, so the caches hold the other form's data) is
  # reported beside it.
  times = {'fp64': [], 'mixed': []}
  last = {}
  for mode in ('fp64', 'mixed'):
    solve_once(mode)
    for _ in range(reps):
      dt, alpha, f, s = solve_once(mode)
      times[mode].append(dt)
      last[mode] = (alpha, f, s)
  inter = {'fp64': [], 'mixed': []}
  for _ in range(3):
    for mode in ('fp64', 'mixed'):
      inter[mode].append(solve_once(mode)[0])
  res = {}
  for mode in ('fp64', 'mixed'):
    alpha, f, s = last[mode]
    res[mode] = (sorted(times[mode])[len(times[mode]) // 2], alpha, f,
                 getattr(s, 'iterations', 0), getattr(s, 'last_residual',
                                                      None), s)
  t64, a64, f64 = res['fp64'][:3]
  tmx, amx, fmx, iters, resid, smx = res['mixed']
  err_a = float(torch.linalg.norm(amx - a64) / torch.linalg.norm(a64))
  err_f = float(torch.linalg.norm(fmx - f64) / torch.linalg.norm(f64))
  n = ntr
  chol_flop = n ** 3 / 3
  log(f'krr: fp64 {t64 * 1e3:.2f} ms, mixed {tmx * 1e3:.2f} ms')
  return {
      'metric': 'KRR solve time vs plain FP64 at <=1e-6 rel err',
      'value': t64 / tmx, 'unit': 'x speed-up (plain FP64 / mixed)',
      'config': {'workload': f'krr_myrtle5_32x32_train{ntr}_test{nte}',
                 'diag_reg': 1e-4, 'get': 'ntk', 'targets': 10},
      'plain_fp64_ms': t64 * 1e3, 'mixed_ms': tmx * 1e3,
      'plain_fp64_tflops': chol_flop / t64 / 1e12,
      'plain_fp64_note': 'torch.linalg.cholesky_ex + cholesky_solve in FP64 '
                         '(rocSOLVER); tflops counts n^3/3 over the whole '
                         'solve + predict time; both forms predict with '
                         'nta_gemm_f32_f64 (fp32 k_td, FP64 alpha)',
      'pcg_iterations': iters, 'final_rel_residual': resid,
      'pcg_tol': tol,
      'pcg_tol_note': 'refinement stops at estimated relative error '
                      '||M^-1 r|| / ||x|| (or relative residual) <= tol',
      'fell_back_to_fp64': bool(smx.impl.fell_back),
      'rel_err_alpha_vs_fp64': err_a, 'rel_err_ftest_vs_fp64': err_f,
      'kernel_seconds': t_kernel,
      'timing': f'median of {reps} consecutive runs of each form after one '
                'untimed run (wall time, inputs resident, kernel build '
                'excluded)',
      'mixed_ms_all': [1e3 * t for t in sorted(times['mixed'])],
      'plain_fp64_ms_all': [1e3 * t for t in sorted(times['fp64'])],
      'interleaved': {
          'note': 'medians of 3 runs of each form, each right after a run '
                  'of the other form (caches holding the other form\'s '
                  'data)',
          'plain_fp64_ms': 1e3 * sorted(inter['fp64'])[1],
          'mixed_ms': 1e3 * sorted(inter['mixed'])[1],
          'value': sorted(inter['fp64'])[1] / sorted(inter['mixed'])[1]},
  }


def m10_result(steps):
  """Myrtle-10 (Conv+Relu blocks [3, 3, 3]) at 32x32, x1 [128], x2 [512]:
  pairs/s and per-kernel GB/s (algorithmic bytes over HIP-event time)."""
  import torch
  from neural_tangents_amd import stax
  from neural_tangents_amd.data import blurred_images
  depths, side, n1, n2 = WORKLOADS['myrtle10_32x32_128x512']
  a, b = blurred_images([n1, n2], side, seed=0)
  dev = torch.device('cuda', 0)
  x1, x2 = torch.from_numpy(a).to(dev), torch.from_numpy(b).to(dev)
  kernel_fn = stax.myrtle(depths=list(depths))[2]
  elapsed, per, _ = timed_kernel_fn(kernel_fn, x1, x2, steps, 1)
  value = n1 * n2 * steps / elapsed
  bpp = canonical_bytes_per_pair(depths, side)
  log(f'myrtle10: {value:.0f} pairs/s')
  table = kernel_table(per, steps)
  dom = max(per, key=lambda n: per[n]['ms'])
  return {'workload': 'myrtle10_32x32_128x512', 'pairs_per_s': value,
          'ms_per_step': elapsed / steps * 1e3, 'steps': steps,
          'path_gbs_canonical': bpp * value / 1e9,
          'path_pct_hbm_peak': 100.0 * bpp * value / 1e9 / HBM_PEAK_GBS,
          'dominant_kernel': dom,
          'dominant_gbs_canonical': table[dom]['gbs'],
          'dominant_frac_hbm_peak': table[dom]['gbs'] / HBM_PEAK_GBS,
          'note': SCHEDULE_NOTES.get(dom, ''),
          'kernels': table}


def grad_result(n_target=2048):
  """Differentiation form: loss 0.5 mean((K_ts (K_ss + lambda I)^-1 Y_s -
  Y_t)^2) of the Myrtle-5 32x32 NTK with 100 support and 2,048 target
  images (n_target: fewer, for a profile), gradients with respect to the
  support images, W_std and b_std (one forward + backward, wall time)."""
  import torch
  from neural_tangents_amd import grad
  from neural_tangents_amd.data import blurred_images, one_hot_targets
  dev = torch.device('cuda', 0)
  xs, xt = blurred_images([100, n_target], 32, seed=5)
  ys = torch.from_numpy(one_hot_targets(100, seed=6)).to(dev)
  yt = torch.from_numpy(one_hot_targets(n_target, seed=7)).to(dev)
  builder = grad.myrtle_builder((2, 1, 1))

  def run(ns, nt):
    xs_t = torch.tensor(xs[:ns], device=dev, requires_grad=True)
    W = torch.tensor(2 ** 0.5, dtype=torch.float64, device=dev,
                     requires_grad=True)
    b = torch.tensor(0.1, dtype=torch.float64, device=dev,
                     requires_grad=True)
    loss = grad.krr_loss(xs_t, ys[:ns], torch.tensor(xt[:nt], device=dev),
                         yt[:nt], W, b, builder, diag_reg=1e-4)
    loss.backward()
    torch.cuda.synchronize()
    return float(loss.detach())

  run(4, 8)                                    # warm-up
  torch.cuda.synchronize()
  t0 = time.perf_counter()
  loss = run(100, n_target)
  dt = time.perf_counter() - t0
  log(f'grad: {dt:.1f} s')
  pairs = 100 * n_target + 100 * 101 // 2
  return {'workload': f'grad_myrtle5_32x32_support100_target{n_target}',
          'seconds': dt, 'loss': loss, 'kernel_pairs': pairs,
          'pairs_per_s_fwd_bwd': pairs / dt}


# How the algorithmic bytes of a launch relate to what it moves.
SCHEDULE_NOTES = {
    'nta_first3_pool_stage':
        'first three stages fused ([Conv-Relu] x 3 - AvgPool, Myrtle-10): '
        'achieved counts the per-layer schedule\'s bytes for those three '
        'kernels (two 32^4 states written and read back, the pooled state '
        'written); the kernel never materialises the first two stages',
    'nta_first_pool_stage':
        'first two stages fused (Conv-Relu-Conv-Relu-AvgPool): achieved '
        'counts the per-layer schedule\'s bytes for those two kernels (the '
        'first stage state written and read back, the pooled state written); '
        'the kernel never materialises the first stage, so traffic (measured '
        'HBM bytes) is far below them; its bound is the vector ALU '
        '(valu_busy_frac, from rocprofv3 counters)',
}


def readme_cnn21_result(n_flat=512, n_gap=16, steps=3):
  """The reference README's performance table (``README.md:370-416``):
  21 x [Conv(1, (3, 3), (1, 1), 'SAME'), Relu] on 32x32x3 inputs, timed as
  milliseconds per NTK entry of an N x N batch (distinct x1, x2, nngp and
  ntk) with a Flatten readout (the diagonal-spatial path, one
  nta_diag_chain launch) and with a GlobalAvgPool readout (21 full-spatial
  stages), next to the README's NVIDIA V100 fp32 figures.  The error
  against the FP64 oracle is taken on a 2x2 block (Flatten) and one entry
  (GlobalAvgPool), outside the timed region."""
  import numpy as np
  import torch
  from neural_tangents_amd import stax
  from neural_tangents_amd.data import blurred_images
  from oracle import stax_oracle as so
  dev = torch.device('cuda', 0)
  layers, olayers = [], []
  for _ in range(21):
    layers += [stax.Conv(1, (3, 3), (1, 1), 'SAME'), stax.Relu()]
    olayers += [so.conv_layer((3, 3), (1, 1), 'SAME', 1.0, None),
                so.relu_layer()]
  res = {'architecture': "21 x [Conv(1, (3, 3), (1, 1), 'SAME'), Relu], "
                         '32x32x3 inputs (README.md:370-378)'}
  for name, n, v100, readout, oreadout, diag in (
      ('flatten', n_flat, 0.0046510, stax.Flatten(), so.flatten_layer(),
       True),
      ('global_avg_pool', n_gap, 2.7001, stax.GlobalAvgPool(),
       so.global_avg_pool_layer(), False)):
    kfn = stax.serial(*layers, readout)[2]
    x1n, x2n = blurred_images([n, n], 32, seed=1)
    x1 = torch.from_numpy(x1n).to(dev)
    x2 = torch.from_numpy(x2n).to(dev)
    k = kfn(x1, x2, ('nngp', 'ntk'))
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
      k = kfn(x1, x2, ('nngp', 'ntk'))
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / steps
    ms_entry = dt * 1e3 / (n * n)
    rows, cols = ([0, n - 1], [0, n - 1]) if diag else ([0], [n - 1])
    ref = so.kernel_fn(olayers + [oreadout], x1n[rows], x2n[cols],
                       ('nngp', 'ntk'), np.float64, diagonal_spatial=diag)
    err = {}
    for g in ('nngp', 'ntk'):
      got = getattr(k, g).cpu().numpy()[np.ix_(rows, cols)].astype(np.float64)
      err[g] = float(np.linalg.norm(got - ref[g]) / np.linalg.norm(ref[g]))
    log(f'readme cnn21 {name}: {ms_entry:.3g} ms/entry (V100 {v100})')
    res[name] = {'batch': [n, n], 'ms_per_entry': ms_entry,
                 'entries_per_s': 1e3 / ms_entry,
                 'readme_v100_fp32_ms_per_entry': v100,
                 'speedup_vs_readme_v100': v100 / ms_entry,
                 'rel_err_vs_fp64_oracle': err}
    del k
    torch.cuda.empty_cache()
  return res


def valu_busy(kernel):
  d = _load_json(VALU_FILE)
  try:
    return d['kernels'][kernel]['valu_busy_frac']
  except (TypeError, KeyError):
    return None


def main():
  ap = argparse.ArgumentParser()
  ap.add_argument('--gpus', type=int, default=1)
  ap.add_argument('--steps', type=int, default=10)
  ap.add_argument('--warmup', type=int, default=2)
  ap.add_argument('--workload', default='myrtle5_32x32_256x256',
                  choices=sorted(WORKLOADS))
  ap.add_argument('--no-cpu', action='store_true')
  ap.add_argument('--no-krr', action='store_true')
  ap.add_argument('--no-m10', action='store_true')
  ap.add_argument('--no-grad', action='store_true')
  ap.add_argument('--no-readme', action='store_true',
                  help='skip the README 21-layer CNN ms/entry lines')
  ap.add_argument('--pmc-calibration', action='store_true',
                  help='first copy 256 MiB with torch (the reference launch '
                  'that scales FETCH_SIZE / WRITE_SIZE counters)')
  ap.add_argument('--krr', action='store_true',
                  help='only the kernel-regression solve benchmark')
  ap.add_argument('--krr-train', type=int, default=8192)
  ap.add_argument('--krr-test', type=int, default=1024)
  ap.add_argument('--krr-reps', type=int, default=7)
  ap.add_argument('--krr-tol', type=float, default=1e-7)
  ap.add_argument('--grad', action='store_true',
                  help='only the differentiation benchmark')
  ap.add_argument('--grad-targets', type=int, default=2048)
  args = ap.parse_args()
  if args.gpus != 1:
    print('bench.py runs on one GPU only: this engine has no multi-GPU path '
          '(pair tiles shard across GPUs without communication; see the '
          'design notes).', file=sys.stderr)
    sys.exit(2)
  if args.krr:
    out = krr_result(args.krr_train, args.krr_test, args.krr_reps,
                     args.krr_tol)
    out.update({'n_gpus': 1, 'higher_is_better': True,
                'dtype': 'fp32+fp64',
                'data': 'synthetic (blurred Gaussian images, seed 0)'})
    print(json.dumps(out), flush=True)
    return
  if args.grad:
    out = grad_result(args.grad_targets)
    out.update({'n_gpus': 1, 'dtype': 'fp32 (FP64 loss and solve)',
                'data': 'synthetic (blurred Gaussian images, seed 5)'})
    print(json.dumps(out), flush=True)
    return

  import numpy as np
  import torch
  from neural_tangents_amd import stax
  from neural_tangents_amd.data import blurred_images

  if args.pmc_calibration:
    src = torch.ones((64 << 20,), dtype=torch.float32, device='cuda')
    dst = torch.empty_like(src)
    dst.copy_(src)
    torch.cuda.synchronize()
    del src, dst

  depths, side, n1, n2 = WORKLOADS[args.workload]
  x1n, x2n = blurred_images([n1, n2], side, seed=0)
  dev = torch.device('cuda', 0)
  x1 = torch.from_numpy(x1n).to(dev)
  x2 = torch.from_numpy(x2n).to(dev)
  _, _, kernel_fn = stax.myrtle(depths=list(depths))

  log(f'{args.workload}: {args.warmup} warm-up + {args.steps} timed steps')
  elapsed, per, k = timed_kernel_fn(kernel_fn, x1, x2, args.steps,
                                    args.warmup)
  dom_name = max(per, key=lambda n: per[n]['ms'])
  dom = per[dom_name]
  avg_ms = dom['ms'] / dom['launches']
  avg_bytes = dom['bytes'] / dom['launches']
  achieved = avg_bytes / (avg_ms * 1e-3) / 1e9
  traffic = measured_traffic(dom_name, dom['pairs'] / dom['launches'])

  pairs = n1 * n2
  ms_per_step = elapsed / args.steps * 1e3
  value = pairs * args.steps / elapsed
  bpp = canonical_bytes_per_pair(depths, side)
  log(f'{args.workload}: {value:.0f} pairs/s')
  fused = dom_name == 'nta_first_pool_stage'
  vb = valu_busy(dom_name) if fused else None

  out = {
      'metric': 'Myrtle NTK pairs/sec + %HBM peak',
      'value': value,
      'unit': 'pairs/s',
      'n_gpus': 1,
      'steps': args.steps,
      'warmup': args.warmup,
      'ms_per_step': ms_per_step,
      'higher_is_better': True,
      'scaling': 'weak',
      'vs_baseline': None,
      'dtype': 'fp32',
      'data': 'synthetic (blurred Gaussian images, seed 0)',
      'config': {'workload': args.workload, 'depths': list(depths),
                 'image': [side, side, 3], 'n1': n1, 'n2': n2,
                 'get': ['nngp', 'ntk'], 'parallelism': 'single GPU'},
      'path_gbs_canonical': bpp * value / 1e9,
      'path_pct_hbm_peak': 100.0 * bpp * value / 1e9 / HBM_PEAK_GBS,
      'roofline': {'bound': 'hbm', 'limiter': 'valu' if fused else 'hbm',
                   'kernel': dom_name,
                   'achieved': achieved, 'peak': HBM_PEAK_GBS,
                   'unit': 'GB/s', 'frac': achieved / HBM_PEAK_GBS,
                   'traffic': traffic, 'avg_launch_ms': avg_ms,
                   'alg_bytes_per_launch': avg_bytes,
                   'valu_busy_frac': vb,
                   'note': SCHEDULE_NOTES.get(dom_name, '')},
      'kernels': kernel_table(per, args.steps),
  }
  got = (k.nngp.cpu().numpy(), k.ntk.cpu().numpy())
  del k
  ti = tile_rows(kernel_fn, x1, x2)

  if not args.no_m10 and args.workload == 'myrtle5_32x32_256x256':
    out['myrtle10'] = m10_result(3)
    torch.cuda.empty_cache()
  if not args.no_krr and args.workload == 'myrtle5_32x32_256x256':
    out['krr'] = krr_result(args.krr_train, args.krr_test, args.krr_reps,
                            args.krr_tol)
    torch.cuda.empty_cache()
  if not args.no_grad and args.workload == 'myrtle5_32x32_256x256':
    out['grad'] = grad_result()
    torch.cuda.empty_cache()
  if not args.no_readme and args.workload == 'myrtle5_32x32_256x256':
    out['readme_cnn21'] = readme_cnn21_result()
  if not args.no_cpu:
    # Error block: rows on both sides of the first row-tile boundary and in
    # the last (tail) row tile; first, middle and last columns.
    er = sorted({0, ti - 1, ti, n1 - ti // 2, n1 - 1})
    ec = sorted({0, n2 // 2, n2 - 2, n2 - 1})
    rows = 2 * host_threads()
    log(f'cpu leg: {rows}x48 FP32 sample, FP64 error block {er} x {ec}')
    base, err = cpu_leg(depths, x1n, x2n, got, rows, 48, er, ec)
    out['cpu_baseline'] = base
    out['rel_err_vs_oracle'] = {
        'nngp': err['nngp'], 'ntk': err['ntk'], 'oracle': 'FP64 numpy',
        'rows': er, 'cols': ec, 'row_tile': ti,
        'note': 'rows span the first pair-tile boundary and the tail tile'}
  print(json.dumps(out), flush=True)


if __name__ == '__main__':
  main()
