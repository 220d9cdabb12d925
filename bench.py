"""Benchmark: Myrtle-5 full-spatial NNGP+NTK on one MI355X.

One step = one ``kernel_fn(x1, x2, ('nngp', 'ntk'))`` over the whole
workload (Myrtle-5, 32x32x3, x1 and x2 each [256, 32, 32, 3] fp32 blurred
Gaussian images, seed 0): 65,536 cross pairs plus the 512 per-image self
chains.  Inputs are resident in HBM before the timed region.

Prints ONE JSON line.  ``value`` is pairs/s; ``roofline`` describes the
dominant kernel, timed with HIP events on its launch stream over the timed
steps; ``cpu_baseline`` is the FP32 oracle (the reference algorithm restated
in numpy, one thread) on a bounded sample of the same workload.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]
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


def measured_traffic(label, pairs_per_launch):
  """HBM bytes per launch of ``label`` from the committed counter summary
  (scaled to the launch's pair count), or None."""
  try:
    with open(PMC_FILE) as f:
      d = json.load(f)
    k = d['kernels'][label]
  except (OSError, KeyError, ValueError):
    return None
  return k['traffic_bytes_per_launch'] * pairs_per_launch / k[
      'pairs_per_launch']


def cpu_baseline(depths, x1, x2, n_sample):
  """Times the oracle (FP32, one thread) on an n_sample x n_sample block."""
  import numpy as np
  from threadpoolctl import threadpool_limits
  from oracle import stax_oracle as so
  layers = so.myrtle_layers(tuple(depths))
  a, b = x1[:n_sample], x2[:n_sample]
  with threadpool_limits(1):
    t0 = time.perf_counter()
    so.kernel_fn(layers, a, b, ('nngp', 'ntk'), dtype=np.float32)
    dt = time.perf_counter() - t0
  pairs = n_sample * n_sample
  return {'value': pairs / dt, 'unit': 'pairs/s', 'cores': 1,
          'kind': 'port',
          'sample': f'{n_sample}x{n_sample} block of the workload '
                    f'({pairs} pairs + {2 * n_sample} self chains), '
                    f'{dt:.1f} s'}


def krr_bench(args):
  """Kernel-regression solve (infinite-time gradient_descent_mse_ensemble):
  Myrtle-5 32x32 NTK of n_train synthetic images, one-hot targets (-0.1),
  diag_reg = 1e-4 (relative to the mean trace).  Times the plain FP64
  Cholesky solve (rocSOLVER through torch) and the mixed solve (FP32 MFMA
  Cholesky + FP64-residual PCG), both on the same kernel, and reports the
  relative error of alpha and of the test predictions against FP64."""
  import torch
  from neural_tangents_amd import _solve, stax
  from neural_tangents_amd.data import krr_workload
  ntr, nte = args.krr_train, args.krr_test
  xtr, xte, ytr, _ = krr_workload(ntr, nte, seed=0)
  dev = torch.device('cuda', 0)
  kernel_fn = stax.myrtle(5)[2]
  t0 = time.perf_counter()
  xtr_d = torch.from_numpy(xtr).to(dev)
  # K_dd by row blocks on and above the block diagonal (mirrored), so that
  # progress is printed every block.
  blk = 2048
  k_dd = torch.empty((ntr, ntr), dtype=torch.float32, device=dev)
  for i0 in range(0, ntr, blk):
    i1 = min(ntr, i0 + blk)
    k_dd[i0:i1, i0:] = kernel_fn(xtr_d[i0:i1], xtr_d[i0:], 'ntk')
    k_dd[i0:, i0:i1] = k_dd[i0:i1, i0:].T
    torch.cuda.synchronize()
    print(f'K_dd rows {i0}-{i1}: {time.perf_counter() - t0:.1f} s',
          file=sys.stderr, flush=True)
  k_dd = 0.5 * (k_dd + k_dd.T)        # exactly symmetric
  k_td = kernel_fn(torch.from_numpy(xte).to(dev), xtr_d, 'ntk')
  torch.cuda.synchronize()
  t_kernel = time.perf_counter() - t0
  print(f'kernels: {t_kernel:.1f} s', file=sys.stderr, flush=True)
  y = torch.from_numpy(ytr).to(dev, torch.float64)
  res = {}
  for mode in ('fp64', 'mixed'):
    times = []
    for rep in range(args.krr_reps):
      torch.cuda.synchronize()
      t = time.perf_counter()
      s = _solve.CholeskySolver(k_dd, 1e-4, False, mode)
      alpha = s.solve(y)
      f = k_td.to(torch.float64) @ alpha
      torch.cuda.synchronize()
      times.append(time.perf_counter() - t)
    res[mode] = (sorted(times)[len(times) // 2], alpha, f,
                 getattr(s, 'iterations', 0), getattr(s, 'last_residual',
                                                      None))
  t64, a64, f64 = res['fp64'][:3]
  tmx, amx, fmx, iters, resid = res['mixed']
  err_a = float(torch.linalg.norm(amx - a64) / torch.linalg.norm(a64))
  err_f = float(torch.linalg.norm(fmx - f64) / torch.linalg.norm(f64))
  out = {
      'metric': 'KRR solve time vs plain FP64 at <=1e-6 rel err',
      'value': t64 / tmx, 'unit': 'x speed-up (plain FP64 / mixed)',
      'n_gpus': 1, 'higher_is_better': True, 'dtype': 'fp32+fp64',
      'data': 'synthetic (blurred Gaussian images, seed 0)',
      'config': {'workload': f'krr_myrtle5_32x32_train{ntr}_test{nte}',
                 'diag_reg': 1e-4, 'get': 'ntk'},
      'plain_fp64_ms': t64 * 1e3, 'mixed_ms': tmx * 1e3,
      'pcg_iterations': iters, 'final_rel_residual': resid,
      'rel_err_alpha_vs_fp64': err_a, 'rel_err_ftest_vs_fp64': err_f,
      'kernel_seconds': t_kernel,
  }
  print(json.dumps(out), flush=True)


# How the algorithmic bytes of a launch relate to what it moves.
SCHEDULE_NOTES = {
    'nta_first_pool_stage':
        'first two stages fused (Conv-Relu-Conv-Relu-AvgPool): achieved '
        'counts the per-layer schedule\'s bytes for those two kernels (the '
        'first stage state written and read back, the pooled state written); '
        'the kernel never materialises the first stage, so traffic (measured '
        'HBM bytes) is far below them and its bound is the vector ALU',
}


def main():
  ap = argparse.ArgumentParser()
  ap.add_argument('--gpus', type=int, default=1)
  ap.add_argument('--steps', type=int, default=3)
  ap.add_argument('--warmup', type=int, default=1)
  ap.add_argument('--workload', default='myrtle5_32x32_256x256',
                  choices=sorted(WORKLOADS))
  ap.add_argument('--no-cpu', action='store_true')
  ap.add_argument('--cpu-sample', type=int, default=14)
  ap.add_argument('--pmc-calibration', action='store_true',
                  help='first copy 256 MiB with torch (the reference launch '
                  'that scales FETCH_SIZE / WRITE_SIZE counters)')
  ap.add_argument('--krr', action='store_true',
                  help='kernel-regression solve benchmark instead')
  ap.add_argument('--krr-train', type=int, default=8192)
  ap.add_argument('--krr-test', type=int, default=1024)
  ap.add_argument('--krr-reps', type=int, default=3)
  args = ap.parse_args()
  if args.gpus != 1:
    print('bench.py runs on one GPU only: this engine has no multi-GPU path '
          '(pair tiles shard across GPUs without communication; see the '
          'design notes).', file=sys.stderr)
    sys.exit(2)
  if args.krr:
    krr_bench(args)
    return

  import torch
  from neural_tangents_amd import _engine, stax
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

  for _ in range(args.warmup):
    kernel_fn(x1, x2, ('nngp', 'ntk'))
  torch.cuda.synchronize()

  _engine.TRACE = []
  torch.cuda.synchronize()
  t0 = time.perf_counter()
  for _ in range(args.steps):
    k = kernel_fn(x1, x2, ('nngp', 'ntk'))
  torch.cuda.synchronize()
  elapsed = time.perf_counter() - t0
  trace, _engine.TRACE = _engine.TRACE, None

  # Per-entry-point launch statistics (HIP events on the launch stream).
  per = {}
  for name, s, e, nbytes, pairs in trace:
    d = per.setdefault(name, {'ms': 0.0, 'launches': 0, 'bytes': 0,
                              'pairs': 0})
    d['ms'] += s.elapsed_time(e)
    d['launches'] += 1
    d['bytes'] += nbytes
    d['pairs'] += pairs
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
      'roofline': {'bound': 'hbm', 'kernel': dom_name,
                   'achieved': achieved, 'peak': HBM_PEAK_GBS,
                   'unit': 'GB/s', 'frac': achieved / HBM_PEAK_GBS,
                   'traffic': traffic, 'avg_launch_ms': avg_ms,
                   'alg_bytes_per_launch': avg_bytes,
                   'note': SCHEDULE_NOTES.get(dom_name, '')},
      'kernels': {n: {'ms_per_step': v['ms'] / args.steps,
                      'launches_per_step': v['launches'] / args.steps,
                      'avg_launch_ms': v['ms'] / v['launches'],
                      'gbs': v['bytes'] / (v['ms'] * 1e-3) / 1e9
                      if v['ms'] > 0 else None}
                  for n, v in per.items()},
  }
  del k
  if not args.no_cpu:
    out['cpu_baseline'] = cpu_baseline(depths, x1n, x2n, args.cpu_sample)
  print(json.dumps(out), flush=True)


if __name__ == '__main__':
  main()
