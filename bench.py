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


def main():
  ap = argparse.ArgumentParser()
  ap.add_argument('--gpus', type=int, default=1)
  ap.add_argument('--steps', type=int, default=3)
  ap.add_argument('--warmup', type=int, default=1)
  ap.add_argument('--workload', default='myrtle5_32x32_256x256',
                  choices=sorted(WORKLOADS))
  ap.add_argument('--no-cpu', action='store_true')
  ap.add_argument('--cpu-sample', type=int, default=14)
  args = ap.parse_args()
  if args.gpus != 1:
    print('bench.py runs on one GPU only: this engine has no multi-GPU path '
          '(pair tiles shard across GPUs without communication; see the '
          'design notes).', file=sys.stderr)
    sys.exit(2)

  import torch
  from neural_tangents_amd import _engine, stax
  from neural_tangents_amd.data import blurred_images

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
                   'traffic': None, 'avg_launch_ms': avg_ms,
                   'alg_bytes_per_launch': avg_bytes},
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
