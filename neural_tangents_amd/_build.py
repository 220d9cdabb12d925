"""Builds libnta.so (all HIP sources under csrc/) for gfx950, in-tree.

Usage: ``python -m neural_tangents_amd._build`` or ``build()`` from Python.
"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, 'csrc')
LIB = os.path.join(HERE, 'libnta.so')
ROOT = os.path.dirname(HERE)

HIP_FLAGS = [
    '-O3', '-std=c++17', '-fPIC', '-shared', '--offload-arch=gfx950',
    # Packed fp32 ops are kept out of the generated code on purpose.
    '-Xclang', '-target-feature', '-Xclang', '-packed-fp32-ops',
    '-Wno-unused-command-line-argument',
]


def sources():
  return sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC)
                if f.endswith('.hip'))


def build(verbose: bool = False) -> str:
  hipcc = shutil.which('hipcc') or '/opt/rocm/bin/hipcc'
  tmp = LIB + '.tmp'
  cmd = [hipcc] + HIP_FLAGS + ['-I', os.path.join(ROOT, 'include'),
                               '-o', tmp] + sources()
  if verbose:
    print(' '.join(cmd), flush=True)
  subprocess.run(cmd, check=True)
  os.replace(tmp, LIB)
  return LIB


if __name__ == '__main__':
  build(verbose='-v' in sys.argv)
  print(LIB)
