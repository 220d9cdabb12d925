"""Pinned synthetic inputs for the Myrtle workloads.

Images: ``numpy.random.default_rng(seed)`` (PCG64) draws ``x1`` and then
``x2`` as float64 ``standard_normal([N, S, S, 3])``; each channel gets a 3x3
box blur (``scipy.ndimage.uniform_filter``, ``mode='reflect'``) and is then
standardised over ``(N, H, W)`` of its own array (ddof=0); the result is cast
to float32.  Targets: one-hot over 10 classes from ``rng.integers(0, 10, N)``,
centred by -0.1.  The reference's own RNG (``jax.random``) is not used, so
this generator defines the workload.
"""
import hashlib

import numpy as np


def _blur_standardise(x: np.ndarray) -> np.ndarray:
  import scipy.ndimage
  x = scipy.ndimage.uniform_filter(x, size=(1, 3, 3, 1), mode='reflect')
  mean = x.mean(axis=(0, 1, 2), keepdims=True)
  std = x.std(axis=(0, 1, 2), keepdims=True)
  return ((x - mean) / std).astype(np.float32)


def blurred_images(sizes, side, channels=3, seed=0):
  """Returns one float32 ``[n, side, side, channels]`` array per entry of
  ``sizes``, drawn in order from one generator."""
  rng = np.random.default_rng(seed)
  out = []
  for n in sizes:
    x = rng.standard_normal((n, side, side, channels))
    out.append(_blur_standardise(x))
  return out


def one_hot_targets(n, classes=10, seed=0, rng=None):
  """``onehot(rng.integers(0, classes, n)) - 0.1`` as float32."""
  rng = np.random.default_rng(seed) if rng is None else rng
  lab = rng.integers(0, classes, n)
  y = np.zeros((n, classes), np.float32)
  y[np.arange(n), lab] = 1.0
  return y - np.float32(0.1)


def krr_workload(n_train, n_test, side=32, channels=3, classes=10, seed=0):
  """Train/test images and one-hot targets for the kernel-regression form."""
  rng = np.random.default_rng(seed)
  xs = []
  for n in (n_train, n_test):
    xs.append(_blur_standardise(rng.standard_normal((n, side, side,
                                                     channels))))
  y_train = one_hot_targets(n_train, classes, rng=rng)
  y_test = one_hot_targets(n_test, classes, rng=rng)
  return xs[0], xs[1], y_train, y_test


def sha256(a: np.ndarray) -> str:
  return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()
