"""Pair-tile shape of the backward (host logic, no GPU work)."""
from neural_tangents_amd.grad import _pair_tile


def test_pair_tiles_take_whole_rows_within_the_budget():
  assert _pair_tile(2048, 100, 512, 32) == (5, 100)
  assert _pair_tile(100, 100, 512, 32) == (5, 100)
  assert _pair_tile(5, 3, 2, 8) == (1, 2)
  assert _pair_tile(3, 4, 512, 8) == (3, 4)
  assert _pair_tile(1, 5000, 512, 32) == (1, 512)


def test_pair_tiles_keep_states_within_32_bit_indexing():
  # a tile's [pairs, side^4] state holds at most 2^30 elements
  assert _pair_tile(100, 100, 512, 64) == (1, 64)
  assert _pair_tile(100, 100, 512, 128) == (1, 4)
  assert _pair_tile(10, 10, 512, 256) == (1, 1)
  for n1, n2, tp, side in ((2048, 100, 512, 32), (7, 9, 1024, 64),
                           (1, 1, 512, 512), (37, 5, 3, 16)):
    ti, tj = _pair_tile(n1, n2, tp, side)
    assert 1 <= ti <= n1 and 1 <= tj <= n2
    assert ti * tj <= max(1, min(tp, (1 << 30) // side ** 4))
