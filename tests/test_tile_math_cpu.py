"""CPU: the integer and rounding facts the tiled kernels rely on.

- The tile flush splits a tile index i into (row, column) with
  row = mulhi32(i, ceil(2^32 / tw)). That is exact for every i < 2^13 and
  2 <= tw <= 2^13 (checked exhaustively here). For tw = 1 the multiplier
  does not fit 32 bits, which is why one-column rectangles take the direct
  path.
- The interior projection leaves a quotient below 2^-40 / 2^-13 = 2^-27 in
  magnitude unrounded. Every q with |q| < 2^-26 gives fma(q, 0.5, 0.5) == 0.5
  in fp32, so the pixel does not depend on how such a quotient rounds; a
  negative q just past -2^-25 already moves the result, so the bound is
  within a factor of two of tight.
"""
import numpy as np


def test_row_split_is_exact_for_all_tile_indices_and_widths():
    i = np.arange(1 << 13, dtype=np.uint64)
    for tw in range(2, (1 << 13) + 1):
        magic = ((1 << 32) + tw - 1) // tw
        assert magic < (1 << 32)
        row = (i * np.uint64(magic)) >> np.uint64(32)
        assert np.array_equal(row, i // np.uint64(tw)), tw


def test_row_split_multiplier_overflows_for_one_column():
    # ceil(2^32 / 1) = 2^32, which truncates to 0 in 32 bits
    assert ((1 << 32) + 1 - 1) // 1 == 1 << 32
    assert np.uint32(((1 << 32) + 1 - 1) // 1 & 0xFFFFFFFF) == 0


def fma_half(q):
    """fp32 fma(q, 0.5, 0.5): q * 0.5 is exact, and for |q| < 2^-20 the sum
    q / 2 + 1 / 2 is exact in fp64, so one rounding to fp32 follows."""
    return np.float32(np.float64(np.float32(q)) * 0.5 + 0.5)


def test_tiny_quotients_land_on_the_centre():
    rng = np.random.default_rng(5)
    lim = np.float32(2.0 ** -26)
    below = np.nextafter(lim, np.float32(0))
    qs = np.concatenate([rng.uniform(-1, 1, 20000).astype(np.float32) * below,
                         np.array([0.0, -0.0, below, -below, 2.0 ** -27, -(2.0 ** -27), 1e-30, -1e-30],
                                  np.float32)])
    assert all(fma_half(q) == np.float32(0.5) for q in qs)
    # just past the rounding midpoints (0.5 - 2^-26 below, 0.5 + 2^-25 above)
    # the result moves
    assert fma_half(np.float32(-(2.0 ** -25) * (1 + 2.0 ** -23))) != np.float32(0.5)
    assert fma_half(np.float32(2.0 ** -24 * (1 + 2.0 ** -23))) != np.float32(0.5)
    # exactly at a midpoint, ties to even keep 0.5
    assert fma_half(np.float32(-(2.0 ** -25))) == np.float32(0.5)
