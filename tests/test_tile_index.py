"""The row of a fragment pair in k_tile (csrc/bin_kernels.h, tile_body): a
face's clipped bbox in a 16x16 tile is h rows of pw = ceil(w / 2) pixel pairs,
and pair li (0 <= li < h * pw) lies in row (li * ceil(2**16 / pw)) >> 16, which
must equal li // pw for every width and height a tile allows."""
import numpy as np


def test_pair_row_by_multiplier_is_exact():
    for w in range(1, 17):
        pw = (w + 1) // 2
        mult = (65536 + pw - 1) // pw
        li = np.arange(16 * pw, dtype=np.int64)
        assert mult < (1 << 17)  # fits bits 14-30 of Lstart
        assert np.array_equal((li * mult) >> 16, li // pw), w


def test_pair_columns_cover_the_row_once():
    # pair p of a row covers columns 2p and 2p + 1, the second only when inside the width
    for w in range(1, 17):
        pw = (w + 1) // 2
        cols = [c for p in range(pw) for c in (2 * p, 2 * p + 1) if c < w]
        assert cols == list(range(w))
