"""The autotuner's work model (CPU): brick counts and z-trimmed plane counts
against closed forms."""
import numpy as np

from ramses2_amd.autotune import count_bricks, level_work


def _block(n, z=None):
    ii = np.arange(n)
    k = np.stack(np.meshgrid(ii, ii, ii, indexing="ij"), -1).reshape(-1, 3)
    return k if z is None else k[k[:, 2] == z]


def test_full_block():
    k = _block(8)
    assert count_bricks(k, (7, 7, 8)) == 4
    # 4 bricks, each marching its 16 cell planes plus halos
    want = 4 * (0.35 * 18 * 18 * (16 + 4) + 16 * 16 * (16 + 2))
    assert level_work(k, (7, 7, 8)) == want


def test_single_layer_is_trimmed():
    k = _block(6, z=3)                      # one oct layer, inside z-brick 0 of shape (3,3,4)
    assert count_bricks(k, (3, 3, 4)) == 4
    want = 4 * (0.35 * 10 * 10 * (2 + 4) + 8 * 8 * (2 + 2))
    assert level_work(k, (3, 3, 4)) == want
    # the untrimmed march would cover all 8 planes of each brick
    assert level_work(k, (3, 3, 4)) < 4 * (0.35 * 10 * 10 * 12 + 8 * 8 * 10)
