"""The autotuner's work model (CPU): brick counts and z-trimmed plane counts
against closed forms."""
import numpy as np
import pytest

from ramses2_amd.autotune import brick_offset, count_bricks, level_work


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


def test_lattice_offset_follows_the_octs():
    """A 4^3 block of octs at (2,2,2) spans 8 aligned bricks of 4^3 and one
    brick of the lattice shifted by (2,2,2), which the library picks."""
    k = _block(4) + 2
    off, w = brick_offset(k, (4, 4, 4), (0, 0, 0), (64, 64, 64))
    assert off == (2, 2, 2)
    assert w == pytest.approx(0.35 * 12 * 12 * (8 + 4) + 10 * 10 * (8 + 2), rel=1e-12)


def test_dense_level_keeps_the_aligned_lattice():
    """A 6^3 block at (1,1,1) fills more than half of a 7^3 box: offset 0
    (4 bricks of 2 octs per dimension); in a large box the lattice shifted by
    one holds it in 3 per dimension."""
    k = _block(6) + 1
    off, w_dense = brick_offset(k, (2, 2, 2), (0, 0, 0), (7, 7, 7))
    assert off == (0, 0, 0)
    off, w_sparse = brick_offset(k, (2, 2, 2), (0, 0, 0), (64, 64, 64))
    assert off == (1, 1, 1)
    assert w_sparse == pytest.approx(27 * (0.35 * 8 * 8 * (4 + 4) + 6 * 6 * (4 + 2)), rel=1e-12)
    assert w_sparse < w_dense


def test_lattice_offset_keeps_origins_in_the_box():
    """An oct at x = box_min rules out every x offset but 0 (the brick holding
    it would start below box_min)."""
    k = np.concatenate([_block(4) + 2, np.array([[0, 3, 3]])])
    off, _ = brick_offset(k, (4, 4, 4), (0, 0, 0), (64, 64, 64))
    assert off[0] == 0 and off[1:] == (2, 2)
    off, _ = brick_offset(k + 5, (4, 4, 4), (5, 5, 5), (69, 69, 69))  # the same relative to box_min
    assert off == (0, 2, 2)
