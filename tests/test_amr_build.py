"""The static AMR mesh builder: the gradient criterion of hydro_refine, the
nesting the coarse-fine path needs, and the restricted initial state."""
import numpy as np

from ramses2_amd.amr_build import build_levels, refine_error, sedov_amr, sedov_profile


def test_refine_error_closed_forms():
    # uniform: no error; a jump 1 -> 2 on one side: 2*(1/3); symmetric in g<->d
    assert refine_error(1.0, 1.0, 1.0, 1e-10) == 0.0
    assert abs(refine_error(1.0, 1.0, 2.0, 0.0) - 2.0 / 3.0) < 1e-15
    assert refine_error(2.0, 1.0, 1.0, 0.0) == refine_error(1.0, 1.0, 2.0, 0.0)
    # the floor keeps vacuum finite
    assert np.isfinite(refine_error(0.0, 0.0, 0.0, 1e-10))
    # scale invariance (a relative criterion)
    assert abs(refine_error(3.0, 5.0, 7.0, 0.0) - refine_error(30.0, 50.0, 70.0, 0.0)) < 1e-15


def test_levels_are_nested_for_the_coarse_fine_path():
    prof = lambda x, y, z: sedov_profile(x, y, z, radius=0.3)  # noqa: E731
    keys = build_levels(3, 6, prof)
    assert set(keys) == {3, 4, 5, 6}
    for l in range(3, 6):
        coarse = {tuple(k) for k in keys[l]}
        n = 2 ** (l - 1)
        for k in keys[l + 1]:
            p = tuple(int(c) for c in np.asarray(k) // 2)
            assert p in coarse
            for dz in (-1, 0, 1):
                for dy in (-1, 0, 1):
                    for dx in (-1, 0, 1):
                        q = ((p[0] + dx) % n, (p[1] + dy) % n, (p[2] + dz) % n)
                        assert q in coarse, (l, k, q)


def test_flags_follow_the_criterion():
    """Every cell of level 4 whose density or pressure error exceeds 0.1 and
    whose oct has all its neighbours is refined."""
    prof = lambda x, y, z: sedov_profile(x, y, z, radius=0.3)  # noqa: E731
    keys = build_levels(3, 5, prof)
    fine = {tuple(k) for k in keys[5]}
    lvl4 = {tuple(k) for k in keys[4]}
    dx = 1.0 / 16
    checked = 0
    for k in keys[4]:
        if not all(((k[0] + a) % 8, (k[1] + b) % 8, (k[2] + c) % 8) in lvl4
                   for a in (-1, 0, 1) for b in (-1, 0, 1) for c in (-1, 0, 1)):
            continue
        for ind in range(8):
            cell = (2 * k[0] + (ind & 1), 2 * k[1] + ((ind >> 1) & 1), 2 * k[2] + ((ind >> 2) & 1))
            x = (np.array(cell) + 0.5) * dx
            m = prof(*x)
            err = 0.0
            for d in range(3):
                e = np.zeros(3)
                e[d] = dx
                lo, hi = prof(*(x - e)), prof(*(x + e))
                err = max(err, refine_error(lo[0], m[0], hi[0], 1e-10), refine_error(lo[4], m[4], hi[4], 1e-10))
            if err > 0.1:
                assert cell in fine
                checked += 1
    assert checked > 0


def test_initial_state_is_restricted():
    am, u = sedov_amr(levelmin=3, levelmax=5, radius=0.3)
    for l in am.levels[1:]:
        i0 = am.head[l] - 1
        kf = am.ckey[i0:i0 + am.noct[l]].astype(np.int64)
        par, ind = am.parent_cells(l, kf)
        np.testing.assert_allclose(u[par, :, ind], u[i0:i0 + am.noct[l]].mean(axis=2), rtol=0, atol=0)
        assert np.all(am.refined[par, ind] == 1)
    # leaf cells carry positive density and pressure
    assert np.all(u[:, 0] > 0)
