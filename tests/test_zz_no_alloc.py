"""rh_cmpdt_noalloc (courant_fine with its partial sums in a static device
array of the library, for callers that cannot allocate device memory) against
rh_cmpdt on the same device state: identical results on two calls in a row,
the second over a smaller oct range, and the oracle's courant_fine."""
import ctypes

import numpy as np
import pytest

from ramses2_amd import HydroParams, UniformLevel, UniformMesh, ic
from ramses2_amd._lib import call, ptr
from tests.helpers import oracle_params, oracle_uniform

pytestmark = pytest.mark.gpu


def cmpdt(name, lev, head, num, p):
    g = np.zeros(3)
    out = [ctypes.c_double() for _ in range(5)]
    call(name, head, num, lev.mesh.dx, p.gamma, p.smallr, p.smallc ** 2, p.courant_factor, ptr(g),
         *[ctypes.byref(o) for o in out])
    return np.array([o.value for o in out])


def test_noalloc_courant_matches_cmpdt():
    mesh = UniformMesh(3, 5)
    p = HydroParams(gamma=1.4, courant_factor=0.5, slope_type=2, riemann="hllc")
    lev = UniformLevel(mesh, ic.sedov_3d(32), p)
    lev.run_device(3)
    print("case: 32^3 Sedov after 3 steps, whole level then an eighth of it", flush=True)
    for head, num in ((1, mesh.noct), (1, mesh.noct // 8)):
        a = cmpdt("rh_cmpdt", lev, head, num, p)
        b = cmpdt("rh_cmpdt_noalloc", lev, head, num, p)
        assert np.array_equal(a, b), (head, num, a, b)
    om = oracle_uniform(mesh, lev.state())
    c = om.courant_fine(5, oracle_params(p))
    b = cmpdt("rh_cmpdt_noalloc", lev, 1, mesh.noct, p)
    assert abs(b[4] - c[4]) <= 1e-14 * c[4]
    # sums of 32768 cells in another order: agreement to a few ulp of the total
    assert np.all(np.abs(b[:3] - c[:3]) <= 1e-12 * np.abs(c[:3]))
