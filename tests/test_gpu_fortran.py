"""The Fortran side of the boundary: ramses2_amd/fortran/sod_driver.f90 calls the
library through hip_interface (iso_c_binding) in the reference's order of level
calls; its Sod run must agree with the CPU oracle run of the same problem."""
import os
import subprocess

import numpy as np
import pytest

from ramses2_amd import HydroParams, UniformMesh, ic
from ramses2_amd import build as B
from ramses2_amd.hydro import newdt
from tests.helpers import oracle_params, oracle_uniform, rel_err_per_var

pytestmark = pytest.mark.gpu


def test_fortran_sod_driver(tmp_path):
    exe = B.fortran_driver_path()
    if not os.path.exists(exe):
        pytest.skip("Fortran driver not built (flang absent at build time)")
    out = tmp_path / "sod.bin"
    print("running", exe, flush=True)
    subprocess.run([exe, str(out)], check=True, timeout=120)
    raw = out.read_bytes()
    nstep = int(np.frombuffer(raw[:4], dtype=np.int32)[0])
    t = float(np.frombuffer(raw[4:12], dtype=np.float64)[0])
    u = np.frombuffer(raw[12:], dtype=np.float64).reshape(512, 5, 2)

    mesh = UniformMesh(1, 10)
    p = HydroParams(gamma=1.4, courant_factor=0.8, slope_type=1, riemann="hllc")
    po = oracle_params(p)
    om = oracle_uniform(mesh, mesh.from_cells(ic.sod_1d(1024)), periodic=False)
    tt, dtold, n = 0.0, 0.0, 0
    while tt < 0.2:
        c = om.courant_fine(10, po)
        dt = newdt(mesh.dx, p, c[4], tt, 0.2, 3, dtold)
        om.set_unew(10)
        om.godunov_fine(10, dt, po)
        om.set_uold(10)
        tt += dt
        dtold = dt
        n += 1
    assert nstep == n
    assert abs(t - tt) < 1e-12
    assert np.all(rel_err_per_var(u, om.uold) < 1e-9)
