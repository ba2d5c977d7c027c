"""Initial conditions of the benchmark and test problems (cell arrays).

All functions return conservative variables (5, N[,N,N]) indexed [var, (k, j,) i]
for uniform cells of size 1/N on the unit box.
"""
import numpy as np

from .mesh import prim_to_cons


def sod_1d(ncell=1024, gamma=1.4, x0=0.5, left=(1.0, 0.0, 1.0), right=(0.125, 0.0, 0.1)):
    """Sod shock tube: (rho,u,p) = left for x < x0, right otherwise (namelist/tube1d.nml:31-40)."""
    x = (np.arange(ncell) + 0.5) / ncell
    L = x < x0
    rho = np.where(L, left[0], right[0])
    u = np.where(L, left[1], right[1])
    p = np.where(L, left[2], right[2])
    z = np.zeros(ncell)
    return prim_to_cons(rho, (u, z, z), p, gamma)


def sedov_3d(ncell, gamma=1.4, e_blast=1.0, rho0=1.0, p0=1e-5, nblast=2):
    """Sedov blast on the periodic unit box: ambient (rho0, 0, p0); the energy
    e_blast deposited uniformly in the central nblast^3 cells, i.e. pressure
    (gamma-1)*e_blast/V there, V the volume of those cells."""
    n = ncell
    rho = np.full((n, n, n), rho0)
    p = np.full((n, n, n), p0)
    dx = 1.0 / n
    lo = n // 2 - nblast // 2
    sl = slice(lo, lo + nblast)
    p[sl, sl, sl] = (gamma - 1.0) * e_blast / (nblast ** 3 * dx ** 3)
    z = np.zeros((n, n, n))
    return prim_to_cons(rho, (z, z, z), p, gamma)


def turbulence_3d(ncell, seed=20261202, gamma=5.0 / 3.0, rho0=1.0, p0=1.0, mach=5.0, kmin=2, kmax=8,
                  slope=-2.0):
    """Supersonic decaying turbulence: Gaussian random solenoidal velocity with
    shell spectrum P(k) ~ k^slope on kmin <= |k| <= kmax (projected so that
    k.v(k) = 0), normalised to rms Mach `mach` in an isothermal-start gas."""
    n = ncell
    rng = np.random.default_rng(seed)
    k1 = np.fft.fftfreq(n, d=1.0 / n)
    kz, ky, kx = np.meshgrid(k1, k1, k1, indexing="ij")
    kk = np.sqrt(kx ** 2 + ky ** 2 + kz ** 2)
    shell = (kk >= kmin) & (kk <= kmax)
    amp = np.zeros_like(kk)
    # |v(k)|^2 ~ P(k) / (4 pi k^2) so that the shell-integrated spectrum is k^slope
    amp[shell] = np.sqrt(kk[shell] ** slope / (4 * np.pi * kk[shell] ** 2))
    vk = []
    for _ in range(3):
        re = rng.standard_normal(kk.shape)
        im = rng.standard_normal(kk.shape)
        vk.append((re + 1j * im) * amp)
    # solenoidal projection: v -= k (k.v)/k^2
    with np.errstate(invalid="ignore", divide="ignore"):
        kdotv = kx * vk[0] + ky * vk[1] + kz * vk[2]
        inv = np.where(kk > 0, 1.0 / np.maximum(kk, 1e-30) ** 2, 0.0)
    vk = [vk[0] - kx * kdotv * inv, vk[1] - ky * kdotv * inv, vk[2] - kz * kdotv * inv]
    vel = [np.real(np.fft.ifftn(v)) for v in vk]
    vrms = np.sqrt(np.mean(vel[0] ** 2 + vel[1] ** 2 + vel[2] ** 2))
    cs = np.sqrt(gamma * p0 / rho0)
    scale = mach * cs / vrms
    vel = [v * scale for v in vel]
    rho = np.full((n, n, n), rho0)
    p = np.full((n, n, n), p0)
    return prim_to_cons(rho, (vel[0], vel[1], vel[2]), p, gamma)


def blast_3d(ncell, rng, gamma=1.4):
    """Seeded random blast: E in [0.1, 10], ambient p log-uniform in [1e-6, 1e-2]."""
    e = rng.uniform(0.1, 10.0)
    p0 = 10 ** rng.uniform(-6, -2)
    return sedov_3d(ncell, gamma=gamma, e_blast=e, p0=p0), {"kind": "blast", "E": e, "p0": p0}


def riemann_3d(ncell, rng, gamma=1.4):
    """Seeded planar two-state Riemann problem: rho, p log-uniform in [0.1, 10]
    per side, interface at the box centre along a random axis."""
    n = ncell
    rl, rr = 10 ** rng.uniform(-1, 1, 2)
    pl, pr = 10 ** rng.uniform(-1, 1, 2)
    axis = int(rng.integers(0, 3))
    idx = np.arange(n)
    left = (idx + 0.5) / n < 0.5
    shape = [1, 1, 1]
    shape[2 - axis] = n
    L = np.broadcast_to(left.reshape(shape), (n, n, n))
    rho = np.where(L, rl, rr)
    p = np.where(L, pl, pr)
    z = np.zeros((n, n, n))
    return prim_to_cons(rho, (z, z, z), p, gamma), {"kind": "riemann", "axis": axis,
                                                     "rho": (rl, rr), "p": (pl, pr)}


def random_state(shape, rng, gamma=1.4, vscale=1.0, near_vacuum=0.0):
    """Smooth-ish random state for parity tests (positive density and pressure)."""
    rho = np.exp(rng.normal(0, 0.5, shape))
    p = np.exp(rng.normal(0, 0.5, shape))
    vel = [rng.normal(0, vscale, shape) for _ in range(3)]
    if near_vacuum > 0:
        m = rng.random(shape) < near_vacuum
        rho[m] = 1e-12
        m = rng.random(shape) < near_vacuum
        p[m] = 1e-14
    return prim_to_cons(rho, vel, p, gamma)
