"""The initial conditions of the benchmark workloads, pinned by what defines
them: the decaying-turbulence field (solenoidal, Gaussian, P(k) ~ k^-2 on
2 <= |k| <= 8, rms Mach 5 at rho = p = 1, gamma = 5/3) and the Sedov blast
(E deposited uniformly in the central 2^3 cells)."""
import numpy as np

from ramses2_amd import ic


def prim(u, gamma):
    rho = u[0]
    v = u[1:4] / rho
    p = (gamma - 1.0) * (u[4] - 0.5 * rho * (v ** 2).sum(0))
    return rho, v, p


def test_turbulence_field():
    n, gamma = 64, 5.0 / 3.0
    u = ic.turbulence_3d(n, seed=7, gamma=gamma)
    rho, v, p = prim(u, gamma)
    assert np.allclose(rho, 1.0, rtol=0, atol=0)
    assert np.allclose(p, 1.0, rtol=1e-9)
    cs = np.sqrt(gamma)
    assert abs(np.sqrt((v ** 2).sum(0).mean()) / cs - 5.0) < 1e-12
    # spectral content: only 2 <= |k| <= 8, and k . v(k) = 0 (solenoidal)
    k1 = np.fft.fftfreq(n, d=1.0 / n)
    kz, ky, kx = np.meshgrid(k1, k1, k1, indexing="ij")
    kk = np.sqrt(kx ** 2 + ky ** 2 + kz ** 2)
    vk = [np.fft.fftn(c) for c in v]
    amp = np.sqrt(sum(np.abs(c) ** 2 for c in vk))
    shell = (kk >= 2) & (kk <= 8)
    assert amp[~shell].max() <= 1e-10 * amp.max()
    div = np.abs(kx * vk[0] + ky * vk[1] + kz * vk[2])
    assert div.max() <= 1e-9 * (kk * amp).max()
    # shell spectrum E(k) ~ k^-2 (log-log fit over the integer shells)
    ks = np.arange(2, 9)
    ek = np.array([(amp[(kk >= k - 0.5) & (kk < k + 0.5)] ** 2).sum() for k in ks])
    slope = np.polyfit(np.log(ks), np.log(ek), 1)[0]
    assert abs(slope + 2.0) < 0.3, slope
    # the same seed gives the same field
    assert np.array_equal(u, ic.turbulence_3d(n, seed=7, gamma=gamma))


def test_sedov_blast():
    n, gamma = 16, 1.4
    u = ic.sedov_3d(n, gamma=gamma)
    rho, v, p = prim(u, gamma)
    assert np.all(rho == 1.0) and np.all(v == 0.0)
    dx = 1.0 / n
    e = u[4] * dx ** 3
    hot = np.zeros((n, n, n), bool)
    hot[n // 2 - 1:n // 2 + 1, n // 2 - 1:n // 2 + 1, n // 2 - 1:n // 2 + 1] = True
    assert abs(e[hot].sum() - 1.0) < 1e-12  # the blast pressure replaces the ambient one there
    assert np.allclose(p[~hot], 1e-5, rtol=1e-12)
    assert np.allclose(p[hot], (gamma - 1) * 1.0 / (8 * dx ** 3), rtol=1e-12)
