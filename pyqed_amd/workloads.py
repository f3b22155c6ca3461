"""Seeded synthetic inputs for the open-system workloads this package is measured on.

Every generator draws from ``numpy.random.default_rng(seed)`` in the order
written below, so the same seed gives the same inputs on every machine.
Units follow pyqed/units.py:1-25 (atomic units; cm^-1, fs and K converted with
the constants below).

* spin_boson       two-level system, Drude bath (1 Drude + 1 Pade pole), depth 5.
* exciton_c2       7-site Frenkel exciton, per-site Drude baths (K=14), depth 4.
* batched_c3       4096 uneven exciton chains, d in 2..16, depth in 2..4.
* lambda_c4        3-level Lambda system with Lindblad dephasing and decay,
                   piecewise-constant laser field of 200 segments.
* chain_c5         24-site exciton chain, per-site Drude baths (K=48), depth 3.
"""
import numpy as np

AU2FS = 2.41888432651e-2          # pyqed/units.py:1
AU2K = 315775.13                  # pyqed/units.py:5
WAVENUMBER2HARTREE = 4.55633525277e-06  # pyqed/units.py:19


def _site_projectors(d):
    Q = np.zeros((d, d, d), np.complex128)
    for j in range(d):
        Q[j, j, j] = 1.0
    return Q


def _ground(d):
    rho0 = np.zeros((d, d), np.complex128)
    rho0[0, 0] = 1.0
    return rho0


def spin_boson():
    """H = (eps/2) sz + (Delta/2) sx with eps = Delta = 1, V = sz, Drude lambda=0.1, gamma=0.5, kT=1."""
    sx = np.array([[0, 1], [1, 0]], np.complex128)
    sz = np.array([[1, 0], [0, -1]], np.complex128)
    return dict(H=0.5 * sz + 0.5 * sx, Q=np.array([sz]), mode=np.array([0, 0]), lam=0.1, gam=0.5, beta=1.0,
                npsd=1, L=5, rho0=_ground(2), dt=0.01, nt=1000, p1=sz)


def exciton_c2(seed=2):
    """7 sites; eps_j ~ N(0, 100 cm^-1); J_ij = -50 cm^-1 / |i-j|^3 (unit spacing); per-site Drude
    lambda = 35 cm^-1, gamma^-1 = 50 fs, T = 300 K, npsd = 1; depth 4; dt = 1 fs."""
    rng = np.random.default_rng(seed)
    d = 7
    eps = rng.normal(0.0, 100.0, d)
    H = np.diag(eps).astype(np.complex128)
    for i in range(d):
        for j in range(d):
            if i != j:
                H[i, j] = -50.0 / abs(i - j) ** 3
    H *= WAVENUMBER2HARTREE
    return dict(H=H, Q=_site_projectors(d), mode=np.repeat(np.arange(d), 2), lam=35.0 * WAVENUMBER2HARTREE,
                gam=AU2FS / 50.0, beta=AU2K / 300.0, npsd=1, L=4, rho0=_ground(d), dt=1.0 / AU2FS, nt=1000)


def batched_c3(n=4096, seed=3):
    """n chains; per instance, in this order: d ~ U{2..16}, L ~ U{2,3,4}, eps ~ N(0,1)^d,
    J ~ U(-0.5,0.5)^(d-1) (tridiagonal), lambda ~ U(0.05,0.5), gamma ~ U(0.2,2), kT ~ U(0.5,2).
    The bath parameters are drawn once per instance and shared by all of its sites; npsd = 1 so K = 2d."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        d = int(rng.integers(2, 17))
        L = int(rng.integers(2, 5))
        eps = rng.normal(0.0, 1.0, d)
        J = rng.uniform(-0.5, 0.5, d - 1)
        lam = float(rng.uniform(0.05, 0.5))
        gam = float(rng.uniform(0.2, 2.0))
        kT = float(rng.uniform(0.5, 2.0))
        H = np.diag(eps).astype(np.complex128)
        H += np.diag(J, 1) + np.diag(J, -1)
        out.append(dict(d=d, L=L, H=H, Q=_site_projectors(d), mode=np.repeat(np.arange(d), 2), lam=lam,
                        gam=gam, beta=1.0 / kT, npsd=1, rho0=_ground(d)))
    return out


def lambda_c4(seed=4, nseg=200, steps_per_seg=20, dt=0.05):
    """Levels (0, 1, 2.5), dipole mu01 = 1, mu12 = 0.7, H(t) = H0 - mu E(t); dephasing sqrt(0.02)|j><j|
    on every level plus decay sqrt(0.01)|0><1|; amplitudes a_s ~ N(0, 0.1); objective rho_22(T)."""
    rng = np.random.default_rng(seed)
    amps = rng.normal(0.0, 0.1, nseg)
    eps = np.array([0.0, 1.0, 2.5])
    mu = np.array([[0, 1, 0], [1, 0, 0.7], [0, 0.7, 0]], np.complex128)
    c_ops = []
    for j in range(3):
        c = np.zeros((3, 3), np.complex128)
        c[j, j] = np.sqrt(0.02)
        c_ops.append(c)
    c = np.zeros((3, 3), np.complex128)
    c[0, 1] = np.sqrt(0.01)
    c_ops.append(c)
    return dict(eps=eps, mu=mu, c_ops=np.array(c_ops), amps=amps, nseg=nseg, steps_per_seg=steps_per_seg, dt=dt,
                nt=nseg * steps_per_seg, rho0=_ground(3), checkpoint_every=50)


def chain_c5(seed=5):
    """24 sites; eps ~ N(0, 0.2); J = -1 nearest neighbour; per-site Drude lambda=0.2, gamma=1, kT=1,
    npsd=1 (K=48); depth 3; dt = 0.005, 5000 steps."""
    rng = np.random.default_rng(seed)
    d = 24
    eps = rng.normal(0.0, 0.2, d)
    H = np.diag(eps).astype(np.complex128)
    H += np.diag(-np.ones(d - 1), 1) + np.diag(-np.ones(d - 1), -1)
    return dict(H=H, Q=_site_projectors(d), mode=np.repeat(np.arange(d), 2), lam=0.2, gam=1.0, beta=1.0,
                npsd=1, L=3, rho0=_ground(d), dt=0.005, nt=5000)


def nmax_of(K, L):
    from math import comb
    return comb(K + L, L)
