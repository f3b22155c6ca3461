"""CPU oracle for the legacy single-exponential HEOM (oqs.HEOMSolver and its second copy
pyqed.HEOM.heom.HEOMSolver) -- TEST INFRASTRUCTURE ONLY.

Imported only by tests/ as the checker.

Reference: pyqed/oqs.py (binggu56/pyqed 0.1.3), pyqed/HEOM/heom.py, pyqed/phys.py,
pyqed/superoperator.py, pyqed/units.py.

  run_params         oqs.py:1826-1845 (_heom): gamma = cutoff, T = temperature (already k_B T),
                     D0 = reorg * gamma * (coth(gamma / 2T) - i)
  propagator_params  oqs.py:1893-1908 (_heom_propagator): T = temperature / au2k (units.py:6),
                     a = pi * reorg * T, b = 0
  sweep              one forward-Euler step with the in-place sweep over the hierarchy
                     (oqs.py:1849-1862 and 1920-1930): ADO 0 first, then n = n_first .. nado-2, each
                     update reading ADO n-1 of this step and ADO n+1 of the previous one; the last
                     ADO is never updated
  heom_run           oqs.py:1808-1874: commutator form on d x d ADOs, n_first = 1, observables
                     Tr(e rho) (superoperator.py:313-314, obs = vdot(vec(e^dag), vec(rho))) after every
                     step, shape (n_e, nt)
  heom_propagator    oqs.py:1876-1935: left products of the N x N blocks u[n] (N = d^2) with
                     L0 = op2sop(H), Sm = op2sop(S), Sp = op2sop(S, 'anticommutator')
                     (superoperator.py:200-246); n_first = 0, so u[0] is updated twice per step;
                     n_first = 1 is HEOM/heom.py:349-413, which updates u[0] once
  run_params_rk      HEOM/heom.py:295-313: gamma = cutoff, T = temperature, D0 = reorg (2T - i gamma)
  rhs_all            HEOM/heom.py:319-331: F_n for n = 0..nado-2 from one hierarchy, F_{nado-1} = 0
  heom_run_rk4       HEOM/heom.py:275-346 with rk4 of phys.py:1051-1064:
                     y += (k1 + 2 k2 + 2 k3 + k4) / 6 dt, observables after every step

What pins them (tests/test_oracle_legacy_heom.py):
  * reorganization 0: heom_run's rho follows (1 - i dt L0)^k vec(rho0) and the propagator's u[0] is
    (1 - i dt L0)^(2k) (the double update), both from numpy matrix powers;
  * pure dephasing ([H, S] = 0), where the single-exponential HEOM is exact: rho_ab(t) =
    rho_ab(0) exp(-i w_ab t - (q_a - q_b)(q_a D0 - q_b conj(D0)) G(t)), G(t) = t/gamma -
    (1 - exp(-gamma t))/gamma^2, reached at first order in dt (checked at two step sizes) with a
    deep enough hierarchy; with asymmetric eigenvalues of S this fixes Re D0, Im D0, gamma, the
    factor n and every sign;
  * the sweep order, from closed forms of one step: for d = 1 (all commutators vanish) the run gives
    ADO n = n! (2 i s Im D0 dt)^n (a Jacobi sweep gives zero beyond n = 1); for the propagator on a
    diagonal (pure-dephasing) element, u[n] = n! (a x dt)^n (1 - i w dt)^2;
  * trace conservation of ADO 0 (Tr of every commutator vanishes);
  * heom_run_rk4: one RK4 step of a linear system is the degree-4 Taylor polynomial of dt times its
    generator, so k steps equal P(dt G)^k y0 with G the dense hierarchy generator assembled from the
    Kronecker superoperators (a route that shares nothing with the commutator code); the pure
    dephasing closed form at D0 = reorg (2T - i gamma), reached with fourth-order convergence;
    reorganization 0 gives P(-i dt L0)^k vec(rho0).
The reference module cannot be imported here (opt_einsum and the pyqed package chain are absent), so
there are no golden vectors for this path; the checks above are its pins.
"""
import numpy as np

AU2K = 315775.13


def comm(a, b):
    return a @ b - b @ a


def run_params(temperature, cutoff, reorganization):
    """(gamma, D0) of _heom (oqs.py:1826-1845)."""
    gamma, T = float(cutoff), float(temperature)
    D0 = reorganization * gamma * (1.0 / np.tanh(gamma / (2.0 * T)) - 1j)
    return gamma, complex(D0)


def propagator_params(temperature, cutoff, reorganization):
    """(gamma, a, b) of _heom_propagator (oqs.py:1893-1908)."""
    T = temperature / AU2K
    return float(cutoff), float(np.pi * reorganization * T), 0.0


def sweep(X, dt, gamma, cr, ci, n_first, prod_A, prod_S, prod_down):
    """One in-place Euler step over X[0..nado-1].  prod_A(Y) and prod_S(Y) give the coherent and
    up-coupling products, prod_down(Y) the (commutator, anticommutator) pair of the down-coupling.
    ADO 0 is updated first; the loop over n then starts at n_first."""
    nado = X.shape[0]

    def update(n):
        rhs = -1j * prod_A(X[n]) - prod_S(X[n + 1]) - n * gamma * X[n]
        if n > 0:
            zc, za = prod_down(X[n - 1])
            rhs = rhs + n * (cr * zc + 1j * ci * za)
        X[n] = X[n] + rhs * dt

    update(0)
    for n in range(n_first, nado - 1):
        update(n)


def heom_run(H, rho0, S, e_ops, gamma, D0, nado, dt, nt, return_ados=False):
    """Observables (n_e, nt) of _heom (oqs.py:1808-1874); with return_ados also the final ADOs."""
    H = np.asarray(H, np.complex128)
    S = np.asarray(S, np.complex128)
    d = H.shape[0]
    X = np.zeros((nado, d, d), np.complex128)
    X[0] = rho0
    obs = np.zeros((len(e_ops), nt), np.complex128)
    for k in range(nt):
        sweep(X, dt, gamma, D0.real, D0.imag, 1, lambda Y: comm(H, Y), lambda Y: comm(S, Y),
              lambda Y: (comm(S, Y), S @ Y + Y @ S))
        for o, e in enumerate(e_ops):
            obs[o, k] = np.trace(np.asarray(e) @ X[0])
    return (obs, X) if return_ados else obs


def superops(H, S):
    """L0, Sm, Sp of _heom_propagator (superoperator.py:200-246, row-major vec)."""
    d = H.shape[0]
    I = np.eye(d)
    L0 = np.kron(H, I) - np.kron(I, H.T)
    Sm = np.kron(S, I) - np.kron(I, S.T)
    Sp = np.kron(S, I) + np.kron(I, S.T)
    return L0, Sm, Sp


def heom_propagator(H, S, gamma, a, b, nado, dt, nt, n_first=0):
    """u (nado, N, N) of _heom_propagator (oqs.py:1876-1935, n_first = 0; HEOM/heom.py:349-413,
    n_first = 1), u[0] starting at the identity."""
    H = np.asarray(H, np.complex128)
    S = np.asarray(S, np.complex128)
    L0, Sm, Sp = superops(H, S)
    N = L0.shape[0]
    u = np.zeros((nado, N, N), np.complex128)
    u[0] = np.eye(N)
    for k in range(nt):
        sweep(u, dt, gamma, a, b, n_first, lambda Y: L0 @ Y, lambda Y: Sm @ Y, lambda Y: (Sm @ Y, Sp @ Y))
    return u


def run_params_rk(temperature, cutoff, reorganization):
    """(gamma, D0) of HEOM/heom.py:_heom (lines 295-313)."""
    gamma, T = float(cutoff), float(temperature)
    return gamma, complex(reorganization * (2.0 * T - 1j * gamma))


def rhs_all(X, H, S, gamma, D0):
    """F_n of HEOM/heom.py:319-331 for every n: -i[H, X_n] - [S, X_{n+1}] - n gamma X_n
    + n (Re D0 [S, X_{n-1}] + i Im D0 {S, X_{n-1}}) for n <= nado-2 (no X_{-1} term at n = 0),
    and 0 for the last ADO."""
    nado = X.shape[0]
    F = np.zeros_like(X)
    for n in range(nado - 1):
        F[n] = -1j * comm(H, X[n]) - comm(S, X[n + 1]) - n * gamma * X[n]
        if n > 0:
            F[n] += n * (D0.real * comm(S, X[n - 1]) + 1j * D0.imag * (S @ X[n - 1] + X[n - 1] @ S))
    return F


def heom_run_rk4(H, rho0, S, e_ops, gamma, D0, nado, dt, nt, return_ados=False):
    """Observables (n_e, nt) of HEOM/heom.py:_heom; with return_ados also the final ADOs."""
    H = np.asarray(H, np.complex128)
    S = np.asarray(S, np.complex128)
    d = H.shape[0]
    X = np.zeros((nado, d, d), np.complex128)
    X[0] = rho0
    obs = np.zeros((len(e_ops), nt), np.complex128)
    h = dt / 2.0
    for k in range(nt):
        k1 = rhs_all(X, H, S, gamma, D0)
        k2 = rhs_all(X + k1 * h, H, S, gamma, D0)
        k3 = rhs_all(X + k2 * h, H, S, gamma, D0)
        k4 = rhs_all(X + k3 * dt, H, S, gamma, D0)
        X = X + (k1 + 2 * k2 + 2 * k3 + k4) / 6.0 * dt
        for o, e in enumerate(e_ops):
            obs[o, k] = np.trace(np.asarray(e) @ X[0])
    return (obs, X) if return_ados else obs


def dephasing_exact(rho0, H, S, D0, gamma, t):
    """Closed form of the pure-dephasing single-exponential HEOM for diagonal H and S."""
    e = np.real(np.diag(H))
    q = np.real(np.diag(S))
    G = t / gamma - (1.0 - np.exp(-gamma * t)) / gamma ** 2
    x = q[:, None] - q[None, :]
    w = e[:, None] - e[None, :]
    return rho0 * np.exp(-1j * w * t - x * (q[:, None] * D0 - q[None, :] * np.conj(D0)) * G)
