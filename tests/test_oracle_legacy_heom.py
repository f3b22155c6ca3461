"""The legacy single-exponential HEOM oracle (oracle/legacy_heom.py) against mathematics.  CPU only."""
import math

import numpy as np

from oracle import legacy_heom as LH


def _rand_herm(d, rng):
    A = rng.normal(size=(d, d)) + 1j * rng.normal(size=(d, d))
    return (A + A.conj().T) / 2


def _rand_rho(d, rng):
    A = rng.normal(size=(d, d)) + 1j * rng.normal(size=(d, d))
    r = A @ A.conj().T
    return r / np.trace(r)


def test_run_without_bath_is_euler_matrix_power():
    rng = np.random.default_rng(0)
    d, dt, nt = 3, 0.01, 60
    H, S, rho0 = _rand_herm(d, rng), _rand_herm(d, rng), _rand_rho(d, rng)
    gamma, D0 = LH.run_params(1.5, 0.8, 0.0)
    e = _rand_herm(d, rng)
    obs, X = LH.heom_run(H, rho0, S, [e], gamma, D0, 4, dt, nt, return_ados=True)
    L0 = np.kron(H, np.eye(d)) - np.kron(np.eye(d), H.T)
    step = np.eye(d * d) - 1j * dt * L0
    ref = (np.linalg.matrix_power(step, nt) @ rho0.reshape(-1)).reshape(d, d)
    assert np.abs(X[0] - ref).max() < 1e-13
    assert np.abs(X[1:]).max() == 0.0
    assert abs(obs[0, -1] - np.trace(e @ ref)) < 1e-13
    assert obs.shape == (1, nt)


def test_propagator_without_bath_updates_u0_twice():
    rng = np.random.default_rng(1)
    d, dt, nt = 2, 0.02, 25
    H, S = _rand_herm(d, rng), _rand_herm(d, rng)
    gamma, a, b = LH.propagator_params(300.0, 0.5, 0.0)
    u = LH.heom_propagator(H, S, gamma, a, b, 4, dt, nt)
    L0 = np.kron(H, np.eye(d)) - np.kron(np.eye(d), H.T)
    step = np.eye(d * d) - 1j * dt * L0
    assert np.abs(u[0] - np.linalg.matrix_power(step, 2 * nt)).max() < 1e-13
    # a single update per step would give the nt-th power, far from this
    assert np.abs(u[0] - np.linalg.matrix_power(step, nt)).max() > 1e-2


def test_run_pure_dephasing_matches_closed_form_at_first_order():
    H = np.diag([0.0, 0.5, 1.2]).astype(complex)
    S = np.diag([1.0, 0.0, -0.5]).astype(complex)          # asymmetric: Im D0 enters the phase
    rho0 = _rand_rho(3, np.random.default_rng(2))
    gamma, D0 = LH.run_params(2.0, 1.0, 0.1)
    T = 4.0
    exact = LH.dephasing_exact(rho0, H, S, D0, gamma, T)
    errs = []
    for dt in (2e-3, 1e-3):
        _, X = LH.heom_run(H, rho0, S, [], gamma, D0, 10, dt, int(round(T / dt)), return_ados=True)
        errs.append(np.abs(X[0] - exact).max())
    assert errs[0] < 4e-4 and errs[1] < 2e-4
    assert 1.8 < errs[0] / errs[1] < 2.2                    # first order in dt
    # the decay and the bath phase are large compared with the tolerance
    assert np.abs(exact - rho0 * np.exp(-1j * (np.diag(H)[:, None] - np.diag(H)[None, :]) * T)).max() > 0.05


def test_run_sweep_order_closed_form_scalar():
    # d = 1: commutators vanish, the anticommutator is 2 s X; one step from X_0 = 1 gives
    # X_n = n! (2 i s Im D0 dt)^n for n <= nado - 2 when ADO n reads the updated ADO n - 1
    s, dt, nado = 0.7, 0.05, 6
    gamma, D0 = LH.run_params(0.9, 1.3, 0.4)
    _, X = LH.heom_run(np.array([[0.3]]), np.array([[1.0]]), np.array([[s]]), [], gamma, D0, nado, dt, 1,
                       return_ados=True)
    c = 2j * s * D0.imag * dt
    expect = [1.0] + [math.factorial(n) * c ** n for n in range(1, nado - 1)] + [0.0]
    assert np.allclose(X[:, 0, 0], expect, rtol=1e-14, atol=0)


def test_propagator_sweep_order_closed_form_dephasing():
    # diagonal H and S: each Liouville element p = (i, j) is a scalar hierarchy with
    # w = e_i - e_j, x = q_i - q_j; one step gives u[n]_pp = n! (a x dt)^n (1 - i w dt)^2
    e, q, dt, nado = np.array([0.2, -0.5]), np.array([0.8, -0.3]), 0.03, 5
    H, S = np.diag(e).astype(complex), np.diag(q).astype(complex)
    gamma, a, b = LH.propagator_params(2.0e5, 0.7, 0.4)
    u = LH.heom_propagator(H, S, gamma, a, b, nado, dt, 1)
    for i in range(2):
        for j in range(2):
            p = 2 * i + j
            w, x = e[i] - e[j], q[i] - q[j]
            for n in range(nado - 1):
                assert abs(u[n][p, p] - math.factorial(n) * (a * x * dt) ** n * (1 - 1j * w * dt) ** 2) < 1e-15
    assert np.abs(u[nado - 1]).max() == 0.0
    off = u - np.einsum('nii,ij->nij', u, np.eye(4))
    assert np.abs(off).max() == 0.0


def test_run_conserves_trace():
    rng = np.random.default_rng(3)
    d = 4
    H, S, rho0 = _rand_herm(d, rng), _rand_herm(d, rng), _rand_rho(d, rng)
    gamma, D0 = LH.run_params(1.0, 0.5, 0.3)
    obs = LH.heom_run(H, rho0, S, [np.eye(d)], gamma, D0, 5, 0.01, 200)
    assert np.abs(obs - 1.0).max() < 1e-12


def test_parameters():
    gamma, D0 = LH.run_params(2.0, 1.0, 0.1)
    assert gamma == 1.0
    assert abs(D0 - 0.1 * (math.cosh(0.25) / math.sinh(0.25) - 1j)) < 1e-15
    gamma, a, b = LH.propagator_params(315775.13, 0.4, 0.2)
    assert gamma == 0.4 and abs(a - math.pi * 0.2) < 1e-15 and b == 0.0


def test_host_copies_are_row_major():
    # the kernels index raw row-major buffers; a broadcast instance stack must not keep the view's
    # axis order when copied for upload
    from pyqed_amd.oqs import host_c_order
    H = np.arange(4.0).reshape(2, 2) + 1j
    a = host_c_order(np.broadcast_to(H, (3, 2, 2)))
    assert a.flags["C_CONTIGUOUS"] and a.strides == (64, 32, 16)
    assert not np.array(np.broadcast_to(H, (3, 2, 2)), np.complex128).flags["C_CONTIGUOUS"]


def _taylor4(Z):
    I = np.eye(Z.shape[0])
    Z2 = Z @ Z
    Z3 = Z2 @ Z
    return I + Z + Z2 / 2 + Z3 / 6 + Z3 @ Z / 24


def _dense_hierarchy_generator(H, S, gamma, D0, nado):
    """dvec(X)/dt of HEOM/heom.py:_heom as one matrix over the stacked row-major vec(X_n), built from
    Kronecker products: vec(A X) = (A x I) vec X, vec(X A) = (I x A^T) vec X."""
    d = H.shape[0]
    I = np.eye(d)
    Lc = lambda A: np.kron(A, I) - np.kron(I, A.T)
    La = lambda A: np.kron(A, I) + np.kron(I, A.T)
    N = d * d
    G = np.zeros((nado * N, nado * N), complex)
    for n in range(nado - 1):
        blk = lambda m: (slice(n * N, (n + 1) * N), slice(m * N, (m + 1) * N))
        G[blk(n)] = -1j * Lc(H) - n * gamma * np.eye(N)
        G[blk(n + 1)] = -Lc(S)
        if n > 0:
            G[blk(n - 1)] = n * (D0.real * Lc(S) + 1j * D0.imag * La(S))
    return G


def test_rk4_run_is_taylor_polynomial_power_of_hierarchy_generator():
    rng = np.random.default_rng(5)
    d, nado, dt, nt = 3, 5, 0.03, 40
    H, S, rho0 = _rand_herm(d, rng), _rand_herm(d, rng), _rand_rho(d, rng)
    gamma, D0 = LH.run_params_rk(0.8, 1.1, 0.3)
    e = _rand_herm(d, rng)
    obs, X = LH.heom_run_rk4(H, rho0, S, [e], gamma, D0, nado, dt, nt, return_ados=True)
    G = _dense_hierarchy_generator(H, S, gamma, D0, nado)
    y0 = np.zeros(nado * d * d, complex)
    y0[:d * d] = rho0.reshape(-1)
    ref = np.linalg.matrix_power(_taylor4(dt * G), nt) @ y0
    assert np.abs(X.reshape(-1) - ref).max() < 1e-13
    assert abs(obs[0, -1] - np.trace(e @ ref[:d * d].reshape(d, d))) < 1e-13
    # the upper ADOs are populated and the last one stays at zero (F_{nado-1} = 0)
    assert np.abs(X[1:nado - 1]).max() > 1e-3 and np.abs(X[-1]).max() == 0.0
    # an Euler sweep on the same inputs is far from this
    _, Xe = LH.heom_run(H, rho0, S, [], gamma, D0, nado, dt, nt, return_ados=True)
    assert np.abs(Xe - X).max() > 1e-4


def test_rk4_run_without_bath_and_params():
    rng = np.random.default_rng(6)
    d, dt, nt = 2, 0.05, 30
    H, S, rho0 = _rand_herm(d, rng), _rand_herm(d, rng), _rand_rho(d, rng)
    gamma, D0 = LH.run_params_rk(0.7, 0.9, 0.0)
    _, X = LH.heom_run_rk4(H, rho0, S, [], gamma, D0, 4, dt, nt, return_ados=True)
    L0 = np.kron(H, np.eye(d)) - np.kron(np.eye(d), H.T)
    ref = np.linalg.matrix_power(_taylor4(-1j * dt * L0), nt) @ rho0.reshape(-1)
    assert np.abs(X[0].reshape(-1) - ref).max() < 1e-13
    # D0 = reorg (2 T - i gamma), temperature taken as k_B T
    assert LH.run_params_rk(0.5, 1.0, 0.1) == (1.0, complex(0.1, -0.1))


def test_rk4_run_pure_dephasing_fourth_order():
    H = np.diag([0.0, 0.5, 1.2]).astype(complex)
    S = np.diag([1.0, 0.0, -0.5]).astype(complex)
    rho0 = _rand_rho(3, np.random.default_rng(2))
    gamma, D0 = LH.run_params_rk(0.5, 1.0, 0.1)
    T = 4.0
    exact = LH.dephasing_exact(rho0, H, S, D0, gamma, T)
    errs = []
    for dt in (0.1, 0.05):
        _, X = LH.heom_run_rk4(H, rho0, S, [], gamma, D0, 12, dt, int(round(T / dt)), return_ados=True)
        errs.append(np.abs(X[0] - exact).max())
    assert errs[0] < 1e-6 and errs[1] < 5e-8
    assert 14.0 < errs[0] / errs[1] < 18.0                  # fourth order in dt
    assert np.abs(exact - rho0 * np.exp(-1j * (np.diag(H)[:, None] - np.diag(H)[None, :]) * T)).max() > 0.05


def test_propagator_heom_module_updates_u0_once():
    rng = np.random.default_rng(7)
    d, dt, nt = 2, 0.02, 25
    H, S = _rand_herm(d, rng), _rand_herm(d, rng)
    gamma, a, b = LH.propagator_params(300.0, 0.5, 0.0)
    u = LH.heom_propagator(H, S, gamma, a, b, 4, dt, nt, n_first=1)
    L0 = np.kron(H, np.eye(d)) - np.kron(np.eye(d), H.T)
    step = np.eye(d * d) - 1j * dt * L0
    assert np.abs(u[0] - np.linalg.matrix_power(step, nt)).max() < 1e-13
    # with a bath, u[1] after one step is a x dt Sm (from the updated u[0]); the oqs copy's double
    # update of u[0] changes u[1] at first order
    gamma, a, b = LH.propagator_params(300.0, 0.5, 0.02)
    u1 = LH.heom_propagator(H, S, gamma, a, b, 4, dt, 1, n_first=1)
    Sm = np.kron(S, np.eye(d)) - np.kron(np.eye(d), S.T)
    assert np.abs(u1[1] - a * dt * Sm @ step).max() < 1e-15
