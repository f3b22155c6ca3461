"""Eigen-decomposition solvers (superoperator.Lindblad_solver, RedfieldSolver.propagator and its
correlation functions) on the GPU spectral-sum kernel against the oracle (oracle/eigen.py).
Tolerances: 1e-12 for the kernel against numpy; 1e-10 max-abs for the solvers (values of order 1,
eigenvector conditioning enters through the host decomposition)."""
import numpy as np
import pytest

from oracle import eigen as OE

pytestmark = pytest.mark.gpu


def _rand(shape, rng):
    return rng.normal(size=shape) + 1j * rng.normal(size=shape)


def _system(d=3, seed=0):
    rng = np.random.default_rng(seed)
    A = _rand((d, d), rng)
    H = (A + A.conj().T) / 2
    c = [0.4 * _rand((d, d), rng) for _ in range(2)]
    B = _rand((d, d), rng)
    rho0 = B @ B.conj().T
    rho0 /= np.trace(rho0)
    ops = [_rand((d, d), rng) for _ in range(4)]
    return H, c, rho0, ops


@pytest.mark.parametrize("nrow,k,ncol,ns", [(1, 7, 1, 5), (3, 37, 21, 130), (70, 16, 1, 64), (9, 81, 81, 33),
                                             (2, 50, 100, 70), (1, 20, 48, 200)])
def test_spectral_sum_kernel(nrow, k, ncol, ns):
    from pyqed_amd.spectral import FREQ, TIME, spectral_sum
    rng = np.random.default_rng(nrow + k)
    A, B = _rand((nrow, k), rng), _rand((k, ncol), rng)
    lam = -np.abs(rng.normal(size=k)) + 1j * rng.normal(size=k)
    s = np.linspace(0.0, 3.0, ns)
    ref_t = np.einsum("aj,tj,jb->tab", A, np.exp(lam[None, :] * s[:, None]), B)
    ref_w = np.einsum("aj,tj,jb->tab", A, -1.0 / (lam[None, :] + 1j * s[:, None]), B)
    out = spectral_sum(TIME, A, lam, B, s)
    assert np.abs(out - ref_t).max() < 1e-12 * max(1.0, np.abs(ref_t).max())
    out = spectral_sum(FREQ, A, lam, B, s, layout="abt")
    assert np.abs(out - np.transpose(ref_w, (1, 2, 0))).max() < 1e-12 * max(1.0, np.abs(ref_w).max())


@pytest.mark.parametrize("biorthogonal", [False, True])
def test_lindblad_solver_matches_oracle(biorthogonal):
    from pyqed_amd.superoperator import Lindblad_solver
    H, c, rho0, (a, b, cc, dd) = _system(seed=1)
    L = OE.liouvillian(H, c)
    sol = Lindblad_solver(H, c, biorthogonal=biorthogonal)
    assert np.abs(sol.liouvillian() - L).max() < 1e-14
    sol.eigenstates()
    ex = dict(exact=biorthogonal)
    tl = np.linspace(0.0, 3.0, 41)
    ws = np.linspace(-2.0, 2.0, 16)           # avoids w = 0, the steady-state pole
    res = sol.evolve(rho0, tl, [a, np.eye(3)])
    assert res.observables.shape == (41, 2) and np.allclose(res.times, tl)
    assert np.abs(res.observables - OE.lindblad_evolve(L, rho0, tl, [a, np.eye(3)], **ex)).max() < 1e-10
    assert np.abs(sol.correlation_2op_1t(rho0, [a, b], tl) - OE.correlation_2op_1t(L, rho0, a, b, tl, **ex)).max() < 1e-10
    assert np.abs(sol.correlation_2op_1w(rho0, [a, b], ws) - OE.correlation_2op_1w(L, rho0, a, b, ws, **ex)).max() < 1e-10
    assert np.abs(sol.correlation_3op_1t(rho0, [a, b, cc], tl)
                  - OE.correlation_3op_1t(L, rho0, a, b, cc, tl, **ex)).max() < 1e-10
    assert np.abs(sol.correlation_3op_1w(rho0, [a, b, cc], ws)
                  - OE.correlation_3op_1w(L, rho0, a, b, cc, ws, **ex)).max() < 1e-10
    taus = np.linspace(0.0, 2.0, 23)
    c32 = sol.correlation_3op_2t(rho0, [a, b, cc], tl, taus)
    assert c32.shape == (23, 41)
    assert np.abs(c32 - OE.correlation_3op_2t(L, rho0, a, b, cc, tl, taus, **ex)).max() < 1e-10
    c42 = sol.correlation_4op_2t(rho0, [a, b, cc, dd], tl, taus)
    assert np.abs(c42 - OE.correlation_3op_2t(L, rho0, a, b @ cc, dd, tl, taus, **ex)).max() < 1e-10


def test_two_level_emission_closed_form():
    from pyqed_amd.superoperator import Lindblad_solver
    w, G = 1.3, 0.4
    H = np.diag([0.0, w]).astype(complex)
    sm = np.array([[0, 1], [0, 0]], complex)
    tl = np.linspace(0.0, 4.0, 101)
    cor = Lindblad_solver(H, [np.sqrt(G) * sm]).correlation_2op_1t(np.diag([0.0, 1.0]).astype(complex),
                                                                   [sm.conj().T, sm], tl)
    assert np.abs(cor - np.exp((1j * w - G / 2) * tl)).max() < 1e-12


def _redfield():
    from pyqed_amd.oqs import RedfieldSolver
    rng = np.random.default_rng(5)
    d = 3
    A = _rand((d, d), rng)
    H = np.diag([0.0, 0.6, 1.4]) + 0.05 * (A + A.conj().T)
    a_ops = [np.diag([1.0, -1.0, 0.5]).astype(complex)]
    S = lambda w: 0.05 * w / (1 - np.exp(-2.0 * w)) if w != 0 else 0.025
    sol = RedfieldSolver(H, a_ops, [S])
    sol.redfield_tensor()
    return sol, H


@pytest.mark.parametrize("method", ["SOS", "EOM"])
def test_redfield_propagator_expect_and_correlations(method):
    sol, H = _redfield()
    t = np.arange(40) * 0.05
    U = sol.propagator(t, method=method)
    ref = OE.redfield_propagator(sol.R, t, method)
    assert U.shape == (9, 9, 40)
    assert np.abs(U - ref).max() < 1e-10
    assert np.abs(sol.G + 1j * U).max() == 0.0
    rho0 = np.diag([0.2, 0.3, 0.5]).astype(complex)
    e_ops = [np.diag([1.0, 0, 0]).astype(complex), np.array([[0, 1, 0], [0, 0, 0], [0, 0, 0]], complex)]
    assert np.abs(sol.expect(rho0, e_ops) - OE.redfield_expect(ref, sol.evecs, rho0, e_ops)).max() < 1e-10
    rng = np.random.default_rng(6)
    a, b = _rand((9, 9), rng), _rand((9, 9), rng)
    corr = sol.correlation_2op_1t(rho0, a, b, t)
    G = -1j * ref
    assert np.abs(corr - np.eye(3).reshape(-1) @ (a @ np.tensordot(G, b @ rho0.reshape(-1), axes=([1], [0])))).max() < 1e-9
    ops = [_rand((3, 3), rng) for _ in range(4)]
    sol.propagator(t[:12], method=method)
    c4 = sol.correlation_4op_3t(rho0, ops, "lr+-", t[:12])
    assert c4.shape == (12, 12, 12)
    ref4 = OE.correlation_4op_3t(-1j * OE.redfield_propagator(sol.R, t[:12], method), rho0, ops, "lr+-")
    assert np.abs(c4 - ref4).max() < 1e-9 * max(1.0, np.abs(ref4).max())
