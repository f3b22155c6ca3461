"""The legacy single-exponential HEOM (oqs.HEOMSolver) on the GPU against the oracle
(oracle/legacy_heom.py).  FP64 throughout; tolerance 1e-12 max-abs on observables, ADOs and
propagator blocks (entries of order 1)."""
import math

import numpy as np
import pytest

from oracle import legacy_heom as LH

pytestmark = pytest.mark.gpu
TOL = 1e-12


def _rand_herm(d, rng):
    A = rng.normal(size=(d, d)) + 1j * rng.normal(size=(d, d))
    return (A + A.conj().T) / 2


def _rand_rho(d, rng):
    A = rng.normal(size=(d, d)) + 1j * rng.normal(size=(d, d))
    r = A @ A.conj().T
    return r / np.trace(r)


@pytest.mark.parametrize("d,nado,nt", [(2, 8, 400), (3, 6, 300), (9, 4, 60), (20, 4, 20)])
def test_run_matches_oracle(d, nado, nt):
    from pyqed_amd.oqs import HEOMSolver
    rng = np.random.default_rng(d)
    H, S, rho0 = _rand_herm(d, rng), 0.5 * _rand_herm(d, rng), _rand_rho(d, rng)
    e_ops = [np.diag(np.eye(d)[0]).astype(complex), _rand_herm(d, rng)]
    dt, T, cutoff, lam = 0.01, 1.5, 0.8, 0.3
    obs = HEOMSolver(H, [S], e_ops).run(rho0, dt, nt, T, cutoff, lam, nado, verbose=False)
    gamma, D0 = LH.run_params(T, cutoff, lam)
    ref = LH.heom_run(H, rho0, S, e_ops, gamma, D0, nado, dt, nt)
    assert obs.shape == ref.shape == (2, nt)
    assert np.abs(obs - ref).max() < TOL


def test_run_batch_parameter_sweep_and_ados():
    from pyqed_amd.oqs import HEOMSolver
    rng = np.random.default_rng(7)
    d, nado, nt, dt, n = 2, 10, 250, 0.02, 37
    H, S = _rand_herm(d, rng), np.diag([1.0, -1.0]).astype(complex)
    rho0s = np.array([_rand_rho(d, rng) for _ in range(n)])
    T = np.linspace(0.5, 5.0, n)
    lam = np.linspace(0.01, 0.4, n)
    e_ops = [np.diag([1.0, 0.0]).astype(complex), np.array([[0, 1], [0, 0]], complex)]
    obs, fin = HEOMSolver(H, [S], e_ops).run_batch(rho0s, dt, nt, T, 1.0, lam, nado, return_ados=True)
    assert obs.shape == (n, 2, nt) and fin.shape == (n, nado, d, d)
    for i in range(n):
        gamma, D0 = LH.run_params(T[i], 1.0, lam[i])
        ref, X = LH.heom_run(H, rho0s[i], S, e_ops, gamma, D0, nado, dt, nt, return_ados=True)
        assert np.abs(obs[i] - ref).max() < TOL
        assert np.abs(fin[i] - X).max() < TOL


def test_run_sweep_order_scalar_closed_form():
    # d = 1: one step from X_0 = 1 gives X_n = n! (2 i s Im D0 dt)^n (the in-place sweep)
    from pyqed_amd.oqs import HEOMSolver
    s, dt, nado = 0.7, 0.05, 6
    _, fin = HEOMSolver(np.array([[0.3]]), [np.array([[s]])], []).run_batch(
        np.ones((1, 1, 1)), dt, 1, 0.9, 1.3, 0.4, nado, return_ados=True)
    gamma, D0 = LH.run_params(0.9, 1.3, 0.4)
    c = 2j * s * D0.imag * dt
    expect = [1.0] + [math.factorial(n) * c ** n for n in range(1, nado - 1)] + [0.0]
    assert np.allclose(fin[0, :, 0, 0], expect, rtol=1e-14, atol=0)


@pytest.mark.parametrize("d,nado,nt", [(2, 6, 300), (3, 4, 100)])
def test_propagator_matches_oracle(d, nado, nt):
    from pyqed_amd.oqs import HEOMSolver
    rng = np.random.default_rng(10 + d)
    H, S = _rand_herm(d, rng), 0.5 * _rand_herm(d, rng)
    dt, temp, cutoff, lam = 0.01, 1.0e5, 1.0, 0.5
    u = HEOMSolver(H, [S], None).propagator(dt, nt, temp, cutoff, lam, nado, verbose=False)
    gamma, a, b = LH.propagator_params(temp, cutoff, lam)
    ref = LH.heom_propagator(H, S, gamma, a, b, nado, dt, nt)
    assert u.shape == ref.shape == (nado, d * d, d * d)
    assert np.abs(u - ref).max() < TOL
    assert np.abs(u[1]).max() > 1e-3                       # the bath terms are exercised


def test_propagator_kernel_with_anticommutator_term():
    # the reference fixes b = 0; the kernel's i b Sp term is checked through the C ABI with b != 0
    from pyqed_amd.liouville import op2sop
    from pyqed_amd.oqs import legacy_heom_launch
    rng = np.random.default_rng(20)
    d, nado, nt, dt = 2, 5, 150, 0.01
    H, S = _rand_herm(d, rng), 0.5 * _rand_herm(d, rng)
    gamma, a, b = 0.9, 0.4, 0.3
    u0 = np.zeros((1, nado, d * d, d * d), complex)
    u0[0, 0] = np.eye(d * d)
    fin, _ = legacy_heom_launch(1, op2sop(H)[None], op2sop(S)[None], op2sop(S, "anticommutator")[None],
                                [[gamma, a, b]], u0, nt, dt)
    ref = LH.heom_propagator(H, S, gamma, a, b, nado, dt, nt)
    assert np.abs(fin[0] - ref).max() < TOL


def test_diagnostics_and_limits(capsys):
    from pyqed_amd._lib import PyqedError
    from pyqed_amd.oqs import HEOMSolver
    d = 2
    H, S = np.diag([0.0, 1.0]).astype(complex), np.diag([1.0, -1.0]).astype(complex)
    sol = HEOMSolver(H, [S], [np.eye(d)])
    obs = sol.run(np.diag([1.0, 0.0]).astype(complex), 0.01, 5, 1.0, 2.0, 0.1, 3)
    assert np.abs(obs - 1.0).max() < 1e-14
    gamma, D0 = LH.run_params(1.0, 2.0, 0.1)
    out = capsys.readouterr().out.splitlines()
    assert out == ['Temperature of the environment = 1.0', 'Cutoff gamma/(kT) = 2.0',
                   'WARNING: High-Temperature Approximation may fail.', 'Reorganization energy = 0.1',
                   'Amplitude of the fluctuations = {}'.format(D0)]
    # a hierarchy that does not fit one workgroup's LDS is refused with a message
    with pytest.raises(PyqedError, match="LDS"):
        HEOMSolver(np.eye(16), [np.eye(16)], []).run(np.eye(16) / 16, 0.01, 1, 1.0, 0.1, 0.1, 200, verbose=False)


def test_reference_example_spin_boson():
    # the call of pyqed's examples/heom.py: H = -sx/2 - sz/2, c_ops = [sz], e_ops = [sz], rho0 = |1><1|
    from pyqed_amd.oqs import HEOMSolver
    sx = np.array([[0, 1], [1, 0]], complex)
    sz = np.diag([1.0, -1.0]).astype(complex)
    H = -0.5 * sx - 0.5 * sz
    rho0 = np.zeros((2, 2))
    rho0[1, 1] = 1
    obs = HEOMSolver(H, c_ops=[sz], e_ops=[sz]).run(rho0=rho0, dt=0.02, nt=100, temperature=600, cutoff=5,
                                                     reorganization=0.2, nado=5, verbose=False)
    gamma, D0 = LH.run_params(600, 5, 0.2)
    ref = LH.heom_run(H, rho0.astype(complex), sz, [sz], gamma, D0, 5, 0.02, 100)
    assert obs.shape == (1, 100)
    assert np.abs(obs - ref).max() < TOL
