"""The Lindblad / Redfield oracle (oracle/lindblad.py) against mathematics, and the FP64 autograd
reference against central finite differences.  CPU only."""
import numpy as np
import pytest
import scipy.linalg as sla

from oracle import lindblad as OL


def _rand_herm(d, rng):
    A = rng.normal(size=(d, d)) + 1j * rng.normal(size=(d, d))
    return (A + A.conj().T) / 2


def test_rhs_equals_kron_superoperator():
    rng = np.random.default_rng(0)
    d = 4
    H = _rand_herm(d, rng)
    c = [rng.normal(size=(d, d)) + 1j * rng.normal(size=(d, d)) for _ in range(2)]
    L = OL.superoperator(H, c)
    rho = rng.normal(size=(d, d)) + 1j * rng.normal(size=(d, d))
    assert np.abs((L @ rho.reshape(-1)).reshape(d, d) - OL.liouvillian_rhs(rho, H, c)).max() < 1e-12


def test_rk4_trajectory_matches_expm():
    rng = np.random.default_rng(1)
    d = 3
    H = _rand_herm(d, rng)
    c = [0.3 * (rng.normal(size=(d, d)) + 1j * rng.normal(size=(d, d)))]
    rho0 = np.diag([0.6, 0.3, 0.1]).astype(complex)
    out = OL.run(H, c, rho0, 0.01, 200)
    exact = OL.expm_evolve(OL.superoperator(H, c), rho0, 2.0)
    assert np.abs(out["rholist"][-1] - exact).max() < 1e-9


def test_two_level_decay_and_dephasing_closed_forms():
    G, gphi = 0.3, 0.2
    rho0 = np.array([[0.3, 0.4 - 0.1j], [0.4 + 0.1j, 0.7]])
    t = np.arange(1, 401) * 0.01
    # decay |1> -> |0>, H = 0
    c = [np.sqrt(G) * np.array([[0, 1], [0, 0]], complex)]
    r = np.array(OL.run(np.zeros((2, 2)), c, rho0, 0.01, 400)["rholist"])
    assert np.abs(r[:, 1, 1] - 0.7 * np.exp(-G * t)).max() < 1e-10
    assert np.abs(np.abs(r[:, 0, 1]) - abs(rho0[0, 1]) * np.exp(-G * t / 2)).max() < 1e-10
    # pure dephasing on both levels
    c = [np.sqrt(gphi) * np.diag([1.0, 0]).astype(complex), np.sqrt(gphi) * np.diag([0, 1.0]).astype(complex)]
    r = np.array(OL.run(np.zeros((2, 2)), c, rho0, 0.01, 400)["rholist"])
    assert np.abs(np.abs(r[:, 0, 1]) - abs(rho0[0, 1]) * np.exp(-gphi * t)).max() < 1e-10


def test_driven_legacy_accumulation():
    H0 = np.diag([0.0, 1.0]).astype(complex)
    H1 = np.array([[0, 1], [1, 0]], complex)
    f = lambda t: 0.1
    rho0 = np.diag([1.0, 0]).astype(complex)
    a = OL.run_driven([H0, [H1, f]], rho0, [], 0.01, 3, legacy=True)["rholist"]
    b = OL.run_driven([H0, [H1, f]], rho0, [], 0.01, 3)["rholist"]
    # step k of the legacy run uses H0 - (k+1) f H1
    rho = rho0
    for k in range(3):
        rho = OL.rk4(rho, OL.liouvillian_rhs, 0.01, H0 - (k + 1) * 0.1 * H1, [])
    assert np.abs(a[-1] - rho).max() < 1e-15 and np.abs(a[-1] - b[-1]).max() > 1e-6


def test_redfield_trace_hermiticity_detailed_balance():
    rng = np.random.default_rng(3)
    d = 3
    H = np.diag([0.0, 0.7, 1.6]) + 0.2 * _rand_herm(d, rng)
    a = _rand_herm(d, rng)
    beta, lam, wc = 2.0, 0.05, 1.0

    def S(w):  # ohmic with detailed balance S(-w)/S(w) = exp(-beta w)
        if w == 0:
            return lam / beta
        return lam * w * np.exp(-abs(w) / wc) / (1 - np.exp(-beta * w))

    R, ev = OL.redfield_tensor(H, [a], [S])
    N = d * d
    tr = np.eye(d).reshape(-1)
    assert np.abs(tr @ R).max() < 1e-12  # trace preserved
    rho = np.diag([0.2, 0.5, 0.3]).astype(complex) + 0.05j * (np.eye(d, k=1) - np.eye(d, k=-1))
    out = OL.run_redfield(R, rho, ev, 0.05, 400)["rholist"]
    last = out[-1]
    assert abs(np.trace(last) - 1) < 1e-10
    assert np.abs(last - last.conj().T).max() < 1e-12
    # the stationary state (null vector of R, eigenbasis) has Gibbs populations when
    # S(-w)/S(w) = exp(-beta w): pins the sign and argument convention of C[n, m] = S(-(E_n - E_m))
    w, V = np.linalg.eig(R)
    i = np.argmin(np.abs(w))
    ss = V[:, i].reshape(d, d)
    ss = ss / np.trace(ss)
    E = np.linalg.eigvalsh(H)
    boltz = np.exp(-beta * E) / np.exp(-beta * E).sum()
    assert abs(w[i]) < 1e-12 and np.abs(np.real(np.diag(ss)) - boltz).max() < 1e-12


def test_autograd_reference_against_finite_differences():
    import torch
    from oracle.lindblad_torch import final_state
    rng = np.random.default_rng(4)
    d = 3
    c = [np.sqrt(0.02) * np.diag(np.eye(d)[j]).astype(complex) for j in range(d)]
    base = OL.superoperator(np.zeros((d, d)), c)
    G = []
    for j in range(d):
        P = np.zeros((d, d), complex)
        P[j, j] = 1
        G.append(-1j * OL.op2sop(P))
    mu = np.array([[0, 1, 0], [1, 0, 0.7], [0, 0.7, 0]], complex)
    B = [1j * OL.op2sop(mu)]
    rho0 = np.diag([1.0, 0, 0]).astype(complex)
    theta = torch.tensor([0.0, 1.0, 2.5], dtype=torch.float64, requires_grad=True)
    amps = torch.tensor(rng.normal(0, 0.3, (1, 6)), dtype=torch.float64, requires_grad=True)
    J = final_state(theta, amps, base, G, B, rho0, 0.05, 5)[2, 2].real
    J.backward()
    h = 1e-6
    for j in range(3):
        tp, tm = theta.detach().clone(), theta.detach().clone()
        tp[j] += h
        tm[j] -= h
        fd = (final_state(tp, amps.detach(), base, G, B, rho0, 0.05, 5)[2, 2].real -
              final_state(tm, amps.detach(), base, G, B, rho0, 0.05, 5)[2, 2].real) / (2 * h)
        assert abs(fd - theta.grad[j]) <= 1e-6 * max(1e-3, abs(theta.grad[j]))
    for s in range(6):
        ap, am = amps.detach().clone(), amps.detach().clone()
        ap[0, s] += h
        am[0, s] -= h
        fd = (final_state(theta.detach(), ap, base, G, B, rho0, 0.05, 5)[2, 2].real -
              final_state(theta.detach(), am, base, G, B, rho0, 0.05, 5)[2, 2].real) / (2 * h)
        assert abs(fd - amps.grad[0, s]) <= 1e-6 * max(1e-3, abs(amps.grad[0, s]))


def test_product_generators_match_oracle():
    """pyqed_amd's host-side generator construction (no GPU) against the oracle's kron forms."""
    from pyqed_amd import oqs
    from pyqed_amd.liouville import lindblad_generator
    rng = np.random.default_rng(5)
    d = 3
    H = _rand_herm(d, rng)
    c = [rng.normal(size=(d, d)) + 1j * rng.normal(size=(d, d))]
    assert np.abs(lindblad_generator(H, c) - OL.superoperator(H, c)).max() < 1e-13
    a = _rand_herm(d, rng)
    S = lambda w: 0.1 / (1 + w * w)
    R1, e1 = oqs.redfield_tensor(H, [a], [S])
    R2, e2 = OL.redfield_tensor(H, [a], [S])
    assert np.abs(R1 - R2).max() < 1e-13 and np.abs(e1 - e2).max() < 1e-13


def test_correlation_closed_form_two_level():
    """<sigma+(t) sigma-(0)> for H = diag(0, w), c = sqrt(G)|0><1|, rho0 = |1><1|: exp((i w - G/2) t)."""
    w, G, dt, Nt = 1.3, 0.4, 0.01, 300
    H = np.diag([0.0, w]).astype(complex)
    sm = np.array([[0, 1], [0, 0]], complex)
    rho0 = np.diag([0.0, 1.0]).astype(complex)
    g = OL.correlation_2op_1t(H, [np.sqrt(G) * sm], rho0, sm.conj().T, sm, dt, Nt)
    t = np.arange(1, Nt + 1) * dt
    assert np.abs(g - np.exp((1j * w - G / 2) * t)).max() < 2e-9  # RK4 global error at dt = 0.01
    # the three-operator form with A = 1 is the same correlation including t = 0
    g3 = OL.correlation_3op_1t(H, [np.sqrt(G) * sm], rho0, np.eye(2), sm.conj().T, sm, dt, Nt)
    assert abs(g3[0] - 1) < 1e-15 and np.abs(g3[1:] - g).max() < 1e-15


def test_correlation_3op_2t_against_expm():
    rng = np.random.default_rng(7)
    d, dt, Nt, Ntau = 3, 0.02, 6, 9
    H = _rand_herm(d, rng)
    c = [0.3 * (rng.normal(size=(d, d)) + 1j * rng.normal(size=(d, d)))]
    a, b, cc = (rng.normal(size=(d, d)) + 1j * rng.normal(size=(d, d)) for _ in range(3))
    rho0 = np.diag([0.5, 0.3, 0.2]).astype(complex)
    got = OL.correlation_3op_2t(H, c, rho0, a, b, cc, dt, Nt, Ntau)
    L = OL.superoperator(H, c)
    for j in range(Nt):
        rho_t = OL.expm_evolve(L, rho0, (j + 1) * dt)
        for m in range(Ntau):
            ref = np.trace(b @ OL.expm_evolve(L, cc @ rho_t @ a, m * dt))
            assert abs(got[j, m] - ref) < 1e-8 * max(1.0, abs(ref))  # RK4 error at dt = 0.02


def test_redfield_two_level_rates_closed_form():
    # H = diag(0, w), a = sigma_x, any spectrum S: with C[n, m] = S(E_m - E_n), l = C * a and
    # D(rho) = -[a, l rho - rho l^dag] (oqs.py:519-570), the excited population decays as
    # d rho_11/dt = -2 S(w) and the ground population fills as d rho_00/dt = +2 S(w); from the ground
    # state the up rate is 2 S(-w).  These fix the scale, the sign and the argument of the rates, which
    # the trace and Gibbs checks leave open.
    w0 = 0.8
    H = np.diag([0.0, w0]).astype(complex)
    sx = np.array([[0, 1], [1, 0]], complex)
    S = lambda w: 0.03 + 0.01 * w + 0.002 * w * w    # no symmetry: S(w) != S(-w)
    R, ev = OL.redfield_tensor(H, [sx], [S])
    assert np.allclose(np.abs(ev), np.eye(2))        # eigenbasis = computational basis up to phases
    v = lambda r: np.asarray(r, complex).reshape(-1)
    d1 = (R @ v(np.diag([0.0, 1.0]))).reshape(2, 2)
    assert abs(d1[1, 1] - (-2 * S(w0))) < 1e-14 and abs(d1[0, 0] - 2 * S(w0)) < 1e-14
    d0 = (R @ v(np.diag([1.0, 0.0]))).reshape(2, 2)
    assert abs(d0[0, 0] - (-2 * S(-w0))) < 1e-14 and abs(d0[1, 1] - 2 * S(-w0)) < 1e-14
