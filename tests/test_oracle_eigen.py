"""The eigen-decomposition oracle (oracle/eigen.py) against direct matrix exponentials, resolvent
solves and closed forms.  CPU only."""
import numpy as np
import scipy.linalg as sla

from oracle import eigen as OE


def _system(d=3, seed=0):
    rng = np.random.default_rng(seed)
    A = rng.normal(size=(d, d)) + 1j * rng.normal(size=(d, d))
    H = (A + A.conj().T) / 2
    c = [0.4 * (rng.normal(size=(d, d)) + 1j * rng.normal(size=(d, d))) for _ in range(2)]
    B = rng.normal(size=(d, d)) + 1j * rng.normal(size=(d, d))
    rho0 = B @ B.conj().T
    rho0 /= np.trace(rho0)
    ops = [rng.normal(size=(d, d)) + 1j * rng.normal(size=(d, d)) for _ in range(4)]
    return H, c, rho0, ops


def _expm_apply(L, X, t):
    d = X.shape[0]
    return (sla.expm(L * t) @ X.reshape(-1)).reshape(d, d)


def test_liouvillian_matches_master_equation():
    H, c, rho0, _ = _system()
    L = OE.liouvillian(H, c)
    rhs = -1j * (H @ rho0 - rho0 @ H)
    for cc in c:
        rhs += cc @ rho0 @ cc.conj().T - 0.5 * (cc.conj().T @ cc @ rho0 + rho0 @ cc.conj().T @ cc)
    assert np.abs((L @ rho0.reshape(-1)).reshape(3, 3) - rhs).max() < 1e-13


def test_evolve_matches_expm():
    H, c, rho0, ops = _system()
    L = OE.liouvillian(H, c)
    tl = np.linspace(0.0, 3.0, 7)
    out = OE.lindblad_evolve(L, rho0, tl, [ops[0], np.eye(3)], exact=True)
    for i, t in enumerate(tl):
        r = _expm_apply(L, rho0, t)
        assert abs(out[i, 0] - np.trace(ops[0] @ r)) < 1e-11
        assert abs(out[i, 1] - 1.0) < 1e-11


def test_regression_correlations_match_expm():
    H, c, rho0, (a, b, cc, dd) = _system(seed=1)
    L = OE.liouvillian(H, c)
    tl = np.linspace(0.0, 2.0, 5)
    c2 = OE.correlation_2op_1t(L, rho0, a, b, tl, exact=True)
    c3 = OE.correlation_3op_1t(L, rho0, a, b, cc, tl, exact=True)
    for i, t in enumerate(tl):
        assert abs(c2[i] - np.trace(a @ _expm_apply(L, b @ rho0, t))) < 1e-10
        assert abs(c3[i] - np.trace(b @ _expm_apply(L, cc @ rho0 @ a, t))) < 1e-10
    taus = np.linspace(0.0, 1.5, 4)
    c32 = OE.correlation_3op_2t(L, rho0, a, b, cc, tl, taus, exact=True)
    assert c32.shape == (4, 5)
    for j, t in enumerate(tl):
        X = cc @ _expm_apply(L, rho0, t) @ a
        for m, tau in enumerate(taus):
            assert abs(c32[m, j] - np.trace(b @ _expm_apply(L, X, tau))) < 1e-10


def test_frequency_domain_is_the_resolvent():
    H, c, rho0, (a, b, cc, _) = _system(seed=2)
    L = OE.liouvillian(H, c)
    d = 3
    idv = np.eye(d).reshape(-1)
    ws = np.array([-1.3, 0.4, 2.2])
    s2 = OE.correlation_2op_1w(L, rho0, a, b, ws, exact=True)
    s3 = OE.correlation_3op_1w(L, rho0, a, b, cc, ws, exact=True)
    for i, w in enumerate(ws):
        M = L + 1j * w * np.eye(d * d)
        x2 = np.linalg.solve(M, (b @ rho0).reshape(-1))
        x3 = np.linalg.solve(M, (cc @ rho0 @ a).reshape(-1))
        assert abs(s2[i] - (-idv @ OE.sop(a, "l") @ x2)) < 1e-10
        assert abs(s3[i] - (-idv @ OE.sop(b, "l") @ x3)) < 1e-10


def test_two_level_closed_form():
    w, G = 1.3, 0.4
    H = np.diag([0.0, w]).astype(complex)
    sm = np.array([[0, 1], [0, 0]], complex)     # |0><1|
    L = OE.liouvillian(H, [np.sqrt(G) * sm])
    rho0 = np.diag([0.0, 1.0]).astype(complex)
    tl = np.linspace(0.0, 4.0, 9)
    cor = OE.correlation_2op_1t(L, rho0, sm.conj().T, sm, tl)   # <sigma+(t) sigma->
    assert np.abs(cor - np.exp((1j * w - G / 2) * tl)).max() < 1e-12


def test_reference_norm_is_the_real_part_of_the_overlap():
    # the reference divides by Re <L_n|R_n>; that equals the exact expansion once every overlap is
    # real (rephased left vectors), and differs from it when the overlaps are complex
    H, c, rho0, (a, b, cc, _) = _system(seed=5)
    L = OE.liouvillian(H, c)
    tl = np.linspace(0.0, 2.0, 5)
    ws = np.array([-0.7, 1.1])
    exact = OE.correlation_2op_1t(L, rho0, a, b, tl, exact=True)
    assert np.abs(OE.correlation_2op_1t(L, rho0, a, b, tl, rephase=True) - exact).max() < 1e-10
    assert np.abs(OE.correlation_3op_1w(L, rho0, a, b, cc, ws, rephase=True)
                  - OE.correlation_3op_1w(L, rho0, a, b, cc, ws, exact=True)).max() < 1e-10
    assert np.abs(OE.lindblad_evolve(L, rho0, tl, [a], rephase=True)
                  - OE.lindblad_evolve(L, rho0, tl, [a], exact=True)).max() < 1e-10
    assert np.abs(OE.correlation_3op_2t(L, rho0, a, b, cc, tl, tl, rephase=True)
                  - OE.correlation_3op_2t(L, rho0, a, b, cc, tl, tl, exact=True)).max() < 1e-10
    assert np.abs(OE.correlation_2op_1t(L, rho0, a, b, tl) - exact).max() > 1e-3


def test_redfield_propagators():
    rng = np.random.default_rng(3)
    N = 9
    R = 0.3 * (rng.normal(size=(N, N)) + 1j * rng.normal(size=(N, N))) - 0.5 * np.eye(N)
    t = np.arange(6) * 0.1
    U = OE.redfield_propagator(R, t, "SOS")
    for k, tk in enumerate(t):
        assert np.abs(U[:, :, k] - sla.expm(R * tk)).max() < 1e-12
    errs = []
    for n in (10, 20):
        tt = np.arange(n + 1) * (0.5 / n)
        errs.append(np.abs(OE.redfield_propagator(R, tt, "EOM")[:, :, -1] - sla.expm(R * 0.5)).max())
    assert errs[1] < 1e-7 and 12 < errs[0] / errs[1] < 20          # fourth order


def test_four_point_function_matches_expm():
    H, c, rho0, ops = _system(d=2, seed=4)
    L = OE.liouvillian(H, c)
    tau = np.array([0.0, 0.3, 0.8])
    G = -1j * OE.redfield_propagator(L, tau, "SOS")
    sig = "lr+-"
    cor = OE.correlation_4op_3t(G, rho0, ops, sig)
    A, B, C, D = (OE.sop(o, s) for o, s in zip(ops, sig))
    idm = np.eye(2).reshape(-1)
    g = [-1j * sla.expm(L * x) for x in tau]
    for i in range(3):
        for j in range(3):
            for k in range(3):
                ref = idm @ A @ g[i] @ B @ g[j] @ C @ g[k] @ D @ rho0.reshape(-1)
                assert abs(cor[i, j, k] - ref) < 1e-11
