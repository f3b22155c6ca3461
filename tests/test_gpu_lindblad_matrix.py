"""The matrix-form Lindblad kernel (pyqed_lindblad_run: any d, rho as a d x d matrix) against the
oracle's restatement of _lindblad / _lindblad_driven (oqs.py:697-714, 1596-1806) and against the
Liouville-space kernel where both apply.  FP64 throughout: 1e-12 max-abs on unit-trace states."""
import numpy as np
import pytest

from oracle import lindblad as OL

pytestmark = pytest.mark.gpu


def _rand_herm(d, rng):
    A = rng.normal(size=(d, d)) + 1j * rng.normal(size=(d, d))
    return (A + A.conj().T) / 2


def _system(d, seed, nc=2):
    rng = np.random.default_rng(seed)
    H = _rand_herm(d, rng) / np.sqrt(d)
    c = [0.3 / np.sqrt(d) * (rng.normal(size=(d, d)) + 1j * rng.normal(size=(d, d))) for _ in range(nc)]
    p = rng.dirichlet(np.ones(d))
    U = np.linalg.qr(rng.normal(size=(d, d)) + 1j * rng.normal(size=(d, d)))[0]
    rho0 = U @ np.diag(p) @ U.conj().T
    e_ops = [np.diag(np.eye(d)[j]).astype(complex) for j in (0, d // 2, d - 1)] + [_rand_herm(d, rng)]
    return H, c, rho0, e_ops


@pytest.mark.parametrize("d,nc", [(33, 2), (40, 1), (70, 2), (129, 1)])
def test_static_against_oracle(d, nc):
    """d beyond the Liouville kernel's limit (d*d > 1024), tile edges at 64 and 128 crossed."""
    from pyqed_amd.oqs import LindbladSolver
    H, c, rho0, e_ops = _system(d, seed=d, nc=nc)
    nt = 30 if d < 100 else 8
    res = LindbladSolver(H, c).run(rho0, 0.02, nt, e_ops=e_ops)
    ref = OL.run(H, c, rho0, 0.02, nt, e_ops)
    err_o = np.abs(res.observables - ref["observables"]).max()
    err_r = np.abs(np.array(res.rholist) - np.array(ref["rholist"])).max()
    print(f"d={d} nc={nc}: observables {err_o:.1e}, states {err_r:.1e}")
    assert res.observables.shape == (nt + 1, len(e_ops))
    assert err_o < 1e-12 and err_r < 1e-12


@pytest.mark.parametrize("d", [2, 3, 5])
def test_matrix_form_equals_liouville_form(d):
    """Both kernels on the same small problem (the Liouville form is the one pinned against the oracle
    at these sizes), 400 steps, with and without collapse operators."""
    from pyqed_amd.oqs import LindbladSolver
    for nc in (0, 2):
        H, c, rho0, e_ops = _system(d, seed=10 + d, nc=nc)
        a = LindbladSolver(H, c, form="matrix").run(rho0, 0.01, 400, e_ops=e_ops)
        b = LindbladSolver(H, c, form="liouville").run(rho0, 0.01, 400, e_ops=e_ops)
        assert np.abs(a.observables - b.observables).max() < 1e-12
        assert np.abs(np.array(a.rholist) - np.array(b.rholist)).max() < 1e-12


@pytest.mark.parametrize("legacy", [False, True])
def test_driven_against_oracle(legacy):
    from pyqed_amd.oqs import LindbladSolver
    d = 36
    H, c, rho0, e_ops = _system(d, seed=5)
    rng = np.random.default_rng(6)
    H1, H2 = _rand_herm(d, rng) / d, _rand_herm(d, rng) / d
    Hd = [H, [H1, lambda t: 0.8 * np.sin(3.0 * t)], [H2, lambda t: 0.5 * np.cos(t)]]
    res = LindbladSolver(Hd, c).run(rho0, 0.02, 25, e_ops=e_ops, legacy_accumulate=legacy)
    ref = OL.run_driven(Hd, rho0, c, 0.02, 25, e_ops=e_ops, legacy=legacy)
    assert np.abs(res.observables - ref["observables"]).max() < 1e-12
    assert np.abs(np.array(res.rholist) - np.array(ref["rholist"])).max() < 1e-12


def test_batch_and_correlation_against_oracle():
    from pyqed_amd.oqs import LindbladSolver
    d = 40
    H, c, rho0, e_ops = _system(d, seed=8)
    rng = np.random.default_rng(9)
    rhos = []
    for _ in range(3):
        p = rng.dirichlet(np.ones(d))
        rhos.append(np.diag(p).astype(complex))
    s = LindbladSolver(H, c)
    final = s.run_batch(np.array(rhos), 0.02, 20)
    for k in range(3):
        ref = OL.run(H, c, rhos[k], 0.02, 20)["rholist"][-1]
        assert np.abs(final[k] - ref).max() < 1e-12
    a_op, b_op = e_ops[3], rng.normal(size=(d, d)) + 0j
    cor = s.correlation_2op_1t(rho0, a_op, b_op, 0.02, 15, output=None)
    ref = OL.correlation_2op_1t(H, c, rho0, a_op, b_op, 0.02, 15)
    assert np.abs(cor - ref).max() < 1e-12 * max(1.0, np.abs(ref).max())


def test_unitary_limit_keeps_trace_and_purity():
    """No collapse operators: trace kept to rounding, purity to RK4's own truncation (the oracle's RK4
    drifts the same way: compared to 1e-12), over many steps."""
    from pyqed_amd.oqs import LindbladSolver
    d = 48
    rng = np.random.default_rng(2)
    H = _rand_herm(d, rng) / np.sqrt(d)
    v = rng.normal(size=d) + 1j * rng.normal(size=d)
    v /= np.linalg.norm(v)
    rho0 = np.outer(v, v.conj())
    res = LindbladSolver(H, []).run(rho0, 0.01, 200)
    rT = res.rholist[-1]
    ref = OL.run(H, [], rho0, 0.01, 200)["rholist"][-1]
    assert abs(np.trace(rT) - 1) < 1e-12
    assert abs(np.trace(rT @ rT).real - np.trace(ref @ ref).real) < 1e-12
    assert abs(np.trace(rT @ rT).real - 1) < 1e-8


def test_redfield_beyond_the_liouville_kernel():
    """RedfieldSolver.evolve at d = 40 (N = 1600 > 1024) on the matrix-form kernel with the tensor's
    operator form, against the oracle's Liouville-space RK4 with the dense tensor."""
    from pyqed_amd.oqs import RedfieldSolver
    rng = np.random.default_rng(12)
    d = 40
    H = np.diag(np.sort(rng.uniform(0, 2, d))) + 0.05 * _rand_herm(d, rng)
    a_ops = [_rand_herm(d, rng) / np.sqrt(d), np.diag(rng.normal(size=d)).astype(complex)]
    S = lambda w: 0.05 * w / (1 - np.exp(-2.0 * w)) if w != 0 else 0.025  # noqa: E731
    rho0 = np.diag(rng.dirichlet(np.ones(d))).astype(complex)
    e_ops = [np.diag(np.eye(d)[j]).astype(complex) for j in (0, 7, d - 1)]
    res = RedfieldSolver(H, a_ops, [S, S]).evolve(rho0, 0.05, 20, e_ops=e_ops)
    R, ev = OL.redfield_tensor(H, a_ops, [S, S])
    ref = OL.run_redfield(R, rho0, ev, 0.05, 20, e_ops)
    assert np.abs(res.observables - ref["observables"]).max() < 1e-12
    assert np.abs(np.array(res.rholist) - np.array(ref["rholist"])).max() < 1e-12


def test_redfield_matrix_form_equals_liouville_form():
    from pyqed_amd.oqs import RedfieldSolver
    rng = np.random.default_rng(13)
    d = 4
    H = np.diag([0.0, 0.5, 1.1, 1.9]) + 0.1 * _rand_herm(d, rng)
    a_ops = [_rand_herm(d, rng), np.diag([1.0, -1, 0.5, 0]).astype(complex)]
    S = lambda w: 0.05 * w / (1 - np.exp(-2.0 * w)) if w != 0 else 0.025  # noqa: E731
    rho0 = np.diag([0.1, 0.2, 0.3, 0.4]).astype(complex)
    e_ops = [np.diag(np.eye(d)[j]).astype(complex) for j in range(d)]
    a = RedfieldSolver(H, a_ops, [S, S], form="matrix").evolve(rho0, 0.05, 300, e_ops=e_ops)
    b = RedfieldSolver(H, a_ops, [S, S], form="liouville").evolve(rho0, 0.05, 300, e_ops=e_ops)
    assert np.abs(a.observables - b.observables).max() < 1e-12
    assert np.abs(np.array(a.rholist) - np.array(b.rholist)).max() < 1e-12
