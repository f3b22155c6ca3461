"""Which kernel LindbladSolver and RedfieldSolver choose (host logic, no device): form="auto" steps
d*d <= 1024 on the Liouville-space kernel and larger d on the matrix-form kernel; "liouville" and
"matrix" force one.  The two propagators are replaced by recorders."""
import numpy as np
import pytest

from pyqed_amd import oqs


def _fake(calls, name):
    def f(*args, **kw):
        # propagate(gens, rho0s, dt, nsteps, ...); propagate_matrix(H0, c_ops, drives, rho0s, dt, nsteps, ...)
        rho0s, nt = (args[3], args[5]) if name == "matrix" else (args[1], args[3])
        n, d = rho0s.shape[0], rho0s.shape[-1]
        calls.append((name, d))
        return np.zeros((n, nt + 1, 0), complex), np.zeros((n, nt, d, d), complex), rho0s
    return f


@pytest.fixture
def calls(monkeypatch):
    c = []
    monkeypatch.setattr(oqs, "propagate", _fake(c, "liouville"))
    monkeypatch.setattr(oqs, "propagate_matrix", _fake(c, "matrix"))
    return c


@pytest.mark.parametrize("d,form,want", [(32, "auto", "liouville"), (33, "auto", "matrix"),
                                         (8, "matrix", "matrix"), (40, "liouville", "liouville")])
def test_lindblad_form_selection(calls, d, form, want):
    H = np.diag(np.arange(d, dtype=float))
    c = np.zeros((d, d))
    c[0, 1] = 0.1
    rho0 = np.zeros((d, d), complex)
    rho0[1, 1] = 1
    oqs.LindbladSolver(H, [c], form=form).run(rho0, 0.01, 3)
    assert calls == [(want, d)]


def test_unknown_form_is_rejected():
    with pytest.raises(ValueError):
        oqs.LindbladSolver(np.eye(2), [], form="dense")
    with pytest.raises(ValueError):
        oqs.RedfieldSolver(np.eye(2), [], [], form="dense")


@pytest.mark.parametrize("d,form,want", [(6, "auto", "liouville"), (33, "auto", "matrix"), (6, "matrix", "matrix")])
def test_redfield_form_selection(monkeypatch, d, form, want):
    calls = []

    def liou(gens, rho0s, dt, nt, *a, **k):
        calls.append(("liouville", rho0s.shape[-1]))
        return np.zeros((1, nt + 1, 0), complex), np.zeros((1, nt, d, d), complex), rho0s

    def matrix(heff, lefts, rights, drives, rho0s, dt, nt, *a, **k):
        calls.append(("matrix", rho0s.shape[-1]))
        return np.zeros((1, nt + 1, 0), complex), np.zeros((1, nt, d, d), complex), rho0s

    monkeypatch.setattr(oqs, "propagate", liou)
    monkeypatch.setattr(oqs, "run_matrix", matrix)
    H = np.diag(np.linspace(0.0, 1.0, d))
    x = np.zeros((d, d))
    x[0, 1] = x[1, 0] = 1.0
    rho0 = np.zeros((d, d), complex)
    rho0[0, 0] = 1
    s = oqs.RedfieldSolver(H, [x], [lambda w: 0.1 if w > 0 else 0.05], form=form)
    s.evolve(rho0, 0.01, 3)
    assert calls == [(want, d)]
