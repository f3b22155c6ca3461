"""The operator forms the matrix-form kernel integrates, against the dense Liouville-space generators
they replace (row-major vec): the Lindblad generator (oqs.py:697-714) as -i(Heff rho - rho Heff^dag) +
sum_k c_k rho c_k^dag with Heff = H - (i/2) sum c_k^dag c_k, and the Bloch-Redfield tensor
(oqs.py:519-570) as -i(Heff rho - rho Heff^dag) + sum_a (A rho L^dag + L rho A) in the eigenbasis.
Host only."""
import numpy as np
import pytest


def _rand_herm(d, rng):
    A = rng.normal(size=(d, d)) + 1j * rng.normal(size=(d, d))
    return (A + A.conj().T) / 2


def _apply(Heff, lefts, rights, rho):
    out = -1j * (Heff @ rho - rho @ Heff.conj().T)
    for lft, rgt in zip(lefts, rights):
        out = out + lft @ rho @ rgt
    return out


@pytest.mark.parametrize("d", [2, 3, 6])
def test_lindblad_operator_form(d):
    from pyqed_amd.liouville import dag, lindblad_generator
    rng = np.random.default_rng(d)
    H = _rand_herm(d, rng)
    cs = [rng.normal(size=(d, d)) + 1j * rng.normal(size=(d, d)) for _ in range(2)]
    Heff = H - 0.5j * sum(dag(c) @ c for c in cs)
    rho = rng.normal(size=(d, d)) + 1j * rng.normal(size=(d, d))
    G = lindblad_generator(H, cs)
    np.testing.assert_allclose(_apply(Heff, cs, [dag(c) for c in cs], rho).reshape(-1), G @ rho.reshape(-1),
                               rtol=0, atol=1e-12)


@pytest.mark.parametrize("d", [3, 5])
def test_redfield_operator_form(d):
    from pyqed_amd.liouville import redfield_operators
    from pyqed_amd.oqs import redfield_tensor
    rng = np.random.default_rng(10 + d)
    H = np.diag(np.sort(rng.uniform(0, 2, d))) + 0.1 * _rand_herm(d, rng)
    a_ops = [_rand_herm(d, rng), np.diag(rng.normal(size=d)).astype(complex)]
    S1 = lambda w: 0.05 * w / (1 - np.exp(-2.0 * w)) if w != 0 else 0.025  # noqa: E731
    S2 = lambda w: 0.02 * np.exp(-abs(w))  # noqa: E731
    R, ev = redfield_tensor(H, a_ops, [S1, S2])
    Heff, lefts, rights, ev2 = redfield_operators(H, a_ops, [S1, S2])
    assert np.array_equal(ev, ev2)
    rho = rng.normal(size=(d, d)) + 1j * rng.normal(size=(d, d))
    np.testing.assert_allclose(_apply(Heff, lefts, rights, rho).reshape(-1), R @ rho.reshape(-1), rtol=0, atol=1e-12)
