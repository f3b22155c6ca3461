"""CPU oracle for the eigen-decomposition solvers -- TEST INFRASTRUCTURE ONLY.

Imported only by tests/ as the checker.

Reference: pyqed/superoperator.py and pyqed/oqs.py (binggu56/pyqed 0.1.3).

  liouvillian            superoperator.py:29-57 with lindblad_dissipator 249-251 (row-major vec)
  lindblad_eig           Lindblad_solver.eigenstates, superoperator.py:489-523 (scipy.linalg.eig,
                         left and right vectors, norm = Re diag(vl^dag vr)); exact=True keeps the
                         complex overlap diag(vl^dag vr), which the expansion needs in general
  lindblad_evolve        Lindblad_solver.evolve, superoperator.py:525-565, per-eigenvector loop
  correlation_2op_1t/1w  superoperator.py:567-637
  correlation_3op_1t/1w  superoperator.py:639-703
  correlation_3op_2t     superoperator.py:705-752, the double loop over (m, n)
  redfield_propagator    RedfieldSolver.propagator, oqs.py:160-209: 'SOS' as
                         einsum('aj,jk,jb->abk', U1, E, U1^-1), 'EOM' as the RK4 of phys.expm
                         (phys.py:2049-2091) with phys.rk4's summation order (phys.py:1051-1063)
  redfield_expect        oqs.py:211-242 (sum(vec(e) * rho), no conjugation, as the reference)
  correlation_4op_3t     oqs.py:270-360 with dense operators

What pins them (tests/test_oracle_eigen.py): with exact=True, scipy.linalg.expm of the generator
applied directly (evolution, regression correlations) and the resolvent as a linear solve of
(L + i w) for the frequency-domain functions; the reference's real-part norm (exact=False, the
default) against the closed-form two-level correlation <sigma+(t) sigma-> = exp((i w - G/2) t), where
the overlaps are real, and against exact=True after the eigenvectors are rephased so that every
overlap is real; the propagators against expm and RK4's fourth-order convergence for 'EOM'; the
four-point function against products of expm.
"""
import numpy as np
import scipy.linalg as sla


def dag(a):
    return np.conj(np.asarray(a)).T


def sop(a, kind):
    a = np.asarray(a, np.complex128)
    I = np.eye(a.shape[-1])
    return {"c": np.kron(a, I) - np.kron(I, a.T), "-": np.kron(a, I) - np.kron(I, a.T),
            "l": np.kron(a, I), "r": np.kron(I, a.T),
            "a": np.kron(a, I) + np.kron(I, a.T), "+": np.kron(a, I) + np.kron(I, a.T)}[kind]


def liouvillian(H, c_ops):
    """The reference's association: L = -i sop(H) + (kron(c, c*) - 1/2 sop+(c^dag c)) per c.  The
    real-part norm makes the expansion depend on LAPACK's eigenvector phases, so the generator is
    built with the reference's exact sequence of operations."""
    L = -1j * sop(H, "c")
    for c in c_ops:
        c = np.asarray(c, np.complex128)
        L = L + (np.kron(c, c.conj()) - 0.5 * sop(dag(c) @ c, "a"))
    return L


def lindblad_eig(L, exact=False, rephase=False):
    w, vl, vr = sla.eig(L, left=True, right=True)
    ov = np.diagonal(dag(vl) @ vr)
    if rephase:
        # multiply each left vector by a phase that makes its overlap real and positive
        vl = vl * (ov / np.abs(ov))[None, :]
        ov = np.diagonal(dag(vl) @ vr)
    norm = ov if exact else ov.real
    return w, vr, vl, norm


def lindblad_evolve(L, rho0, tlist, e_ops, exact=False, rephase=False):
    evals, U1, U2, norm = lindblad_eig(L, exact, rephase)
    r0 = np.asarray(rho0).flatten()
    k = U1.shape[-1]
    coeff = np.array([np.vdot(U2[:, n], r0) / norm[n] for n in range(k)])
    out = np.zeros((len(tlist), len(e_ops)), complex)
    for i, t in enumerate(tlist):
        rho = U1.dot(coeff * np.exp(evals * t))
        out[i, :] = [np.vdot(dag(np.asarray(e)).flatten(), rho) for e in e_ops]
    return out


def _coeff(L, sup, start, exact=False, rephase=False):
    evals, U1, U2, norm = lindblad_eig(L, exact, rephase)
    d = int(round(np.sqrt(L.shape[0])))
    idv = np.eye(d).flatten()
    k = U1.shape[-1]
    s = np.asarray(start).flatten()
    return evals, np.array([np.vdot(idv, sup.dot(U1[:, n])) * np.vdot(U2[:, n], s) / norm[n] for n in range(k)])


def correlation_2op_1t(L, rho0, a, b, tlist, exact=False, rephase=False):
    evals, c = _coeff(L, sop(a, "l"), np.asarray(b) @ np.asarray(rho0), exact, rephase)
    return np.array([np.sum(np.exp(evals * t) * c) for t in tlist])


def correlation_2op_1w(L, rho0, a, b, w, exact=False, rephase=False):
    evals, c = _coeff(L, sop(a, "l"), np.asarray(b) @ np.asarray(rho0), exact, rephase)
    return np.array([np.sum(-1.0 / (evals + 1j * x) * c) for x in w])


def correlation_3op_1t(L, rho0, a, b, c, tlist, exact=False, rephase=False):
    evals, co = _coeff(L, sop(b, "l"), np.asarray(c) @ np.asarray(rho0) @ np.asarray(a), exact, rephase)
    return np.array([np.sum(np.exp(evals * t) * co) for t in tlist])


def correlation_3op_1w(L, rho0, a, b, c, w, exact=False, rephase=False):
    evals, co = _coeff(L, sop(b, "l"), np.asarray(c) @ np.asarray(rho0) @ np.asarray(a), exact, rephase)
    return np.array([np.sum(-1.0 / (evals + 1j * x) * co) for x in w])


def correlation_3op_2t(L, rho0, a, b, c, tlist, taulist, exact=False, rephase=False):
    evals, U1, U2, norm = lindblad_eig(L, exact, rephase)
    d = int(round(np.sqrt(L.shape[0])))
    idv = np.eye(d).flatten()
    r0 = np.asarray(rho0).flatten()
    k = U1.shape[-1]
    coeff = np.zeros((k, k), complex)
    for m in range(k):
        for n in range(k):
            coeff[m, n] = (np.vdot(idv, sop(b, "l").dot(U1[:, m]))
                           * np.vdot(U2[:, m], sop(a, "r").dot(sop(c, "l").dot(U1[:, n]))) / norm[m]
                           * np.vdot(U2[:, n], r0) / norm[n])
    return np.exp(np.outer(evals, taulist)).T @ coeff @ np.exp(np.outer(evals, tlist))


def redfield_propagator(R, t, method="SOS"):
    R = np.asarray(R, np.complex128)
    t = np.asarray(t, float)
    if method == "SOS":
        evals, U1 = sla.eig(R)
        U2 = sla.inv(U1)
        E = np.exp(evals[:, None] * t[None, :])
        return np.einsum("aj,jk,jb->abk", U1, E, U2)
    N = R.shape[0]
    dt = t[1] - t[0] if len(t) > 1 else 0.0
    U = np.eye(N, dtype=complex)
    out = []
    for _ in range(len(t)):
        out.append(U.copy())
        k1 = R @ U
        k2 = R @ (U + k1 * dt / 2)
        k3 = R @ (U + k2 * dt / 2)
        k4 = R @ (U + k3 * dt)
        U = U + (k1 + 2 * k2 + 2 * k3 + k4) / 6.0 * dt
    return np.dstack(out)


def redfield_expect(U, evecs, rho0, e_ops):
    r0 = (dag(evecs) @ np.asarray(rho0) @ evecs).flatten()
    rho = np.tensordot(U, r0, axes=([1], [0]))
    return np.array([(dag(evecs) @ np.asarray(e) @ evecs).flatten().dot(rho) for e in e_ops]).T


def correlation_4op_3t(G, rho0, ops, signature):
    a, b, c, d = (sop(o, s) for o, s in zip(ops, signature))
    n = int(round(np.sqrt(G.shape[0])))
    idm = np.eye(n).flatten()
    rho = d @ np.asarray(rho0).flatten()
    tmp = np.tensordot(G, rho, axes=((1), (0)))
    tmp = c @ tmp
    tmp = np.tensordot(G, tmp, axes=([1], [0]))
    tmp = np.tensordot(b, tmp, axes=([1], [0]))
    tmp = np.tensordot(G, tmp, axes=([1], [0]))
    return np.einsum("a,ab,bijk->ijk", idm, a, tmp)
