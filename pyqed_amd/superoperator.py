"""Eigen-decomposition Lindblad solver (pyqed/superoperator.py:455-772) on the GPU spectral-sum kernel.

Lindblad_solver(H, c_ops) diagonalises the Liouvillian once on the host (scipy.linalg.eig with left
and right eigenvectors, as the reference) and evaluates every time or frequency series
    sum_n x_n f(lambda_n, s) y_n,   f = exp(lambda t) or -1/(lambda + i w)
with pyqed_spectral_sum.  The coefficient vectors are formed with matrix products instead of the
reference's per-eigenvector Python loops; the values agree to rounding.

Documented differences from the reference:
  * evolve returns its Result with the given time grid (the reference's Result(times=...) raises);
  * evolve and the correlation functions diagonalise on first use instead of failing when
    eigenstates() was not called;
  * eigenstates(k) with k set raises NotImplementedError (the reference's partial-spectrum branch
    returns undefined names);
  * liouvillian() is a dense array (the reference builds a scipy.sparse matrix).

Kept as the reference has it: every expansion divides by norm_n = Re <L_n|R_n>, the real part of the
left/right overlap, which is exact only when the overlaps are real.  The result then depends on the
phases LAPACK gives the eigenvectors, so liouvillian() repeats the reference's exact sequence of
floating-point operations, and the decomposition sees bitwise the reference's matrix.
Lindblad_solver(H, c_ops, biorthogonal=True) divides by the complex overlap instead, which gives
e^{Lt} exactly for any diagonalisable L.
"""
import numpy as np
import scipy.linalg

from .liouville import dag
from .mol import Result
from .spectral import FREQ, TIME, spectral_sum

_KINDS = {"commutator": "c", "c": "c", "-": "c", "left": "l", "l": "l", "right": "r", "r": "r",
          "anticommutator": "a", "a": "a", "+": "a"}


def operator_to_vector(rho):
    """Row-major vec (superoperator.py:111-128)."""
    return np.asarray(rho.toarray() if hasattr(rho, "toarray") else rho).flatten()


dm2vec = operator_to_vector


def operator_to_superoperator(a, kind="commutator"):
    """Row-major superoperators (superoperator.py:200-246): kron(a, I) -/+ kron(I, a^T), left, right."""
    a = np.asarray(a.toarray() if hasattr(a, "toarray") else a, np.complex128)
    if kind not in _KINDS:
        raise ValueError('Error: superoperator {} does not exist.'.format(kind))
    I = np.eye(a.shape[-1])
    k = _KINDS[kind]
    if k == "c":
        return np.kron(a, I) - np.kron(I, a.T)
    if k == "l":
        return np.kron(a, I)
    if k == "r":
        return np.kron(I, a.T)
    return np.kron(a, I) + np.kron(I, a.T)


def left(a):
    return operator_to_superoperator(a, "left")


def right(a):
    return operator_to_superoperator(a, "right")


def lindblad_dissipator(l):
    """kron(l, l*) - 1/2 (kron(l^dag l, I) + kron(I, (l^dag l)^T)) (superoperator.py:249-251)."""
    l = np.asarray(l.toarray() if hasattr(l, "toarray") else l, np.complex128)
    return np.kron(l, l.conj()) - 0.5 * operator_to_superoperator(dag(l) @ l, "anticommutator")


def liouvillian(H, c_ops):
    """-i [H, .] + sum_c D[c] as a dense N x N matrix (superoperator.py:29-57)."""
    L = -1j * operator_to_superoperator(H)
    for c in c_ops or []:
        L = L + lindblad_dissipator(c)
    return L


def obs(rho, a):
    """vdot(vec(a^dag), vec(rho)) = Tr(a rho) (superoperator.py:313-314)."""
    return np.vdot(operator_to_vector(dag(np.asarray(a))), np.asarray(rho).reshape(-1))


class Lindblad_solver:
    def __init__(self, H, c_ops=None, *, biorthogonal=False):
        self.biorthogonal = biorthogonal
        self.H = H
        self.c_ops = c_ops
        self.L = None
        self.dim = H.shape[-1] ** 2
        self.idv = operator_to_vector(np.identity(H.shape[-1]))
        self.left_eigvecs = None
        self.right_eigvecs = None
        self.eigvals = None
        self.norm = None

    def liouvillian(self):
        L = liouvillian(self.H, self.c_ops)
        self.L = L
        return L

    def eigenstates(self, k=None):
        """Full right / left eigen-decomposition of L (superoperator.py:489-523); returns
        (eigvals, right eigenvectors, left eigenvectors)."""
        if k is not None:
            raise NotImplementedError("eigenstates(k): only the full decomposition (k=None) is supported")
        L = self.L if self.L is not None else self.liouvillian()
        w, vl, vr = scipy.linalg.eig(np.asarray(L), left=True, right=True)
        self.eigvals = w
        self.left_eigvecs = vl
        self.right_eigvecs = vr
        ov = np.diagonal(dag(vl) @ vr)
        self.norm = ov if self.biorthogonal else ov.real
        return w, vr, vl

    def _spectrum(self):
        if self.eigvals is None:
            self.eigenstates()
        return self.eigvals, self.right_eigvecs, self.left_eigvecs, self.norm

    def _project(self, vec):
        """(U2^dag vec) / norm: expansion coefficients of a Liouville vector."""
        _, _, U2, norm = self._spectrum()
        return (dag(U2) @ vec) / norm

    def evolve(self, rho0, tlist, e_ops):
        """Observables Tr(e rho(t)) at every t of tlist (superoperator.py:525-565): Result with
        observables [len(tlist), len(e_ops)]."""
        evals, U1, _, _ = self._spectrum()
        result = Result(times=tlist)
        coeff = self._project(operator_to_vector(rho0))
        # obs(rho, e) = vdot(vec(e^dag), rho): fold the observable rows into the expansion
        rows = np.array([np.conj(operator_to_vector(dag(np.asarray(e)))) @ U1 for e in e_ops]).reshape(len(e_ops), -1)
        if len(e_ops):
            out = spectral_sum(TIME, rows, evals, coeff[:, None], tlist)
            result.observables = out[:, :, 0]
        else:
            result.observables = np.zeros((len(tlist), 0), complex)
        return result

    def _two_op_coeff(self, a_sup, start):
        evals, U1, _, _ = self._spectrum()
        x = self.idv @ (a_sup @ U1)
        y = self._project(operator_to_vector(start))
        return evals, x, y

    def correlation_2op_1t(self, rho0, ops, tlist):
        """<A(t)B> = sum_n e^{lambda_n t} <<I|A|R_n>> <<L_n|B rho0>> / norm_n (superoperator.py:567-601)."""
        a, b = ops
        evals, x, y = self._two_op_coeff(left(a), np.asarray(b) @ np.asarray(rho0))
        return spectral_sum(TIME, x[None], evals, y[:, None], tlist)[:, 0, 0]

    def correlation_2op_1w(self, rho0, ops, w):
        """S(w) = sum_n -1/(lambda_n + i w) coeff_n (superoperator.py:603-637)."""
        a, b = ops
        evals, x, y = self._two_op_coeff(left(a), np.asarray(b) @ np.asarray(rho0))
        return spectral_sum(FREQ, x[None], evals, y[:, None], w)[:, 0, 0]

    def correlation_3op_1t(self, rho0, ops, t):
        """<A B(t) C> with coefficients <<I|B|R_n>> <<L_n|C rho0 A>> / norm_n (superoperator.py:639-670)."""
        a, b, c = (np.asarray(o) for o in ops)
        evals, x, y = self._two_op_coeff(left(b), c @ np.asarray(rho0) @ a)
        return spectral_sum(TIME, x[None], evals, y[:, None], t)[:, 0, 0]

    def correlation_3op_1w(self, rho0, ops, w):
        """Frequency-domain correlation_3op_1t (superoperator.py:672-703)."""
        a, b, c = (np.asarray(o) for o in ops)
        evals, x, y = self._two_op_coeff(left(b), c @ np.asarray(rho0) @ a)
        return spectral_sum(FREQ, x[None], evals, y[:, None], w)[:, 0, 0]

    def correlation_3op_2t(self, rho0, ops, tlist, taulist, k=None):
        """<A(t) B(t+tau) C(t)> as [len(taulist), len(tlist)] (superoperator.py:705-752):
        e^{lambda tau}^T coeff e^{lambda t}, coeff[m, n] = <<I|B|R_m>> <<L_m|right(A) left(C)|R_n>>
        <<L_n|rho0>> / (norm_m norm_n).  Two spectral sums: over n for every t, then over m for
        every tau."""
        a, b, c = ops
        evals, U1, U2, norm = self._spectrum()
        x = self.idv @ (left(b) @ U1)
        M = dag(U2) @ (right(a) @ (left(c) @ U1))
        y = (dag(U2) @ operator_to_vector(rho0)) / norm
        coeff = (x / norm)[:, None] * M * y[None, :]
        ones = np.ones((len(evals), 1), np.complex128)
        P = spectral_sum(TIME, coeff, evals, ones, tlist)[:, :, 0]       # [Nt, k]: sum_n coeff[m, n] e^{lambda_n t}
        return spectral_sum(TIME, P, evals, ones, taulist)[:, :, 0]     # [Ntau, Nt]

    def correlation_4op_2t(self, rho0, ops, tlist, taulist, k=None):
        """<A(t) B(t+tau) C(t+tau) D(t)> = correlation_3op_2t with [A, BC, D] (superoperator.py:754-772)."""
        if len(ops) != 4:
            raise ValueError('Number of operators is not 4.')
        a, b, c, d = ops
        return self.correlation_3op_2t(rho0=rho0, ops=[a, np.asarray(b) @ np.asarray(c), d], tlist=tlist,
                                       taulist=taulist, k=k)
