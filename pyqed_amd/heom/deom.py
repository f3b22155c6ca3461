"""Drop-in for pyqed.heom.deom: Bath and DEOMSolver on the MI355X propagator.

DEOMSolver keeps the reference's constructor, setters and
``run(rho0, dt, nt, p1=None)`` contract (pyqed/heom/deom.py:953-1114):
returns ``(t_save, ddos_save)`` with t_save = k*dt (k = 0..nt) and ddos_save
either the list of rho_0(t) matrices or, with p1, the array Tr(p1 rho_0(t)).
The time loop runs on the GPU through libpyqed_amd (there is no CPU path).

Documented differences from the reference:
  * rho0 is not modified (the reference aliases it as ADO 0 and updates it in
    place, deom.py:1092,766); pass ``inplace=True`` to get that behaviour.
  * no tqdm progress bar.
  * pulse functions must be real-valued (complex fields raise).
Keyword-only options: precision ("fp64" plain FP64, or "mixed" FP32
right-hand side with FP64 state), save_every (keep every n-th rho_0).
"""
import numpy as np

from .bath import (Bath, decompose_spectrum_pade, decompose_spectrum_pade_imag, decompose_spectrum_pade_real,
                   decompose_spectrum_prony, drude, fit_t, function_bose, matsubara_approximation_distribution,
                   pade_approximation_distribution, prony_find_gamma, prony_fitting, single_oscillator,
                   sort_symmetry, spectrum_exp)
from .engine import HeomEngine, field_table

__all__ = ["Bath", "DEOMSolver", "decompose_spectrum_pade", "pade_approximation_distribution",
           "matsubara_approximation_distribution", "function_bose", "single_oscillator", "drude",
           "decompose_spectrum_prony", "prony_fitting", "prony_find_gamma", "decompose_spectrum_pade_real",
           "decompose_spectrum_pade_imag", "sort_symmetry", "fit_t", "spectrum_exp"]


class DEOMSolver:
    """Hierarchical equations of motion solver (deom.py:953-1114)."""

    def __init__(self, system=None, system_dipole=None, bath=None, coupling=None, coupling_dipole=None,
                 pulse_system_func=None, pulse_coupling_func=None, lmax=None):
        self.system = system
        self.system_dipole = system_dipole
        self.coupling = coupling
        self.coupling_dipole = coupling_dipole
        self.pulse_system_func = pulse_system_func
        self.pulse_coupling_func = pulse_coupling_func
        self.lmax = lmax
        self.nsys = 0
        self.nmax = 1
        self.nind = 0
        self.nmod = 0
        self.bath = bath
        self.comb_list = []
        self.engine = None

    # ---------------------------------------------------------------- setters (deom.py:992-1033)
    def set_hierarchy(self, lmax):
        self.lmax = lmax

    def set_system(self, system):
        self.system = np.array(system, dtype=np.complex128)

    def set_system_dipole(self, system_dipole):
        self.system_dipole = np.array(system_dipole, dtype=np.complex128)

    def set_coupling(self, coupling):
        self.coupling = np.array(coupling, dtype=np.complex128)

    def set_coupling_dipole(self, coupling_dipole):
        self.coupling_dipole = np.array(coupling_dipole, dtype=np.complex128)

    def set_pulse_system_func(self, pulse_system_func):
        self.pulse_system_func = pulse_system_func

    def set_pulse_coupling_func(self, pulse_coupling_func):
        self.pulse_coupling_func = pulse_coupling_func

    # ---------------------------------------------------------------- checks (deom.py:1035-1064)
    def check_(self):
        if self.system is None:
            raise ValueError("System Hamiltonian is not set.")
        if self.coupling is None:
            raise ValueError("system bath interaction operator is not set.")
        if self.bath is None:
            raise ValueError("bath is not set.")
        if self.lmax is None:
            raise ValueError("hierarchy depth (lmax) is not set.")
        self.nsys = np.shape(self.system)[0]
        self.nind = len(self.bath.expn)
        self.nmod = int(np.max(self.bath.mode)) + 1 if self.nind else 0

    def init_(self):
        """Pascal table and nmax as the reference keeps them (deom.py:1048-1060); the keys come from
        the library's tables."""
        from .. import _lib
        combmax = self.nind + self.lmax + 1
        comb = np.zeros((combmax, combmax), np.int64)
        comb[0, 0] = 1
        for i in range(1, combmax):
            comb[i, 0] = 1
            for j in range(1, combmax):
                comb[i, j] = comb[i - 1, j] + comb[i - 1, j - 1]
        self.comb_list = comb
        self.nmax = int(comb[self.lmax + self.nind, self.lmax])
        self.keys = _lib.hierarchy_tables(self.nind, self.lmax)[0]

    # ---------------------------------------------------------------- time-dependent couplings
    def _coupling_dipoles(self, ncoup, d):
        """Matrices C_m with Q_m(t) = Q_m + C_m f_c(t), as numpy evaluates generate_time
        (deom.py:676-687): with a 2-D coupling_dipole, coupling_dip[m] is its row m, broadcast
        over the rows of Q_m."""
        cd = np.asarray(self.coupling_dipole, np.complex128)
        if cd.ndim == 3:
            return cd[:ncoup]
        if cd.ndim == 2:
            return np.array([np.broadcast_to(cd[m], (d, d)) for m in range(ncoup)])
        raise ValueError("coupling_dipole must be 2-D or 3-D")

    def _instance(self, rho0):
        H = np.asarray(self.system, np.complex128)
        d = H.shape[0]
        Q = np.array([np.asarray(q, np.complex128) for q in self.coupling]).reshape(-1, d, d)
        b = self.bath
        inst = dict(H=H, Q=Q, L=int(self.lmax), rho0=np.asarray(rho0, np.complex128),
                    bath=(b.etal, b.etar, b.etaa, b.expn, np.asarray(b.mode)))
        return inst

    def run(self, rho0, dt, nt, p1=None, *, precision="fp64", save_every=1, inplace=False, trace_tol=None):
        """Propagate nt RK4 steps; returns (t_save, ddos_save) as deom.py:1072-1114.

        trace_tol: when given, raise RuntimeError if |Tr rho(t) - Tr rho0| exceeds it at any saved time
        (every coupling term of the hierarchy leaves the trace of rho unchanged, so drift means a broken
        run: a non-finite value, or a step far beyond stability)."""
        self.check_()
        self.init_()
        inst = self._instance(rho0)
        d = inst["H"].shape[0]
        fs = fc = None
        if self.system_dipole is not None and self.pulse_system_func is not None:
            sd = np.asarray(self.system_dipole, np.complex128)
            if np.any(sd != 0):
                inst["sdip"] = sd
                fs = self.pulse_system_func
        if self.coupling_dipole is not None and self.pulse_coupling_func is not None:
            cd = self._coupling_dipoles(inst["Q"].shape[0], d)
            if np.any(cd != 0):
                inst["cdip"] = cd
                fc = self.pulse_coupling_func
        field = None
        tab = None
        if fs is not None or fc is not None:
            tab = field_table(nt, dt, fs, fc)
        eng = HeomEngine([inst])
        self.engine = eng
        if tab is not None:
            field = eng.torch.as_tensor(tab, device=eng.device)
        traj = eng.run(dt, nt, precision=precision, field=field, save_every=save_every)[0]
        t_save = np.arange(traj.shape[0], dtype=np.float64) * dt * save_every
        if trace_tol is not None:
            tr = np.trace(traj, axis1=1, axis2=2)
            drift = np.abs(tr - tr[0])
            bad = ~np.isfinite(drift) | (drift > trace_tol)
            if np.any(bad):
                k = int(np.argmax(bad))
                raise RuntimeError(f"trace drift {drift[k]:.3e} > {trace_tol:.1e} at t = {t_save[k]:.6g} "
                                   f"(saved step {k})")
        if inplace:
            np.copyto(rho0, traj[-1])
        if p1 is None:
            return t_save, [traj[i] for i in range(traj.shape[0])]
        P1 = np.asarray(p1, np.complex128)
        return t_save, np.einsum("ij,tji->t", P1, traj)

    # ---------------------------------------------------------------- generator and spectra
    def gen_generate_propgator(self):
        """Dense generator self.propgator, vec(RHS(rho)) = propgator vec(rho), row-major ADO blocks
        (deom.py:1116-1125, 769-884), assembled from the hierarchy tables."""
        from .generator import heom_generator
        self.check_()
        self.init_()
        b = self.bath
        Q = [np.asarray(q, np.complex128) for q in self.coupling]
        self.propgator = heom_generator(np.asarray(self.system, np.complex128), Q,
                                        (b.etal, b.etar, b.etaa, b.expn), b.mode, self.lmax, sparse=False)
        return self.propgator

    def correlation_4op_3t(self, operator_a, operator_b, operator_c, operator_d, rho0, T, w_x, w_y, if_full=True,
                           cut_off_min=0.5, cut_off_max=1.1, if_load=False, if_save=False, lcr="llll", *,
                           method="eig", tol=1e-10, krylov_dim=400):
        """Two-frequency four-point correlation (deom.py:1127-1210):
        C(wx, wy) = Tr_0[ d G(wx) c e^{L T} b G(wy) a rho0 ],  G(w) = (-L - i w)^-1,
        with operator x acting as lcr[.] ('l' x., 'r' .x, 'c' x. + .x; a uses lcr[0], d lcr[3]),
        evaluated in the eigenbasis of the dense generator L; with if_full=False only eigenvalues with
        cut_off_min*min(Re) < Re < cut_off_max*max(Re) are kept.  The eigen-decomposition runs on the
        host, as in the reference; the frequency grid, the reference's double loop, runs on the GPU
        spectral-sum kernel.

        method="krylov" avoids the dense generator altogether: every product with L is the GPU
        right-hand-side kernel, the resolvents of all frequencies share one Arnoldi basis per starting
        vector and e^{L T} is a Krylov exponential (pyqed_amd.heom.krylov), each basis grown until its
        residual is below tol relative to its starting vector (at most krylov_dim vectors).  It
        evaluates the same expression as if_full=True, for hierarchies too large for la.eig."""
        if method == "krylov":
            return self._correlation_krylov((operator_a, operator_b, operator_c, operator_d), rho0, T, w_x, w_y,
                                            lcr, tol, krylov_dim)
        if method != "eig":
            raise ValueError("method must be 'eig' or 'krylov'")
        import os
        import scipy.linalg as la
        from .generator import system_action
        if getattr(self, "propgator", None) is None:
            self.gen_generate_propgator()
        if if_load and os.path.exists("correlation_4op_3t.npz"):
            data = np.load("correlation_4op_3t.npz")
            self._delta, self._V, self._Vinv = data["Δ"], data["V"], data["V_inv"]
        if getattr(self, "_delta", None) is None:
            self._delta, self._V = la.eig(self.propgator)
            self._Vinv = la.pinv(self._V)
        if if_save:
            np.savez("correlation_4op_3t.npz", Δ=self._delta, V=self._V, V_inv=self._Vinv)
        n = self.nsys
        act = {name: system_action(op, self.nmax, c).toarray()
               for name, op, c in zip("abcd", (operator_a, operator_b, operator_c, operator_d), lcr)}
        rho = np.zeros(self.nmax * n * n, complex)
        rho[:n * n] = np.asarray(rho0, complex).reshape(-1)
        lam, V, Vi = self._delta, self._V, self._Vinv
        if not if_full:
            re = np.real(lam)
            keep = (re > np.min(re) * cut_off_min) & (re < np.max(re) * cut_off_max)
            lam, V, Vi = lam[keep], V[:, keep], Vi[keep, :]
        left = (act["d"] @ V)[:n * n]                                              # d V (ADO 0 rows)
        mid = ((Vi @ act["c"] @ V) * np.exp(lam * T)[None, :]) @ (Vi @ act["b"] @ V)  # c e^{LT} b
        right = Vi @ (act["a"] @ rho)                                              # a rho
        trace_row = left.reshape(n, n, -1)[np.arange(n), np.arange(n)].sum(axis=0)  # Tr over ADO 0
        # the (w_x, w_y) grid as two resolvent spectral sums on the GPU, 1/(-lam - i w) = -1/(lam + i w):
        #   P[j, m] = sum_k mid[m, k] g_k(w_y[j]) right[k];  C[i, j] = sum_m trace_row[m] g_m(w_x[i]) P[j, m]
        from ..spectral import FREQ, spectral_sum
        P = spectral_sum(FREQ, mid, lam, right[:, None], w_y)[:, :, 0]
        return spectral_sum(FREQ, trace_row[None, :], lam, P.T, w_x)[:, 0, :]

    def _correlation_krylov(self, ops, rho0, T, w_x, w_y, lcr, tol, krylov_dim):
        from .. import _lib
        from .krylov import correlation_grid
        torch = _lib.require_gpu()
        self.check_()
        self.init_()
        eng = HeomEngine([self._instance(rho0)], packed=False)
        try:
            n, nmax = self.nsys, int(eng.nmax[0])
            M = nmax * n * n
            dev = eng.device
            buf = torch.zeros(((M + 15) // 16) * 16, dtype=torch.complex128, device=dev)

            def matvec(v):
                buf[:M] = v
                return eng.rhs(buf)[:M]

            def on_blocks(op, side):
                """The action of op on a stack of d x d blocks [..., n, n] ('l' A., 'r' .A, 'c' A. + .A)."""
                A = torch.as_tensor(np.array(op, np.complex128, order="C"), device=dev)
                if side not in ("l", "r", "c"):
                    raise ValueError(side)

                def f(X):
                    if side == "l":
                        return torch.matmul(A, X)
                    if side == "r":
                        return torch.matmul(X, A)
                    return torch.matmul(A, X) + torch.matmul(X, A)
                return f

            def on_vector(op, side):
                g = on_blocks(op, side)
                return lambda v: g(v.reshape(nmax, n, n)).reshape(-1)

            act = {k: on_vector(op, c) for k, op, c in zip("abc", ops[:3], lcr[:3])}
            d_blocks = on_blocks(ops[3], lcr[3])

            def trace_d(V):
                """Tr of d applied to ADO 0, for every row of V."""
                Z = d_blocks(V[:, :n * n].reshape(V.shape[0], n, n))
                return torch.diagonal(Z, dim1=-2, dim2=-1).sum(-1).cpu().numpy()

            rho = torch.zeros(M, dtype=torch.complex128, device=dev)
            rho[:n * n] = torch.as_tensor(np.asarray(rho0, np.complex128).reshape(-1), device=dev)
            return correlation_grid(matvec, act, trace_d, rho, float(T), w_x, w_y, torch, tol=tol, m_max=krylov_dim)
        finally:
            eng.close()

    @property
    def ddos(self):
        """Final hierarchy state [nmax, d, d] of the last run (the reference's self.ddos)."""
        if self.engine is None:
            return None
        return self.engine.hierarchy(0)
