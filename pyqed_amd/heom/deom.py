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

from .bath import (Bath, decompose_spectrum_pade, drude, function_bose, matsubara_approximation_distribution,
                   pade_approximation_distribution, single_oscillator)
from .engine import HeomEngine, field_table

__all__ = ["Bath", "DEOMSolver", "decompose_spectrum_pade", "pade_approximation_distribution",
           "matsubara_approximation_distribution", "function_bose", "single_oscillator", "drude"]


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

    def run(self, rho0, dt, nt, p1=None, *, precision="fp64", save_every=1, inplace=False):
        """Propagate nt RK4 steps; returns (t_save, ddos_save) as deom.py:1072-1114."""
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
        if inplace:
            np.copyto(rho0, traj[-1])
        if p1 is None:
            return t_save, [traj[i] for i in range(traj.shape[0])]
        P1 = np.asarray(p1, np.complex128)
        return t_save, np.einsum("ij,tji->t", P1, traj)

    @property
    def ddos(self):
        """Final hierarchy state [nmax, d, d] of the last run (the reference's self.ddos)."""
        if self.engine is None:
            return None
        return self.engine.hierarchy(0)
