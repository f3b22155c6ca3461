"""Result container of the Lindblad / Redfield propagators (drop-in for pyqed.mol.Result,
pyqed/mol.py:98-171)."""
import numpy as np


class Result:
    def __init__(self, description=None, psi0=None, rho0=None, dt=None, Nt=None, times=None, t0=0, nout=1):
        self.description = description
        self.dt = dt
        self.timesteps = self.nt = Nt
        self.observables = None
        self.rholist = None
        self.psilist = []
        self.psi = None
        self.rho0 = rho0
        self.psi0 = psi0
        self.nout = nout
        if Nt is None and times is not None:
            # an explicit time grid (the eigen-decomposition solvers pass one); the reference's
            # Result(times=...) raises on Nt // nout with Nt None
            self.times = np.asarray(times)
        else:
            self.times = t0 + np.arange(Nt // nout + 1) * dt * nout

    def expect(self):
        return self.observables

    def dump(self, fname):
        """Save observables, times and states as an .npz archive (no pickling)."""
        arrs = dict(times=self.times, dt=np.array(self.dt), nt=np.array(self.nt))
        if self.observables is not None:
            arrs["observables"] = np.asarray(self.observables)
        if self.rholist:
            arrs["rholist"] = np.asarray(self.rholist)
        np.savez(fname, **arrs)

    def save(self, fname):
        self.dump(fname)
