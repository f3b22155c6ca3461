"""pyqed.HEOM.heom on the GPU: the second copy of the legacy single-exponential HEOM solver
(pyqed/HEOM/heom.py:161-414).  It differs from oqs.HEOMSolver in two places, both kept:

  run         classic RK4 over the whole hierarchy (HEOM/heom.py:275-346; rk4 of phys.py:1051-1064)
              with D0 = reorg (2T - i gamma), T = temperature in energy units; kernel mode 2
  propagator  the forward-Euler sweep with U_0 updated once per step (the loop over n starts at 1,
              HEOM/heom.py:349-413); kernel mode 3

Both run in pyqed_legacy_heom_run, one workgroup per hierarchy with every ADO in LDS.  The module's
enumeration helpers (state_number_enumerate, enr_state_dictionaries) and the Matsubara expansion of
the Drude-Lorentz correlation function (_calc_matsubara_params) are host code, as in the reference.
"""
import itertools

import numpy as np

from .. import oqs as _oqs
from ..oqs import AU2K, _dense, legacy_heom_launch
from ..liouville import op2sop


def state_number_enumerate(dims, excitations=None, state=None, idx=0):
    """Every occupation tuple [n_1, n_2, ...] with 0 <= n_i < dims[i], in lexicographic order (the
    last index fastest), as HEOM/heom.py:40-94 yields them: arrays when excitations is None, tuples
    otherwise; a truthy excitations drops states whose total exceeds it.  state and idx fix the
    first idx entries (the reference's recursion arguments)."""
    dims = list(dims)
    head = [] if state is None else [int(v) for v in list(state)[:idx]]
    for tail in itertools.product(*(range(int(n)) for n in dims[len(head):])):
        s = tuple(head) + tuple(tail)
        if excitations and sum(s) > excitations:
            continue
        yield np.array(s) if excitations is None else s


def enr_state_dictionaries(dims, excitations):
    """(number of states, state -> index, index -> state) over state_number_enumerate(dims,
    excitations) (HEOM/heom.py:97-126)."""
    state2idx, idx2state = {}, {}
    for i, s in enumerate(state_number_enumerate(dims, excitations)):
        state2idx[s] = i
        idx2state[i] = s
    return len(state2idx), state2idx, idx2state


def _calc_matsubara_params(N_exp, coup_strength, cut_freq, temperature):
    """Coefficients c_k and rates nu_k of the Drude-Lorentz bath correlation function
    C(t) = sum_k c_k e^{-nu_k t} (HEOM/heom.py:129-158), hbar = 1, beta = 1/temperature:
    nu_0 = gamma, c_0 = lambda gamma (cot(gamma beta / 2) - i);
    nu_k = 2 pi k / beta, c_k = 4 lambda gamma nu_k / ((nu_k^2 - gamma^2) beta) for k >= 1."""
    lam, gam, beta = coup_strength, cut_freq, 1.0 / temperature
    c, nu = [], []
    for k in range(N_exp):
        if k == 0:
            nu.append(gam)
            c.append(lam * gam * (1.0 / np.tan(gam * beta / 2.0) - 1j))
        else:
            v = 2 * np.pi * k / beta
            nu.append(v)
            c.append(4 * lam * gam * v / ((v ** 2 - gam ** 2) * beta))
    return c, nu


def rk_run_params(temperature, cutoff, reorganization):
    """(gamma, D0) of HEOM/heom.py:_heom (lines 295-313): gamma = cutoff, D0 = reorg (2T - i gamma);
    arrays broadcast."""
    T, gamma, lam = (np.asarray(v, np.float64) for v in (temperature, cutoff, reorganization))
    return gamma, lam * (2.0 * T - 1j * gamma)


def _heom(H, rho0, c_ops, e_ops, temperature, cutoff, reorganization, nado, dt, nt, fname=None,
          return_result=True, verbose=True):
    """Observables (n_e, nt) after every RK4 step (HEOM/heom.py:275-346)."""
    return HEOMSolver(H, c_ops, e_ops).run(rho0, dt, nt, temperature, cutoff, reorganization, nado,
                                           verbose=verbose)


def _heom_propagator(H, c_ops, e_ops, temperature, cutoff, reorganization, nado, dt, nt, fname=None,
                     verbose=True):
    """u (nado, d^2, d^2) of HEOM/heom.py:349-413."""
    return HEOMSolver(H, c_ops, e_ops).propagator(dt, nt, temperature, cutoff, reorganization, nado,
                                                  verbose=verbose)


class HEOMSolver(_oqs.HEOMSolver):
    """pyqed.HEOM.heom.HEOMSolver (HEOM/heom.py:161-272) on the GPU.

    run(rho0, dt, nt, temperature, cutoff, reorganization, nado) -> observables (n_e, nt), RK4;
    propagator(dt, nt, temperature, cutoff, reorganization, nado) -> u (nado, d^2, d^2), temperature
    in kelvin, U_0 updated once per step.  The printed diagnostics are the reference's.
    Documented differences: correlation_2op_1t works (the reference's module calls a name it never
    imports and raises NameError); it is the Lindblad regression <A(t)B> of oqs.py:717-791, the
    routine the oqs copy calls.  correlation_3op_2t returns None, as the reference's body is `pass`.
    The whole hierarchy of one instance must fit one workgroup's LDS ((4 nado + 2) d^2 complex
    doubles for run); larger cases raise."""

    def run(self, rho0, dt, nt, temperature, cutoff, reorganization, nado, *, verbose=True):
        H, S = self._ops()
        gamma, D0 = rk_run_params(temperature, cutoff, reorganization)
        gamma, D0, T = float(gamma), complex(D0), float(temperature)
        if verbose:
            self._announce(T, gamma, reorganization, D0, 'Cutoff gamma/(kT)')
        d = H.shape[0]
        ados = np.zeros((1, nado, d, d), np.complex128)
        ados[0, 0] = _dense(rho0)
        e_ops = list(self.e_ops) if self.e_ops is not None else []
        _, obs = legacy_heom_launch(2, H[None], S[None], None, [[gamma, D0.real, D0.imag]], ados, nt, dt, e_ops)
        return obs[0].T.copy()

    def run_batch(self, rho0s, dt, nt, temperature, cutoff, reorganization, nado, return_ados=False):
        """Many RK4 hierarchies in one launch: rho0s [n, d, d]; temperature, cutoff and reorganization
        are scalars or length-n arrays.  Returns observables [n, n_e, nt] (and the final ADOs
        [n, nado, d, d] with return_ados)."""
        H, S = self._ops()
        rho0s = np.asarray(rho0s, np.complex128)
        n, d = rho0s.shape[0], H.shape[0]
        gamma, D0 = (np.broadcast_to(v, (n,)) for v in rk_run_params(temperature, cutoff, reorganization))
        params = np.stack([gamma, D0.real, D0.imag], axis=1)
        ados = np.zeros((n, nado, d, d), np.complex128)
        ados[:, 0] = rho0s
        e_ops = list(self.e_ops) if self.e_ops is not None else []
        fin, obs = legacy_heom_launch(2, np.broadcast_to(H, (n, d, d)), np.broadcast_to(S, (n, d, d)), None, params,
                                      ados, nt, dt, e_ops)
        obs = np.transpose(obs, (0, 2, 1)).copy()
        return (obs, fin) if return_ados else obs

    def propagator(self, dt, nt, temperature, cutoff, reorganization, nado, *, verbose=True):
        H, S = self._ops()
        gamma = float(cutoff)
        T = temperature / AU2K
        a, b = np.pi * reorganization * T, 0.0
        if verbose:
            self._announce(T, gamma, reorganization, a, 'High-Temperature check gamma/(kT)')
        N = H.shape[0] ** 2
        u = np.zeros((1, nado, N, N), np.complex128)
        u[0, 0] = np.eye(N)
        fin, _ = legacy_heom_launch(3, op2sop(H)[None], op2sop(S)[None], op2sop(S, "anticommutator")[None],
                                    [[gamma, a, b]], u, nt, dt)
        return fin[0]

    def correlation_3op_2t(self, rho0, ops, dt, Nt, Ntau):
        return None
