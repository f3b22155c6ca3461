"""Drop-in Lindblad, Redfield and legacy HEOM solvers (pyqed/oqs.py) on the MI355X kernels.

LindbladSolver(H, c_ops, e_ops).run(rho0, dt, Nt, t0=0., e_ops=None, return_result=True) and
RedfieldSolver(H, c_ops, spectra, e_ops).evolve(rho0, dt, Nt, ...) return the reference's Result
(observables, rholist, times).  The RK4 is that of pyqed/phys.py:1051-1063 and runs on the GPU.
HEOMSolver(H, c_ops, e_ops).run / propagator are the single-exponential HEOM of oqs.py:1332-1372.

Documented differences from the reference:
  * the state is dense: the reference's _lindblad converts rho to csr and, with numpy 2, the
    coherent term -i[H, rho] then evaluates to an array of empty matrices (lost); here it is kept;
  * a driven Hamiltonian H = [H0, [H1, f1], ...] means H0 - sum_i f_i(t) H_i at t = t0 + (k+1) dt
    for step k (the reference's evaluation point, oqs.py:1792-1796) without the reference's in-place
    accumulation of every past drive into H0 (oqs.py:1725-1732); run(..., legacy_accumulate=True)
    reproduces that accumulation;
  * rholist holds dense numpy arrays.
"""
import numpy as np
import scipy.linalg

from . import _lib
from .liouville import dag, lindblad_generator, op2sop, propagate, propagate_matrix, redfield_operators, run_matrix
from .mol import Result
from .spectral import TIME, spectral_sum


class LindbladSolver:
    """form (keyword only): "auto" runs d*d <= 1024 on the Liouville-space kernel (vec(rho) and its
    generator in LDS) and larger d on the matrix-form kernel (pyqed_lindblad_run); "liouville" or
    "matrix" force one of them."""

    def __init__(self, H=None, c_ops=None, e_ops=None, *, form="auto"):
        if form not in ("auto", "liouville", "matrix"):
            raise ValueError("form must be 'auto', 'liouville' or 'matrix'")
        self.c_ops = c_ops
        self.e_ops = e_ops
        self.H = H
        self.form = form

    def set_c_ops(self, c_ops):
        self.c_ops = c_ops

    def set_e_ops(self, e_ops):
        self.e_ops = e_ops

    def setH(self, H):
        self.H = H

    def configure(self, c_ops, e_ops):
        self.c_ops = c_ops
        self.e_ops = e_ops

    def liouvillian(self):
        """Liouvillian superoperator (superoperator.liouvillian), row-major vec."""
        return lindblad_generator(self.H, self.c_ops or [])

    def _prop(self, rho0s, dt, Nt, field=None, e_ops=(), keep_states=True, static=False):
        """Propagate initial states [n, d, d] (field [n, Nt, n_drive] for a driven H) on the kernel that
        `form` selects; returns (observables, states, final) as liouville.propagate."""
        c_ops = self.c_ops or []
        rho0s = np.asarray(rho0s, np.complex128)
        d = rho0s.shape[-1]
        if isinstance(self.H, list):
            if static:
                raise NotImplementedError("correlation functions need a static Hamiltonian")
            H0 = np.asarray(self.H[0], np.complex128)
            drives = [np.asarray(Hi, np.complex128) for Hi, _ in self.H[1:]]
        else:
            H0 = np.asarray(self.H, np.complex128)
            drives = []
        if self.form == "matrix" or (self.form == "auto" and d * d > 1024):
            return propagate_matrix(H0, c_ops, drives, rho0s, dt, Nt, field if drives else None, e_ops,
                                    keep_states=keep_states)
        gens = [lindblad_generator(H0, c_ops)] + [1j * op2sop(Hi) for Hi in drives]
        return propagate(np.array(gens), rho0s, dt, Nt, field if drives else None, e_ops, keep_states=keep_states)

    def run(self, rho0, dt, Nt, t0=0.0, e_ops=None, return_result=True, *, legacy_accumulate=False):
        e_ops = [] if e_ops is None else list(e_ops)
        rho0 = np.asarray(rho0, np.complex128)
        result = Result(dt=dt, Nt=Nt, rho0=rho0)
        if isinstance(self.H, list):
            terms = self.H[1:]
            field = np.zeros((Nt, len(terms)))
            for k in range(Nt):
                t = t0 + (k + 1) * dt
                for i, (_, fi) in enumerate(terms):
                    v = complex(fi(t))
                    if v.imag != 0.0:
                        raise NotImplementedError("complex-valued drive amplitudes are not supported")
                    field[k, i] = v.real
            if legacy_accumulate:
                field = np.cumsum(field, axis=0)
            obs, states, _ = self._prop(rho0[None], dt, Nt, field[None], e_ops)
            result.observables = obs[0, 1:]
        else:
            obs, states, _ = self._prop(rho0[None], dt, Nt, None, e_ops)
            result.observables = obs[0]
        result.rholist = [states[0, k] for k in range(states.shape[1])]
        return result

    def run_batch(self, rho0s, dt, Nt, fields=None):
        """Propagate many initial states (and, for a driven H, per-instance fields [n, Nt, n_terms])
        in one launch; returns the final states [n, d, d]."""
        _, _, final = self._prop(rho0s, dt, Nt, fields, (), keep_states=False)
        return final

    # ---------------------------------------------------------------- correlation functions
    # Quantum regression on the GPU Liouville kernel (oqs.py:717-791, 1193-1327).  The propagations
    # of a two-time correlation (one per t) run as one batched launch.

    def correlation_2op_1t(self, rho0, a_op, b_op, dt, Nt, output='cor.dat'):
        """<A(t)B> = Tr[A e^{Lt}(B rho0)] at t = dt, ..., Nt*dt (oqs.py:1193-1222 -> 717-791): Nt
        values.  As the reference, lines 't value' are written to `output` (None: no file)."""
        A = np.asarray(a_op, np.complex128)
        start = np.asarray(b_op, np.complex128) @ np.asarray(rho0, np.complex128)
        obs, _, _ = self._prop(start[None], dt, Nt, None, [A], keep_states=False, static=True)
        cor = obs[0, 1:, 0].copy()
        if output is not None:
            with open(output, 'w') as f:
                t = 0.0
                for k in range(Nt):
                    t += dt
                    f.write('{} {} \n'.format(t, cor[k]))
        return cor

    def correlation_3op_1t(self, rho0, oplist, dt=0.005, Nt=1):
        """<A B(t) C> = Tr[B e^{Lt}(C rho0 A)] at t = 0, dt, ..., Nt*dt (oqs.py:1224-1242): Nt+1 values."""
        a_op, b_op, c_op = (np.asarray(o, np.complex128) for o in oplist)
        start = c_op @ np.asarray(rho0, np.complex128) @ a_op
        obs, _, _ = self._prop(start[None], dt, Nt, None, [b_op], keep_states=False, static=True)
        return obs[0, :, 0]

    def correlation_4op_1t(self, rho0, ops, dt, nt):
        """<A B(t) C(t) D> (oqs.py:1299-1313, the definition that takes effect): correlation_3op_1t
        with [A, BC, D]."""
        if len(ops) != 4:
            raise ValueError('Number of operators is not 4.')
        a, b, c, d = ops
        return self.correlation_3op_1t(rho0, [a, np.asarray(b) @ np.asarray(c), d], dt, nt)

    def correlation_3op_2t(self, rho0, ops, dt, Nt, Ntau):
        """<A(t) B(t+tau) C(t)> as [Nt, Ntau] (oqs.py:1263-1297): row j is t = (j+1) dt (the states
        after each step), column m is tau = m dt.  The Nt regression propagations C rho(t) A run in
        one launch.  (The reference stores the Ntau+1 values of each propagation in a row of width
        Ntau, which raises; the first Ntau values, tau = 0 .. (Ntau-1) dt, are kept here.)"""
        a_op, b_op, c_op = (np.asarray(o, np.complex128) for o in ops)
        _, states, _ = self._prop(np.asarray(rho0, np.complex128)[None], dt, Nt, None, (), static=True)
        starts = c_op @ states[0] @ a_op
        obs, _, _ = self._prop(starts, dt, Ntau, None, [b_op], keep_states=False, static=True)
        return obs[:, :Ntau, 0]

    def correlation_4op_2t(self, rho0, ops, dt, nt, ntau):
        """<A(t) B(t+tau) C(t+tau) D(t)> (oqs.py:1315-1327): correlation_3op_2t with [A, BC, D]."""
        if len(ops) != 4:
            raise ValueError('Number of operators is not 4.')
        a, b, c, d = ops
        return self.correlation_3op_2t(rho0, [a, np.asarray(b) @ np.asarray(c), d], dt, nt, ntau)


def redfield_tensor(H, a_ops, spectra, secular=False):
    """Bloch-Redfield generator in the eigenbasis of H (oqs.py:519-570); returns (R, evecs) with R a
    dense N x N matrix such that d vec(rho)/dt = R vec(rho).  `secular` is accepted and ignored, as
    in the reference."""
    for a in a_ops:
        a = a.toarray() if hasattr(a, "toarray") else np.asarray(a)
        if not np.allclose(a, dag(a)):
            raise TypeError("Operators in a_ops must be Hermitian.")
    H = H.toarray() if hasattr(H, "toarray") else np.asarray(H)
    evals, evecs = np.linalg.eigh(H)
    W = np.real(evals[:, None] - evals[None, :])
    N = len(evals)
    R = 0
    for a, s in zip(a_ops, spectra):
        a = a.toarray() if hasattr(a, "toarray") else np.asarray(a, np.complex128)
        C = np.array([[s(-W[n, m]) for m in range(N)] for n in range(N)], dtype=float)
        A = dag(evecs) @ a @ evecs
        Lk = C * A
        R = R + op2sop(A) @ (op2sop(Lk, "left") - op2sop(dag(Lk), "right"))
    return -1j * op2sop(np.diag(evals)) - R, evecs


class RedfieldSolver:
    """form (keyword only): "auto" runs evolve for d*d <= 1024 on the Liouville-space kernel with the
    dense Redfield tensor and larger d on the matrix-form kernel with the tensor's operator form
    (liouville.redfield_operators); "liouville" or "matrix" force one of them."""

    def __init__(self, H, c_ops=None, spectra=None, e_ops=None, *, form="auto"):
        if form not in ("auto", "liouville", "matrix"):
            raise ValueError("form must be 'auto', 'liouville' or 'matrix'")
        self.form = form
        self.H = H
        self.c_ops = c_ops
        self.R = None
        self.spectra = spectra
        self.evecs = None
        self.dim = H.shape[0]
        self.U = None
        self.G = None
        self.e_ops = e_ops

    def idm(self, sp=True):
        """vec of the identity operator (oqs.py:45-49), dense for either value of sp."""
        return np.eye(self.dim).reshape(-1)

    def configure(self, H, c_ops, e_ops):
        self.c_ops = c_ops
        self.e_ops = e_ops
        self.H = H

    def redfield_tensor(self, secular=False):
        if self.spectra is None:
            raise TypeError("Specify the bath spectral function.")
        R, evecs = redfield_tensor(self.H, self.c_ops, self.spectra, secular)
        self.R = R
        self.evecs = evecs
        return R, evecs

    def propagator(self, t, method='SOS'):
        """Liouville-space propagator U(t) = exp(R t) on the grid t, shape [N, N, len(t)]
        (oqs.py:160-209).  'SOS' / 'eseries': R = U1 diag(lambda) U1^-1 on the host, the expansion
        on the GPU spectral-sum kernel.  'EOM': the reference's RK4 of dU/dt = R U from U = 1 with
        dt = t[1] - t[0] (phys.expm, phys.py:2049-2091), every column of U one instance of the
        Liouville kernel; entry k is U after k steps.  Sets self.U and self.G = -i U."""
        if self.R is None:
            raise TypeError('Redfield tensor is not computed. Please call redfield_tensor()')
        t = np.asarray(t, np.float64).reshape(-1)
        R = np.asarray(self.R, np.complex128)
        N = R.shape[0]
        if method == 'EOM':
            print('Computing the propagator. Please make sure that the ground-state energy is 0.')
            U = np.zeros((N, N, len(t)), np.complex128)
            U[:, :, 0] = np.eye(N)
            if len(t) > 1:
                d = self.dim
                cols = np.eye(N, dtype=np.complex128).reshape(N, d, d)
                _, states, _ = propagate(R[None], cols, t[1] - t[0], len(t) - 1, None, ())
                U[:, :, 1:] = np.transpose(states.reshape(N, len(t) - 1, N), (2, 0, 1))
            self.U = U
        elif method in ['eseries', 'SOS']:
            evals, U1 = scipy.linalg.eig(R)
            U2 = scipy.linalg.inv(U1)
            self.U = spectral_sum(TIME, U1, evals, U2, t, layout="abt")
        else:
            raise ValueError("method must be 'EOM', 'SOS' or 'eseries'")
        self.G = -1j * self.U
        return self.U

    def expect(self, rho0, e_ops):
        """Observables from the stored propagator (oqs.py:211-242): rho0 and e_ops to the eigenbasis,
        out[k, j] = sum(vec(e_j) * U[:, :, k] vec(rho0)) with vec row-major, as the reference."""
        ev = self.evecs
        U = self.U
        rho0_eb = (dag(ev) @ np.asarray(rho0) @ ev).reshape(-1)
        e_eb = [dag(ev) @ np.asarray(e) @ ev for e in e_ops]
        rho = np.tensordot(U, rho0_eb, axes=([1], [0]))
        out = np.zeros((U.shape[-1], len(e_ops)), dtype=complex)
        for j, e in enumerate(e_eb):
            out[:, j] = e.reshape(-1).dot(rho)
        return out

    def correlation_2op_1t(self, rho0, a, b, tau):
        """<<I| a G(tau) b |rho0>> with Liouville-space operators a, b (oqs.py:244-268).  The
        reference's sparse column idm makes this product raise; the row-vector contraction it
        intends is computed here."""
        if self.G is None:
            self.propagator(tau)
        G = self.G
        idm = self.idm()
        rho0 = np.asarray(rho0)
        if rho0.ndim == 2:
            rho0 = rho0.reshape(-1)
        a = np.asarray(a.toarray() if hasattr(a, "toarray") else a)
        b = np.asarray(b.toarray() if hasattr(b, "toarray") else b)
        return idm @ (a @ np.tensordot(G, b @ rho0, axes=([1], [0])))

    def correlation_4op_3t(self, rho0, oplist, signature, tau):
        """<<I| A G(tau3) B G(tau2) C G(tau1) D |rho0>> as [tau3, tau2, tau1] (oqs.py:270-360), the
        operators promoted with signature characters from 'lr+-' (or 'c', 'a')."""
        from .superoperator import operator_to_superoperator
        if len(oplist) != 4:
            raise ValueError('Number of operators is not 4.')
        a, b, c, d = (operator_to_superoperator(op, sg) for op, sg in zip(oplist, signature))
        if self.G is None:
            self.propagator(tau)
        G = self.G
        idm = self.idm(sp=False)
        rho0 = np.asarray(rho0.toarray() if hasattr(rho0, "toarray") else rho0)
        rho = d @ rho0.reshape(-1)
        tmp = np.tensordot(G, rho, axes=((1), (0)))
        tmp = c @ tmp
        tmp = np.tensordot(G, tmp, axes=([1], [0]))
        tmp = np.tensordot(b, tmp, axes=([1], [0]))
        tmp = np.tensordot(G, tmp, axes=([1], [0]))
        return np.einsum('a, ab, bijk -> ijk', idm, a, tmp)

    def evolve(self, rho0, dt, Nt, evecs=None, e_ops=[], store_states=False, t0=0, nout=1):
        """RK4 on vec(rho) in the eigenbasis (oqs.py:57-81, 364-459): observables [Nt, n_e] in the
        eigenbasis (no t0 row), rholist back-transformed to the original basis."""
        rho0 = np.asarray(rho0, np.complex128)
        d = rho0.shape[-1]
        if self.form == "matrix" or (self.form == "auto" and d * d > 1024):
            if self.spectra is None:
                raise TypeError("Specify the bath spectral function.")
            Heff, lefts, rights, ev = redfield_operators(self.H, self.c_ops, self.spectra)
            self.evecs = ev
            e_eig = [dag(ev) @ np.asarray(e, np.complex128) @ ev for e in (e_ops or [])]
            obs, states, _ = run_matrix(Heff, lefts, rights, (), (dag(ev) @ rho0 @ ev)[None], dt, Nt, None, e_eig)
        else:
            if self.R is None:
                self.redfield_tensor()
            ev = self.evecs
            e_eig = [dag(ev) @ np.asarray(e, np.complex128) @ ev for e in (e_ops or [])]
            obs, states, _ = propagate(np.asarray(self.R, np.complex128)[None], (dag(ev) @ rho0 @ ev)[None], dt, Nt,
                                       None, e_eig)
        result = Result(dt=dt, Nt=Nt, rho0=rho0)
        result.observables = obs[0, 1:]
        result.rholist = [ev @ states[0, k] @ dag(ev) for k in range(states.shape[1])]
        return result


# ------------------------------------------------------------------ legacy single-exponential HEOM

AU2K = 315775.13   # K per hartree (pyqed/units.py:6)


def _dense(a):
    return np.asarray(a.toarray() if hasattr(a, "toarray") else a, np.complex128)


def legacy_run_params(temperature, cutoff, reorganization):
    """(gamma, D0) of _heom (oqs.py:1826-1845): gamma = cutoff, D0 = reorg gamma (coth(gamma/2T) - i),
    T in energy units; arrays broadcast."""
    T, gamma, lam = (np.asarray(v, np.float64) for v in (temperature, cutoff, reorganization))
    return gamma, lam * gamma * (1.0 / np.tanh(gamma / (2.0 * T)) - 1j)


def host_c_order(x, dtype=np.complex128):
    """A C-ordered copy of x.  The kernels index raw row-major buffers, and np.array of a broadcast
    or transposed view keeps the view's axis order (order='K'), which torch.as_tensor then keeps."""
    return np.array(x, dtype, order="C")


def legacy_heom_launch(mode, A, S, Sp, params, ados, nt, dt, e_ops=()):
    """Call pyqed_legacy_heom_run.  A, S, Sp: [n_inst, M, M]; params: [n_inst, 3]; ados:
    [n_inst, nado, M, M] (initial hierarchy).  Returns (final ados, observables [n_inst, nt, n_e])."""
    torch = _lib.require_gpu()
    L = _lib.lib()
    dev = torch.device("cuda", torch.cuda.current_device())

    def dev_c(x):
        return torch.as_tensor(host_c_order(x), device=dev)

    A_t, S_t = dev_c(A), dev_c(S)
    P_t = dev_c(Sp) if Sp is not None else None
    par = torch.as_tensor(host_c_order(params, np.float64), device=dev)
    ado = dev_c(ados).clone()
    n_inst, nado, M = ado.shape[0], ado.shape[1], ado.shape[2]
    if not all(t.is_contiguous() for t in (A_t, S_t, par, ado) + ((P_t,) if P_t is not None else ())):
        raise ValueError("legacy_heom_launch: device buffers must be contiguous")
    n_obs = len(e_ops)
    E = obs = None
    if n_obs:
        E = dev_c(np.array([_dense(e) for e in e_ops]))
        obs = torch.empty((n_inst, max(nt, 1), n_obs), dtype=torch.complex128, device=dev)
    _lib.check(L.pyqed_legacy_heom_run(mode, M, nado, n_inst, A_t.data_ptr(), S_t.data_ptr(), _lib.ptr(P_t),
                                       par.data_ptr(), ado.data_ptr(), int(nt), float(dt), n_obs, _lib.ptr(E),
                                       _lib.ptr(obs), _lib.stream_ptr(torch)))
    torch.cuda.synchronize()
    o = obs.cpu().numpy()[:, :nt] if n_obs else np.zeros((n_inst, nt, 0), complex)
    return ado.cpu().numpy(), o


class HEOMSolver:
    """Single-exponential high-temperature HEOM of pyqed's oqs.HEOMSolver (oqs.py:1332-1372) on the GPU.

    run(rho0, dt, nt, temperature, cutoff, reorganization, nado) -> observables (n_e, nt), the
    forward-Euler hierarchy of _heom (oqs.py:1808-1874) with temperature in energy units;
    propagator(dt, nt, temperature, cutoff, reorganization, nado) -> u (nado, d^2, d^2), the
    superoperator recursion of _heom_propagator (oqs.py:1876-1935) with temperature in kelvin.
    Both keep the reference's in-place sweep, truncation and printed diagnostics.  Documented
    differences: e_ops=None means no observables (the reference raises on len(None)); the whole
    hierarchy of one instance must fit one workgroup's LDS ((nado + 2) d^2 complex doubles for run,
    (nado + 3) d^4 for propagator, one block more when the block exceeds 256 elements), and larger
    cases raise."""

    def __init__(self, H=None, c_ops=None, e_ops=None):
        self.c_ops = c_ops
        self.e_ops = e_ops
        self.H = H

    def set_c_ops(self, c_ops):
        self.c_ops = c_ops

    def set_e_ops(self, e_ops):
        self.e_ops = e_ops

    def setH(self, H):
        self.H = H

    def configure(self, c_ops, e_ops):
        self.c_ops = c_ops
        self.e_ops = e_ops

    @staticmethod
    def _announce(T, gamma, reorg, amp, check_label):
        print('Temperature of the environment = {}'.format(T))
        print('{} = {}'.format(check_label, gamma / T))
        if gamma / T > 0.8:
            print('WARNING: High-Temperature Approximation may fail.')
        print('Reorganization energy = {}'.format(reorg))
        print('Amplitude of the fluctuations = {}'.format(amp))

    def _ops(self):
        return _dense(self.H), _dense(self.c_ops[0])

    def run(self, rho0, dt, nt, temperature, cutoff, reorganization, nado, *, verbose=True):
        H, S = self._ops()
        gamma, D0 = legacy_run_params(temperature, cutoff, reorganization)
        gamma, D0, T = float(gamma), complex(D0), float(temperature)
        if verbose:
            self._announce(T, gamma, reorganization, D0, 'Cutoff gamma/(kT)')
        d = H.shape[0]
        ados = np.zeros((1, nado, d, d), np.complex128)
        ados[0, 0] = _dense(rho0)
        e_ops = list(self.e_ops) if self.e_ops is not None else []
        _, obs = legacy_heom_launch(0, H[None], S[None], None, [[gamma, D0.real, D0.imag]], ados, nt, dt, e_ops)
        return obs[0].T.copy()

    def run_batch(self, rho0s, dt, nt, temperature, cutoff, reorganization, nado, return_ados=False):
        """Many hierarchies in one launch: rho0s [n, d, d]; temperature, cutoff and reorganization are
        scalars or length-n arrays.  Returns observables [n, n_e, nt] (and the final ADOs
        [n, nado, d, d] with return_ados)."""
        H, S = self._ops()
        rho0s = np.asarray(rho0s, np.complex128)
        n, d = rho0s.shape[0], H.shape[0]
        gamma, D0 = (np.broadcast_to(v, (n,)) for v in legacy_run_params(temperature, cutoff, reorganization))
        params = np.stack([gamma, D0.real, D0.imag], axis=1)
        ados = np.zeros((n, nado, d, d), np.complex128)
        ados[:, 0] = rho0s
        e_ops = list(self.e_ops) if self.e_ops is not None else []
        fin, obs = legacy_heom_launch(0, np.broadcast_to(H, (n, d, d)), np.broadcast_to(S, (n, d, d)), None, params,
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
        fin, _ = legacy_heom_launch(1, op2sop(H)[None], op2sop(S)[None], op2sop(S, "anticommutator")[None],
                                    [[gamma, a, b]], u, nt, dt)
        return fin[0]

    def correlation_2op_1t(self, rho0, a_op, b_op, dt, Nt, output='cor.dat'):
        """<A(t)B> of the Lindblad equation with H and c_ops (oqs.py:1381-1403 calls the Lindblad
        regression routine of oqs.py:717-791): LindbladSolver.correlation_2op_1t."""
        return LindbladSolver(self.H, self.c_ops).correlation_2op_1t(rho0, a_op, b_op, dt, Nt, output)

    def correlation_3op_2t(self, rho0, ops, dt, Nt, Ntau):
        """<A(t)B(t+tau)C(t)> of the Lindblad equation with H and c_ops (oqs.py:1405-1435, the same
        regression as LindbladSolver.correlation_3op_2t)."""
        return LindbladSolver(self.H, self.c_ops).correlation_3op_2t(rho0, ops, dt, Nt, Ntau)
