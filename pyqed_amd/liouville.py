"""Liouville-space propagation on the GPU: d vec(rho)/dt = (A_0 + sum_i f_i A_i) vec(rho).

Builds row-major generators (vec(A X B) = kron(A, B^T) vec(X), pyqed/superoperator.py:200-246)
and calls pyqed_liouville_run / pyqed_liouville_adjoint of libpyqed_amd.
"""
import numpy as np

from . import _lib


def dag(a):
    return np.conj(np.asarray(a)).T


def op2sop(a, kind="commutator"):
    a = np.asarray(a, np.complex128)
    I = np.eye(a.shape[-1])
    if kind == "commutator":
        return np.kron(a, I) - np.kron(I, a.T)
    if kind == "left":
        return np.kron(a, I)
    if kind == "right":
        return np.kron(I, a.T)
    if kind == "anticommutator":
        return np.kron(a, I) + np.kron(I, a.T)
    raise ValueError(f"Error: superoperator {kind} does not exist.")


def lindblad_generator(H, c_ops):
    """-i[H, .] + sum_c (c . c^dag - 1/2 {c^dag c, .}) as an N x N matrix (N = d^2)."""
    L = -1j * op2sop(H)
    for c in c_ops or []:
        c = np.asarray(c, np.complex128)
        L = L + np.kron(c, c.conj()) - 0.5 * op2sop(dag(c) @ c, "anticommutator")
    return L


def propagate(gens, rho0s, dt, nsteps, field=None, e_ops=(), save_every=1, keep_states=True):
    """Run pyqed_liouville_run.  gens: [nterm, N, N]; rho0s: [n_inst, d, d]; field: [n_inst, nsteps,
    nterm-1] or None.  Returns (observables [n_inst, nsteps+1, n_obs], states [n_inst, nsave, d, d],
    final [n_inst, d, d]) as host arrays."""
    torch = _lib.require_gpu()
    L = _lib.lib()
    gens = np.ascontiguousarray(np.asarray(gens, np.complex128))
    nterm, N, _ = gens.shape
    rho0s = np.asarray(rho0s, np.complex128)
    n_inst, d = rho0s.shape[0], rho0s.shape[1]
    dev = torch.device("cuda", torch.cuda.current_device())
    A = torch.as_tensor(gens, device=dev)
    rho = torch.as_tensor(np.ascontiguousarray(rho0s.reshape(n_inst, N)), device=dev).clone()
    f = None
    if nterm > 1:
        f = torch.as_tensor(np.ascontiguousarray(np.asarray(field, np.float64).reshape(n_inst, nsteps, nterm - 1)),
                            device=dev)
    n_obs = len(e_ops)
    obs = obs_out = None
    if n_obs:
        obs = torch.as_tensor(np.ascontiguousarray(np.array([np.asarray(e, np.complex128).T.reshape(-1)
                                                              for e in e_ops])), device=dev)
        obs_out = torch.empty((n_inst, nsteps + 1, n_obs), dtype=torch.complex128, device=dev)
    nsave = nsteps // save_every if keep_states else 0
    rho_out = torch.empty((n_inst, max(nsave, 1), N), dtype=torch.complex128, device=dev) if nsave else None
    _lib.check(L.pyqed_liouville_run(N, nterm, A.data_ptr(), n_inst, rho.data_ptr(), _lib.ptr(f), nsteps, float(dt),
                                     n_obs, _lib.ptr(obs), _lib.ptr(obs_out), _lib.ptr(rho_out), save_every,
                                     _lib.stream_ptr(torch)))
    torch.cuda.synchronize()
    o = obs_out.cpu().numpy() if n_obs else np.zeros((n_inst, nsteps + 1, 0), complex)
    st = rho_out.cpu().numpy().reshape(n_inst, -1, d, d)[:, :nsave] if nsave else np.zeros((n_inst, 0, d, d), complex)
    return o, st, rho.cpu().numpy().reshape(n_inst, d, d)


def propagate_matrix(H0, c_ops, drives, rho0s, dt, nsteps, field=None, e_ops=(), save_every=1, keep_states=True):
    """Run pyqed_lindblad_run: the Lindblad equation in matrix form (any d), with H(t) = H0 -
    sum_i f_i H_i for drives H_i and field [n_inst, nsteps, len(drives)].  Returns the same
    (observables [n_inst, nsteps+1, n_obs], states [n_inst, nsave, d, d], final [n_inst, d, d]) as
    propagate."""
    H0 = np.asarray(H0, np.complex128)
    cs = [np.asarray(c, np.complex128) for c in (c_ops or [])]
    Heff = H0 - 0.5j * sum((dag(c) @ c for c in cs), np.zeros_like(H0))
    return run_matrix(Heff, cs, [dag(c) for c in cs], drives, rho0s, dt, nsteps, field, e_ops, save_every, keep_states)


def redfield_operators(H, a_ops, spectra):
    """The Bloch-Redfield generator of redfield_tensor (oqs.py:519-570) in operator form, in the
    eigenbasis of H: d rho/dt = -i (Heff rho - rho Heff^dag) + sum_a (A_a rho L_a^dag + L_a rho A_a),
    A_a = V^dag a V, L_a = C_a * A_a (elementwise, C_a[n, m] = S_a(-(E_n - E_m))), Heff = diag(E) -
    i sum_a A_a L_a.  Returns (Heff, lefts, rights, evecs): the pairs (lefts[k], rights[k]) of the
    terms lefts[k] rho rights[k]."""
    H = H.toarray() if hasattr(H, "toarray") else np.asarray(H)
    evals, evecs = np.linalg.eigh(H)
    W = np.real(evals[:, None] - evals[None, :])
    N = len(evals)
    Heff = np.diag(evals).astype(np.complex128)
    lefts, rights = [], []
    for a, sp in zip(a_ops, spectra):
        a = a.toarray() if hasattr(a, "toarray") else np.asarray(a, np.complex128)
        if not np.allclose(a, dag(a)):
            raise TypeError("Operators in a_ops must be Hermitian.")
        C = np.array([[sp(-W[n, m]) for m in range(N)] for n in range(N)], dtype=float)
        A = dag(evecs) @ a @ evecs
        Lk = C * A
        Heff = Heff - 1j * (A @ Lk)
        lefts += [A, Lk]
        rights += [dag(Lk), A]
    return Heff, lefts, rights, evecs


def run_matrix(Heff, lefts, rights, drives, rho0s, dt, nsteps, field=None, e_ops=(), save_every=1, keep_states=True):
    """pyqed_lindblad_run for d rho/dt = -i (Hs rho - rho Hs^dag) + sum_k lefts[k] rho rights[k],
    Hs = Heff - sum_i f_i drives[i]."""
    torch = _lib.require_gpu()
    L = _lib.lib()
    Heff = np.asarray(Heff, np.complex128)
    d = Heff.shape[0]
    ls = [np.asarray(x, np.complex128) for x in lefts]
    rs = [np.asarray(x, np.complex128) for x in rights]
    hs = [np.asarray(h, np.complex128) for h in (drives or [])]
    rho0s = np.asarray(rho0s, np.complex128)
    n_inst = rho0s.shape[0]
    dev = torch.device("cuda", torch.cuda.current_device())

    def dev_arr(a):
        return torch.as_tensor(np.ascontiguousarray(np.asarray(a, np.complex128)), device=dev)

    He, HeD = dev_arr(Heff), dev_arr(dag(Heff))
    Lm = dev_arr(np.array(ls)) if ls else None
    Rm = dev_arr(np.array(rs)) if rs else None
    Hd = dev_arr(np.array(hs)) if hs else None
    HdD = dev_arr(np.array([dag(h) for h in hs])) if hs else None
    f = None
    if hs:
        f = torch.as_tensor(np.ascontiguousarray(np.asarray(field, np.float64).reshape(n_inst, nsteps, len(hs))),
                            device=dev)
    rho = dev_arr(rho0s.reshape(n_inst, d, d)).clone()
    n_obs = len(e_ops)
    obs = obs_out = None
    if n_obs:
        obs = dev_arr(np.array([np.asarray(e, np.complex128).T.reshape(-1) for e in e_ops]))
        obs_out = torch.empty((n_inst, nsteps + 1, n_obs), dtype=torch.complex128, device=dev)
    nsave = nsteps // save_every if keep_states else 0
    rho_out = torch.empty((n_inst, max(nsave, 1), d * d), dtype=torch.complex128, device=dev) if nsave else None
    wb = _lib.C.c_int64()
    _lib.check(L.pyqed_lindblad_work_bytes(d, len(ls), len(hs), n_inst, _lib.C.byref(wb)))
    work = torch.empty(max(wb.value, 16), dtype=torch.uint8, device=dev)
    _lib.check(L.pyqed_lindblad_run(d, len(ls), He.data_ptr(), HeD.data_ptr(), _lib.ptr(Lm), _lib.ptr(Rm), len(hs),
                                    _lib.ptr(Hd), _lib.ptr(HdD), n_inst, rho.data_ptr(), _lib.ptr(f), nsteps, float(dt),
                                    n_obs, _lib.ptr(obs), _lib.ptr(obs_out), _lib.ptr(rho_out), save_every,
                                    work.data_ptr(), work.numel(), _lib.stream_ptr(torch)))
    torch.cuda.synchronize()
    o = obs_out.cpu().numpy() if n_obs else np.zeros((n_inst, nsteps + 1, 0), complex)
    st = rho_out.cpu().numpy().reshape(n_inst, -1, d, d)[:, :nsave] if nsave else np.zeros((n_inst, 0, d, d), complex)
    return o, st, rho.cpu().numpy().reshape(n_inst, d, d)
