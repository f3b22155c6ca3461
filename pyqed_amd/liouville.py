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
