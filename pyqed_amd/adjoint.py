"""Gradients of Lindblad propagation with respect to Hamiltonian parameters and piecewise-constant
field amplitudes, through the adjoint (backward-in-time) kernel pyqed_liouville_adjoint.

The reference has no gradients; this module adds them for the Lindblad path
(pyqed/oqs.py:1699-1806 semantics: H(t) = H0 - sum_f E_f(t) mu_f, field held over each RK4 step).

Generator of step k:  A_k = A_0(theta) + sum_f E_f[seg(k)] B_f,
    A_0(theta) = A_base + sum_j theta_j G_j       (parameters entering linearly, e.g. energies)
    seg(k) = k // steps_per_seg                   (step k uses segment floor(k / steps_per_seg))
dJ/dtheta_j = Re sum conj(gA_0) * G_j and dJ/dE_f[s] = sum_{k in s} dJ/dE_f(k), where gA_0 and the
per-step field gradients come from one backward sweep of the kernel (forward states checkpointed
every `ckpt_every` steps and recomputed per segment).
"""
import numpy as np

from . import _lib
from .liouville import lindblad_generator, op2sop


class LindbladAdjoint:
    """Propagator + adjoint for A_0(theta) + sum_f E_f B_f on the GPU."""

    def __init__(self, base_gen, param_gens, field_gens, dt, steps_per_seg, nseg, ckpt_every=50):
        self.base = np.asarray(base_gen, np.complex128)
        self.G = np.asarray(param_gens, np.complex128).reshape(-1, *self.base.shape)
        self.B = np.asarray(field_gens, np.complex128).reshape(-1, *self.base.shape)
        self.N = self.base.shape[0]
        self.d = int(round(np.sqrt(self.N)))
        # a d x d density matrix when N = d^2, otherwise a plain state vector (e.g. a whole hierarchy)
        self.shape = (self.d, self.d) if self.d * self.d == self.N else (self.N,)
        self.dt = float(dt)
        self.sps = int(steps_per_seg)
        self.nseg = int(nseg)
        self.nsteps = self.sps * self.nseg
        self.ckpt_every = int(ckpt_every)

    def _gens(self, theta):
        A0 = self.base + np.tensordot(np.asarray(theta, float), self.G, axes=1) if len(self.G) else self.base
        return np.ascontiguousarray(np.concatenate([A0[None], self.B], axis=0))

    def _field(self, amps):
        amps = np.asarray(amps, np.float64).reshape(len(self.B), self.nseg)
        return np.ascontiguousarray(np.repeat(amps, self.sps, axis=1).T)  # [nsteps, F]

    def forward(self, theta, amps, rho0):
        """rho(T) as a host complex array [d, d]."""
        torch = _lib.require_gpu()
        L = _lib.lib()
        dev = torch.device("cuda", torch.cuda.current_device())
        A = torch.as_tensor(self._gens(theta), device=dev)
        rho = torch.as_tensor(np.asarray(rho0, np.complex128).reshape(1, -1).copy(), device=dev)
        f = torch.as_tensor(self._field(amps)[None], device=dev) if len(self.B) else None
        _lib.check(L.pyqed_liouville_run(self.N, 1 + len(self.B), A.data_ptr(), 1, rho.data_ptr(), _lib.ptr(f),
                                         self.nsteps, self.dt, 0, None, None, None, 1, _lib.stream_ptr(torch)))
        torch.cuda.synchronize()
        return rho.cpu().numpy().reshape(self.shape)

    def gradient(self, theta, amps, rho0, rhoT_bar):
        """(dJ/dtheta [P], dJ/damps [F, nseg], dJ/drho0) for rhoT_bar = dJ/dRe rho(T) + i dJ/dIm rho(T)."""
        torch = _lib.require_gpu()
        L = _lib.lib()
        dev = torch.device("cuda", torch.cuda.current_device())
        nterm = 1 + len(self.B)
        A = torch.as_tensor(self._gens(theta), device=dev)
        rho = torch.as_tensor(np.asarray(rho0, np.complex128).reshape(1, -1).copy(), device=dev)
        f = torch.as_tensor(self._field(amps)[None], device=dev) if len(self.B) else None
        bar = torch.as_tensor(np.asarray(rhoT_bar, np.complex128).reshape(1, -1).copy(), device=dev)
        gf = torch.zeros((1, self.nsteps, max(nterm - 1, 1)), dtype=torch.float64, device=dev)
        gA = torch.zeros((1, nterm, self.N, self.N), dtype=torch.complex128, device=dev)
        r0b = torch.zeros((1, self.N), dtype=torch.complex128, device=dev)
        nck = (self.nsteps + self.ckpt_every - 1) // self.ckpt_every + 1
        ck_bytes = (nck + self.ckpt_every) * self.N * 16
        ck = torch.empty(ck_bytes, dtype=torch.uint8, device=dev)
        _lib.check(L.pyqed_liouville_adjoint(self.N, nterm, A.data_ptr(), 1, rho.data_ptr(), _lib.ptr(f), self.nsteps,
                                             self.dt, self.ckpt_every, bar.data_ptr(),
                                             gf.data_ptr() if nterm > 1 else None, gA.data_ptr(), r0b.data_ptr(),
                                             ck.data_ptr(), ck_bytes, _lib.stream_ptr(torch)))
        torch.cuda.synchronize()
        gA0 = gA[0, 0].cpu().numpy()
        g_theta = np.array([np.real(np.sum(np.conj(gA0) * Gj)) for Gj in self.G])
        g_steps = gf[0].cpu().numpy()[:, :nterm - 1]
        g_amps = g_steps.reshape(self.nseg, self.sps, -1).sum(axis=1).T
        return g_theta, g_amps, r0b[0].cpu().numpy().reshape(self.shape)


def lambda_system(eps, mu, c_ops, dt, steps_per_seg, nseg, ckpt_every=50):
    """H(t) = diag(eps) - E(t) mu with Lindblad c_ops; theta = eps (level energies), one field."""
    d = len(eps)
    base = lindblad_generator(np.zeros((d, d)), c_ops)
    G = []
    for j in range(d):
        P = np.zeros((d, d), complex)
        P[j, j] = 1.0
        G.append(-1j * op2sop(P))
    B = [1j * op2sop(np.asarray(mu, np.complex128))]
    return LindbladAdjoint(base, G, B, dt, steps_per_seg, nseg, ckpt_every)


def heom_problem(H, Q, bath, mode, L, dt, steps_per_seg, nseg, param_ops=(), field_ops=(), ckpt_every=50):
    """HEOM gradients for small hierarchies (nmax*d^2 <= 512): the dense HEOM generator (sparse
    assembly from the library's tables) propagated by the Liouville-space kernels.
    H(t) = H + sum_j theta_j P_j + sum_f E_f(t) M_f (the HEOM sign, deom.py:676-687); theta
    multiplies param_ops P_j, the field amplitudes multiply field_ops M_f.  The state is the whole
    hierarchy; the reduced density matrix is its first d*d block."""
    from .heom.generator import heom_generator
    H = np.asarray(H, np.complex128)
    d = H.shape[0]
    base = heom_generator(H, Q, bath, mode, L, sparse=False)
    nmax = base.shape[0] // (d * d)
    if base.shape[0] > 512:
        raise ValueError("heom_problem handles nmax*d^2 <= 512 (the Liouville adjoint kernel's LDS limit)")
    eye = np.eye(nmax)
    G = [np.kron(eye, -1j * op2sop(np.asarray(P, np.complex128))) for P in param_ops]
    B = [np.kron(eye, -1j * op2sop(np.asarray(M, np.complex128))) for M in field_ops]
    prob = LindbladAdjoint(base, np.array(G) if G else np.zeros((0,) + base.shape),
                           np.array(B) if B else np.zeros((0,) + base.shape), dt, steps_per_seg, nseg, ckpt_every)
    prob.hier_d = d
    return prob


def embed_rho0(prob, rho0):
    """Hierarchy initial state (vector [nmax*d*d]) with rho0 in ADO 0 (for heom_problem)."""
    v = np.zeros(prob.N, np.complex128)
    d = prob.hier_d
    v[:d * d] = np.asarray(rho0, np.complex128).reshape(-1)
    return v.reshape(prob.shape)


def _autograd_function():
    import torch

    class LindbladFieldFunction(torch.autograd.Function):
        """rho(T) (as a real view [d, d, 2]) of a LindbladAdjoint problem, differentiable with respect
        to theta [P] and field amplitudes [F, nseg] (float64 tensors)."""

        @staticmethod
        def forward(ctx, theta, amps, problem, rho0):
            ctx.problem = problem
            ctx.rho0 = rho0
            ctx.save_for_backward(theta, amps)
            rT = problem.forward(theta.detach().cpu().numpy(), amps.detach().cpu().numpy(), rho0)
            return torch.view_as_real(torch.as_tensor(rT))

        @staticmethod
        def backward(ctx, grad):
            theta, amps = ctx.saved_tensors
            g = grad.detach().cpu().numpy()
            bar = g[..., 0] + 1j * g[..., 1]
            gt, ga, _ = ctx.problem.gradient(theta.detach().cpu().numpy(), amps.detach().cpu().numpy(), ctx.rho0, bar)
            return (torch.as_tensor(gt, dtype=theta.dtype), torch.as_tensor(ga.reshape(amps.shape), dtype=amps.dtype),
                    None, None)

    return LindbladFieldFunction


def final_state(theta, amps, problem, rho0):
    """Differentiable rho(T) as a real view [d, d, 2]; theta, amps are float64 torch tensors."""
    return _autograd_function().apply(theta, amps, problem, np.asarray(rho0, np.complex128))
