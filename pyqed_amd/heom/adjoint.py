"""Gradients of HEOM propagation (DEOMSolver.run, pyqed/heom/deom.py:725-766, 1072-1114) by a streamed
discrete adjoint on the GPU.  The reference has no gradients.

Forward: the RK4 step of deom.py:725-766 with H(t) = H + mu f(t) (generate_time, deom.py:676-687),
the field tabulated per step at t, t + dt/2 and t + dt (columns 0, 1, 2 of the field table).

Adjoint of one step.  For y' = y + dt/6 (k1 + 2 k2 + 2 k3 + k4), k_s = L_s x_s with stage inputs
x1 = y, x2 = y + dt/2 k1, x3 = y + dt/2 k2, x4 = y + dt k3 and L_1 = L(t), L_2 = L(t + dt/2),
L_3 = L(t + dt), the cotangent u of y' maps to the cotangent of y by
    K1 = L_3^H u,  K2 = L_2^H (u + dt/2 K1),  K3 = L_2^H (u + dt/2 K2),  K4 = L_1^H (u + dt K3),
    lambda = u + dt/6 (K1 + 2 K2 + 2 K3 + K4),
i.e. one RK4 step of the adjoint operator with the stage times reversed.  Every L^H application is the
library's FP64 right-hand-side kernel on a PYQED_PLAN_ADJOINT plan (the same hierarchy tables with
-H^H, Q^H and the neighbour coefficients swapped and conjugated): the whole hierarchy streams through
HBM, as in the forward run.

Gradients.  The cotangents of the right-hand sides are kbar_1 = dt/6 (u + dt K3), kbar_2 = dt/3 (u +
dt/2 K2), kbar_3 = dt/3 (u + dt/2 K1), kbar_4 = dt/6 u, and with dL/dtheta X = -i [dH/dtheta, X]
    dJ/dtheta = Re(-i Tr(dH/dtheta G)),   G = sum_steps sum_s sum_ADOs [x_s, kbar_s^H],
accumulated per step and field column by pyqed_heom_adjoint_contract.  J is real with
dJ = Re sum conj(rhoT_bar) d rho(T) (rhoT_bar = dJ/dRe rho + i dJ/dIm rho; for J = Re Tr(O rho(T)),
rhoT_bar = O^H).

Checkpoints: the forward run keeps the state every `ckpt_every` steps; the backward sweep recomputes one
segment of states at a time, so memory is (nt / ckpt_every + ckpt_every) states.
"""
import ctypes as C

import numpy as np

from .. import _lib


class HeomAdjoint:
    """Forward run and streamed adjoint of one or more HEOM instances (HeomEngine instance dicts)."""

    def __init__(self, instances, dt, nt, *, ckpt_every=50):
        from .engine import HeomEngine
        self.fwd = HeomEngine(instances, hermitian=False, packed=False)
        self.adj = HeomEngine(instances, hermitian=False, packed=False, adjoint=True)
        for e in (self.fwd, self.adj):
            e.autotune_cache = False
        self.torch = self.fwd.torch
        self.L = self.fwd.L
        self.dt = float(dt)
        self.nt = int(nt)
        self.ckpt_every = max(1, int(ckpt_every))
        self.n = self.fwd.n
        wb = C.c_int64()
        _lib.check(self.L.pyqed_heom_adjoint_work_bytes(self.adj.plan, C.byref(wb)))
        self._work = self.torch.empty(max(wb.value, 16), dtype=self.torch.uint8, device=self.fwd.device)
        self._wb = wb.value

    # ------------------------------------------------------------------ helpers
    def _field(self, field):
        if field is None:
            if self.fwd.driven:
                raise ValueError("an instance has a system or coupling dipole: pass the field table [nt, 6]")
            return None
        f = np.ascontiguousarray(np.asarray(field, np.float64))
        if f.ndim != 2 or f.shape[1] != 6 or f.shape[0] < self.nt:
            raise ValueError(f"field table must be [>= {self.nt}, 6], got {f.shape}")
        return self.torch.as_tensor(f, device=self.fwd.device)

    def _rhs(self, eng, x, out, field, step, col):
        _lib.check(self.L.pyqed_heom_rhs(eng.plan, x.data_ptr(), out.data_ptr(),
                                         None if field is None else field.data_ptr(), int(step), int(col),
                                         _lib.stream_ptr(self.torch)))

    def _comb(self, eng, x, out, base, alpha, acc, beta, init, field, step, col):
        """out = base + alpha L x (out, base optional) and acc = acc + beta L x (or beta L x when init),
        in the right-hand-side launch."""
        def ptr(t):
            return None if t is None else t.data_ptr()
        _lib.check(self.L.pyqed_heom_rhs_combine(eng.plan, x.data_ptr(), ptr(out), ptr(base), float(alpha), ptr(acc),
                                                 float(beta), 1 if init else 0,
                                                 None if field is None else field.data_ptr(), int(step), int(col),
                                                 _lib.stream_ptr(self.torch)))

    def _to_work(self, k, m):
        """Caller-basis matrix -> working basis of instance k (the engine may propagate in a rotated basis)."""
        U = self.fwd.prep[k]["U"]
        m = np.asarray(m, np.complex128)
        return m if U is None else U.conj().T @ m @ U

    def _from_work(self, k, m):
        U = self.fwd.prep[k]["U"]
        return m if U is None else U @ m @ U.conj().T

    # ------------------------------------------------------------------ forward
    def forward(self, field=None):
        """Final reduced states rho(T) (host, caller's basis) of every instance."""
        f = self._field(field)
        self.fwd.reset()
        if self.nt:
            self.fwd.advance(self.nt, self.dt, "fp64", field=f)
        self.torch.cuda.synchronize()
        return self.fwd.reduced_states()

    # ------------------------------------------------------------------ gradient
    def gradient(self, rhoT_bar, field=None):
        """Backward sweep.  rhoT_bar: list (one per instance) of d x d cotangents of rho(T).

        Returns a dict with, per instance (lists):
          rho      rho(T)
          J        Re sum conj(rhoT_bar) rho(T)
          gH       cotangent of H: dJ = Re sum conj(gH) dH (so dJ/dtheta = Re sum conj(gH) dH/dtheta)
          grho0    cotangent of rho(0)
          G        [nt, 3, d, d] per-step contraction matrices (working basis), for field gradients
          gfield   [nt, 3] dJ/df at t, t + dt/2, t + dt of every step when the instance has a system dipole
        """
        torch = self.torch
        fwd, adj = self.fwd, self.adj
        dt, nt, ck = self.dt, self.nt, self.ckpt_every
        f = self._field(field)
        dev = fwd.device
        n = fwd.n_elems
        z = lambda: torch.zeros(n, dtype=torch.complex128, device=dev)  # noqa: E731
        # forward with checkpoints
        fwd.reset()
        ckpts = []
        for c0 in range(0, nt, ck):
            ckpts.append(fwd.y.clone())
            fwd.advance(min(ck, nt - c0), dt, "fp64", field=f)
        torch.cuda.synchronize()
        rhoT = fwd.reduced_states()
        # terminal cotangent: ADO 0 of every instance
        lam = z()
        J = []
        for k in range(self.n):
            bar = np.asarray(rhoT_bar[k], np.complex128)
            J.append(float(np.real(np.sum(np.conj(bar) * rhoT[k]))))
            o = int(fwd.state_offset[k])
            d = fwd.prep[k]["d"]
            lam[o:o + d * d] = torch.as_tensor(self._to_work(k, bar).reshape(-1), device=dev)
        G = torch.zeros((max(nt, 1), 3, fwd.traj_stride), dtype=torch.complex128, device=dev)
        Ys = torch.zeros((min(ck, max(nt, 1)), n), dtype=torch.complex128, device=dev)
        x2, x3, x4 = z(), z(), z()
        X2, X3, X4, acc = z(), z(), z(), z()
        w = (C.c_double * 4)(dt / 6, dt / 3, dt / 3, dt / 6)
        xp = (C.c_void_p * 4)()
        up = (C.c_void_p * 4)()
        stream = _lib.stream_ptr(torch)
        for seg in reversed(range(len(ckpts))):
            c0 = seg * ck
            m = min(ck, nt - c0)
            fwd.y.copy_(ckpts[seg])
            fwd.step = c0
            for j in range(m):
                Ys[j].copy_(fwd.y)
                if j < m - 1:
                    fwd.advance(1, dt, "fp64", field=f)
            for j in reversed(range(m)):
                s = c0 + j
                y = Ys[j]
                # forward stage inputs of step s: x2 = y + dt/2 k1, x3 = y + dt/2 k2, x4 = y + dt k3
                self._comb(fwd, y, x2, y, dt / 2, None, 0.0, False, f, s, 0)
                self._comb(fwd, x2, x3, y, dt / 2, None, 0.0, False, f, s, 1)
                self._comb(fwd, x3, x4, y, dt, None, 0.0, False, f, s, 1)
                # adjoint stages, times reversed: K1 = L^H u, X2 = u + dt/2 K1, ..., acc = K1 + 2 K2 + 2 K3 + K4
                u = lam
                self._comb(adj, u, X2, u, dt / 2, acc, 1.0, True, f, s, 2)
                self._comb(adj, X2, X3, u, dt / 2, acc, 2.0, False, f, s, 1)
                self._comb(adj, X3, X4, u, dt, acc, 2.0, False, f, s, 1)
                self._comb(adj, X4, None, None, 0.0, acc, 1.0, False, f, s, 0)
                # kbar_1 = dt/6 X4 pairs with x1 = y, kbar_2 = dt/3 X3 with x2, kbar_3 = dt/3 X2 with x3,
                # kbar_4 = dt/6 u with x4
                for i, (a, b) in enumerate(((y, X4), (x2, X3), (x3, X2), (x4, u))):
                    xp[i], up[i] = a.data_ptr(), b.data_ptr()
                _lib.check(self.L.pyqed_heom_adjoint_contract(adj.plan, xp, up, w, G[s].data_ptr(),
                                                              self._work.data_ptr(), self._wb, stream))
                lam.add_(acc, alpha=dt / 6)
        torch.cuda.synchronize()
        Gh = G.cpu().numpy()
        lam_h = lam.cpu().numpy()
        out = dict(rho=rhoT, J=J, gH=[], grho0=[], G=[], gfield=[])
        for k in range(self.n):
            p = fwd.prep[k]
            d = p["d"]
            to = int(fwd.traj_offset[k])
            Gk = Gh[:nt, :, to:to + d * d].reshape(nt, 3, d, d)
            Gtot = self._from_work(k, Gk.sum(axis=(0, 1)))
            out["gH"].append(1j * Gtot.conj().T)
            o = int(fwd.state_offset[k])
            out["grho0"].append(self._from_work(k, lam_h[o:o + d * d].reshape(d, d)))
            out["G"].append(Gk)
            mu = p["sdip"]
            out["gfield"].append(None if mu is None else np.real(-1j * np.einsum("ab,tcba->tc", mu, Gk)))
        return out

    def close(self):
        self.fwd.close()
        self.adj.close()


def piecewise_field(amps, steps_per_seg, nt=None):
    """Field table [nt, 6] of a piecewise-constant system field: step k uses amps[k // steps_per_seg] at
    all three of its evaluation points (columns 0-2); no coupling field (columns 3-5)."""
    amps = np.asarray(amps, np.float64).reshape(-1)
    per_step = np.repeat(amps, int(steps_per_seg))
    if nt is not None:
        per_step = per_step[:nt]
    tab = np.zeros((per_step.size, 6))
    tab[:, :3] = per_step[:, None]
    return tab


class HeomControl:
    """One HEOM instance with H(theta) = H0 + sum_j theta_j P_j and a piecewise-constant field entering
    H(t) = H(theta) + mu E(t) (deom.py:676-687), E held over steps_per_seg steps per amplitude.

    final_state(theta, amps) -> rho(T); gradient(theta, amps, rhoT_bar) -> (dJ/dtheta, dJ/damps,
    dJ/drho0) for J with cotangent rhoT_bar, from one forward run and one backward sweep."""

    def __init__(self, H0, param_ops, mu, Q, bath, L, rho0, dt, steps_per_seg, nseg, *, ckpt_every=50):
        self.H0 = np.asarray(H0, np.complex128)
        self.P = [np.asarray(p, np.complex128) for p in param_ops]
        self.mu = np.asarray(mu, np.complex128)
        self.Q = np.asarray(Q, np.complex128)
        self.bath = bath
        self.Ldepth = int(L)
        self.rho0 = np.asarray(rho0, np.complex128)
        self.dt = float(dt)
        self.sps = int(steps_per_seg)
        self.nseg = int(nseg)
        self.nt = self.sps * self.nseg
        self.ckpt_every = int(ckpt_every)

    def _problem(self, theta):
        H = self.H0 + sum(float(t) * p for t, p in zip(np.asarray(theta, float).reshape(-1), self.P))
        inst = dict(H=H, Q=self.Q, bath=self.bath, L=self.Ldepth, rho0=self.rho0, sdip=self.mu)
        return HeomAdjoint([inst], self.dt, self.nt, ckpt_every=self.ckpt_every)

    def final_state(self, theta, amps):
        prob = self._problem(theta)
        try:
            return prob.forward(piecewise_field(amps, self.sps))[0]
        finally:
            prob.close()

    def gradient(self, theta, amps, rhoT_bar):
        prob = self._problem(theta)
        try:
            r = prob.gradient([rhoT_bar], piecewise_field(amps, self.sps))
        finally:
            prob.close()
        gH = r["gH"][0]
        g_theta = np.array([np.real(np.sum(np.conj(gH) * p)) for p in self.P])
        g_amps = r["gfield"][0].sum(axis=1).reshape(self.nseg, self.sps).sum(axis=1)
        return g_theta, g_amps, r["grho0"][0], r


def _autograd_function():
    import torch

    class HeomFieldFunction(torch.autograd.Function):
        """rho(T) (real view [d, d, 2]) of a HeomControl problem, differentiable with respect to theta [P]
        and the field amplitudes [nseg] (float64 tensors); the backward pass is the streamed adjoint."""

        @staticmethod
        def forward(ctx, theta, amps, problem):
            ctx.problem = problem
            ctx.save_for_backward(theta, amps)
            rT = problem.final_state(theta.detach().cpu().numpy(), amps.detach().cpu().numpy())
            return torch.view_as_real(torch.as_tensor(rT))

        @staticmethod
        def backward(ctx, grad):
            theta, amps = ctx.saved_tensors
            g = grad.detach().cpu().numpy()
            bar = g[..., 0] + 1j * g[..., 1]
            gt, ga, _, _ = ctx.problem.gradient(theta.detach().cpu().numpy(), amps.detach().cpu().numpy(), bar)
            return torch.as_tensor(gt, dtype=theta.dtype), torch.as_tensor(ga, dtype=amps.dtype), None

    return HeomFieldFunction


def final_state(theta, amps, problem):
    """Differentiable rho(T) of a HeomControl problem as a real view [d, d, 2]."""
    return _autograd_function().apply(theta, amps, problem)
