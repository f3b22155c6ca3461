"""FP64 autograd reference for HEOM gradients -- TEST INFRASTRUCTURE ONLY.

The HEOM right-hand side of pyqed/heom/deom.py:641-673 and the RK4 step of deom.py:725-766, written
with torch complex128 operations on the CPU so that torch.autograd supplies the gradients the
streamed adjoint is checked against.  H(t) = H0 + sum_j theta_j P_j + mu E(t) (generate_time,
deom.py:676-687) with E piecewise constant: step k uses amps[k // steps_per_seg] at all of its stage
times.  Tables come from oracle.heom.Hierarchy."""
import numpy as np
import torch

from .heom import Hierarchy


def _rhs(rho, H, Q, hier, bath, mode, minus_sel, plus_sel):
    etal, etar, etaa, expn = bath
    keys = torch.as_tensor(hier.keys)
    damp = (keys.to(torch.complex128) * expn[None, :]).sum(1)
    out = -damp[:, None, None] * rho - 1j * (torch.matmul(H, rho) - torch.matmul(rho, H))
    sq = torch.sqrt(etaa)
    for k in range(hier.K):
        q = Q[int(mode[k])]
        sel, nb, n = minus_sel[k]
        if sel.numel():
            p = rho[nb]
            c = (-1j * torch.sqrt(n) / sq[k])[:, None, None]
            out = out.index_add(0, sel, c * (etal[k] * torch.matmul(q, p) - etar[k] * torch.matmul(p, q)))
        sel, nb, n = plus_sel[k]
        if sel.numel():
            p = rho[nb]
            c = (-1j * torch.sqrt(n + 1) * sq[k])[:, None, None]
            out = out.index_add(0, sel, c * (torch.matmul(q, p) - torch.matmul(p, q)))
    return out


def final_state(theta, amps, H0, P, mu, Q, bath, mode, L, rho0, dt, steps_per_seg):
    """rho(T) (torch complex128 [d, d]); theta [nP] and amps [nseg] are float64 tensors."""
    H0 = torch.as_tensor(np.asarray(H0, np.complex128))
    d = H0.shape[0]
    P = torch.as_tensor(np.asarray(P, np.complex128).reshape(-1, d, d))
    mu = torch.as_tensor(np.asarray(mu, np.complex128))
    Q = torch.as_tensor(np.asarray(Q, np.complex128).reshape(-1, d, d))
    bath = tuple(torch.as_tensor(np.asarray(a, np.complex128)) for a in bath)
    hier = Hierarchy(len(mode), L)
    minus_sel, plus_sel = [], []
    for k in range(hier.K):
        s = np.nonzero(hier.minus[:, k] >= 0)[0]
        minus_sel.append((torch.as_tensor(s), torch.as_tensor(hier.minus[s, k]),
                          torch.as_tensor(hier.keys[s, k].astype(np.float64)).to(torch.complex128)))
        s = np.nonzero(hier.plus[:, k] >= 0)[0]
        plus_sel.append((torch.as_tensor(s), torch.as_tensor(hier.plus[s, k]),
                         torch.as_tensor(hier.keys[s, k].astype(np.float64)).to(torch.complex128)))
    H = H0 + torch.einsum("j,jab->ab", theta.to(torch.complex128), P) if P.shape[0] else H0
    y = torch.zeros((hier.nmax, d, d), dtype=torch.complex128)
    r0 = rho0 if isinstance(rho0, torch.Tensor) else torch.as_tensor(np.asarray(rho0, np.complex128))
    y = y.index_put((torch.tensor([0]),), r0[None])
    for s in range(amps.shape[0]):
        Ht = H + amps[s].to(torch.complex128) * mu
        f = lambda x: _rhs(x, Ht, Q, hier, bath, mode, minus_sel, plus_sel)  # noqa: E731
        for _ in range(steps_per_seg):
            k1 = f(y)
            k2 = f(y + k1 * (dt / 2))
            k3 = f(y + k2 * (dt / 2))
            k4 = f(y + k3 * dt)
            y = y + (k1 + 2 * k2 + 2 * k3 + k4) * (dt / 6)
    return y[0]
