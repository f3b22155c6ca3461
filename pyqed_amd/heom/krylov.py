"""Krylov-subspace spectra of a HEOM generator on the GPU, for hierarchies whose dense generator is
too large to diagonalise (deom.py:1127-1210 needs la.eig of an (nmax d^2)^2 matrix).

Every product L x is one call of the library's FP64 right-hand-side kernel (pyqed_heom_rhs); the
Arnoldi vectors live on the device.  Resolvents of all frequencies share one Arnoldi basis, since
K_m(L + s, v) = K_m(L, v): (L + s) x = r is solved in the basis by the small least-squares problem
min ||beta e1 - (Hbar + s Ibar) y|| (GMRES).  e^{L T} v uses the same kind of basis: beta V expm(T H) e1.
Bases grow until every requested shift (or the exponential) meets the residual tolerance.
"""
import numpy as np
import scipy.linalg as sla


class Arnoldi:
    """Incremental Arnoldi factorisation L V_k = V_{k+1} Hbar_k with reorthogonalised classical
    Gram-Schmidt.  Basis and Hessenberg matrix stay on the device; the host reads them once per
    block of steps (no synchronisation inside a block)."""

    def __init__(self, matvec, v0, torch, m_max):
        self.matvec, self.torch = matvec, torch
        self.M = v0.numel()
        self.beta = float(torch.linalg.vector_norm(v0))
        dev = v0.device
        self.V = torch.zeros((m_max + 1, self.M), dtype=torch.complex128, device=dev)
        self.Hd = torch.zeros((m_max + 1, m_max), dtype=torch.complex128, device=dev)
        self.H = np.zeros((m_max + 1, m_max), complex)
        self.k = 0
        self.exhausted = self.beta == 0.0  # invariant subspace reached
        if not self.exhausted:
            self.V[0] = v0 / self.beta

    def extend(self, n):
        torch = self.torch
        k0 = self.k
        stop = min(self.k + n, self.V.shape[0] - 1)
        for k in range(k0, stop if not self.exhausted else k0):
            w = self.matvec(self.V[k])
            Vk = self.V[:k + 1]
            for _ in range(2):
                h = torch.mv(Vk, w.conj()).conj()       # V^dag w without a conjugated copy of V
                w = w - torch.mv(Vk.transpose(0, 1), h)
                self.Hd[:k + 1, k] += h
            hn = torch.linalg.vector_norm(w)
            self.Hd[k + 1, k] = hn
            self.V[k + 1] = w / torch.clamp(hn, min=1e-300)
        self.H[:stop + 1, :stop] = self.Hd[:stop + 1, :stop].cpu().numpy()
        # a vanishing subdiagonal entry means L V_j = V_j H_j: keep the basis up to there
        scale = max(1.0, np.abs(self.H[:k0 + 1, :k0]).max()) if k0 > 0 else 1.0
        for k in range(k0, stop):
            scale = max(scale, np.abs(self.H[:k + 2, k]).max())
            if abs(self.H[k + 1, k]) <= 1e-13 * scale:
                self.k = k + 1
                self.exhausted = True
                return self
        self.k = stop
        return self

    def hbar(self):
        k = self.k
        return self.H[:k + 1, :k] if not self.exhausted else self.H[:k, :k]

    def shifted_solve(self, shifts, rhs_sign=1.0):
        """y_s and residual norms of (L + s) x = rhs_sign * v0 in the basis, for every shift s: the
        least-squares problems min ||beta e1 - (Hbar + s Ibar) y|| of GMRES, solved for all shifts at
        once by Givens rotations of the Hessenberg matrix (O(m^2) per shift)."""
        Hb = self.hbar()
        rows, cols = Hb.shape
        shifts = np.atleast_1d(np.asarray(shifts, complex))
        A = np.repeat(Hb[None].astype(complex), len(shifts), axis=0)
        idx = np.arange(min(rows, cols))
        A[:, idx, idx] += shifts[:, None]
        g = np.zeros((len(shifts), rows), complex)
        g[:, 0] = rhs_sign * self.beta
        for j in range(min(cols, rows - 1)):
            a, b = A[:, j, j].copy(), A[:, j + 1, j].copy()
            r = np.sqrt(np.abs(a) ** 2 + np.abs(b) ** 2)
            safe = np.where(r > 0, r, 1.0)
            c = np.where(r > 0, a / safe, 1.0)
            sn = np.where(r > 0, b / safe, 0.0)
            top, bot = A[:, j, j:].copy(), A[:, j + 1, j:].copy()
            A[:, j, j:] = np.conj(c)[:, None] * top + np.conj(sn)[:, None] * bot
            A[:, j + 1, j:] = -sn[:, None] * top + c[:, None] * bot
            gt, gb = g[:, j].copy(), g[:, j + 1].copy()
            g[:, j] = np.conj(c) * gt + np.conj(sn) * gb
            g[:, j + 1] = -sn * gt + c * gb
        # back substitution on the triangular part (a zero pivot, from an exact breakdown, drops its unknown)
        y = np.zeros((len(shifts), cols), complex)
        for i in range(cols - 1, -1, -1):
            acc = g[:, i] - np.einsum("sk,sk->s", A[:, i, i + 1:cols], y[:, i + 1:cols])
            piv = A[:, i, i]
            y[:, i] = np.where(piv != 0, acc / np.where(piv != 0, piv, 1.0), 0.0)
        e1 = np.zeros(rows, complex)
        e1[0] = rhs_sign * self.beta
        res = np.array([np.linalg.norm(e1 - (Hb @ y[q] + shifts[q] * np.concatenate([y[q], np.zeros(rows - cols)])))
                        for q in range(len(shifts))])
        return y, res

    def expm_coeffs(self, T):
        """Coefficients c of e^{L T} v0 ~ V_k c and an a-posteriori error estimate."""
        k = self.k
        Hm = self.H[:k, :k]
        E = sla.expm(T * Hm)
        c = self.beta * E[:, 0]
        err = 0.0 if self.exhausted else self.beta * abs(self.H[k, k - 1]) * abs(E[k - 1, 0])
        return c, err

    def combine(self, coeffs):
        """V_k @ coeffs for coefficient vectors as columns ([k] or [k, n])."""
        torch = self.torch
        c = torch.as_tensor(np.ascontiguousarray(coeffs), device=self.V.device)
        Vk = self.V[:coeffs.shape[0]]
        return Vk.transpose(0, 1) @ c


def grow_until(arn, check, first=16):
    """Extend the basis until check(arn) holds or no growth is possible.  The basis grows by half
    between checks (16, 24, 36, 54, ...): a check costs small dense work of the basis size (a matrix
    exponential is cubic in it), far more than the device work of one Arnoldi step, and a smaller
    growth factor overshoots the size the tolerance needs by less."""
    while True:
        if arn.k > 0 and check(arn):
            return arn
        if arn.exhausted or arn.k >= arn.V.shape[0] - 1:
            if not check(arn):
                raise RuntimeError(f"Krylov basis of {arn.k} vectors did not reach the tolerance; "
                                   "raise krylov_dim or tol")
            return arn
        arn.extend(max(first, arn.k // 2))


def correlation_grid(matvec, act, trace_d, rho_vec, T, w_x, w_y, torch, tol=1e-10, m_max=400):
    """C[i, j] = Tr0[ d G(w_x[i]) c e^{L T} b G(w_y[j]) a rho ],  G(w) = (-L - i w)^-1.
    act: dict of device functions 'a', 'b', 'c' (system actions on every ADO); trace_d(V) gives
    Tr0[d v] for every row v of V."""
    w_x = np.asarray(w_x, float)
    w_y = np.asarray(w_y, float)
    v0 = act["a"](rho_vec)
    # (-L - i w) x = v  <=>  (L + i w) x = -v
    shifts_y = 1j * w_y

    def converged(shifts, scale):
        def ok(arn):
            _, r = arn.shifted_solve(shifts, -1.0)
            return np.all(r <= tol * max(scale, 1e-300))
        return ok

    A1 = grow_until(Arnoldi(matvec, v0, torch, m_max), converged(shifts_y, float(torch.linalg.vector_norm(v0))))
    Y, _ = A1.shifted_solve(shifts_y, -1.0)
    X = A1.combine(Y.T)                                       # [M, Ny]: G(w_y[j]) a rho
    out = np.zeros((len(w_x), len(w_y)), complex)
    for j in range(len(w_y)):
        u = act["b"](X[:, j].contiguous())
        nu = float(torch.linalg.vector_norm(u))
        Ae = grow_until(Arnoldi(matvec, u, torch, m_max), lambda a: a.expm_coeffs(T)[1] <= tol * max(nu, 1e-300))
        c, _ = Ae.expm_coeffs(T)
        z = act["c"](Ae.combine(c[:, None])[:, 0].contiguous())  # c e^{LT} b G(w_y) a rho
        nz = float(torch.linalg.vector_norm(z))
        A2 = grow_until(Arnoldi(matvec, z, torch, m_max), converged(1j * w_x, nz))
        Y2, _ = A2.shifted_solve(1j * w_x, -1.0)
        ell = trace_d(A2.V[:Y2.shape[1]])                         # Tr0[d v_k] for every basis vector
        out[:, j] = Y2 @ ell
    return out
