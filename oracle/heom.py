"""CPU oracle for the HEOM path -- TEST INFRASTRUCTURE ONLY.

This module restates, in numpy, the reference solver's algorithm so that the
GPU propagator can be checked against it.  It is imported only by tests/,
__graft_entry__.smoke() and the cpu_baseline leg of bench.py, always as the
checker and never as the thing measured or shipped.  The product package
(pyqed_amd) never imports it.

Reference: pyqed/heom/deom.py (binggu56/pyqed 0.1.3).

Functions and what pins each one (tests/test_oracle_*.py):
  comb_table, ado_rank          deom.py:1048-1060, 555-565
      pinned by brute-force enumeration for every (K<=6, L<=6), by the
      reference's own tables (tests/golden/tables_small.npz) and SHA-256
      digests for K=14/L=4, K=48/L=3, K=32/L=4 (tables_digest.npz).
  gen_keys, neighbour_tables    deom.py:609-638, 588-605   (same fixtures)
  psd_poles                     deom.py:104-206 (pade=1)
      pinned by the closed forms N=1 (xi = 2 sqrt 15, r = 5/2), the sum rule
      sum r = N(2N+3)/2, and the reference's Bath fixtures.
  drude_bath                    deom.py:226-307 for J = 2 lam gam w/(w^2+gam^2)
      pinned by tests/golden/bath_*.npz and the high-order limit
      Re eta_0 -> lam gam cot(beta gam / 2).
  rhs_loop, rhs_vec             deom.py:641-673
      pinned by the reference trajectories (traj_*.npz), by the independently
      assembled dense generator (dense_generator, deom.py:769-884) and by the
      pure-dephasing closed form.
  rk4_step, run                 deom.py:725-766, 1072-1114
      pinned by traj_c1.npz (1000 steps), traj_c2.npz, traj_driven.npz,
      traj_example_sx.npz, traj_c3_small.npz and by expm of the generator.
"""
import math

import numpy as np


# ----------------------------------------------------------------------------- hierarchy tables
def comb_table(K, L):
    """Pascal table C(i, j) for i, j < K+L+1 (deom.py:1048-1059)."""
    n = K + L + 1
    c = np.zeros((n, n), np.int64)
    c[0, 0] = 1
    for i in range(1, n):
        c[i, 0] = 1
        for j in range(1, n):
            c[i, j] = c[i - 1, j] + c[i - 1, j - 1]
    return c


def ado_rank(key, comb):
    """Rank of multi-index ``key``: sum_i C(s_i + i, i + 1), s_i the prefix sums (deom.py:555-565)."""
    s = 0
    r = 0
    for i, n in enumerate(key):
        s += int(n)
        r += int(comb[s + i, i + 1])
    return r


def gen_keys(K, L):
    """keys[nmax, K]: multi-index of every rank, filled tier by tier as deom.py:609-638 does."""
    comb = comb_table(K, L)
    nmax = int(comb[L + K, L])
    keys = np.zeros((nmax, K), np.int64)
    lo, hi = 0, 1
    for tier in range(L + 1):
        for a in range(lo, hi):
            key = keys[a]
            for k in range(K):
                if key.sum() < L:
                    kp = key.copy()
                    kp[k] += 1
                    keys[ado_rank(kp, comb)] = kp
                if key[k] > 0:
                    km = key.copy()
                    km[k] -= 1
                    keys[ado_rank(km, comb)] = km
        lo, hi = hi, int(comb[K + tier + 1, K]) if tier + 1 <= L else hi
    return keys


def neighbour_tables(keys, L):
    """plus[n, k] = rank(n + e_k) (-1 at tier L); minus[n, k] = rank(n - e_k) (-1 where n_k = 0).

    These are what hash_plus / hash_minus return (deom.py:588-605) at the
    call sites deom.py:652-664."""
    nmax, K = keys.shape
    comb = comb_table(K, L)
    plus = -np.ones((nmax, K), np.int64)
    minus = -np.ones((nmax, K), np.int64)
    for a in range(nmax):
        key = keys[a]
        tier = key.sum()
        for k in range(K):
            if tier < L:
                kp = key.copy()
                kp[k] += 1
                plus[a, k] = ado_rank(kp, comb)
            if key[k] > 0:
                km = key.copy()
                km[k] -= 1
                minus[a, k] = ado_rank(km, comb)
    return plus, minus


# ----------------------------------------------------------------------------- bath coefficients
def psd_poles(N):
    """[N-1/N] Pade spectrum decomposition of the Bose function (deom.py:104-162, pade=1, Bose).

    Returns poles xi_j and residues r_j such that
    1/(1-exp(-x)) ~ 1/x + 1/2 + sum_j 2 r_j x / (x^2 + xi_j^2)."""
    if N == 0:
        return np.zeros(0), np.zeros(0)
    M = 2 * N
    off = np.array([1.0 / math.sqrt((3.0 + 2.0 * i) * (5.0 + 2.0 * i)) for i in range(M - 1)])
    ev = np.linalg.eigvalsh(np.diag(off, 1) + np.diag(off, -1))
    ev = -np.sort(-ev)
    xi = 2.0 / ev[:N]
    xi2 = xi * xi
    M2 = M - 1
    off2 = np.array([1.0 / math.sqrt((5.0 + 2.0 * i) * (7.0 + 2.0 * i)) for i in range(M2 - 1)])
    ev2 = np.linalg.eigvalsh(np.diag(off2, 1) + np.diag(off2, -1)) if M2 > 1 else np.zeros(1)
    ev2 = -np.sort(-ev2)
    Mh = M2 // 2
    zeta2 = (2.0 / ev2[:Mh]) ** 2
    scale = N * (2.0 * N + 3.0)
    r = np.zeros(N)
    for j in range(N):
        t = 0.5 * scale if j == N - 1 else 0.5 * scale * (zeta2[j] - xi2[j]) / (xi2[N - 1] - xi2[j])
        for k in range(Mh):
            if k != j:
                t *= (zeta2[k] - xi2[j]) / (xi2[k] - xi2[j])
        r[j] = t
    return xi, r


def bose_pade(x, xi, r):
    """Pade-approximated Bose factor (deom.py:68-73)."""
    return 1.0 / x + 0.5 + sum(2.0 * r[j] * x / (x * x + xi[j] ** 2) for j in range(len(xi)))


def drude_bath(lam, gam, beta, npsd):
    """Exponential decomposition of the Drude-Lorentz bath J(w) = 2 lam gam w / (w^2 + gam^2).

    Follows deom.py:226-307 for this spectrum: one spectral pole gamma with
    eta_0 = -2i * Res * n_B(-i gamma beta) and npsd Pade poles xi_j/beta with
    eta_j = -2i r_j / beta * J-part at -i xi_j / beta.  Returns
    (etal, etar, etaa, expn) as in Bath (etaa = |eta|)."""
    xi, r = psd_poles(npsd)
    # spectral pole w = -i gam: residue of numer/(w - i gam) is lam*gam (times -2i) times Bose factor
    eta0 = -2j * lam * gam * bose_pade(-1j * gam * beta, xi, r)
    etal = [eta0]
    expn = [gam + 0j]
    for j in range(npsd):
        nu = xi[j] / beta
        w = -1j * nu
        f = 2.0 * lam * gam * w / ((w - 1j * gam) * (w + 1j * gam))
        etal.append(-2j * r[j] / beta * f)
        expn.append(nu + 0j)
    etal = np.array(etal, np.complex128)
    etar = np.conj(etal)
    etaa = np.abs(etal)
    return etal, etar, etaa.astype(np.complex128), np.array(expn, np.complex128)


def multi_drude_bath(lams, gams, betas, npsd):
    """Concatenate one Drude bath per site (Bath.__init__, deom.py:920-931)."""
    parts = [drude_bath(l, g, b, npsd) for l, g, b in zip(lams, gams, betas)]
    return tuple(np.concatenate([p[i] for p in parts]) for i in range(4))


# ----------------------------------------------------------------------------- right-hand side
def ranks(keys, comb):
    """ado_rank of every row of keys [N, K] at once: sum_i C(s_i + i, i + 1) over prefix sums s_i."""
    keys = np.asarray(keys, np.int64)
    K = keys.shape[1]
    s = np.cumsum(keys, axis=1)
    i = np.arange(K)
    return comb[s + i[None, :], (i + 1)[None, :]].sum(axis=1)


def tables_vectorised(K, L):
    """keys, plus and minus as gen_keys / neighbour_tables give them, built with array operations:
    every multi-index of total <= L is placed at its rank, and the neighbours are the ranks of
    keys +- e_k."""
    comb = comb_table(K, L)
    nmax = int(comb[L + K, L])
    rows = [np.zeros((1, 0), np.int64)]
    for k in range(K):  # all prefixes of length k + 1 with total <= L
        prev = rows[-1]
        tot = prev.sum(1) if prev.shape[1] else np.zeros(prev.shape[0], np.int64)
        reps = L - tot + 1
        new = np.repeat(prev, reps, axis=0)
        val = np.concatenate([np.arange(r) for r in reps])
        rows.append(np.concatenate([new, val[:, None]], axis=1))
    allk = rows[-1] if K else np.zeros((1, 0), np.int64)
    keys = np.zeros((nmax, K), np.int64)
    if K:
        keys[ranks(allk, comb)] = allk
    plus = -np.ones((nmax, K), np.int64)
    minus = -np.ones((nmax, K), np.int64)
    tier = keys.sum(1)
    for k in range(K):
        e = np.zeros(K, np.int64)
        e[k] = 1
        up = tier < L
        plus[up, k] = ranks(keys[up] + e, comb)
        dn = keys[:, k] > 0
        minus[dn, k] = ranks(keys[dn] - e, comb)
    return keys, plus, minus


class Hierarchy:
    """Tables and coefficients for one HEOM instance (tables from tables_vectorised, which equal the
    loop restatements gen_keys / neighbour_tables)."""

    def __init__(self, K, L):
        self.K, self.L = K, L
        self.keys, self.plus, self.minus = tables_vectorised(K, L)
        self.nmax = self.keys.shape[0]
        self.tier = self.keys.sum(1)


def rhs_loop(rho, H, Q, hier, bath, mode):
    """Per-ADO loop in the reference's order (deom.py:641-673); Q is a list of d x d couplings."""
    etal, etar, etaa, expn = bath
    out = np.empty_like(rho)
    K, L = hier.K, hier.L
    for a in range(hier.nmax):
        key = hier.keys[a]
        x = rho[a]
        dot = -np.sum(key * expn) * x
        dot = dot - 1j * (H @ x - x @ H)
        for k in range(K):
            n = key[k]
            q = Q[mode[k]]
            if n > 0:
                p = rho[hier.minus[a, k]]
                dot = dot - 1j * np.sqrt(n) / np.sqrt(etaa[k]) * (etal[k] * q @ p - etar[k] * p @ q)
            if key.sum() < L:
                p = rho[hier.plus[a, k]]
                dot = dot - 1j * np.sqrt(n + 1) * np.sqrt(etaa[k]) * (q @ p - p @ q)
        out[a] = dot
    return out


def rhs_vec(rho, H, Q, hier, bath, mode):
    """Same operator as rhs_loop, vectorised over ADOs (summation order differs at rounding level)."""
    etal, etar, etaa, expn = bath
    keys = hier.keys
    damp = (keys * expn[None, :]).sum(1)
    out = -damp[:, None, None] * rho
    out -= 1j * (np.matmul(H[None], rho) - np.matmul(rho, H[None]))
    sq_eta = np.sqrt(etaa)
    for k in range(hier.K):
        q = Q[mode[k]]
        sel = np.nonzero(hier.minus[:, k] >= 0)[0]
        if sel.size:
            p = rho[hier.minus[sel, k]]
            n = keys[sel, k].astype(np.float64)
            c = (-1j * np.sqrt(n) / sq_eta[k])[:, None, None]
            out[sel] += c * (etal[k] * np.matmul(q[None], p) - etar[k] * np.matmul(p, q[None]))
        sel = np.nonzero(hier.plus[:, k] >= 0)[0]
        if sel.size:
            p = rho[hier.plus[sel, k]]
            n = keys[sel, k].astype(np.float64)
            c = (-1j * np.sqrt(n + 1) * sq_eta[k])[:, None, None]
            out[sel] += c * (np.matmul(q[None], p) - np.matmul(p, q[None]))
    return out


class DiagTerms:
    """Per-exponential gathers and coefficient matrices for rhs_vec_diag (diagonal couplings)."""

    def __init__(self, hier, bath, mode, qdiag):
        etal, etar, etaa, expn = bath
        q = np.asarray(qdiag)
        self.terms = []
        sq = np.sqrt(etaa)
        for k in range(hier.K):
            qk = q[mode[k]]
            sel = np.nonzero(hier.minus[:, k] >= 0)[0]
            if sel.size:
                n = hier.keys[sel, k].astype(np.float64)
                cm = etal[k] * qk[:, None] - etar[k] * qk[None, :]  # (Q X)_ij = q_i X_ij, (X Q)_ij = X_ij q_j
                self.terms.append((sel, hier.minus[sel, k], (-1j * np.sqrt(n) / sq[k])[:, None, None], cm))
            sel = np.nonzero(hier.plus[:, k] >= 0)[0]
            if sel.size:
                n = hier.keys[sel, k].astype(np.float64)
                cp = qk[:, None] - qk[None, :]
                self.terms.append((sel, hier.plus[sel, k], (-1j * np.sqrt(n + 1) * sq[k])[:, None, None], cp))
        self.damp = (hier.keys * np.asarray(expn)[None, :]).sum(1)[:, None, None]


def rhs_vec_diag(rho, H, terms):
    """rhs_vec for diagonal couplings Q_m = diag(q_m): every coupling product is a row or column scaling,
    so each neighbour term is one gather and one elementwise multiply-add.  The all-core CPU baseline
    uses it (same operator as rhs_vec; summation order differs at rounding level)."""
    out = -terms.damp * rho - 1j * (np.matmul(H[None], rho) - np.matmul(rho, H[None]))
    for sel, nb, c, cm in terms.terms:
        out[sel] += c * cm[None] * rho[nb]
    return out


def rk4_step(y, rhs, dt, t, Ht=None):
    """One classic RK4 step with the reference's buffer order (deom.py:725-766).

    ``rhs(x, t)`` returns dX/dt.  Summation: acc = k1; X = y + acc*dt/2;
    acc += 2 k2; X = y + k2*dt/2; acc += 2 k3; X = y + k3*dt; acc += k4;
    y += acc*dt/6."""
    d1 = rhs(y, t)
    x = y + d1 * dt / 2
    d2 = rhs(x, t + dt / 2)
    d1 = d1 + d2 * 2
    x = y + d2 * dt / 2
    d2 = rhs(x, t + dt / 2)
    d1 = d1 + d2 * 2
    x = y + d2 * dt
    d2 = rhs(x, t + dt)
    d1 = d1 + d2
    return y + d1 * dt / 6


def run(H, Q, bath, mode, L, rho0, dt, nt, sdip=None, fs=None, cdip=None, fc=None, vectorised=True,
        save_every=1, state_out=None):
    """DEOMSolver.run (deom.py:1072-1114): returns (t, rho_0 trajectory [nt//save_every+1, d, d]).

    Time dependence follows generate_time (deom.py:676-687): H(t) = H + sdip f_s(t),
    Q_m(t) = Q_m + cdip[m] f_c(t) (with a 2-D cdip, cdip[m] is its row m broadcast over rows,
    as numpy evaluates the reference's expression)."""
    H = np.asarray(H, np.complex128)
    d = H.shape[0]
    Q = [np.asarray(q, np.complex128) for q in Q]
    K = len(bath[3])
    hier = Hierarchy(K, L)
    f = rhs_vec if vectorised else rhs_loop

    def rhs(x, t):
        Ht = H if fs is None else H + np.asarray(sdip) * fs(t)
        Qt = Q if fc is None else [Q[m] + np.asarray(cdip)[m] * fc(t) for m in range(len(Q))]
        return f(x, Ht, Qt, hier, bath, mode)

    y = np.zeros((hier.nmax, d, d), np.complex128)
    y[0] = rho0
    out = [y[0].copy()]
    for i in range(nt):
        y = rk4_step(y, rhs, dt, i * dt)
        if (i + 1) % save_every == 0:
            out.append(y[0].copy())
    if state_out is not None:
        state_out.append(y)
    t = np.arange(len(out)) * dt * save_every
    return t, np.array(out)


# ----------------------------------------------------------------------------- generator / expm cross-check
def dense_generator(H, Q, hier, bath, mode):
    """Dense HEOM generator G with vec(RHS(rho)) = G vec(rho), row-major index a*d*d + i*d + j,
    assembled entry by entry like generate_propgator / propgator_element (deom.py:769-884)."""
    etal, etar, etaa, expn = bath
    d = H.shape[0]
    n = hier.nmax * d * d
    G = np.zeros((n, n), np.complex128)
    I = np.eye(d)

    def blk(a, b, M):
        G[a * d * d:(a + 1) * d * d, b * d * d:(b + 1) * d * d] += M

    # row-major vec: vec(A X B) = (A kron B^T) vec(X)
    comm_H = np.kron(H, I) - np.kron(I, H.T)
    for a in range(hier.nmax):
        key = hier.keys[a]
        blk(a, a, -np.sum(key * expn) * np.eye(d * d) - 1j * comm_H)
        for k in range(hier.K):
            q = Q[mode[k]]
            nk = key[k]
            if nk > 0:
                c = -1j * np.sqrt(nk) / np.sqrt(etaa[k])
                blk(a, hier.minus[a, k], c * (etal[k] * np.kron(q, I) - etar[k] * np.kron(I, q.T)))
            if key.sum() < hier.L:
                c = -1j * np.sqrt(nk + 1) * np.sqrt(etaa[k])
                blk(a, hier.plus[a, k], c * (np.kron(q, I) - np.kron(I, q.T)))
    return G


def pure_dephasing_coherence(eps, bath, t, rho01_0=0.5):
    """rho_01(t) for H = eps/2 sz, V = sz: rho01(0) e^{-i eps t} exp(-4 Re g(t)) with
    g(t) = sum_k eta_k / gamma_k^2 (gamma_k t - 1 + e^{-gamma_k t}) (exact for the exponential bath)."""
    etal, _, _, expn = bath
    t = np.asarray(t, np.float64)[:, None]
    g = (etal[None, :] / expn[None, :] ** 2 * (expn[None, :] * t - 1 + np.exp(-expn[None, :] * t))).sum(1)
    return rho01_0 * np.exp(-1j * eps * t[:, 0]) * np.exp(-4 * g.real)
