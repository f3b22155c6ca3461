"""CPU oracle for the Lindblad and Redfield paths -- TEST INFRASTRUCTURE ONLY.

Imported only by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg, as the checker.

Reference: pyqed/oqs.py, pyqed/phys.py, pyqed/superoperator.py (binggu56/pyqed 0.1.3).

  liouvillian_rhs     oqs.py:697-714 (liouvillian / lindbladian), dense semantics
  rk4                 phys.py:1051-1063 (summation order (k1 + 2k2 + 2k3 + k4)/6 * dt)
  run                 oqs.py:1596-1696 (_lindblad): observables (Nt+1, n_e) with the t0 row
  run_driven          oqs.py:1699-1806 (_lindblad_driven): H_k = H0 - sum_i f_i(t_{k+1}) H_i held
                      over step k, observables (Nt, n_e); legacy=True reproduces the reference's
                      in-place accumulation of the drive into H0 (oqs.py:1725-1732)
  superoperator       superoperator.py:29-58, 200-251 (kron(a, I) - kron(I, a^T) etc.)
  redfield_tensor     oqs.py:519-570
  run_redfield        oqs.py:364-459 (_redfield): observables (Nt, n_e) in the eigenbasis
  correlation_*       oqs.py:717-791 (_correlation_2p_1t), 1193-1327 (LindbladSolver.correlation_*):
                      quantum regression with the same RK4; 3op_2t keeps tau = 0..(Ntau-1)dt (the
                      reference's assignment of Ntau+1 values into Ntau columns raises)

What pins them (tests/test_oracle_oqs.py): expm of the kron-form Liouvillian, the analytic
two-level decay and dephasing laws, RK4's fourth-order convergence, trace and Hermiticity
preservation, detailed balance of the Redfield populations; for the correlation functions, expm of the
Liouvillian applied to the regression state and the closed-form two-level correlation
<sigma+(t) sigma->  = exp((i w - G/2) t).  The reference's _lindblad cannot be
used for golden vectors: with numpy 2 / scipy 1.15 its csr state makes np.dot(dense, csr) return
object arrays and the coherent term vanishes; the dense semantics restated here are the intended
ones.  redfield_tensor has no golden vector either; besides trace preservation and the Gibbs
steady state it is pinned by the closed-form two-level rates (H = diag(0, w), a = sigma_x: the excited
population decays at 2 S(w), the ground one rises at 2 S(-w)), which fix the scale, the sign and
the argument convention of the rates."""
import numpy as np
import scipy.linalg as sla


def dag(a):
    return np.conj(np.asarray(a)).T


def liouvillian_rhs(rho, H, c_ops):
    """-i[H, rho] + sum_c (c rho c^dag - 1/2 {c^dag c, rho})  (oqs.py:697-714)."""
    out = -1j * (H @ rho - rho @ H)
    for c in c_ops:
        cd = dag(c)
        cdc = cd @ c
        out = out + (c @ (rho @ cd) - 0.5 * (cdc @ rho + rho @ cdc))
    return out


def rk4(rho, fun, dt, *args):
    """phys.py:1051-1063."""
    dt2 = dt / 2.0
    k1 = fun(rho, *args)
    k2 = fun(rho + k1 * dt2, *args)
    k3 = fun(rho + k2 * dt2, *args)
    k4 = fun(rho + k3 * dt, *args)
    return rho + (k1 + 2 * k2 + 2 * k3 + k4) / 6.0 * dt


def obs_dm(rho, op):
    """Tr(op rho)  (phys.py:1257-1263)."""
    return np.trace(op @ rho)


def run(H, c_ops, rho0, dt, Nt, e_ops=()):
    rho = np.array(rho0, np.complex128)
    obs = np.zeros((Nt + 1, len(e_ops)), complex)
    obs[0] = [obs_dm(rho, e) for e in e_ops]
    rholist = []
    for k in range(Nt):
        rho = rk4(rho, liouvillian_rhs, dt, H, c_ops)
        rholist.append(rho.copy())
        obs[k + 1] = [obs_dm(rho, e) for e in e_ops]
    return dict(observables=obs, rholist=rholist, times=np.arange(Nt + 1) * dt)


def correlation_2op_1t(H, c_ops, rho0, a_op, b_op, dt, Nt):
    """<A(t)B> = Tr[A rho(t)], rho(0) = B rho0, at t = dt..Nt*dt (oqs.py:717-791)."""
    return run(H, c_ops, np.asarray(b_op) @ np.asarray(rho0), dt, Nt, [a_op])["observables"][1:, 0]


def correlation_3op_1t(H, c_ops, rho0, a_op, b_op, c_op, dt, Nt):
    """<A B(t) C> = Tr[B rho(t)], rho(0) = C rho0 A, at t = 0..Nt*dt (oqs.py:1224-1242)."""
    start = np.asarray(c_op) @ np.asarray(rho0) @ np.asarray(a_op)
    return run(H, c_ops, start, dt, Nt, [b_op])["observables"][:, 0]


def correlation_3op_2t(H, c_ops, rho0, a_op, b_op, c_op, dt, Nt, Ntau):
    """<A(t) B(t+tau) C(t)>, [Nt, Ntau]: row j at t = (j+1)dt, column m at tau = m*dt (oqs.py:1263-1297)."""
    rhos = run(H, c_ops, rho0, dt, Nt)["rholist"]
    out = np.zeros((Nt, Ntau), complex)
    for j, rho in enumerate(rhos):
        out[j] = correlation_3op_1t(H, c_ops, rho, a_op, b_op, c_op, dt, Ntau)[:Ntau]
    return out


def run_driven(H, rho0, c_ops, dt, Nt, t0=0.0, e_ops=(), legacy=False):
    """H = [H0, [H1, f1], ...]: step k uses H0 - sum_i f_i(t0 + (k+1) dt) H_i."""
    H0 = np.array(H[0], np.complex128)
    acc = H0.copy()
    rho = np.array(rho0, np.complex128)
    obs = np.zeros((Nt, len(e_ops)), complex)
    rholist = []
    t = t0
    for k in range(Nt):
        t += dt
        if legacy:
            for Hi, fi in H[1:]:
                acc = acc - fi(t) * np.asarray(Hi)
            Ht = acc
        else:
            Ht = H0.copy()
            for Hi, fi in H[1:]:
                Ht = Ht - fi(t) * np.asarray(Hi)
        rho = rk4(rho, liouvillian_rhs, dt, Ht, c_ops)
        rholist.append(rho.copy())
        obs[k] = [obs_dm(rho, e) for e in e_ops]
    return dict(observables=obs, rholist=rholist)


# ----------------------------------------------------------------------------- superoperators
def op2sop(a, kind="commutator"):
    """Row-major Liouville space: vec(A X B) = kron(A, B^T) vec(X)  (superoperator.py:200-246)."""
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
    raise ValueError(kind)


def lindblad_dissipator(c):
    """kron(c, c*) - 1/2 {c^dag c, .}  (superoperator.py:249-251)."""
    c = np.asarray(c, np.complex128)
    return np.kron(c, c.conj()) - 0.5 * op2sop(dag(c) @ c, "anticommutator")


def superoperator(H, c_ops):
    """superoperator.liouvillian (superoperator.py:29-58)."""
    L = -1j * op2sop(H)
    for c in c_ops:
        L = L + lindblad_dissipator(c)
    return L


def expm_evolve(L, rho0, t):
    return (sla.expm(L * t) @ np.asarray(rho0, np.complex128).reshape(-1)).reshape(np.shape(rho0))


# ----------------------------------------------------------------------------- Redfield
def transform(A, v):
    """v^dag A v  (phys.py:1121-1138)."""
    return dag(v) @ A @ v


def redfield_tensor(H, a_ops, spectra):
    """oqs.py:519-570: R = -i op2sop(diag E) - sum_k op2sop(A_k)(left(L_k) - right(L_k^dag)),
    L_k = C_k * A_k, C_k[n, m] = S_k(-(E_n - E_m)), everything in the eigenbasis of H."""
    for a in a_ops:
        if not np.allclose(a, dag(a)):
            raise TypeError("Operators in a_ops must be Hermitian.")
    evals, evecs = np.linalg.eigh(H)
    W = np.real(evals[:, None] - evals[None, :])
    N = len(evals)
    R = 0
    for a, s in zip(a_ops, spectra):
        C = np.zeros((N, N))
        for n in range(N):
            for m in range(N):
                C[n, m] = s(-W[n, m])
        A = transform(np.asarray(a, np.complex128), evecs)
        Lk = C * A
        R = R + op2sop(A) @ (op2sop(Lk, "left") - op2sop(dag(Lk), "right"))
    return -1j * op2sop(np.diag(evals)) - R, evecs


def run_redfield(R, rho0, evecs, dt, Nt, e_ops=()):
    """oqs.py:364-459 with return_result=True."""
    N = rho0.shape[0]
    rho = transform(np.asarray(rho0, np.complex128), evecs).reshape(-1)
    e_eig = [transform(np.asarray(e, np.complex128), evecs) for e in e_ops]
    obs = np.zeros((Nt, len(e_ops)), complex)
    rholist = []
    for k in range(Nt):
        rho = rk4(rho, lambda x, M: M @ x, dt, R)
        tmp = rho.reshape(N, N)
        rholist.append(transform(tmp, dag(evecs)))
        obs[k] = [obs_dm(tmp, e) for e in e_eig]
    return dict(observables=obs, rholist=rholist)
