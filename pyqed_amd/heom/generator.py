"""Sparse / dense HEOM generator (host side): the operator G with vec(RHS(rho)) = G vec(rho).

Drop-in for DEOMSolver.gen_generate_propgator (pyqed/heom/deom.py:769-884, 1116-1125): the same
row-major index a*d*d + i*d + j, assembled from the library's hierarchy tables with vectorised
Kronecker blocks instead of element-by-element loops.  Used for spectra / propagators of small
hierarchies and for the HEOM adjoint through the Liouville-space kernels.
"""
import numpy as np
import scipy.sparse as sps

from .. import _lib


def heom_generator(H, Q, bath, mode, L, sparse=True):
    """G for system H, couplings Q[m], bath (etal, etar, etaa, expn), mode[k], depth L."""
    etal, etar, etaa, expn = (np.asarray(a, np.complex128) for a in bath)
    mode = np.asarray(mode, int)
    H = np.asarray(H, np.complex128)
    d = H.shape[0]
    K = len(expn)
    keys, plus, minus = _lib.hierarchy_tables(K, L)
    nmax = keys.shape[0]
    I = np.eye(d)
    comm_H = sps.csr_matrix(np.kron(H, I) - np.kron(I, H.T))
    damp = (keys * expn[None, :]).sum(1)
    blocks = [sps.kron(sps.identity(nmax, format="csr"), -1j * comm_H),
              sps.kron(sps.diags(-damp), sps.identity(d * d))]
    sq_eta = np.sqrt(etaa)
    for k in range(K):
        q = np.asarray(Q[mode[k]], np.complex128)
        lq = np.kron(q, I)
        rq = np.kron(I, q.T)
        rows = np.nonzero(minus[:, k] >= 0)[0]
        if rows.size:
            c = -1j * np.sqrt(keys[rows, k]) / sq_eta[k]
            S = sps.csr_matrix((c, (rows, minus[rows, k])), shape=(nmax, nmax))
            blocks.append(sps.kron(S, sps.csr_matrix(etal[k] * lq - etar[k] * rq)))
        rows = np.nonzero(plus[:, k] >= 0)[0]
        if rows.size:
            c = -1j * np.sqrt(keys[rows, k] + 1) * sq_eta[k]
            S = sps.csr_matrix((c, (rows, plus[rows, k])), shape=(nmax, nmax))
            blocks.append(sps.kron(S, sps.csr_matrix(lq - rq)))
    G = sum(blocks[1:], blocks[0]).tocsr()
    return G if sparse else G.toarray()


def system_action(A, nmax, lcr="l"):
    """Action of a system operator on every ADO (generate_actions, deom.py:887-892): 'l' A.,
    'r' .A, 'c' A. + .A (the reference adds both sides for 'c')."""
    A = np.asarray(A, np.complex128)
    d = A.shape[0]
    I = np.eye(d)
    if lcr == "l":
        blk = np.kron(A, I)
    elif lcr == "r":
        blk = np.kron(I, A.T)
    elif lcr == "c":
        blk = np.kron(A, I) + np.kron(I, A.T)
    else:
        raise ValueError(lcr)
    return sps.kron(sps.identity(nmax, format="csr"), sps.csr_matrix(blk)).tocsr()
