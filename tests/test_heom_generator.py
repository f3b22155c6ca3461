"""Host-side HEOM generator and spectra (DEOMSolver.gen_generate_propgator, correlation_4op_3t):
the table-based sparse assembly against the oracle's entry-by-entry assembly (deom.py:769-884),
and the eigen-decomposition correlation against dense linear solves.  CPU only."""
import numpy as np
import scipy.linalg as sla

from conftest import bath_of, golden
from oracle import heom as O


def _solver(g, L):
    from pyqed_amd.heom.deom import Bath, DEOMSolver
    return DEOMSolver(g["H"], None, Bath.from_arrays(*bath_of(g)), list(g["Q"]), None, None, None, L)


def test_generator_matches_oracle_assembly():
    g = golden("traj_example_sx.npz")
    s = _solver(g, 3)
    G = s.gen_generate_propgator()
    hier = O.Hierarchy(3, 3)
    ref = O.dense_generator(g["H"], list(g["Q"]), hier, bath_of(g)[:4], g["bath_mode"])
    assert np.array_equal(G, ref)


def test_generator_propagates_the_reference_trajectory():
    """expm(G t) applied to rho0 reproduces the reference's RK4 trajectory to RK4 accuracy."""
    g = golden("traj_c1.npz")
    s = _solver(g, int(g["L"]))
    G = s.gen_generate_propgator()
    v = np.zeros(G.shape[0], complex)
    v[:4] = g["rho0"].reshape(-1)
    for k in (100, 500, 1000):
        r = (sla.expm(G * k * float(g["dt"])) @ v)[:4].reshape(2, 2)
        assert np.abs(r - g["rho"][k]).max() < 1e-9


def test_correlation_4op_3t_against_linear_solves():
    from pyqed_amd.heom.generator import system_action
    g = golden("traj_c1.npz")
    s = _solver(g, 3)
    sx = np.array([[0, 1], [1, 0]], complex)
    sz = np.diag([1.0, -1.0]).astype(complex)
    rho0 = g["rho0"]
    wx, wy = np.array([-0.5, 0.3]), np.array([0.1, 0.7, 1.2])
    T = 0.8
    for lcr in ("llll", "lrcl"):
        C = s.correlation_4op_3t(sx, sz, sx, sz, rho0, T, wx, wy, lcr=lcr)
        G = s.propgator
        n = G.shape[0]
        acts = [system_action(op, s.nmax, c).toarray() for op, c in zip((sx, sz, sx, sz), lcr)]
        v0 = np.zeros(n, complex)
        v0[:4] = rho0.reshape(-1)
        E = sla.expm(G * T)
        for i, a in enumerate(wx):
            for j, b in enumerate(wy):
                x = np.linalg.solve(-G - 1j * b * np.eye(n), acts[0] @ v0)
                x = acts[2] @ (E @ (acts[1] @ x))
                x = acts[3] @ np.linalg.solve(-G - 1j * a * np.eye(n), x)
                assert abs(np.trace(x[:4].reshape(2, 2)) - C[i, j]) < 1e-9 * max(1.0, abs(C[i, j]))
