"""Host-side HEOM generator and spectra (DEOMSolver.gen_generate_propgator, correlation_4op_3t):
the table-based sparse assembly against the oracle's entry-by-entry assembly (deom.py:769-884).
CPU only; the correlation (its frequency grid runs on the GPU) is in test_gpu_heom_spectra.py."""
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
