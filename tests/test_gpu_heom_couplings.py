"""DEOMSolver on the GPU with time-dependent couplings Q_m(t) = Q_m + coupling_dipole[m] f_c(t)
(deom.py:676-687) and with couplings that do not commute (deom.py:651-664 with general Q[m]), against
the reference's trajectories (tests/golden/traj_cdip.npz, traj_dense.npz); and the product's own bath
routes (Bath over sympy spectra, the closed-form drude) feeding the solver, against the spin-boson and
7-site exciton trajectories.

Tolerances: plain FP64 kernels <= 1e-12 max-abs on rho_0(t); mixed (FP32 right-hand side, FP64 state)
<= 1e-6."""
import numpy as np
import pytest

from conftest import bath_of, golden
from test_heom_couplings_oracle import CDIP, DENSE, case, fields

pytestmark = pytest.mark.gpu


def _run(c, prec):
    from pyqed_amd.heom.deom import Bath, DEOMSolver
    fs, fc = fields(c)
    b = Bath.from_arrays(*(c["bath_" + k] for k in ("etal", "etar", "etaa", "expn", "mode")))
    s = DEOMSolver(c["H"], c.get("sdip"), b, list(c["Q"]), c.get("cdip"), fs, fc, int(c["L"]))
    t, traj = s.run(c["rho0"], float(c["dt"]), int(c["nt"]), precision=prec)
    return s, np.array(traj)


@pytest.mark.parametrize("prec,tol", [("fp64", 1e-12), ("mixed", 1e-6)])
@pytest.mark.parametrize("name", CDIP + DENSE)
def test_against_reference(name, prec, tol):
    c = case("traj_cdip.npz" if name.startswith("cdip") else "traj_dense.npz", name)
    s, traj = _run(c, prec)
    err = np.abs(traj - c["rho"]).max()
    print(f"{name} {prec}: max |rho - reference| = {err:.2e}")
    assert err <= tol


def test_dense_cases_take_the_dense_coupling_path():
    """Non-commuting couplings (and a coupling dipole that does not commute with its coupling) use the
    general-Q kernel branch; diagonal coupling dipoles keep the diagonal branch."""
    for npz, name, dense in [("traj_dense.npz", n, True) for n in DENSE] + [
            ("traj_cdip.npz", "cdip_dense", True), ("traj_cdip.npz", "cdip_rows", True),
            ("traj_cdip.npz", "cdip_diag", False), ("traj_cdip.npz", "cdip_sites", False)]:
        s, _ = _run(case(npz, name), "fp64")
        assert s.engine.dense == dense, name
        assert not s.engine.hermitian, name


def _c1_bath_routes():
    import sympy as sp
    from pyqed_amd.heom.bath import Bath, drude
    w = sp.symbols(r"\omega", real=True)
    spe = 2 * 0.1 * 0.5 * w / (0.5 ** 2 + w ** 2)
    yield "sympy Bath", Bath([spe], w, [1.0], [1], [0, 0])
    yield "Bath.drude", Bath.drude(0.1, 0.5, 1.0, 1)
    e = drude(0.1, 0.5, 1.0, 1)
    yield "drude arrays", Bath.from_arrays(*e, [0, 0])


@pytest.mark.parametrize("prec,tol", [("fp64", 1e-12), ("mixed", 1e-6)])
def test_product_bath_spin_boson(prec, tol):
    """The spin-boson trajectory (1000 steps) from baths the product computes itself."""
    from pyqed_amd.heom.deom import DEOMSolver
    g = golden("traj_c1.npz")
    for label, b in _c1_bath_routes():
        s = DEOMSolver(g["H"], None, b, list(g["Q"]), None, None, None, int(g["L"]))
        _, traj = s.run(g["rho0"], float(g["dt"]), int(g["nt"]), precision=prec)
        err = np.abs(np.array(traj) - g["rho"]).max()
        print(f"spin-boson, {label}, {prec}: {err:.2e}")
        assert err <= tol, label


def test_product_bath_exciton7():
    """The 7-site exciton (20 steps; gamma = 1/50 fs < 1 in atomic units, so the reference's Bath and the
    closed form agree) from the product's sympy Bath and from its closed form."""
    import sympy as sp
    from pyqed_amd import workloads as W
    from pyqed_amd.heom.bath import Bath
    from pyqed_amd.heom.deom import DEOMSolver
    g = golden("traj_c2.npz")
    w = W.exciton_c2()
    d = w["H"].shape[0]
    wsym = sp.symbols(r"\omega", real=True)
    spe = 2 * w["lam"] * w["gam"] * wsym / (w["gam"] ** 2 + wsym ** 2)
    mode = list(np.repeat(np.arange(d), 2))
    for label, b in (("sympy Bath", Bath([spe] * d, wsym, [w["beta"]] * d, [1] * d, mode)),
                     ("Bath.drude", Bath.drude(w["lam"], w["gam"], w["beta"], 1, nbath=d))):
        for k in ("etal", "etar", "etaa", "expn"):
            assert np.abs(getattr(b, k) - g["bath_" + k]).max() <= 1e-13 * np.abs(g["bath_" + k]).max(), (label, k)
        s = DEOMSolver(g["H"], None, b, list(g["Q"]), None, None, None, int(g["L"]))
        _, traj = s.run(g["rho0"], float(g["dt"]), int(g["nt"]))
        err = np.abs(np.array(traj) - g["rho"]).max()
        print(f"exciton7, {label}: {err:.2e}")
        assert err <= 1e-12, label
