"""The oracle on the reference's trajectories with time-dependent couplings (coupling dipoles,
deom.py:676-687) and with couplings that do not commute (deom.py:651-664 with general Q[m]).
Golden vectors: tests/golden/traj_cdip.npz and traj_dense.npz from make_golden.py."""
import numpy as np
import pytest

from conftest import golden
from oracle import heom as O

CDIP = ["cdip_diag", "cdip_dense", "cdip_rows", "cdip_sites"]
DENSE = ["dense_sxsz", "dense_three", "dense_six"]


def case(npz, name):
    g = golden(npz)
    c = {k.split("__", 1)[1]: g[k] for k in g.files if k.startswith(name + "__")}
    assert c, name
    return c


def fields(c):
    fs = fc = None
    if "fs" in c:
        fs = lambda t, a=float(c["fs"][0]), w=float(c["fs"][1]): a * np.sin(w * t)  # noqa: E731
    if "fc" in c:
        fc = lambda t, a=float(c["fc"][0]), w=float(c["fc"][1]): a * np.cos(w * t)  # noqa: E731
    return fs, fc


def bath(c):
    return tuple(c["bath_" + k] for k in ("etal", "etar", "etaa", "expn")), c["bath_mode"]


@pytest.mark.parametrize("name", CDIP + DENSE)
def test_oracle_matches_reference(name):
    c = case("traj_cdip.npz" if name.startswith("cdip") else "traj_dense.npz", name)
    fs, fc = fields(c)
    b, mode = bath(c)
    _, rho = O.run(c["H"], list(c["Q"]), b, mode, int(c["L"]), c["rho0"], float(c["dt"]), int(c["nt"]),
                   sdip=c.get("sdip"), fs=fs, cdip=c.get("cdip"), fc=fc)
    err = np.abs(rho - c["rho"]).max()
    print(f"{name}: oracle vs reference {err:.2e}")
    assert err <= 1e-12


def test_dense_cases_do_not_commute():
    """The dense fixtures exercise the general-Q branch: no common eigenbasis of the couplings exists."""
    from pyqed_amd.heom.engine import common_eigenbasis
    for name in DENSE:
        c = case("traj_dense.npz", name)
        assert common_eigenbasis(list(c["Q"])) is None, name
    c = case("traj_cdip.npz", "cdip_dense")
    assert common_eigenbasis([c["Q"][0], c["cdip"][0]]) is None


def test_dropping_the_coupling_dipole_is_caught():
    """Planted slip in the oracle: ignoring Q_m(t)'s field term moves the trajectory far past tolerance."""
    for name in ("cdip_diag", "cdip_dense", "cdip_rows"):
        c = case("traj_cdip.npz", name)
        fs, _ = fields(c)
        b, mode = bath(c)
        _, rho = O.run(c["H"], list(c["Q"]), b, mode, int(c["L"]), c["rho0"], float(c["dt"]), int(c["nt"]),
                       sdip=c.get("sdip"), fs=fs)
        assert np.abs(rho - c["rho"]).max() > 1e-6, name
