"""Prony bath decomposition (pyqed_amd.heom.bath, deom.py:21-66, 310-542) against the reference's
outputs (tests/golden/prony.npz, written by tests/golden/make_golden.py) and against the correlation
function it fits.  CPU only."""
import numpy as np
import pytest
import sympy as sp

from conftest import golden

NAMES = ("etal", "etar", "etaa", "expn")


def _drude(lam, gam):
    w = sp.symbols(r"\omega", real=True)
    return 2 * lam * gam * w / (gam ** 2 + w ** 2), w


def _brownian():
    w = sp.symbols(r"\omega", real=True)
    lam, w0, g = 0.5, 1.0, 0.6
    return 2 * lam * w0 ** 2 * g * w / ((w0 ** 2 - w ** 2) ** 2 + g ** 2 * w ** 2), w


def _close(out, g, prefix, rtol):
    for k, v in zip(NAMES, out):
        ref = g[f"{prefix}_{k}"]
        assert np.asarray(v).shape == ref.shape, (prefix, k)
        assert np.abs(np.asarray(v) - ref).max() <= rtol * max(1.0, np.abs(ref).max()), (prefix, k)


def test_small_helpers_match_reference():
    from pyqed_amd.heom.bath import fit_t, sort_symmetry, spectrum_exp
    g = golden("prony.npz")
    out = sort_symmetry(g["sort_in_etal"].copy(), g["sort_in_expn"].copy())
    for k, v in zip(NAMES, out):
        assert np.array_equal(np.asarray(v), g[f"sort_{k}"]), k
    res = np.zeros(7, complex)
    fit_t(g["fit_t_t"], res, g["sort_in_expn"], g["sort_in_etal"])
    assert np.array_equal(res, g["fit_t"])
    res = np.zeros(7, complex)
    spectrum_exp(g["fit_t_t"], res, g["sort_in_expn"], g["sort_in_etal"])
    assert np.array_equal(res, g["spectrum_exp"])


def test_pade_real_imag_match_reference():
    from pyqed_amd.heom.bath import decompose_spectrum_pade_imag, decompose_spectrum_pade_real
    g = golden("prony.npz")
    spe, w = _drude(0.2, 1.0)
    _close(decompose_spectrum_pade_real(spe, w), g, "drude_pade_real", 1e-13)
    _close(decompose_spectrum_pade_imag(spe, w), g, "drude_pade_imag", 1e-13)
    spe, w = _brownian()
    _close(decompose_spectrum_pade_real(spe, w), g, "brownian_pade_real", 1e-13)


@pytest.mark.parametrize("case,nind", [("drude", 3), ("brownian", [2, 2])])
def test_prony_fitting_matches_reference_on_its_samples(case, nind):
    from pyqed_amd.heom.bath import prony_fitting
    g = golden("prony.npz")
    out = prony_fitting(g[f"{case}_samples"], np.linspace(0, 1, 401), nind, 20.0, 200)
    _close(out, g, case, 1e-12)


@pytest.mark.parametrize("case", ["drude", "brownian"])
def test_prony_end_to_end_fits_the_samples(case):
    from pyqed_amd.heom.bath import decompose_spectrum_pade, decompose_spectrum_prony, fit_t
    g = golden("prony.npz")
    spe, w = _drude(0.2, 1.0) if case == "drude" else _brownian()
    nind = 3 if case == "drude" else [2, 2]
    out = decompose_spectrum_prony(spe, w, 1.0, nind, scale=20.0, n=200, npsd=10)
    if case == "drude":
        # one well-conditioned fit: the end-to-end result matches the reference's too
        _close(out, g, case, 1e-9)
    # what the fit is for: sum etal exp(-expn t) reproduces the Pade correlation function it sampled
    etal_p, _, _, expn_p = decompose_spectrum_pade(spe, w, 1.0, 10)
    t = np.linspace(0.0, 20.0, 401)
    ref = fit_t(t, np.zeros(401, complex), expn_p, etal_p)
    fit = fit_t(t, np.zeros(401, complex), out[3], out[0])
    assert np.abs(fit - ref).max() < 2e-2 * np.abs(ref).max()
