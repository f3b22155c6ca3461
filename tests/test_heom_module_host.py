"""Host functions of pyqed_amd.HEOM.heom (pyqed/HEOM/heom.py:40-158).  CPU only."""
import numpy as np
import scipy.integrate as si

from pyqed_amd.HEOM import heom as HM


def test_state_number_enumerate_order_and_types():
    states = [list(s) for s in HM.state_number_enumerate([2, 2])]
    assert states == [[0, 0], [0, 1], [1, 0], [1, 1]]          # HEOM/heom.py:45-51
    assert all(isinstance(s, np.ndarray) for s in HM.state_number_enumerate([2, 3]))
    ex = list(HM.state_number_enumerate([3, 3, 2], excitations=2))
    assert all(isinstance(s, tuple) and sum(s) <= 2 for s in ex)
    # every state of total <= 2 appears exactly once, in lexicographic order
    brute = [(a, b, c) for a in range(3) for b in range(3) for c in range(2) if a + b + c <= 2]
    assert ex == brute
    # excitations = 0 is falsy in the reference's test: no restriction, tuples
    assert len(list(HM.state_number_enumerate([2, 3], excitations=0))) == 6


def test_enr_state_dictionaries_count_and_inverse():
    n, s2i, i2s = HM.enr_state_dictionaries([4, 4, 4], 3)
    # states of 3 modes with total <= 3 (each below 4): C(3 + 3, 3) = 20
    assert n == 20 == len(s2i) == len(i2s)
    assert all(s2i[i2s[i]] == i for i in range(n))
    assert i2s[0] == (0, 0, 0) and i2s[n - 1] == (3, 0, 0)


def test_matsubara_expansion_reproduces_drude_correlation():
    # C(t) = (1/pi) int_0^inf J(w) [coth(beta w / 2) cos(w t) - i sin(w t)] dw,
    # J(w) = 2 lam gam w / (w^2 + gam^2); the Matsubara sum converges to it as N grows
    lam, gam, T, t = 0.3, 0.8, 1.5, 0.2
    beta = 1.0 / T
    J = lambda w: 2 * lam * gam * w / (w * w + gam * gam)
    re = si.quad(lambda w: J(w) / np.tanh(beta * w / 2), 0, np.inf, weight="cos", wvar=t)[0] / np.pi
    im = -si.quad(J, 0, np.inf, weight="sin", wvar=t)[0] / np.pi

    def corr(N):
        c, nu = HM._calc_matsubara_params(N, lam, gam, T)
        return sum(ck * np.exp(-vk * t) for ck, vk in zip(c, nu))

    C = corr(200)
    assert abs(C.real - re) < 1e-9 and abs(C.imag - im) < 1e-9
    # the Matsubara terms matter at this t: the leading term alone is off by 2.5e-2, two terms by 2e-3
    assert abs(corr(1).real - re) > 1e-2 and 1e-3 < abs(corr(2).real - re) < 1e-2
    # leading term: nu_0 = gamma, c_0 = lam gam (cot(beta gam / 2) - i)
    c1, nu1 = HM._calc_matsubara_params(1, lam, gam, T)
    assert nu1 == [gam] and abs(c1[0] - lam * gam * (1 / np.tan(beta * gam / 2) - 1j)) < 1e-15


def test_deom_alias_packages():
    # pyqed.deom re-exports the HEOM solver module; pyqed.deom.heom is the legacy copy
    import pyqed_amd.deom as D
    import pyqed_amd.deom.heom as DH
    from pyqed_amd.heom import deom
    assert D.DEOMSolver is deom.DEOMSolver and D.Bath is deom.Bath
    assert D.decompose_spectrum_pade is deom.decompose_spectrum_pade
    assert DH.HEOMSolver is HM.HEOMSolver and DH._heom is HM._heom
