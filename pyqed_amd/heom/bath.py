"""Bath correlation functions as sums of exponentials (host-side preprocessing).

C(t) = sum_k etal_k exp(-expn_k t), with etar_k the coefficient of the
conjugate correlation and etaa_k the magnitude used to scale the hierarchy.

Two routes:
  * decompose_spectrum_pade / Bath(...) follow pyqed/heom/deom.py:104-307 and
    895-942 step by step for a sympy spectral density, so a Bath built from the
    same inputs has the same coefficients as the reference's.  That includes
    one property of the reference worth knowing: the spectral-pole residues are
    formed as numer / prod(w - p) from sympy's numerator/denominator split,
    which assumes a monic denominator.  sympy normalises some rational
    functions to a unit constant term instead (for the Drude form whenever
    gamma > 1), and the coefficients then carry the factor 1/lead(denom).
    Pass ``monic=True`` to divide by the leading coefficient (physical
    normalisation).
  * drude(...) and brownian(...) evaluate the Drude-Lorentz and Brownian-oscillator + Pade
    decompositions in closed form (physical normalisation), with no sympy: the route for large batches.
  * decompose_spectrum_prony(...) and its helpers follow deom.py:21-66, 310-542: C(t) sampled
    from the Pade decomposition and fitted by Prony's method (Hankel eigenvectors, polynomial roots,
    least-squares amplitudes).  Host-only preprocessing, run once per bath.
"""
import math

import numpy as np


# ----------------------------------------------------------------------------- Bose-function poles
def matsubara_approximation_distribution(N, BoseFermi=1):
    """Matsubara poles 2(i+1)pi (Bose) or (2i+1)pi (Fermi), unit residues (deom.py:84-101)."""
    if BoseFermi == 1:
        return np.array([2.0 * (i + 1) * np.pi for i in range(N)]), np.ones(N)
    return np.array([(2.0 * i + 1) * np.pi for i in range(N)]), np.ones(N)


def _tridiag_eigs_desc(n, a0):
    """Eigenvalues (descending) of the zero-diagonal symmetric tridiagonal matrix with
    off-diagonal 1/sqrt((a0 + 2i)(a0 + 2i + 2)), i = 0..n-2."""
    if n <= 0:
        return np.zeros(0)
    off = np.array([1.0 / math.sqrt((a0 + 2.0 * i) * (a0 + 2.0 * (i + 1))) for i in range(n - 1)])
    m = np.diag(off, 1) + np.diag(off, -1) if n > 1 else np.zeros((1, 1))
    return np.sort(np.linalg.eigvalsh(m))[::-1]


def pade_approximation_distribution(N, BoseFermi=1, pade=1):
    """Pade spectrum decomposition of the Bose (1) / Fermi (2) function (deom.py:104-206).

    pade = 0: Matsubara; 1: [N-1/N]; 2: [N/N]; 3: [N+1/N].  Returns (poles, residues)."""
    if N < 0 or BoseFermi < 1 or BoseFermi > 2 or pade < 0 or pade > 3:
        raise ValueError("N or BoseFermi or pade has wrong value!")
    if pade == 0:
        return matsubara_approximation_distribution(N, BoseFermi)
    if pade in (1, 2):
        if N == 0:
            return [], []
        M = 2 * N + pade // 2
        a = 3.0 if BoseFermi == 1 else 1.0
        pole = 2.0 / _tridiag_eigs_desc(M, a)[:N]
        p2 = pole * pole
        M2 = M - 1
        a2 = 5.0 if BoseFermi == 1 else 3.0
        half = M2 // 2
        z2 = (2.0 / _tridiag_eigs_desc(M2, a2)[:half]) ** 2
        if BoseFermi == 1:
            scale = N * (2.0 * N + 3.0) if pade == 1 else 1.0 / (4.0 * (N + 1.0) * (2.0 * N + 3.0))
        else:
            scale = N * (2.0 * N + 1.0) if pade == 1 else 1.0 / (4.0 * (N + 1.0) * (2.0 * N + 1.0))
        resi = np.zeros(N)
        for j in range(N):
            if pade == 2:
                v = 0.5 * scale * (z2[j] - p2[j])
            elif j == N - 1:
                v = 0.5 * scale
            else:
                v = 0.5 * scale * (z2[j] - p2[j]) / (p2[N - 1] - p2[j])
            for k in range(half):
                if k != j:
                    v *= (z2[k] - p2[j]) / (p2[k] - p2[j])
            resi[j] = v
        return pole, resi
    # pade == 3: continued-fraction construction of the [N+1/N] scheme (deom.py:163-206)
    n1 = N + 1
    a = 3.0 if BoseFermi == 1 else 1.0
    dd = np.empty(2 * n1)
    dd[0] = 0.25 / a
    dd[-1] = -4.0 * (N + 1.0) ** 2 * (a + 2 * N) ** 2 * (a + 4 * N + 2.0)
    for i in range(1, n1):
        dd[2 * i - 1] = -4.0 * i * i * (a + 2.0 * i - 2.0) ** 2 * (a + 4.0 * i - 2.0)
        dd[2 * i] = -0.25 * (a + 4.0 * i) / i / (i + 1) / (a + 2.0 * i - 2.0) / (a + 2.0 * i)
    cum = np.cumsum(dd[1::2][:n1])
    M = 2 * N + 1
    off = np.array([1.0 / math.sqrt(dd[i + 1] * dd[i + 2]) for i in range(M - 1)])
    ev = np.sort(np.linalg.eigvalsh(np.diag(off, 1) + np.diag(off, -1)))[::-1]
    pole = 2.0 / ev[:N]
    resi = np.zeros(N)
    for j in range(N):
        s2 = pole[j] ** 2
        r0, t1 = 0.0, 0.25 / dd[1]
        e0, e1, e2 = 0.0, 0.5, 0.0
        for i in range(n1):
            r1 = t1 if (i == j or i == N) else t1 / (pole[i] ** 2 - s2)
            r2 = 2.0 * math.sqrt(abs(r1)) * (1.0 if r1 > 0 else -1.0)
            r1 = 2.0 * math.sqrt(abs(r1))
            e2 = dd[2 * i] * r1 * e1 - 0.25 * r1 * r0 * s2 * e0
            e0, e1 = e1, e2
            e2 = dd[2 * i + 1] * r2 * e1 - 0.25 * r2 * r1 * s2 * e0
            e0, e1 = e1, e2
            r0 = r2
            if i != N:
                t1 = cum[i] / cum[i + 1]
        resi[j] = e2
    return pole, resi


def function_bose(x, pole, resi):
    """Pade-approximated Bose factor 1/x + 1/2 + sum_j 2 r_j x / (x^2 + xi_j^2) (deom.py:68-73)."""
    return 1.0 / x + 0.5 + sum(2.0 * resi[i] * x / (x ** 2 + pole[i] ** 2) for i in range(len(pole)))


# ----------------------------------------------------------------------------- sympy route
def decompose_spectrum_pade(spe, w_sp, beta, npsd, pade=1, bose_fermi=1, monic=False):
    """Exponential decomposition of a sympy spectral density (deom.py:226-307).

    Returns (etal, etar, etaa, expn) as numpy arrays."""
    import sympy as sp

    re_part, im_part = sp.cancel(spe).as_real_imag()
    part = re_part if im_part == 0 else im_part
    numer, denom = sp.cancel(sp.factor(part)).as_numer_denom()
    roots = [complex(r) for r in sp.nroots(denom)]
    lead = complex(sp.Poly(denom, w_sp).LC()) if monic else 1.0
    allp = np.array(roots, np.complex128)
    lower = np.array([1j * r for r in roots if r.imag < 0], np.complex128)  # expn = i * pole
    order = np.argsort(np.abs(np.imag(lower)))[::-1]
    mags = np.sort(np.abs(np.imag(lower)))[::-1]
    ccs = lower[order[mags != 0]]
    reals = lower[order[mags == 0]]
    num_f = sp.lambdify(w_sp, numer, "numpy")
    pole, resi = pade_approximation_distribution(npsd, bose_fermi, pade)
    temp = 1.0 / beta

    def residue_eta(ex):
        p = -1j * ex  # the pole itself
        others = allp[np.abs(allp - p) > 1e-14]
        val = -2j * complex(num_f(p)) / complex(np.prod(p - others)) / lead
        return val * function_bose(p / temp, pole, resi)

    expn = list(lower[order])
    etal, etar, etaa = [], [], []
    for ii in range(0, len(ccs), 2):
        e0, e1 = residue_eta(ccs[ii]), residue_eta(ccs[ii + 1])
        etal += [e0, e1]
        etar += [np.conj(e1), np.conj(e0)]
        etaa += [np.sqrt(abs(e0) * abs(np.conj(e1))), np.sqrt(abs(e1) * abs(np.conj(e0)))]
    for ex in reals:
        e = residue_eta(ex)
        etal.append(e)
        etar.append(np.conj(e))
        etaa.append(np.sqrt(abs(e) * abs(np.conj(e))))
    for j in range(len(pole)):
        z = -1j * pole[j] * temp
        jsum = complex(num_f(z)) / complex(np.prod(z - allp)) / lead
        expn.append(pole[j] * temp)
        e = -2j * resi[j] * temp * jsum
        etal.append(e)
        etar.append(np.conj(e))
        etaa.append(abs(e))
    return (np.array(etal, np.complex128), np.array(etar, np.complex128), np.array(etaa, np.float64),
            np.array(expn, np.complex128))


# ----------------------------------------------------------------------------- Prony route
def spectrum_exp(w, res, expn, etal, sigma=-1):
    """res += sum_k etal_k / (expn_k + sigma i w), in place and term by term (deom.py:21-24)."""
    for e, x in zip(etal, expn):
        res += e / (x + sigma * 1j * w)


def fit_t(t, res, expn, etal):
    """res += sum_k etal_k exp(-expn_k t), in place and term by term (deom.py:62-66)."""
    for e, x in zip(etal, expn):
        res += e * np.exp(-x * t)
    return res


def sort_symmetry(etal, expn, if_sqrt=True):
    """Order exponentials by decreasing |Im expn| and pair the complex ones (deom.py:37-59): etar of a
    pair is the conjugate of its partner's etal, etaa = |etal| (square-rooted with if_sqrt).  As in
    the reference, etar starts as etal permuted a second time and every entry is then overwritten."""
    order = np.argsort(np.abs(np.imag(expn)))[::-1]
    mags = np.sort(np.abs(np.imag(expn)))[::-1]
    expn = expn[order]
    etal = etal[order]
    etar = etal[order]
    ncc = len(np.where(expn[mags > 1e-10])[0])
    etaa = np.zeros(len(etal), dtype=float)
    for i in range(0, ncc, 2):
        etar[i], etar[i + 1] = np.conj(etal[i + 1]), np.conj(etal[i])
        etaa[i], etaa[i + 1] = np.abs(etal[i]), np.abs(etal[i + 1])
    for i in range(ncc, len(expn)):
        etar[i] = np.conj(etal[i])
        etaa[i] = np.abs(etal[i])
    if if_sqrt:
        etaa = np.sqrt(etaa)
    return etal, etar, etaa, expn


def decompose_spectrum_pade_real(spe, w_sp):
    """Exponentials of one part of a sympy spectral density without the Bose factor
    (deom.py:310-365): its imaginary part if nonzero, else its real part.  etal follows the sorted
    poles (complex pairs by decreasing |Im|, then the rest); expn is returned unsorted, as in the
    reference."""
    import sympy as sp

    re_part, im_part = sp.cancel(spe).as_real_imag()
    part = re_part if im_part == 0 else im_part
    numer, denom = sp.cancel(sp.factor(part)).as_numer_denom()
    roots = [complex(r) for r in sp.nroots(denom)]
    allp = np.array(roots, np.complex128)
    expn = np.array([1j * r for r in roots if r.imag < 0], np.complex128)
    order = np.argsort(np.abs(np.imag(expn)))[::-1]
    mags = np.sort(np.abs(np.imag(expn)))[::-1]
    ccs = expn[order[mags != 0]]
    reals = expn[order[mags == 0]]
    num_f = sp.lambdify(w_sp, numer, "numpy")

    def coefficient(ex):
        p = -1j * ex
        others = allp[np.abs(allp - p) > 1e-14]
        return complex(-1j * complex(num_f(p)) / complex(np.prod(p - others)))

    etal, etar, etaa = [], [], []
    for ii in range(0, len(ccs), 2):
        e0, e1 = coefficient(ccs[ii]), coefficient(ccs[ii + 1])
        etal += [e0, e1]
        etar += [np.conj(e1), np.conj(e0)]
        etaa += [np.sqrt(abs(e0) * abs(np.conj(e1))), np.sqrt(abs(e1) * abs(np.conj(e0)))]
    for ex in reals:
        e = coefficient(ex)
        etal.append(e)
        etar.append(np.conj(e))
        etaa.append(np.sqrt(abs(e) * abs(np.conj(e))))
    return np.array(etal), np.array(etar), np.array(etaa), np.array(expn)


def decompose_spectrum_pade_imag(spe, w_sp):
    """The reference's second copy of decompose_spectrum_pade_real (deom.py:368-424), identical in
    effect."""
    return decompose_spectrum_pade_real(spe, w_sp)


def prony_find_gamma(h, n_sample, nind, *, verbose=False):
    """Prony roots (deom.py:427-444): eigen-decompose the n_sample x n_sample Hankel matrix of h
    (scipy.linalg.eig), rotate each eigenvector by exp(-i arg(sigma)/2), sort by |sigma| descending,
    take the roots of the polynomial whose coefficients are the entries of vector number nind
    (highest degree last) and keep the nind roots of smallest modulus.  verbose=True prints nind, as the
    reference does on every call."""
    import scipy.linalg as la

    mat_h = np.zeros((n_sample, n_sample))
    for i in range(n_sample):
        mat_h[i, :] = h[i:n_sample + i]
    sing_vs, Q = la.eig(mat_h)
    phase_mat = np.diag([np.exp(-1j * np.angle(sv) / 2.0) for sv in sing_vs])
    vs = np.array([np.abs(sv) for sv in sing_vs])
    Qp = np.dot(Q, phase_mat)
    order = np.argsort(vs)[::-1]
    Qp = Qp[:, order]
    if verbose:
        print(nind)
    gamma = np.roots(Qp[:, nind][::-1])
    return gamma[np.argsort(np.abs(gamma))[:nind]]


def prony_fitting(h, t, nind, scale, n, gamma_real=None, gamma_imag=None, *, verbose=False):
    """Least-squares amplitudes for Prony roots (deom.py:447-504): gamma from prony_find_gamma on Re h
    (nind an int) or separately on Re h and Im h (nind = [n_real, n_imag]) unless given; the normal
    equations of the stacked real system; expn = -2 n log(gamma) / scale; then sort_symmetry."""
    import scipy.linalg as la

    if isinstance(nind, list):
        gamma_real = (prony_find_gamma(np.real(h), n, nind[0], verbose=verbose) if gamma_real is None
                      else np.array(gamma_real))
        gamma_imag = (prony_find_gamma(np.imag(h), n, nind[1], verbose=verbose) if gamma_imag is None
                      else np.array(gamma_imag))
        gamma = np.append(gamma_real, gamma_imag)
        n_row = nind[0] + nind[1]
    else:
        gamma = prony_find_gamma(np.real(h), n, nind, verbose=verbose)
        n_row = nind
    t_new = 2 * n * np.log(gamma)
    n_col = len(t)
    gamma_m = np.zeros((2 * n_col, 2 * n_row), dtype=float)
    for i in range(n_row):
        for j in range(n_col):
            gj = gamma[i] ** j
            gamma_m[j, i] = np.real(gj)
            gamma_m[n_col + j, n_row + i] = np.real(gj)
            gamma_m[j, n_row + i] = -np.imag(gj)
            gamma_m[n_col + j, i] = np.imag(gj)
    h_m = np.append(np.real(h), np.imag(h))
    omega = (la.inv(gamma_m.T @ gamma_m) @ gamma_m.T @ h_m).reshape(2, n_row)
    etal_p = omega[0, :] + 1.j * omega[1, :]
    expn_p = -t_new / scale
    return sort_symmetry(etal_p, expn_p)


def decompose_spectrum_prony(spe, w_sp, beta, nind, scale=250000, n=1250, npsd=10, bose_fermi=1, *, verbose=False):
    """Prony decomposition of a sympy spectral density (deom.py:507-542): sample C(t) from its Pade
    decomposition (npsd poles) at scale * linspace(0, 1, 2n + 1) and fit nind exponentials.  With
    nind = ['a', k] (or [k, 'a']) the real (imaginary) part keeps its analytic exponentials; the
    reference ends the program for the Bose real-part case and the Fermi imaginary-part case, and so
    does this function (SystemExit, with the reference's message).  verbose=True prints the first and
    last ten samples and the roots' counts, as the reference always does."""
    etal_pade, _, _, expn_pade = decompose_spectrum_pade(spe, w_sp, beta, npsd, bose_fermi=bose_fermi)
    t = np.linspace(0, 1, 2 * n + 1)
    res_t = np.zeros(len(t), dtype=complex)
    fit_t(scale * t, res_t, expn_pade, etal_pade)
    if verbose:
        print("check the sample points")
        print(res_t[:10])
        print(res_t[-10:])
    if isinstance(nind, list):
        if nind[0] == 'a':
            _, _, _, expn_real = decompose_spectrum_pade_real(spe, w_sp)
            gamma_real = np.exp(-expn_real * scale / (2 * n))
            nind[0] = len(gamma_real)
            if bose_fermi == 1:
                raise SystemExit("For the bose case, C(t) have the analytical imag part")
            return prony_fitting(res_t, t, nind, scale, n, gamma_real=gamma_real, verbose=verbose)
        elif nind[1] == 'a':
            _, _, _, expn_imag = decompose_spectrum_pade_imag(spe, w_sp)
            gamma_imag = np.exp(-expn_imag * scale / (2 * n))
            nind[1] = len(gamma_imag)
            if bose_fermi == 2:
                raise SystemExit("For the fermi case, C(t) have the analytical real part")
            return prony_fitting(res_t, t, nind, scale, n, gamma_imag=gamma_imag, verbose=verbose)
    return prony_fitting(res_t, t, nind, scale, n, verbose=verbose)


def single_oscillator(omega, w_sp, beta, nind):
    """Two-exponential decomposition of one undamped oscillator (deom.py:544-552)."""
    etal = np.array([1 / (2 * (1 - np.exp(-beta * omega))), -1 / (2 * (1 - np.exp(beta * omega)))], complex)
    etar = np.array([-1 / (2 * (1 - np.exp(beta * omega))), 1 / (2 * (1 - np.exp(-beta * omega)))], complex)
    etaa = np.sqrt(np.abs(etal + etar))
    expn = np.array([1j * omega, -1j * omega])
    return etal, etar, etaa, expn


# ----------------------------------------------------------------------------- closed form
def drude(lam, gam, beta, npsd, pade=1):
    """Drude-Lorentz J(w) = 2 lam gam w / (w^2 + gam^2) in closed form (physical normalisation).

    eta_0 = -2i lam gam n(-i gam beta) for the spectral pole gamma, and for each
    Pade pole xi_j: gamma_j = xi_j / beta, eta_j = 4 r_j lam gam nu_j / (beta (nu_j^2 - gam^2)).
    Returns (etal, etar, etaa, expn)."""
    pole, resi = pade_approximation_distribution(npsd, 1, pade)
    eta0 = -2j * lam * gam * function_bose(-1j * gam * beta, pole, resi)
    etal = [eta0]
    expn = [complex(gam)]
    for j in range(len(pole)):
        nu = pole[j] / beta
        etal.append(complex(4.0 * resi[j] * lam * gam * nu / (beta * (nu * nu - gam * gam))))
        expn.append(complex(nu))
    etal = np.array(etal, np.complex128)
    return etal, np.conj(etal), np.abs(etal), np.array(expn, np.complex128)


def brownian(lam, omega0, gam, beta, npsd, pade=1):
    """Brownian oscillator J(w) = 2 lam w0^2 gam w / ((w0^2 - w^2)^2 + gam^2 w^2) in closed form
    (physical normalisation), the decomposition decompose_spectrum_pade makes of it (deom.py:226-307)
    without sympy.

    The four poles of J are +-Omega +- i gam/2 with Omega = sqrt(w0^2 - gam^2/4) (underdamped), or
    +-i kappa_1,2 with kappa = gam/2 -+ sqrt(gam^2/4 - w0^2) (overdamped).  Each pole p in the lower half
    plane gives the exponent i p and the amplitude -2i N(p) / prod_{q != p} (p - q) n_B(beta p), with
    N(w) = 2 lam w0^2 gam w and n_B the Pade Bose function; each Pade pole xi_j gives the exponent
    xi_j / beta and the amplitude -2i r_j / beta J(-i xi_j / beta) (the pole sum of the reference).
    Order as the reference lists them: the complex pair (gam/2 + i Omega, gam/2 - i Omega), whose etar
    are the conjugates of each other's etal, or the real exponents (kappa_2 >= kappa_1), then the
    Pade terms.  The critically damped case (gam = 2 w0, a double pole) is rejected.
    Returns (etal, etar, etaa, expn)."""
    lam, w0, g = float(lam), float(omega0), float(gam)
    if not (w0 > 0 and g > 0):
        raise ValueError("brownian: omega0 and gamma must be positive")
    disc = w0 * w0 - g * g / 4
    if abs(disc) <= 1e-12 * w0 * w0:
        raise ValueError("brownian: critically damped spectrum (gamma = 2 omega0) has a double pole")
    if disc > 0:
        om = math.sqrt(disc)
        lower = [complex(om, -g / 2), complex(-om, -g / 2)]   # exponents g/2 + i om, g/2 - i om
        poles = np.array([complex(om, g / 2), complex(-om, g / 2)] + lower)
    else:
        r = math.sqrt(-disc)
        k1, k2 = g / 2 - r, g / 2 + r
        lower = [complex(0, -k2), complex(0, -k1)]              # exponents k2 >= k1
        poles = np.array([complex(0, k1), complex(0, k2)] + lower)
    pole, resi = pade_approximation_distribution(npsd, 1, pade)

    def numer(w):
        return 2.0 * lam * w0 * w0 * g * w

    etal = []
    for p in lower:
        others = poles[np.abs(poles - p) > 1e-14 * max(1.0, abs(p))]
        etal.append(-2j * numer(p) / complex(np.prod(p - others)) * function_bose(p * beta, pole, resi))
    expn = [1j * p for p in lower]
    if disc > 0:
        etar = [np.conj(etal[1]), np.conj(etal[0])]
        etaa = [math.sqrt(abs(etal[0]) * abs(etal[1]))] * 2
    else:
        etar = [np.conj(e) for e in etal]
        etaa = [abs(e) for e in etal]
    for j in range(len(pole)):
        z = -1j * pole[j] / beta
        e = -2j * resi[j] / beta * numer(z) / complex(np.prod(z - poles))
        expn.append(complex(pole[j] / beta))
        etal.append(e)
        etar.append(np.conj(e))
        etaa.append(abs(e))
    return (np.array(etal, np.complex128), np.array(etar, np.complex128), np.array(etaa, np.float64),
            np.array(expn, np.complex128))


class Bath:
    """Bath of one or more spectral densities, drop-in for pyqed.heom.deom.Bath (deom.py:895-942).

    Bath(spectrum_sp: list[sympy], w_sp, beta: list, npsd: list, mode: list[int], function=None)
    concatenates the decomposition of every spectral density; ``mode[k]`` names
    the coupling operator of exponential k."""

    def __init__(self, spectrum_sp=None, w_sp=None, beta=None, npsd=None, mode=None, function=None, *,
                 monic=False):
        self.bath = spectrum_sp
        self.w_sp = w_sp
        self.beta = beta
        self.npsd = npsd
        self.mode = mode
        if spectrum_sp is None:
            self.etal = self.etar = self.etaa = self.expn = np.zeros(0, np.complex128)
            return
        if function is None:
            function = [decompose_spectrum_pade] * len(spectrum_sp)
        if isinstance(spectrum_sp, list) and isinstance(beta, list) and isinstance(npsd, list):
            if len(spectrum_sp) != len(beta) or len(spectrum_sp) != len(npsd):
                import warnings
                warnings.warn("the length of bath, w_sp, beta, npsd is not equal!")
            parts = []
            for i in range(len(spectrum_sp)):
                f = function[i]
                if f is decompose_spectrum_pade:
                    parts.append(f(spectrum_sp[i], w_sp, beta[i], npsd[i], monic=monic))
                else:
                    parts.append(f(spectrum_sp[i], w_sp, beta[i], npsd[i]))
            self.etal = np.concatenate([np.asarray(p[0]) for p in parts]).astype(np.complex128)
            self.etar = np.concatenate([np.asarray(p[1]) for p in parts]).astype(np.complex128)
            self.etaa = np.concatenate([np.asarray(p[2]) for p in parts]).astype(np.complex128)
            self.expn = np.concatenate([np.asarray(p[3]) for p in parts]).astype(np.complex128)
            if self.mode is None:
                raise ValueError("mode is not set")
            if len(self.mode) != len(self.expn):
                raise ValueError("the length of mode is not equal to the number of dissipatons")
        else:
            # The reference's single-expression branch is unreachable (len() of a sympy
            # expression raises before it); here a single expression is accepted.
            self.etal, self.etar, self.etaa, self.expn = decompose_spectrum_pade(spectrum_sp, w_sp, beta, npsd,
                                                                                  monic=monic)
            self.etaa = self.etaa.astype(np.complex128)
            self.mode = np.zeros(len(self.expn), np.int64)

    @classmethod
    def from_arrays(cls, etal, etar, etaa, expn, mode):
        b = cls()
        b.etal = np.asarray(etal, np.complex128)
        b.etar = np.asarray(etar, np.complex128)
        b.etaa = np.asarray(etaa, np.complex128)
        b.expn = np.asarray(expn, np.complex128)
        b.mode = np.asarray(mode, np.int64)
        return b

    @classmethod
    def brownian(cls, lam, omega0, gam, beta, npsd, mode=None, nbath=1):
        """nbath identical Brownian-oscillator baths in closed form; mode defaults to bath i -> coupling i."""
        e = brownian(lam, omega0, gam, beta, npsd)
        k = len(e[3])
        etal, etar, etaa, expn = (np.tile(a, nbath) for a in e)
        if mode is None:
            mode = np.repeat(np.arange(nbath), k)
        return cls.from_arrays(etal, etar, etaa, expn, mode)

    @classmethod
    def drude(cls, lam, gam, beta, npsd, mode=None, nbath=1):
        """nbath identical Drude-Lorentz baths in closed form; mode defaults to bath i -> coupling i."""
        e = drude(lam, gam, beta, npsd)
        k = len(e[3])
        etal, etar, etaa, expn = (np.tile(a, nbath) for a in e)
        if mode is None:
            mode = np.repeat(np.arange(nbath), k)
        return cls.from_arrays(etal, etar, etaa, expn, mode)
