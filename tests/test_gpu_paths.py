"""Every HEOM kernel path against the CPU oracle on a random full hierarchy state.

One RK4 step from a random (non-physical, all-ADO) state exercises every term of the operator at
non-trivial values: the commutator, the damping, and both neighbour terms of every ADO.
Paths: generic kernel (any diagonal coupling), site-projector kernel with VALU commutator
(d = 6..11) and with FP32 MFMA commutator (one 16 x 16 tile for d = 12..16, 2 x 2 tiles for
d = 17..32; mixed precision only, FP64 runs these classes on the VALU form).
Tolerances (max-abs relative to max |state|): FP64 1e-12; mixed (FP32 right-hand side) 1e-5 per
single step from a random state, whose RHS is O(1) everywhere."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _instance(d, L, seed, weight=1.0, diag_q=False):
    from pyqed_amd.heom.bath import drude
    rng = np.random.default_rng(seed)
    A = rng.normal(size=(d, d)) + 1j * rng.normal(size=(d, d))
    H = (A + A.conj().T) / 4
    Q = np.zeros((d, d, d), complex)
    for j in range(d):
        if diag_q:
            Q[j] = np.diag(rng.normal(size=d))
        else:
            Q[j, j, j] = weight
    e = drude(0.3, 0.8, 1.2, 1)
    bath = tuple(np.tile(x, d) for x in e) + (np.repeat(np.arange(d), 2),)
    rho0 = np.zeros((d, d), complex)
    rho0[0, 0] = 1
    return dict(H=H, Q=Q, L=L, rho0=rho0, bath=bath)


def _oracle_step(inst, state, dt):
    from oracle import heom as O
    etal, etar, etaa, expn, mode = inst["bath"]
    bath = (etal, etar, etaa, expn)
    hier = O.Hierarchy(len(mode), inst["L"])
    return O.rk4_step(state, lambda x, t: O.rhs_vec(x, inst["H"], list(inst["Q"]), hier, bath, mode), dt, 0.0)


@pytest.mark.parametrize("d,L", [(6, 2), (9, 2), (12, 2), (13, 2), (16, 2), (17, 2), (20, 1), (24, 2), (31, 1)])
@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_paths_one_step_random_state(d, L, prec):
    from pyqed_amd.heom.engine import HeomEngine
    inst = _instance(d, L, seed=d * 10 + L, weight=0.7)
    eng = HeomEngine([inst], resident=False, hermitian=False)
    rng = np.random.default_rng(d)
    nmax = int(eng.nmax[0])
    state = (rng.normal(size=(nmax, d, d)) + 1j * rng.normal(size=(nmax, d, d))) * 0.1
    dt = 0.05
    ref = _oracle_step(inst, state, dt)
    tol = 1e-12 if prec == "fp64" else 1e-5
    for generic in (False, True):
        eng = HeomEngine([inst], resident=False, generic=generic, hermitian=False)
        cls = eng.classes()
        eng.set_state([state])
        eng.advance(1, dt, prec)
        got = eng.hierarchy(0)
        err = np.abs(got - ref).max() / np.abs(ref).max()
        print(f"d={d} L={L} {prec} generic={generic} classes={cls['items']} err={err:.2e}")
        assert err <= tol
        if not generic:
            # d <= 16: one of the two projector classes (VALU or MFMA commutator; the boundary is a
            # build option); d > 16: the wide class
            proj = cls["items"][3] if d > 16 else cls["items"][1] + cls["items"][2]
            assert proj > 0 and cls["items"][0] == 0


@pytest.mark.parametrize("d,L", [(7, 2), (12, 2), (16, 2), (19, 2), (24, 2), (31, 2)])
@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_hermitian_mode_random_hermitian_state(d, L, prec):
    """Hermitian mode reads column terms as conjugated rows; exact for a Hermitian hierarchy."""
    from pyqed_amd.heom.engine import HeomEngine
    inst = _instance(d, L, seed=d + 7, weight=0.9)
    rng = np.random.default_rng(d + 1)
    probe = HeomEngine([inst], resident=False)
    assert probe.hermitian
    nmax = int(probe.nmax[0])
    A = (rng.normal(size=(nmax, d, d)) + 1j * rng.normal(size=(nmax, d, d))) * 0.1
    state = A + np.conj(np.transpose(A, (0, 2, 1)))
    ref = _oracle_step(inst, state, 0.05)
    tol = 1e-12 if prec == "fp64" else 1e-5
    for herm in (True, False):
        eng = HeomEngine([inst], resident=False, hermitian=herm)
        eng.set_state([state])
        eng.advance(1, 0.05, prec)
        err = np.abs(eng.hierarchy(0) - ref).max() / np.abs(ref).max()
        print(f"d={d} {prec} hermitian={herm} err={err:.2e}")
        assert err <= tol


@pytest.mark.parametrize("d", [9, 14, 21])
@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_unsorted_site_map(d, prec):
    """Exponentials whose sites are not in increasing order of k: the projector kernels sort each
    neighbour list by site (the in-order case only counts the site groups)."""
    from pyqed_amd.heom.engine import HeomEngine
    inst = _instance(d, 2, seed=d + 31, weight=0.8)
    etal, etar, etaa, expn, mode = inst["bath"]
    perm = np.random.default_rng(d).permutation(d)
    inst["bath"] = (etal, etar, etaa, expn, perm[mode])
    assert np.any(np.diff(perm[mode]) < 0)
    rng = np.random.default_rng(d + 2)
    nmax = int(HeomEngine([inst], resident=False).nmax[0])
    A = (rng.normal(size=(nmax, d, d)) + 1j * rng.normal(size=(nmax, d, d))) * 0.1
    state = A + np.conj(np.transpose(A, (0, 2, 1)))
    ref = _oracle_step(inst, state, 0.05)
    tol = 1e-12 if prec == "fp64" else 1e-5
    for herm in (True, False):
        eng = HeomEngine([inst], resident=False, hermitian=herm)
        assert eng.classes()["items"][0] == 0
        eng.set_state([state])
        eng.advance(1, 0.05, prec)
        err = np.abs(eng.hierarchy(0) - ref).max() / np.abs(ref).max()
        print(f"d={d} {prec} hermitian={herm} err={err:.2e}")
        assert err <= tol


def test_hermitian_flag_rejected_for_complex_exponents():
    from pyqed_amd import _lib
    from pyqed_amd.heom.engine import HeomEngine
    inst = _instance(6, 1, seed=1)
    etal, etar, etaa, expn, mode = inst["bath"]
    inst["bath"] = (etal, etar, etaa, expn + 0.1j, mode)
    with pytest.raises(_lib.PyqedError, match="HERMITIAN"):
        HeomEngine([inst], hermitian=True)
    assert not HeomEngine([inst]).hermitian


def test_general_diagonal_couplings_use_generic_path():
    from pyqed_amd.heom.engine import HeomEngine
    d, L = 7, 2
    inst = _instance(d, L, seed=3, diag_q=True)
    eng = HeomEngine([inst], resident=False, hermitian=False)
    assert eng.classes()["items"][0] > 0 and sum(eng.classes()["items"][1:]) == 0
    rng = np.random.default_rng(5)
    nmax = int(eng.nmax[0])
    state = (rng.normal(size=(nmax, d, d)) + 1j * rng.normal(size=(nmax, d, d))) * 0.1
    ref = _oracle_step(inst, state, 0.05)
    eng.set_state([state])
    eng.advance(1, 0.05, "fp64")
    assert np.abs(eng.hierarchy(0) - ref).max() <= 1e-12 * np.abs(ref).max()


def test_mixed_batch_of_all_paths():
    """One plan holding instances of every class (and resident ones), 50 steps, against FP64 per
    instance runs of the generic kernel."""
    from pyqed_amd.heom.engine import HeomEngine
    insts = [_instance(d, 2, seed=100 + d) for d in (2, 3, 5, 6, 8, 12, 14, 16, 18)]
    eng = HeomEngine(insts)
    out = eng.run(0.02, 50, precision="mixed")
    for k, inst in enumerate(insts):
        ref = HeomEngine([inst], generic=True, resident=False).run(0.02, 50, precision="fp64")[0]
        err = np.abs(out[k] - ref).max()
        print(f"instance d={inst['H'].shape[0]}: mixed batched vs fp64 generic {err:.2e}")
        assert err <= 1e-6


@pytest.mark.parametrize("stage_prec", ["mixed"])
def test_pipelined_variant_bitwise_equal(stage_prec):
    """The two-deep pipelined kernel computes the same arithmetic in the same order as the base
    projector kernel: identical states after several steps, across many work items."""
    import torch
    from pyqed_amd.heom.engine import HeomEngine
    insts = [_instance(d, L, seed=200 + d) for d in (9, 12, 14, 16) for L in (2, 3)]
    outs = []
    for variant in (0, 1):
        eng = HeomEngine(insts, resident=False, packed=False)  # the pipelined kernel has no packed form
        eng.autotune_cache = False
        eng.set_tiling(1024, 4, variant)
        assert eng.hermitian and eng.classes()["items"][2] > 0
        eng.advance(3, 0.02, stage_prec)
        torch.cuda.synchronize()
        outs.append(eng.y.clone())
    assert torch.equal(outs[0], outs[1])


@pytest.mark.parametrize("d", [8, 16])
def test_hermitian_mode_conserves_trace_like_exact_path(d):
    """Regression: in Hermitian mode the diagonal element must use X'(s, s) itself; conj(X'(s, s))
    turns the rounding-level imaginary part of a neighbour's diagonal into a real contribution to
    Tr rho_0 (c_L + c_R = 0 for the n + e_k terms only with the unconjugated value)."""
    from pyqed_amd.heom.engine import HeomEngine
    inst = _instance(d, 2, seed=d + 40, weight=1.0)
    rng = np.random.default_rng(d + 3)
    probe = HeomEngine([inst], resident=False)
    nmax = int(probe.nmax[0])
    A = (rng.normal(size=(nmax, d, d)) + 1j * rng.normal(size=(nmax, d, d))) * 0.1
    state = A + np.conj(np.transpose(A, (0, 2, 1)))
    state[:, np.arange(d), np.arange(d)] += 1e-3j  # a non-Hermitian residue on every diagonal
    traces = {}
    for herm in (True, False):
        eng = HeomEngine([inst], resident=False, hermitian=herm)
        eng.set_state([state])
        eng.advance(1, 0.05, "fp64")
        traces[herm] = np.trace(eng.hierarchy(0)[0])
    assert abs(traces[True] - traces[False]) <= 1e-13, traces


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_neighbour_entry_limit(prec):
    """K + min(K, L) = 64 neighbour entries per ADO is the projector path's limit (d = 31, K = 62,
    L = 2: the wide class at its largest row buffer); one site more (d = 32, K = 64, L = 2: 66 entries)
    takes the generic kernel.  Both match the oracle's RK4 step from a random full hierarchy state."""
    from pyqed_amd.heom.engine import HeomEngine
    tol = 1e-12 if prec == "fp64" else 1e-5
    for d, L, want in ((31, 2, 3), (32, 2, 0)):
        inst = _instance(d, L, seed=900 + d, weight=0.7)
        eng = HeomEngine([inst], resident=False, hermitian=False)
        items = eng.classes()["items"]
        assert items[want] > 0 and sum(items) == items[want], (d, items)
        rng = np.random.default_rng(d)
        nmax = int(eng.nmax[0])
        state = (rng.normal(size=(nmax, d, d)) + 1j * rng.normal(size=(nmax, d, d))) * 0.1
        ref = _oracle_step(inst, state, 0.05)
        eng.set_state([state])
        eng.advance(1, 0.05, prec)
        err = np.abs(eng.hierarchy(0) - ref).max() / np.abs(ref).max()
        print(f"d={d} L={L} {prec} classes={items} err={err:.2e}")
        assert err <= tol
