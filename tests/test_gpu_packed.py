"""Packed Hermitian state (PYQED_PLAN_PACKED): y and ynew hold upper triangles, the stage inputs full
matrices.  Every kernel class (generic, projector with VALU and MFMA commutator, LDS-resident) is
checked against the CPU oracle from a random Hermitian hierarchy state and against the full-matrix
plan over many steps.  Tolerances as in test_gpu_paths: FP64 1e-12, mixed 1e-5 per step from a random
state (max-abs relative); trajectories 1e-12 (FP64) and 1e-6 (mixed: the full-matrix plan evolves the
lower triangle with its own FP32 right-hand side, which drifts from the conjugate of the upper one)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _inst(d, L, seed, diag_q=False):
    from pyqed_amd.heom.bath import drude
    rng = np.random.default_rng(seed)
    A = rng.normal(size=(d, d)) + 1j * rng.normal(size=(d, d))
    Q = np.zeros((d, d, d), complex)
    for j in range(d):
        Q[j] = np.diag(rng.normal(size=d)) if diag_q else np.diag(np.eye(d)[j] * 0.8)
    e = drude(0.3, 0.8, 1.2, 1)
    bath = tuple(np.tile(x, d) for x in e) + (np.repeat(np.arange(d), 2),)
    rho0 = np.zeros((d, d), complex)
    rho0[0, 0] = 0.7
    rho0[-1, -1] = 0.3
    rho0[0, -1] = rho0[-1, 0] = 0.2
    return dict(H=(A + A.conj().T) / 4, Q=Q, L=L, rho0=rho0, bath=bath)


def _oracle_step(inst, state, dt):
    from oracle import heom as O
    etal, etar, etaa, expn, mode = inst["bath"]
    hier = O.Hierarchy(len(mode), inst["L"])
    return O.rk4_step(state, lambda x, t: O.rhs_vec(x, inst["H"], list(inst["Q"]), hier, (etal, etar, etaa, expn),
                                                    mode), dt, 0.0)


@pytest.mark.parametrize("d,diag_q", [(5, True), (7, False), (13, False), (16, False), (17, False), (20, False),
                                      (24, False), (31, False)])
@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_packed_one_step_random_hermitian_state(d, diag_q, prec):
    from pyqed_amd.heom.engine import HeomEngine
    inst = _inst(d, 2, seed=d + 5, diag_q=diag_q)
    eng = HeomEngine([inst], resident=False)
    assert eng.hermitian and eng.packed
    nmax = int(eng.nmax[0])
    rng = np.random.default_rng(d)
    A = (rng.normal(size=(nmax, d, d)) + 1j * rng.normal(size=(nmax, d, d))) * 0.1
    state = A + np.conj(np.transpose(A, (0, 2, 1)))
    ref = _oracle_step(inst, state, 0.05)
    eng.set_state([state])
    eng.advance(1, 0.05, prec)
    got = eng.hierarchy(0)
    err = np.abs(got - ref).max() / np.abs(ref).max()
    print(f"d={d} diag_q={diag_q} {prec} classes={eng.classes()['items']} err={err:.2e}")
    assert err <= (1e-12 if prec == "fp64" else 1e-5)


@pytest.mark.parametrize("prec,tol", [("fp64", 1e-12), ("mixed", 1e-6)])
def test_packed_batch_matches_full_matrix_plan(prec, tol):
    """One plan with instances of every class, LDS-resident ones included, 60 steps saved every 7:
    packed and full-matrix plans give the same trajectories and final hierarchies."""
    from pyqed_amd.heom.engine import HeomEngine
    insts = [_inst(d, 2, seed=40 + d) for d in (2, 3, 6, 9, 12, 14, 16, 18, 25)] + [_inst(7, 2, seed=3, diag_q=True)]
    a = HeomEngine(insts)
    b = HeomEngine(insts, packed=False)
    assert a.packed and not b.packed and a.hermitian and b.hermitian
    assert a.stats()["resident_instances"] > 0
    ta = a.run(0.02, 60, precision=prec, save_every=7)
    tb = b.run(0.02, 60, precision=prec, save_every=7)
    for k in range(len(insts)):
        e = np.abs(ta[k] - tb[k]).max()
        h = np.abs(a.hierarchy(k) - b.hierarchy(k)).max()
        d = ta[k].shape[-1]
        off = ~np.eye(d, dtype=bool)
        herm = np.abs((ta[k] - np.conj(np.transpose(ta[k], (0, 2, 1))))[:, off]).max() if d > 1 else 0.0
        print(f"instance {k} d={insts[k]['H'].shape[0]}: traj {e:.2e} hierarchy {h:.2e} hermiticity {herm:.1e}")
        assert e <= tol and h <= tol
        assert herm == 0.0  # packed trajectories: (j, i) is written as the conjugate of (i, j)


def test_packed_needs_hermitian_and_sizes():
    from pyqed_amd import _lib
    from pyqed_amd.heom.engine import HeomEngine
    inst = _inst(6, 2, seed=9)
    with pytest.raises(_lib.PyqedError, match="PACKED"):
        HeomEngine([inst], hermitian=False, packed=True)
    a = HeomEngine([inst], resident=False)
    b = HeomEngine([inst], resident=False, packed=False)
    nmax = int(a.nmax[0])
    assert a.y.numel() == nmax * 21 and b.y.numel() == nmax * 36
    mixed = _lib.PREC_MIXED
    assert a.ws_bytes[mixed] < b.ws_bytes[mixed]
    small = a.torch.empty(a.ws_bytes[mixed] - 256, dtype=a.torch.uint8, device=a.device)
    with pytest.raises(_lib.PyqedError, match="workspace too small"):
        _lib.check(a.L.pyqed_heom_run(a.plan, mixed, a.y.data_ptr(), small.data_ptr(), small.numel(), None, 0, 1,
                                      0.01, None, 1, _lib.stream_ptr(a.torch)))


def test_launch_timing_counts_every_stage_launch():
    from pyqed_amd.heom.engine import HeomEngine
    insts = [_inst(d, 2, seed=70 + d) for d in (6, 14, 20)]
    eng = HeomEngine(insts, resident=False)
    eng.set_timing(True)
    eng.advance(3, 0.01, "mixed")
    ms, n = eng.read_timing()
    eng.set_timing(False)
    items = eng.classes()["items"]
    nc = len(items)
    for s in range(4):
        for c in range(nc):
            assert n[nc * s + c] == (3 if items[c] else 0)
            assert (ms[nc * s + c] > 0) == bool(items[c])
    eng.advance(1, 0.01, "mixed")
    assert sum(eng.read_timing()[1]) == 0  # recording stopped


@pytest.mark.parametrize("d,L", [(17, 2), (24, 2), (31, 1)])
def test_wide_block_variant(d, L):
    """The block-cooperative wide kernel (variant bit 1: one ADO per block, one MFMA tile per wavefront)
    against the oracle from a random Hermitian state, and against the default wavefront form over 20
    steps (packed mixed runs)."""
    from pyqed_amd.heom.engine import HeomEngine
    inst = _inst(d, L, seed=d + 11)
    rng = np.random.default_rng(d)
    probe = HeomEngine([inst], resident=False)
    assert probe.packed and probe.classes()["items"][3] > 0
    nmax = int(probe.nmax[0])
    A = (rng.normal(size=(nmax, d, d)) + 1j * rng.normal(size=(nmax, d, d))) * 0.1
    state = A + np.conj(np.transpose(A, (0, 2, 1)))
    ref = _oracle_step(inst, state, 0.05)
    outs = {}
    for variant in (0, 2):
        eng = HeomEngine([inst], resident=False)
        eng.autotune_cache = False
        eng.set_tiling(2048, 8, variant)
        eng.set_state([state])
        eng.advance(1, 0.05, "mixed")
        err = np.abs(eng.hierarchy(0) - ref).max() / np.abs(ref).max()
        print(f"d={d} variant {variant}: one step vs oracle {err:.2e}")
        assert err <= 1e-5
        e2 = HeomEngine([inst], resident=False)
        e2.autotune_cache = False
        e2.set_tiling(2048, 8, variant)
        outs[variant] = e2.run(0.02, 20, precision="mixed")[0]
    assert np.abs(outs[0] - outs[2]).max() <= 1e-6


@pytest.mark.parametrize("d,L", [(7, 3), (20, 2)])
def test_single_instance_form_matches_table_form(d, L):
    """A projector class holding one hierarchy runs the single-instance kernel (record from the kernel
    arguments, computed item ranges, fixed-stride neighbour lists); beside a second hierarchy of the
    same class it runs the table-driven kernel.  Every ADO's arithmetic is the same in both, so the
    trajectories agree bit for bit."""
    from pyqed_amd.heom.engine import HeomEngine
    a, b = _inst(d, L, seed=400 + d), _inst(d, L, seed=500 + d)
    cls = 1 if d <= 11 else 3
    one = HeomEngine([a], resident=False)
    two = HeomEngine([a, b], resident=False)
    assert one.packed and two.packed
    assert one.class_single()[cls] == 0 and two.class_single()[cls] == -1
    ra = one.run(0.01, 20, "mixed", save_every=5)[0]
    rb = two.run(0.01, 20, "mixed", save_every=5)[0]
    assert np.array_equal(ra, rb)


def test_packed_wide_form_at_the_entry_limit():
    """The packed wide MFMA form (single-instance kernel, neighbour rows by LDS DMA) at the projector
    path's 64-entry limit (d = 31, K = 62, depth 2): 20 mixed-precision steps against FP64."""
    from pyqed_amd.heom.engine import HeomEngine
    inst = _inst(31, 2, seed=931)
    eng = HeomEngine([inst], resident=False)
    assert eng.packed and eng.classes()["items"][3] > 0 and eng.class_single()[3] == 0
    mix = eng.run(0.01, 20, "mixed", save_every=5)[0]
    f64 = HeomEngine([inst], resident=False).run(0.01, 20, "fp64", save_every=5)[0]
    err = np.abs(mix - f64).max()
    tr = np.abs(np.trace(mix, axis1=1, axis2=2) - 1).max()
    print(f"d=31 depth 2 packed mixed vs fp64 {err:.2e}, trace {tr:.2e}")
    assert err <= 1e-6 and tr <= 1e-6
