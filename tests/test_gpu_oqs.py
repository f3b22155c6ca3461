"""Lindblad / Redfield propagation and Lindblad adjoint gradients on the GPU against the oracle.

Tolerances: propagation 1e-12 max-abs (FP64 throughout); adjoint gradients within 1e-10 of FP64
autograd (same discretisation) and within 1e-5 relative of central finite differences (h = 1e-6)."""
import numpy as np
import pytest

from oracle import lindblad as OL

pytestmark = pytest.mark.gpu


def _rand_herm(d, rng):
    A = rng.normal(size=(d, d)) + 1j * rng.normal(size=(d, d))
    return (A + A.conj().T) / 2


def _system(d=3, seed=0):
    rng = np.random.default_rng(seed)
    H = _rand_herm(d, rng)
    c = [0.3 * (rng.normal(size=(d, d)) + 1j * rng.normal(size=(d, d))) for _ in range(2)]
    rho0 = np.diag(rng.dirichlet(np.ones(d))).astype(complex)
    e_ops = [np.diag(np.eye(d)[j]).astype(complex) for j in range(d)] + [_rand_herm(d, rng)]
    return H, c, rho0, e_ops


@pytest.mark.parametrize("d", [2, 3, 8])
def test_lindblad_static(d):
    from pyqed_amd.oqs import LindbladSolver
    H, c, rho0, e_ops = _system(d)
    res = LindbladSolver(H, c).run(rho0, 0.01, 500, e_ops=e_ops)
    ref = OL.run(H, c, rho0, 0.01, 500, e_ops)
    assert res.observables.shape == (501, len(e_ops))
    assert np.abs(res.observables - ref["observables"]).max() < 1e-12
    assert len(res.rholist) == 500 and np.abs(np.array(res.rholist) - np.array(ref["rholist"])).max() < 1e-12
    assert np.allclose(res.times, np.arange(501) * 0.01)


@pytest.mark.parametrize("legacy", [False, True])
def test_lindblad_driven(legacy):
    from pyqed_amd.oqs import LindbladSolver
    H, c, rho0, e_ops = _system(3, seed=1)
    H1 = _rand_herm(3, np.random.default_rng(2))
    f = lambda t: 0.3 * np.cos(1.3 * t)
    res = LindbladSolver([H, [H1, f]], c).run(rho0, 0.02, 300, e_ops=e_ops, legacy_accumulate=legacy)
    ref = OL.run_driven([H, [H1, f]], rho0, c, 0.02, 300, e_ops=e_ops, legacy=legacy)
    assert res.observables.shape == (300, len(e_ops))
    assert np.abs(res.observables - ref["observables"]).max() < 1e-12


@pytest.mark.parametrize("d,nfield", [(2, 1), (3, 2), (2, 3)])
def test_lindblad_driven_field_counts(d, nfield):
    # one field: the compile-time N = 4 / 9 kernels keep the generator rows in registers; more than
    # one field takes the general kernel
    from pyqed_amd.oqs import LindbladSolver
    print(f"[case] driven Lindblad d={d} fields={nfield}", flush=True)
    H, c, rho0, e_ops = _system(d, seed=5 + d)
    rng = np.random.default_rng(6)
    terms = [[_rand_herm(d, rng), (lambda t, w=1.0 + 0.4 * j: 0.2 * np.cos(w * t))] for j in range(nfield)]
    res = LindbladSolver([H] + terms, c).run(rho0, 0.02, 250, e_ops=e_ops)
    ref = OL.run_driven([H] + terms, rho0, c, 0.02, 250, e_ops=e_ops)
    assert np.abs(res.observables - ref["observables"]).max() < 1e-12


def test_lindblad_batch_of_initial_states():
    from pyqed_amd.oqs import LindbladSolver
    H, c, _, _ = _system(4, seed=3)
    rng = np.random.default_rng(4)
    rho0s = []
    for _ in range(37):
        v = rng.normal(size=4) + 1j * rng.normal(size=4)
        rho0s.append(np.outer(v, v.conj()) / np.vdot(v, v).real)
    final = LindbladSolver(H, c).run_batch(np.array(rho0s), 0.01, 200)
    for k in range(37):
        ref = OL.run(H, c, rho0s[k], 0.01, 200)["rholist"][-1]
        assert np.abs(final[k] - ref).max() < 1e-12


def test_redfield_evolve():
    from pyqed_amd.oqs import RedfieldSolver
    rng = np.random.default_rng(5)
    d = 4
    H = np.diag([0.0, 0.5, 1.1, 1.9]) + 0.1 * _rand_herm(d, rng)
    a_ops = [_rand_herm(d, rng), np.diag([1.0, -1, 0.5, 0]).astype(complex)]
    S = lambda w: 0.05 * w / (1 - np.exp(-2.0 * w)) if w != 0 else 0.025
    rho0 = np.diag([0.1, 0.2, 0.3, 0.4]).astype(complex)
    e_ops = [np.diag(np.eye(d)[j]).astype(complex) for j in range(d)]
    res = RedfieldSolver(H, a_ops, [S, S]).evolve(rho0, 0.05, 400, e_ops=e_ops)
    R, ev = OL.redfield_tensor(H, a_ops, [S, S])
    ref = OL.run_redfield(R, rho0, ev, 0.05, 400, e_ops)
    assert res.observables.shape == (400, d)
    assert np.abs(res.observables - ref["observables"]).max() < 1e-12
    assert np.abs(np.array(res.rholist) - np.array(ref["rholist"])).max() < 1e-12


def _c4():
    from pyqed_amd import workloads as W
    from pyqed_amd.adjoint import lambda_system
    w = W.lambda_c4()
    prob = lambda_system(w["eps"], w["mu"], list(w["c_ops"]), w["dt"], w["steps_per_seg"], w["nseg"],
                         w["checkpoint_every"])
    return w, prob


def test_c4_adjoint_against_autograd_and_checkpointing():
    import torch
    from oracle.lindblad_torch import final_state as torch_final
    from pyqed_amd.adjoint import final_state
    w, prob = _c4()
    theta = torch.tensor(w["eps"], dtype=torch.float64, requires_grad=True)
    amps = torch.tensor(w["amps"][None], dtype=torch.float64, requires_grad=True)
    rT = final_state(theta, amps, prob, w["rho0"])
    J = rT[2, 2, 0]
    J.backward()
    t2 = theta.detach().clone().requires_grad_(True)
    a2 = amps.detach().clone().requires_grad_(True)
    ref = torch_final(t2, a2, prob.base, prob.G, prob.B, w["rho0"], w["dt"], w["steps_per_seg"])
    assert abs(ref[2, 2].real.item() - J.item()) < 1e-13
    ref[2, 2].real.backward()
    gscale = max(a2.grad.abs().max().item(), t2.grad.abs().max().item())
    print(f"C4 J = {J.item():.6e}; max |grad| = {gscale:.3e}")
    assert (theta.grad - t2.grad).abs().max().item() <= 1e-10 * gscale
    assert (amps.grad - a2.grad).abs().max().item() <= 1e-10 * gscale
    # the result does not depend on the checkpoint interval
    bar = np.zeros((3, 3), complex)
    bar[2, 2] = 1.0
    gt50, ga50, _ = prob.gradient(w["eps"], w["amps"][None], w["rho0"], bar)
    for ce in (1, 7, 4000):
        prob.ckpt_every = ce
        gt, ga, _ = prob.gradient(w["eps"], w["amps"][None], w["rho0"], bar)
        assert np.abs(gt - gt50).max() <= 1e-13 * gscale and np.abs(ga - ga50).max() <= 1e-13 * gscale
    prob.ckpt_every = 50


def test_c4_adjoint_against_central_differences():
    w, prob = _c4()
    bar = np.zeros((3, 3), complex)
    bar[2, 2] = 1.0
    gt, ga, _ = prob.gradient(w["eps"], w["amps"][None], w["rho0"], bar)
    h = 1e-6
    J = lambda th, am: prob.forward(th, am, w["rho0"])[2, 2].real
    fd_t = np.array([(J(w["eps"] + h * np.eye(3)[j], w["amps"][None]) - J(w["eps"] - h * np.eye(3)[j], w["amps"][None]))
                     / (2 * h) for j in range(3)])
    fd_a = np.zeros(w["nseg"])
    for s in range(w["nseg"]):
        e = np.zeros(w["nseg"])
        e[s] = h
        fd_a[s] = (J(w["eps"], (w["amps"] + e)[None]) - J(w["eps"], (w["amps"] - e)[None])) / (2 * h)
    g = np.concatenate([gt, ga[0]])
    fd = np.concatenate([fd_t, fd_a])
    rel = np.abs(g - fd) / np.maximum(np.abs(fd), 1e-3 * np.abs(fd).max())
    print(f"C4 adjoint vs central FD (h=1e-6): max relative error {rel.max():.2e} over {g.size} parameters")
    assert rel.max() <= 1e-5


def test_heom_adjoint_small_hierarchy():
    """HEOM gradients (spin-boson, depth 3: 10 ADOs x 4) w.r.t. H parameters and a driving field,
    from the adjoint kernel on the dense HEOM generator, against FP64 autograd and central FD."""
    import torch
    from oracle import heom as O
    from oracle.lindblad_torch import final_state as torch_final
    from pyqed_amd.adjoint import embed_rho0, final_state, heom_problem
    sx = np.array([[0, 1], [1, 0]], complex)
    sz = np.diag([1.0, -1.0]).astype(complex)
    bath = O.drude_bath(0.1, 0.5, 1.0, 1)
    prob = heom_problem(np.zeros((2, 2)), [sz], bath, [0, 0], 3, 0.05, 10, 8, param_ops=[sz / 2, sx / 2],
                        field_ops=[sx], ckpt_every=7)
    rho0 = embed_rho0(prob, np.diag([1.0, 0.0]))
    theta = torch.tensor([1.0, 1.0], dtype=torch.float64, requires_grad=True)
    amps = torch.tensor(np.random.default_rng(0).normal(0, 0.2, (1, 8)), dtype=torch.float64, requires_grad=True)
    J = final_state(theta, amps, prob, rho0)[0, 0]  # Re rho_00 (element 0 of ADO 0)
    J.backward()
    t2 = theta.detach().clone().requires_grad_(True)
    a2 = amps.detach().clone().requires_grad_(True)
    torch_final(t2, a2, prob.base, prob.G, prob.B, rho0, prob.dt, prob.sps)[0].real.backward()
    assert (theta.grad - t2.grad).abs().max() <= 1e-10 and (amps.grad - a2.grad).abs().max() <= 1e-10
    h = 1e-6
    f = lambda th, am: prob.forward(th, am, rho0)[0].real
    for j in range(2):
        e = np.eye(2)[j] * h
        fd = (f(theta.detach().numpy() + e, amps.detach().numpy()) - f(theta.detach().numpy() - e,
                                                                        amps.detach().numpy())) / (2 * h)
        assert abs(fd - theta.grad[j].item()) <= 1e-5 * max(abs(fd), 1e-3)


def test_lindblad_correlation_functions_against_oracle(tmp_path):
    """Quantum-regression correlations (oqs.py:717-791, 1193-1327) on the GPU against the oracle's
    RK4 restatement: one- and two-time, three and four operators (tolerance 1e-12 max-abs)."""
    from oracle import lindblad as OL
    from pyqed_amd.oqs import LindbladSolver
    rng = np.random.default_rng(11)
    d = 3
    A = rng.normal(size=(d, d)) + 1j * rng.normal(size=(d, d))
    H = (A + A.conj().T) / 2
    c_ops = [0.3 * (rng.normal(size=(d, d)) + 1j * rng.normal(size=(d, d))), 0.2 * np.diag([1.0, -1.0, 0.5])]
    a, b, c, dd = (rng.normal(size=(d, d)) + 1j * rng.normal(size=(d, d)) for _ in range(4))
    rho0 = np.diag([0.5, 0.3, 0.2]).astype(complex)
    s = LindbladSolver(H, c_ops)
    dt, Nt, Ntau = 0.02, 40, 25
    out = tmp_path / "cor.dat"
    g2 = s.correlation_2op_1t(rho0, a, b, dt, Nt, output=str(out))
    assert g2.shape == (Nt,)
    assert np.abs(g2 - OL.correlation_2op_1t(H, c_ops, rho0, a, b, dt, Nt)).max() <= 1e-12 * np.abs(g2).max()
    lines = out.read_text().splitlines()
    assert len(lines) == Nt and abs(complex(lines[-1].split()[1]) - g2[-1]) <= 1e-12 * abs(g2[-1])
    g3 = s.correlation_3op_1t(rho0, [a, b, c], dt, Nt)
    assert g3.shape == (Nt + 1,)
    assert np.abs(g3 - OL.correlation_3op_1t(H, c_ops, rho0, a, b, c, dt, Nt)).max() <= 1e-12 * np.abs(g3).max()
    g4 = s.correlation_4op_1t(rho0, [a, b, c, dd], dt, Nt)
    assert np.abs(g4 - OL.correlation_3op_1t(H, c_ops, rho0, a, b @ c, dd, dt, Nt)).max() <= 1e-12 * np.abs(g4).max()
    m3 = s.correlation_3op_2t(rho0, [a, b, c], dt, Nt, Ntau)
    assert m3.shape == (Nt, Ntau)
    ref = OL.correlation_3op_2t(H, c_ops, rho0, a, b, c, dt, Nt, Ntau)
    assert np.abs(m3 - ref).max() <= 1e-12 * np.abs(ref).max()
    m4 = s.correlation_4op_2t(rho0, [a, b, c, dd], dt, Nt, Ntau)
    ref4 = OL.correlation_3op_2t(H, c_ops, rho0, a, b @ c, dd, dt, Nt, Ntau)
    assert np.abs(m4 - ref4).max() <= 1e-12 * np.abs(ref4).max()


def test_lindblad_correlation_two_level_closed_form():
    """<sigma+(t) sigma-> = exp((i w - G/2) t) for spontaneous decay (RK4 error at dt = 0.01)."""
    from pyqed_amd.oqs import LindbladSolver
    w, G, dt, Nt = 1.3, 0.4, 0.01, 300
    sm = np.array([[0, 1], [0, 0]], complex)
    s = LindbladSolver(np.diag([0.0, w]).astype(complex), [np.sqrt(G) * sm])
    g = s.correlation_2op_1t(np.diag([0.0, 1.0]).astype(complex), sm.conj().T, sm, dt, Nt, output=None)
    t = np.arange(1, Nt + 1) * dt
    assert np.abs(g - np.exp((1j * w - G / 2) * t)).max() < 2e-9
