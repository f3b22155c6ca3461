"""The 4096-instance batched exciton workload at its full size on one GPU (24.7M ADOs, packed
Hermitian state, mixed precision): properties that hold at any size -- trace and Hermiticity of every
reduced density matrix, non-negative populations -- and agreement of a seeded sample of instances with separate FP64 runs of
the same instances (<= 1e-6, the mixed-precision bar)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _instances(n):
    from pyqed_amd import workloads as W
    from pyqed_amd.heom.bath import drude
    out = []
    for p in W.batched_c3(n, seed=3):
        e = drude(p["lam"], p["gam"], p["beta"], p["npsd"])
        d = p["H"].shape[0]
        bath = tuple(np.tile(a, d) for a in e) + (np.repeat(np.arange(d), len(e[3])),)
        out.append(dict(H=p["H"], Q=p["Q"], L=p["L"], rho0=p["rho0"], bath=bath))
    return out


def test_full_batch_trace_hermiticity_and_sample_vs_fp64():
    import torch
    from pyqed_amd.heom.engine import HeomEngine
    insts = _instances(4096)
    nsteps, dt = 5, 0.01
    print(f"[case] 4096-instance batch, {nsteps} mixed steps", flush=True)
    eng = HeomEngine(insts)
    assert eng.packed and eng.n_ados == 24730133
    eng.advance(nsteps, dt, "mixed")
    rho = eng.reduced_states()
    eng.close()
    del eng
    torch.cuda.empty_cache()
    tr = np.array([np.trace(r) for r in rho])
    assert np.abs(tr - 1.0).max() < 1e-6                       # every instance keeps Tr rho = 1
    # the packed state stores upper triangles, so off-diagonal Hermiticity is exact; the diagonal's
    # imaginary part is what FP32 right-hand sides leave there
    assert max(np.abs(np.imag(np.diag(r))).max() for r in rho) < 1e-6
    assert min(np.real(np.diag(r)).min() for r in rho) > -1e-6   # populations stay non-negative
    rng = np.random.default_rng(7)
    sample = sorted(rng.choice(len(insts), 12, replace=False).tolist())
    print(f"[case] FP64 runs of instances {sample}", flush=True)
    worst = 0.0
    for i in sample:
        e = HeomEngine([insts[i]], resident=False, packed=False)
        e.advance(nsteps, dt, "fp64")
        ref = e.reduced_states()[0]
        e.close()
        worst = max(worst, float(np.abs(rho[i] - ref).max()))
    print(f"full batch: max |rho_mixed - rho_fp64| over the sample = {worst:.2e}")
    assert worst < 1e-6
