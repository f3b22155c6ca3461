"""Graph-captured step loops: every replay must reproduce the eagerly launched steps bit for bit,
with eager work and synchronisations between replays."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _insts():
    from pyqed_amd import workloads as W
    from pyqed_amd.heom.bath import drude
    out = []
    for p in W.batched_c3(40, seed=11):
        e = drude(p["lam"], p["gam"], p["beta"], p["npsd"])
        d = p["H"].shape[0]
        out.append(dict(H=p["H"], Q=p["Q"], L=min(p["L"], 3), rho0=p["rho0"],
                        bath=tuple(np.tile(a, d) for a in e) + (np.repeat(np.arange(d), 2),)))
    return out


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_graph_replays_match_eager(prec):
    import torch
    from pyqed_amd.heom.engine import HeomEngine
    insts = _insts()
    eager = HeomEngine(insts)
    graph = HeomEngine(insts)
    chunk = 5
    for rep in range(4):
        eager.advance(chunk, 0.01, prec)
        n = graph.advance_graph(chunk, 0.01, prec, chunk=chunk)
        assert n == 1
        torch.cuda.synchronize()
        _ = (eager.y.abs().sum() + graph.y.abs().sum()).item()  # eager work between replays
        assert torch.equal(eager.y, graph.y), f"replay {rep} differs"
    assert eager.step == graph.step == 4 * chunk
    # a remainder that is not a multiple of the chunk runs eagerly
    eager.advance(7, 0.01, prec)
    graph.advance_graph(7, 0.01, prec, chunk=chunk)
    torch.cuda.synchronize()
    assert torch.equal(eager.y, graph.y)
