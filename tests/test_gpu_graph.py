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


def test_graph_recaptured_after_tiling_and_workspace_change():
    """A captured chunk holds the addresses of the workspace and of the work queue.  Changing the tiling
    (which rebuilds the queue) or running another precision (which can grow the workspace) must drop it,
    and later replays must still equal eager steps."""
    import torch
    from pyqed_amd.heom.engine import HeomEngine
    insts = _insts()
    eager = HeomEngine(insts)
    graph = HeomEngine(insts)
    for e in (eager, graph):
        e.autotune_cache = False
    chunk = 4

    def both(fn_eager, fn_graph):
        fn_eager()
        fn_graph()
        torch.cuda.synchronize()
        assert torch.equal(eager.y, graph.y)

    both(lambda: eager.advance(chunk, 0.01, "mixed"), lambda: graph.advance_graph(chunk, 0.01, "mixed", chunk=chunk))
    assert len(graph._graphs) == 1
    # another precision: the FP64 workspace is larger, so the mixed graph's workspace is replaced
    from pyqed_amd import _lib
    grows = graph.ws_bytes[_lib.PREC_FP64] > graph.ws_bytes[_lib.PREC_MIXED]
    both(lambda: eager.advance(3, 0.01, "fp64"), lambda: graph.advance(3, 0.01, "fp64"))
    assert len(graph._graphs) == (0 if grows else 1)
    both(lambda: eager.advance(chunk, 0.01, "mixed"), lambda: graph.advance_graph(chunk, 0.01, "mixed", chunk=chunk))
    # a new tiling rebuilds the work queue
    for e in (eager, graph):
        e.set_tiling(2048, 4, 0)
    assert len(graph._graphs) == 0
    for _ in range(2):
        both(lambda: eager.advance(chunk, 0.01, "mixed"),
             lambda: graph.advance_graph(chunk, 0.01, "mixed", chunk=chunk))
    graph.close()
    assert not graph._graphs


def test_advance_rejects_bad_field_and_traj():
    """The kernels index the field table by step and the trajectory by saved row with no bounds of their
    own, so advance() checks both; graph capture refuses driven instances."""
    import torch
    from pyqed_amd.heom.engine import HeomEngine
    sx = np.array([[0, 1], [1, 0]], complex)
    inst = dict(H=np.diag([0.5, -0.5]).astype(complex), Q=np.array([np.diag([1.0, -1.0])]), L=2,
                rho0=np.diag([1.0, 0.0]).astype(complex), sdip=sx,
                bath=(np.array([0.1 + 0j]), np.array([0.1 + 0j]), np.array([0.1 + 0j]), np.array([1.0 + 0j]),
                      np.array([0])))
    eng = HeomEngine([inst])
    dev = eng.device
    with pytest.raises(ValueError):
        eng.advance(5, 0.01, "fp64")  # driven, no field
    with pytest.raises(ValueError):
        eng.advance(5, 0.01, "fp64", field=torch.zeros((4, 6), dtype=torch.float64, device=dev))  # too short
    with pytest.raises(ValueError):
        eng.advance(5, 0.01, "fp64", field=torch.zeros((5, 6), dtype=torch.float32, device=dev))
    with pytest.raises(ValueError):
        eng.advance(5, 0.01, "fp64", field=torch.zeros((5, 6), dtype=torch.float64))  # host tensor
    f = torch.zeros((10, 6), dtype=torch.float64, device=dev)
    eng.advance(5, 0.01, "fp64", field=f)
    with pytest.raises(ValueError):
        eng.advance(6, 0.01, "fp64", field=f)  # rows 5..10 needed after 5 steps
    with pytest.raises(ValueError):
        eng.advance(5, 0.01, "fp64", field=f, traj=torch.zeros((2, eng.traj_stride), dtype=torch.complex128,
                                                              device=dev))
    with pytest.raises(ValueError):
        eng.advance_graph(5, 0.01, "fp64", chunk=5)
    torch.cuda.synchronize()


def test_class_tiling_changes_only_the_partition():
    """Per-class work-item sizes and grids (the autotuner's per-class choice) change how the queue is
    cut, not the arithmetic: results are bitwise equal to the plan-wide tiling's."""
    import numpy as np
    from pyqed_amd.heom.engine import HeomEngine
    from pyqed_amd import workloads as W
    from pyqed_amd.heom.bath import drude
    insts = []
    for p in W.batched_c3(24, seed=5):
        e = drude(p["lam"], p["gam"], p["beta"], p["npsd"])
        d = p["H"].shape[0]
        bath = tuple(np.tile(a, d) for a in e) + (np.repeat(np.arange(d), len(e[3])),)
        insts.append(dict(H=p["H"], Q=p["Q"], L=p["L"], rho0=p["rho0"], bath=bath))
    runs = []
    for per_class in (False, True):
        eng = HeomEngine(insts, resident=False)
        eng.autotune_cache = False
        eng.set_tiling(4096, 8, 0)
        items0 = eng.classes()["items"]
        if per_class:
            for c in range(4):
                if items0[c]:
                    eng.set_class_tiling(c, 1024 * (c + 1), 4 + 4 * (c % 2))
            ct = eng.class_tiling()
            assert all(ct[c] == (1024 * (c + 1), 4 + 4 * (c % 2)) for c in range(4) if items0[c])
            items1 = eng.classes()["items"]
            assert any(items1[c] != items0[c] for c in range(4) if items0[c])
        runs.append(eng.run(0.01, 30, precision="mixed", save_every=10))
        eng.close()
    for a, b in zip(*runs):
        assert np.array_equal(a, b)
