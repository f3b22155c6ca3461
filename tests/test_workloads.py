"""The seeded workloads the benchmarks run (host only): hierarchy sizes from the combinatorics
C(K + L, L) per instance, against the figures the committed benchmark records report."""
from math import comb

from pyqed_amd import workloads as W


def test_single_hierarchy_sizes():
    c1, c2, c5 = W.spin_boson(), W.exciton_c2(), W.chain_c5()
    assert comb(len(c1["mode"]) + c1["L"], c1["L"]) == 21
    assert comb(len(c2["mode"]) + c2["L"], c2["L"]) == 3060
    assert comb(len(c5["mode"]) + c5["L"], c5["L"]) == 20825
    assert c2["H"].shape == (7, 7) and c5["H"].shape == (24, 24)


def test_batched_workload_sizes():
    ps = W.batched_c3(4096, seed=3)
    ados = sum(comb(2 * p["d"] + p["L"], p["L"]) for p in ps)
    full = sum(comb(2 * p["d"] + p["L"], p["L"]) * p["d"] ** 2 for p in ps)
    packed = sum(comb(2 * p["d"] + p["L"], p["L"]) * p["d"] * (p["d"] + 1) // 2 for p in ps)
    # the plan's packed state is 2,489,974,890 elements: every instance starts on an 8-element boundary
    assert (ados, full, packed) == (24730133, 4646722250, 2489964568)
    assert {p["d"] for p in ps} == set(range(2, 17)) and {p["L"] for p in ps} == {2, 3, 4}
    assert len({(p["d"], p["L"]) for p in ps}) == 45
    assert all(len(p["mode"]) == 2 * p["d"] for p in ps)
