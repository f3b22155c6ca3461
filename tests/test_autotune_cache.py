"""The autotuner's cache (host logic, no device): workload signatures and the record / lookup round
trip, with a stub engine standing in for a plan on a device."""
import types

from pyqed_amd.heom import autotune as A


def _engine(packed=True, cus=256, shapes=((12, 24, 3), (7, 14, 2))):
    props = types.SimpleNamespace(gcnArchName="gfx950:sramecc+:xnack-", multi_processor_count=cus)
    cuda = types.SimpleNamespace(get_device_properties=lambda dev: props)
    return types.SimpleNamespace(torch=types.SimpleNamespace(cuda=cuda), device=0, hermitian=True, dense=False,
                                 packed=packed, prep=[dict(d=d, mode=[0] * k, L=L) for d, k, L in shapes],
                                 n=len(shapes), n_ados=1234)


def test_signature_separates_layout_precision_device_and_shapes():
    base = A.signature(_engine(), "mixed")
    assert base == A.signature(_engine(), "mixed")
    assert base != A.signature(_engine(packed=False), "mixed")
    assert base != A.signature(_engine(), "fp64")
    assert base != A.signature(_engine(cus=304), "mixed")
    assert base != A.signature(_engine(shapes=((12, 24, 3),)), "mixed")


def test_record_and_lookup_round_trip(tmp_path):
    path = str(tmp_path / "tuning.json")
    eng = _engine()
    assert A.lookup(eng, "mixed", path) is None and A.lookup_classes(eng, "mixed", path) is None
    result = {"best": (8192, 12, 0), "timings": {"8192x12x0": 1.5}}
    classes = {"classes": [(2048, 8), None, (16384, 16), None], "ms_per_step": 1.4}
    A.record(eng, "mixed", result, path, classes=classes, workload="c3")
    assert A.lookup(eng, "mixed", path) == (8192, 12, 0)
    assert A.lookup_classes(eng, "mixed", path) == [(2048, 8), None, (16384, 16), None]
    assert A.lookup(eng, "fp64", path) is None
    # a second entry leaves the first in place
    A.record(_engine(packed=False), "mixed", {"best": (4096, 4, 1), "timings": {"4096x4x1": 2.0}}, path)
    assert A.lookup(eng, "mixed", path) == (8192, 12, 0)
    assert A.lookup(_engine(packed=False), "mixed", path) == (4096, 4, 1)
    assert A.lookup_classes(_engine(packed=False), "mixed", path) is None


def test_cached_entries_use_the_candidate_sets():
    for e in A.load_cache().values():
        assert e["target_elems"] in A.TARGETS and e["blocks_per_cu"] in A.BLOCKS_PER_CU
        assert e.get("variant", 0) in A.VARIANTS
        for c in e.get("class_tiling") or []:
            if c:
                assert c[0] in A.TARGETS and c[1] in A.BLOCKS_PER_CU
