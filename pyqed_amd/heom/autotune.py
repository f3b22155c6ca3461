"""Autotuner for the HEOM work queue: chooses the work-item size (complex elements per item), the
persistent grid (blocks per CU) and the kernel variant of the streamed stage kernels by timing RK4
steps, then refines the item size and grid of each kernel class on its own launch times.

Results are cached per workload signature (instances' (d, K, L), Hermitian mode, state layout,
precision, device architecture and CU count) in pyqed_amd/tuning.json; HeomEngine applies a cached
entry before its first step.

    python -m pyqed_amd.heom.autotune --workload c3 [--out tuning.json]
"""
import argparse
import hashlib
import json
import os
import time

import numpy as np

from .. import _lib

TUNING_FILE = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "tuning.json")
TARGETS = (2048, 4096, 8192, 16384)
BLOCKS_PER_CU = (4, 8, 12, 16)
VARIANTS = (0, 1, 2)  # bit 0: pipelined MFMA-class kernel (Hermitian full-matrix plans); bit 1: block kernel
#                      for the wide class (packed plans with 17 <= d <= 32)


def signature_of(arch, cus, precision, hermitian, dense, shapes, packed):
    """Cache key of a workload: device architecture and CU count, precision, plan flags and every
    instance's (d, K, L)."""
    key = [arch, int(cus), precision, bool(hermitian), bool(dense), [list(map(int, t)) for t in shapes]]
    if packed:
        key.append("packed")  # packed Hermitian state: its own entry
    return hashlib.sha1(json.dumps(key).encode()).hexdigest()[:20]


def signature(engine, precision):
    """The device enters through its gfx architecture and CU count, which stay the same under a
    profiler (the marketing name reads empty there)."""
    props = engine.torch.cuda.get_device_properties(engine.device)
    return signature_of(props.gcnArchName, props.multi_processor_count, precision, engine.hermitian, engine.dense,
                        [(p["d"], len(p["mode"]), p["L"]) for p in engine.prep], getattr(engine, "packed", False))


def load_cache(path=TUNING_FILE):
    if os.path.exists(path):
        with open(path) as f:
            return json.load(f)
    return {}


def lookup(engine, precision, path=TUNING_FILE):
    e = load_cache(path).get(signature(engine, precision))
    return None if e is None else (int(e["target_elems"]), int(e["blocks_per_cu"]), int(e.get("variant", 0)))


def lookup_classes(engine, precision, path=TUNING_FILE):
    """Per-class (target_elems, blocks_per_cu) of the cached entry, or None (classes without streamed
    work keep the plan-wide setting)."""
    e = load_cache(path).get(signature(engine, precision))
    return None if e is None or "class_tiling" not in e else [tuple(x) if x else None for x in e["class_tiling"]]


def set_tiling(engine, target, bpc, variant=0):
    engine.set_tiling(target, bpc, variant)


def tune(engine, dt, precision="mixed", targets=TARGETS, blocks_per_cu=BLOCKS_PER_CU, variants=VARIANTS, steps=2,
         warmup=1):
    """Time every candidate (work-item size x blocks per CU x kernel variant) on `engine` (its state
    advances); leaves the best applied and returns {"best": (target, bpc, variant), "timings": {...}}
    with milliseconds per step."""
    torch = engine.torch
    timings = {}
    wide = engine.classes()["items"][3] > 0
    keep = {0} | ({1} if engine.hermitian and not engine.packed else set()) | ({2} if engine.packed and wide else set())
    variants = tuple(v for v in variants if v in keep) or (0,)
    for v in variants:
        for t in targets:
            for b in blocks_per_cu:
                set_tiling(engine, t, b, v)
                engine.advance(warmup, dt, precision)
                torch.cuda.synchronize()
                t0 = time.perf_counter()
                engine.advance(steps, dt, precision)
                torch.cuda.synchronize()
                timings[f"{t}x{b}x{v}"] = (time.perf_counter() - t0) / steps * 1e3
    best = min(timings, key=timings.get)
    bt, bb, bv = (int(x) for x in best.split("x"))
    set_tiling(engine, bt, bb, bv)
    return {"best": (bt, bb, bv), "timings": timings}


def _step_ms(engine, dt, precision, steps, warmup):
    torch = engine.torch
    engine.advance(warmup, dt, precision)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    engine.advance(steps, dt, precision)
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / steps * 1e3


def tune_classes(engine, dt, precision="mixed", targets=TARGETS, blocks_per_cu=BLOCKS_PER_CU, steps=2, warmup=1):
    """With the plan-wide best applied, time each kernel class's launches (HIP events per class launch)
    over every (item size, blocks per CU) of that class alone and keep its fastest; the refined
    configuration is kept only if the whole step is not slower.  Returns {"classes": [(t, b) or None],
    "class_timings": {...}, "ms_per_step": ...}."""
    items = engine.classes()["items"]
    active = [c for c in range(len(items)) if items[c] > 0]
    base = engine.class_tiling()
    base_ms = _step_ms(engine, dt, precision, steps, warmup)
    out = {"class_timings": {}}
    if len(active) < 2:  # one class: the plan-wide sweep already chose for it
        out.update(classes=[base[c] if c in active else None for c in range(len(items))], ms_per_step=base_ms)
        return out
    nc = len(items)
    for c in active:
        times = {}
        for t in targets:
            for b in blocks_per_cu:
                engine.set_class_tiling(c, t, b)
                engine.advance(warmup, dt, precision)
                engine.set_timing(True)
                engine.advance(steps, dt, precision)
                ms, n = engine.read_timing()
                engine.set_timing(False)
                times[(t, b)] = sum(ms[s * nc + c] / max(n[s * nc + c], 1) for s in range(4))
        best = min(times, key=times.get)
        engine.set_class_tiling(c, *best)
        out["class_timings"][str(c)] = {f"{t}x{b}": v for (t, b), v in times.items()}
    tuned = engine.class_tiling()
    tuned_ms = _step_ms(engine, dt, precision, steps, warmup)
    if tuned_ms > base_ms:
        for c in active:
            engine.set_class_tiling(c, *base[c])
        tuned, tuned_ms = base, base_ms
    out.update(classes=[tuned[c] if c in active else None for c in range(nc)], ms_per_step=tuned_ms,
               plan_wide_ms_per_step=base_ms)
    return out


def record(engine, precision, result, path=TUNING_FILE, classes=None, workload=None):
    cache = load_cache(path)
    bt, bb, bv = result["best"]
    e = {"target_elems": bt, "blocks_per_cu": bb, "variant": bv, "precision": precision,
         "ms_per_step": result["timings"][f"{bt}x{bb}x{bv}"], "instances": engine.n, "ados": int(engine.n_ados)}
    if workload:
        e["workload"] = workload
    if classes is not None:
        e["class_tiling"] = [list(x) if x else None for x in classes["classes"]]
        e["ms_per_step_classes"] = classes["ms_per_step"]
    cache[signature(engine, precision)] = e
    with open(path, "w") as f:
        json.dump(cache, f, indent=1, sort_keys=True)


def main():
    import sys
    sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__)))))
    import bench
    from .engine import HeomEngine
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", default="c3")
    ap.add_argument("--n-inst", type=int, default=4096)
    ap.add_argument("--precision", default="mixed")
    ap.add_argument("--out", default=TUNING_FILE)
    a = ap.parse_args()
    insts, dt, _ = bench.instances(a.workload, a.n_inst)
    eng = HeomEngine(insts)
    res = tune(eng, dt, a.precision)
    cls = tune_classes(eng, dt, a.precision)
    record(eng, a.precision, res, a.out, classes=cls, workload=a.workload)
    print(json.dumps({"workload": a.workload, "precision": a.precision, **res, **cls}), flush=True)


if __name__ == "__main__":
    main()
