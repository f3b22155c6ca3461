"""Benchmark: ADO-updates/s of batched HEOM propagation on one MI355X.

Default workload: 4096 seeded uneven exciton-chain HEOM instances (d in 2..16,
depth 2..4, K = 2d Drude + Pade exponentials, dt = 0.01) advanced together in
one plan with the FP32-right-hand-side / FP64-state kernels.  A "step" is one
RK4 step of every instance.

Printed JSON line:
  value          whole-batch ADO-updates/s of the mixed-precision batched path
  fp64_batched   ADO-updates/s of the same batched kernels in plain FP64 (one plan over every
                 instance, chunked only if it does not fit the device), and ratio_vs_fp64_batched
  fp64_plain     ADO-updates/s of the plain FP64 kernels run one instance per plan (same
                 instances, same steps), and ratio_vs_fp64_plain
  error_vs_fp64  max-abs population / coherence difference of rho between the mixed and plain FP64
                 runs over the workload's full step count (2000 for the batch) on a seeded sample of
                 64 instances covering every (d, depth) stratum; after_timed_steps: the same over
                 every instance after the timed steps
  error_vs_oracle  the same mixed trajectories against the CPU oracle (the reference's FP64 solver
                 restated, oracle.heom.run) over the full step count, for the cheapest instance of as
                 many (d, depth) strata as a 30 s CPU budget allows (with the CPU legs only)
  roofline       the dominant stage kernel (the kernel-class launch with the longest average time,
                 timed per launch with HIP events on its stream, in a pass of up to 10 steps after
                 the timed steps): algorithmic HBM bytes per launch /
                 average launch time, vs 8 TB/s; `launches` gives time, bytes and fraction for every
                 (stage, class) launch, `step` the same for all stage kernels of a step; `traffic`
                 the measured HBM bytes per launch from the committed counter summary when it was
                 taken at this workload, state layout and tiling
  cpu_baseline   numpy FP64 restatement (the oracle) on the host's cores (R1: all cores, vectorised;
                 R0: one core, the reference's per-ADO loop) on a stratified sample (a single
                 hierarchy: consecutive steps for a time budget), with the CPU model; `value` and
                 `leg` are the faster of the two

usage: python bench.py [--steps K] [--warmup W] [--workload c3|c2|c5|c1|c4|lindblad|legacy|...]

Other workloads: c4 (Lindblad adjoint gradients), lindblad (driven Lindblad forward), legacy /
legacy-rk4 (oqs.HEOMSolver), spectral (eigen-decomposition spectral sums), krylov (HEOM spectra),
heom-adjoint (streamed HEOM gradients), lindblad-matrix (Lindblad in matrix form at d = 64).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "ADO-updates/s, FP32-batched vs FP64 plain HEOM, 1 GPU, ≤1e-6 pop. error"
HBM_PEAK_GBS = 8000.0
# Algorithmic bytes per full-matrix element of the four RK4 stages, for classic RK4 with three
# registers (y and the accumulator in complex128, the stage input X in complex64 for mixed precision
# or complex128 for FP64), neighbour re-reads counted once: stage 0 reads y, writes acc and X;
# stages 1-2 read X, read and write acc, read y, write X; stage 3 reads X and acc, writes y.
# Mixed 40 + 64 + 64 + 40 = 208 B per element per RK4 step, FP64 48 + 80 + 80 + 48 = 256 B.
# roofline.achieved prices a launch by these figures times the class's full-matrix elements, the
# same figure for every state layout and register scheme.
ALG = {"mixed": (40, 64, 64, 40), "fp64": (48, 80, 80, 48)}


def alg_bytes(stage, ts, F):
    return F * ALG["mixed" if ts == 8 else "fp64"][stage]


# Bytes the implementation's own layout moves in one stage launch over a class of F stage-buffer elements (full matrices,
# complex of ts bytes) and S RK4-state elements (S = F, or the upper triangles of a packed Hermitian
# plan):  ts * F * NX[stage] + S * SB[stage], with the stage-buffer traffic NX
#   stage 0: write X (packed: also read the full-matrix FP32/FP64 copy of y, the stage input)
#   stage 1, 2: read X, write X'
#   stage 3: read X (packed: also write the full-matrix copy of y)
# and the RK4-state bytes SB per state element.  FP64 keeps the running new state ynew (complex
# double): stage 0 reads y, writes ynew; stages 1-2 read y, ynew, write ynew; stage 3 reads ynew,
# writes y.  Unpacked mixed precision keeps the FP32 increment inc = ynew - y (8 bytes): stage 0
# reads y (its input), writes inc; stages 1-2 read y, inc, write inc; stage 3 reads y, inc, writes y.
# Packed mixed precision keeps two FP32 registers A, B and takes y for stages 0-2 from their FP32
# stage input: stage 0 writes A; stage 1 reads A, writes B; stage 2 reads A, B, writes B; stage 3
# reads y, B, writes y.  Unpacked: mixed 176 B per element per RK4 step, FP64 256 B.
NX = {False: (1, 2, 2, 1), True: (2, 2, 2, 2)}
SB = {("fp64", False): (32, 48, 48, 32), ("fp64", True): (32, 48, 48, 32),
      ("mixed", False): (24, 32, 32, 40), ("mixed", True): (8, 16, 24, 40)}
# committed per-kernel HBM traffic summaries (bench.py --traffic-from), per workload
TRAFFIC_FILES = {"c3": os.path.join(ROOT, "profiles", "heom_proj_kernels_exciton4096_traffic.json"),
                 "c5": os.path.join(ROOT, "profiles", "heom_proj_kernels_chain24_traffic.json")}


def stage_bytes(stage, ts, F, S, packed):
    return ts * F * NX[packed][stage] + S * SB[("mixed" if ts == 8 else "fp64", packed)][stage]


def instances(workload, n_inst):
    from pyqed_amd import workloads as W
    from pyqed_amd.heom.bath import drude

    def chain(p):
        e = drude(p["lam"], p["gam"], p["beta"], p["npsd"])
        d = p["H"].shape[0]
        bath = tuple(np.tile(a, d) for a in e) + (np.repeat(np.arange(d), len(e[3])),)
        return dict(H=p["H"], Q=p["Q"], L=p["L"], rho0=p["rho0"], bath=bath)

    if workload == "c3":
        ps = W.batched_c3(n_inst, seed=3)
        return [chain(p) for p in ps], 0.01, ("batched uneven HEOM: %d seeded exciton chains, d 2..16, depth 2..4, "
                                             "K=2d (Drude + 1 Pade pole per site), dt=0.01" % n_inst)
    if workload == "c2":
        p = W.exciton_c2()
        return [chain(p)], p["dt"], "7-site Frenkel exciton HEOM, K=14, depth 4 (3060 ADOs), dt=1 fs"
    if workload == "c5":
        p = W.chain_c5()
        return [chain(p)], p["dt"], "24-site exciton chain HEOM, K=48, depth 3 (20825 ADOs), dt=0.005"
    if workload == "c1":
        p = W.spin_boson()  # one coupling operator (sigma_z) shared by both exponentials
        e = drude(p["lam"], p["gam"], p["beta"], p["npsd"])
        inst = dict(H=p["H"], Q=p["Q"], L=p["L"], rho0=p["rho0"], bath=tuple(e) + (np.asarray(p["mode"]),))
        return [inst], p["dt"], "spin-boson HEOM, K=2, depth 5 (21 ADOs), dt=0.01"
    raise SystemExit(f"unknown workload {workload}")


def max_errs(a, b):
    d = a.shape[-1]
    dd = np.abs(np.diagonal(a - b, axis1=-2, axis2=-1)).max()
    off = np.abs((a - b)[..., ~np.eye(d, dtype=bool)]).max() if d > 1 else 0.0
    return float(dd), float(off)


def fp64_plain(insts, dt, warmup, steps, torch):
    """Plain FP64 kernel, one instance per plan, full-matrix state, no LDS residency: time `steps` steps
    after `warmup`."""
    from pyqed_amd.heom.engine import HeomEngine
    total = 0.0
    finals = []
    ados = 0
    for inst in insts:
        eng = HeomEngine([inst], resident=False, packed=False)
        eng.advance(warmup, dt, "fp64")
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        eng.advance(steps, dt, "fp64")
        torch.cuda.synchronize()
        total += time.perf_counter() - t0
        finals.append(eng.reduced_states()[0])
        ados += int(eng.n_ados)
        eng.close()
        del eng
    return ados * steps / total, total, finals


def cpu_info():
    """Host CPU model and the cores this process may use."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return model, len(os.sched_getaffinity(0))


def _strata(insts):
    """One instance per (d, depth) stratum: {(d, L): index}."""
    strata = {}
    for i, p in enumerate(insts):
        strata.setdefault((p["H"].shape[0], p["L"]), i)
    return strata


class _ChunkedRHS:
    """The vectorised oracle right-hand side (diagonal couplings as row/column scalings) split over ADO
    ranges, evaluated by a thread pool: numpy releases the GIL inside these array kernels."""

    def __init__(self, H, qd, bath, mode, L, nthreads):
        from oracle import heom as O
        self.O = O
        hier = O.Hierarchy(len(mode), L)
        self.nmax = hier.nmax
        self.H = np.asarray(H, np.complex128)
        terms = O.DiagTerms(hier, bath, mode, qd)
        self.damp = terms.damp
        bounds = np.linspace(0, hier.nmax, min(nthreads * 4, hier.nmax) + 1).astype(int)
        self.chunks = []
        for a0, a1 in zip(bounds[:-1], bounds[1:]):
            if a1 <= a0:
                continue
            tl = []
            for sel, nb, c, cm in terms.terms:
                m = (sel >= a0) & (sel < a1)
                if m.any():
                    tl.append((sel[m] - a0, nb[m], c[m], cm))
            self.chunks.append((a0, a1, tl))

    def __call__(self, rho, pool):
        out = np.empty_like(rho)
        H = self.H

        def work(ch):
            a0, a1, tl = ch
            x = rho[a0:a1]
            o = -self.damp[a0:a1] * x - 1j * (np.matmul(H[None], x) - np.matmul(x, H[None]))
            for sel, nb, c, cm in tl:
                o[sel] += c * cm[None] * rho[nb]
            out[a0:a1] = o
        list(pool.map(work, self.chunks))
        return out


def cpu_baseline(insts, dt, r0_budget_s=10.0, r1_budget_s=5.0, max_steps=1, workers=None):
    """CPU baselines on the box's host cores, on a stratified sample of the workload (one instance per
    (d, depth) stratum, or the one instance):
      R1  all cores: every sampled instance one RK4 step with the vectorised numpy FP64 oracle, each
          stage split over ADO ranges on min(16, usable cores) threads; the headline cpu_baseline value
          (ADO-updates of the sample / time).  A single-instance workload runs consecutive steps until
          r1_budget_s is spent or max_steps (its step count) are done;
      R0  one core: the per-ADO loop with the reference's structure (oracle rhs_loop, deom.py:641-673),
          one RK4 step per stratum, smallest first, until r0_budget_s is spent; a single instance runs
          as many steps as the budget allows, up to max_steps (the whole run for small ones)."""
    from concurrent.futures import ThreadPoolExecutor
    from math import comb
    from threadpoolctl import threadpool_limits
    from oracle import heom as O
    model, host_cores = cpu_info()
    strata = _strata(insts)
    nthr = max(1, min(workers or 16, host_cores))
    ados, wall = 0, 0.0
    with threadpool_limits(limits=1), ThreadPoolExecutor(nthr) as pool:
        for (d, L), i in sorted(strata.items()):
            p = insts[i]
            etal, etar, etaa, expn, mode = p["bath"]
            qd = np.array([np.diag(q) for q in np.asarray(p["Q"])])
            f = _ChunkedRHS(p["H"], qd, (etal, etar, etaa, expn), np.asarray(mode), L, nthr)
            y = np.zeros((f.nmax, d, d), np.complex128)
            y[0] = p["rho0"]
            n1 = 0
            while True:
                t0 = time.perf_counter()
                y = O.rk4_step(y, lambda x, t: f(x, pool), dt, n1 * dt)
                wall += time.perf_counter() - t0
                ados += f.nmax
                n1 += 1
                if len(strata) > 1 or wall >= r1_budget_s or n1 >= max_steps:
                    break
    r1 = dict(value=ados / wall, unit="ADO-updates/s", cores=nthr, seconds=wall, ado_updates=ados,
              strata=len(strata), steps=n1 if len(strata) == 1 else 1)
    # R0: the reference's per-ADO loop structure on one core
    done, ados0, used, steps0 = [], 0, 0.0, 1
    with threadpool_limits(limits=1):
        for (d, L), i in sorted(strata.items()):
            p = insts[i]
            etal, etar, etaa, expn, mode = p["bath"]
            nado = comb(len(mode) + L, L)
            t1 = time.perf_counter()
            O.run(p["H"], list(p["Q"]), (etal, etar, etaa, expn), mode, L, p["rho0"], dt, 1, vectorised=False)
            step_s = time.perf_counter() - t1
            used += step_s
            ados0 += nado
            if len(strata) == 1 and max_steps > 1:
                # one instance: as many further steps as the budget allows, up to the workload's count
                more = int(min(max_steps - 1, max(0.0, r0_budget_s - used) // max(step_s, 1e-9)))
                if more > 0:
                    t1 = time.perf_counter()
                    O.run(p["H"], list(p["Q"]), (etal, etar, etaa, expn), mode, L, p["rho0"], dt, more,
                          vectorised=False)
                    used += time.perf_counter() - t1
                    ados0 += nado * more
                    steps0 += more
            done.append(f"d={d},L={L}")
            if used > r0_budget_s:
                break
    r0 = dict(value=ados0 / used, unit="ADO-updates/s", cores=1, seconds=used, ado_updates=ados0,
              strata=len(done), steps=steps0, sample=f"{done[0]} .. {done[-1]}" if len(done) > 1 else done[0])
    what = (f"{r1['steps']} RK4 steps of the instance" if len(strata) == 1 else
            f"one RK4 step of one instance per (d, depth) stratum ({len(strata)} strata)")
    s1 = (f"R1: {what}, {ados} ADO-updates, numpy FP64 oracle vectorised over ADOs with the diagonal "
          f"couplings as row/column scalings, each stage split over ADO ranges on {nthr} threads, {wall:.1f} s")
    s0 = (f"R0: {steps0} RK4 step(s) of {r0['sample']}, {ados0} ADO-updates, the per-ADO loop with the "
          f"reference's structure on one thread, {used:.1f} s")
    # the faster CPU leg is the baseline (on small hierarchies the threaded leg loses to its own overhead)
    use_r1 = r1["value"] >= r0["value"]
    return dict(value=max(r1["value"], r0["value"]), unit="ADO-updates/s", cores=nthr if use_r1 else 1,
                kind="port", sample=s1 if use_r1 else s0, leg="R1" if use_r1 else "R0",
                cpu_model=model, host_cores=host_cores, R1=r1,
                R0=dict(r0, note="per-ADO loop with the reference's structure (oracle rhs_loop), one thread"))


def kernel_name(stage, cls, precision, eng):
    """Template name of the kernel a (stage, class) launch runs (as a kernel trace lists it)."""
    T = "float" if precision == "mixed" else "double"
    TI = "double" if (stage == 0 and not eng.packed) else T
    pack = "true" if eng.packed else "false"
    if cls == 0:
        return f"heom_stage_kernel<{T}, {TI}, {T}, {stage}, {pack}>"
    mfma = "true" if (cls >= 2 and precision == "mixed") else "false"
    herm = "true" if eng.hermitian else "false"
    nt = 2 if (cls == 3 and precision == "mixed") else 1
    # single-instance form: FP32 packed classes 1 and 3 when the class's streamed work is one instance
    one = precision == "mixed" and eng.packed and cls in (1, 3) and eng.class_single()[cls] >= 0
    return f"heom_proj_kernel<{T}, {TI}, {T}, {stage}, {mfma}, {herm}, {pack}, {nt}, {'true' if one else 'false'}>"


def measured_traffic(name, key, workload, packed, tiling):
    """HBM bytes per launch of `name` (FETCH_SIZE + WRITE_SIZE, copy-calibrated) from the committed
    counter summary of workload `key`, or None unless that summary was taken on the same workload,
    state layout and work-queue tiling as this run."""
    path = TRAFFIC_FILES.get(key)
    if path is None or not os.path.exists(path):
        return None
    with open(path) as f:
        t = json.load(f)
    if t.get("workload") != workload or bool(t.get("packed", False)) != bool(packed):
        return None
    want = {k: tiling[k] for k in ("target_elems", "blocks_per_cu", "variant")}
    if t.get("tiling") != want:
        return None
    # per-class tilings: those of the summary's run, or (older summaries) every class at the plan's
    cls = tiling.get("classes") or []
    if t.get("classes") is not None:
        if t["classes"] != cls:
            return None
    elif any(c != [want["target_elems"], want["blocks_per_cu"]] for c in cls):
        return None
    k = t["per_dispatch"].get(name)
    return None if k is None else k["fetch_bytes"] + k["write_bytes"]


def _pmc_rows(d):
    """(kernel name, counter value) of every dispatch in the rocprofv3 --pmc CSV files under d."""
    import csv
    import glob
    import re
    out = []
    for path in glob.glob(os.path.join(d, "**", "*counter_collection*.csv"), recursive=True):
        with open(path) as f:
            for r in csv.DictReader(f):
                name = r.get("Kernel_Name") or r.get("Kernel-Name") or ""
                name = re.sub(r"^void ", "", name).replace("(anonymous namespace)::", "")
                depth, cut = 0, len(name)
                for i, ch in enumerate(name):  # drop the argument list after the template arguments
                    if ch == "<":
                        depth += 1
                    elif ch == ">":
                        depth -= 1
                    elif ch == "(" and depth == 0:
                        cut = i
                        break
                out.append((name[:cut].strip(), float(r["Counter_Value"])))
    return out


def traffic_from(fetch_dir, write_dir, copy_fetch_dir, copy_write_dir, bench_json, copy_mib=1024, key="c3"):
    """Per-dispatch HBM bytes of every kernel from rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE runs (KiB per
    dispatch), each scaled by what the same counter reads for a device copy of known size
    (bench.py --copy-calibration), written to the traffic summary together with the workload, state
    layout and tiling of the profiled run (bench_json: its printed JSON line)."""
    def per_kernel(d):
        acc = {}
        for name, v in _pmc_rows(d):
            acc.setdefault(name, []).append(v)
        return {k: float(np.mean(v)) for k, v in acc.items()}

    def copy_kib(d):
        vals = [v for n, v in _pmc_rows(d) if "copyBuffer" in n or "copy" in n.lower()]
        return max(vals) if vals else None

    with open(bench_json) as f:
        b = json.loads([ln for ln in f if ln.startswith("{")][-1])
    nbytes = copy_mib * (1 << 20)  # bench.py --copy-calibration copy_mib: one copy of that many bytes
    fs = nbytes / (copy_kib(copy_fetch_dir) * 1024.0)
    ws = nbytes / (copy_kib(copy_write_dir) * 1024.0)
    fk, wk = per_kernel(fetch_dir), per_kernel(write_dir)
    t = {"workload": b["config"]["workload"], "packed": b["state_layout"].startswith("packed"),
         "tiling": {k: b["tiling"][k] for k in ("target_elems", "blocks_per_cu", "variant")},
         "classes": b["tiling"].get("classes"),
         "fetch_scale": fs, "write_scale": ws,
         "per_dispatch": {k: {"fetch_bytes": fk[k] * 1024.0 * fs, "write_bytes": wk.get(k, 0.0) * 1024.0 * ws}
                          for k in fk}}
    with open(TRAFFIC_FILES[key], "w") as f:
        json.dump(t, f, indent=1)
    print(json.dumps({"written": TRAFFIC_FILES[key], "kernels": len(fk), "fetch_scale": fs, "write_scale": ws}))


def stream_probe(torch):
    """Achievable HBM bandwidth on this device (GB/s): torch device copies of 1, 8 and 32 GiB, and a
    stream with the byte mix of a middle RK4 stage (read complex128 ynew, y and complex64 X; write
    complex128 ynew and complex64 X'), as two fused torch additions over 2^30 elements."""
    def timed(fn, reps=5):
        fn()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            fn()
        e1.record()
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / reps / 1e3
    out = {}
    for gib in (1, 8, 32):
        n = gib << 30
        a = torch.ones(n, dtype=torch.uint8, device="cuda")
        b = torch.empty_like(a)
        out[f"copy_{gib}GiB_GBps"] = 2 * n / timed(lambda: b.copy_(a)) / 1e9
        del a, b
    n = 1 << 30
    y = torch.zeros(n, dtype=torch.complex128, device="cuda")
    yn = torch.zeros_like(y)
    x = torch.zeros(n, dtype=torch.complex64, device="cuda")
    x2 = torch.zeros_like(x)

    def stage():
        torch.add(yn, x, alpha=0.5, out=yn)   # read ynew, X; write ynew
        torch.add(y, yn, alpha=0.25, out=x2)  # read y, ynew; write X'
    out["stage_mix_GBps"] = n * (16 + 8 + 16 + 16 + 16 + 8) / timed(stage) / 1e9
    return out


LEGACY_METRIC = "ADO-updates/s, legacy single-exponential HEOM (oqs.HEOMSolver.run, forward Euler, FP64), 1 GPU"
LEGACY_RK_METRIC = "ADO-updates/s, legacy single-exponential HEOM (HEOM.heom.HEOMSolver.run, RK4, FP64), 1 GPU"
FP64_PEAK_TFLOPS = 78.6


def legacy_workload(n_inst, rk=False):
    """The spin-boson of pyqed's examples/heom.py:77-96 (H = -sx/2 - sz/2, S = sz, e_ops = [sz],
    rho0 = |1><1|, dt = 0.02, cutoff 5, nado 5), swept over n_inst (temperature, reorganization) points.
    rk: the RK4 copy (pyqed.HEOM.heom), whose D0 = reorg (2T - i gamma) takes T as k_B T in the
    Hamiltonian's units, swept over kT = 0.5..2 instead of 300..900."""
    sx = np.array([[0, 1], [1, 0]], complex)
    sz = np.diag([1.0, -1.0]).astype(complex)
    H = -0.5 * sx - 0.5 * sz
    rho0 = np.diag([0.0, 1.0]).astype(complex)
    k = np.arange(n_inst)
    temperature = (0.5 + 1.5 * (k % 64) / 63.0) if rk else (300.0 + 600.0 * (k % 64) / 63.0)
    reorg = 0.05 + 0.35 * (k // 64 % 64) / 63.0
    return dict(H=H, S=sz, e_ops=[sz], rho0=rho0, dt=0.02, cutoff=5.0, nado=5, temperature=temperature,
                reorganization=reorg)


def legacy_flops(d, nado):
    """FP64 flops of one step of one instance: ADO 0 takes two commutators (4 complex d x d products),
    ADOs 1..nado-2 three (6 products: [H, X_n], [S, X_n+1], S X_n-1 and X_n-1 S); 8 d^3 each."""
    return 8 * d ** 3 * (4 + 6 * (nado - 2))


def legacy_cpu_baseline(w, budget_s=12.0, rk=False):
    from oracle import legacy_heom as LH
    n_done, t0, nt = 0, time.perf_counter(), 200
    while time.perf_counter() - t0 < budget_s:
        i = (n_done * 997) % len(w["temperature"])
        if rk:
            gamma, D0 = LH.run_params_rk(w["temperature"][i], w["cutoff"], w["reorganization"][i])
            LH.heom_run_rk4(w["H"], w["rho0"], w["S"], w["e_ops"], gamma, D0, w["nado"], w["dt"], nt)
        else:
            gamma, D0 = LH.run_params(w["temperature"][i], w["cutoff"], w["reorganization"][i])
            LH.heom_run(w["H"], w["rho0"], w["S"], w["e_ops"], gamma, D0, w["nado"], w["dt"], nt)
        n_done += 1
    el = time.perf_counter() - t0
    return {"value": n_done * nt * (w["nado"] - 1) / el, "unit": "ADO-updates/s", "cores": 1, "kind": "port",
            "sample": f"{n_done} instances x {nt} steps of the same sweep (numpy FP64 restatement, one thread)"}


def bench_legacy(args, torch):
    from pyqed_amd import _lib
    from pyqed_amd.HEOM.heom import rk_run_params
    from pyqed_amd.oqs import legacy_run_params
    L = _lib.lib()
    rk = args.workload == "legacy-rk4"
    w = legacy_workload(args.n_inst, rk)
    n, d, nado = args.n_inst, w["H"].shape[0], w["nado"]
    gamma, D0 = (rk_run_params if rk else legacy_run_params)(w["temperature"], w["cutoff"], w["reorganization"])
    params = np.stack([np.broadcast_to(gamma, (n,)), D0.real, D0.imag], axis=1)
    dev = torch.device("cuda", 0)
    A = torch.as_tensor(np.array(np.broadcast_to(w["H"], (n, d, d)), order="C"), device=dev)
    S = torch.as_tensor(np.array(np.broadcast_to(w["S"], (n, d, d)), order="C"), device=dev)
    E = torch.as_tensor(np.array(w["e_ops"]), device=dev)
    par = torch.as_tensor(params, device=dev)
    ado0 = np.zeros((n, nado, d, d), complex)
    ado0[:, 0] = w["rho0"]
    ado = torch.as_tensor(ado0, device=dev)
    obs = torch.empty((n, max(args.steps, args.warmup, 1), 1), dtype=torch.complex128, device=dev)
    stream = torch.cuda.current_stream()

    def launch(nt):
        _lib.check(L.pyqed_legacy_heom_run(2 if rk else 0, d, nado, n, A.data_ptr(), S.data_ptr(), None, par.data_ptr(),
                                           ado.data_ptr(), nt, w["dt"], 1, E.data_ptr(), obs.data_ptr(),
                                           stream.cuda_stream))

    launch(args.warmup)
    torch.cuda.synchronize()
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
    t0 = time.perf_counter()
    ev[0].record(stream)
    launch(args.steps)
    ev[1].record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    launch_ms = ev[0].elapsed_time(ev[1])
    flops = legacy_flops(d, nado) * n * args.steps * (4 if rk else 1)  # RK4: four right-hand sides per step
    achieved = flops / (launch_ms / 1e3) / 1e12
    out = {
        "metric": LEGACY_RK_METRIC if rk else LEGACY_METRIC, "value": n * (nado - 1) * args.steps / wall,
        "unit": "ADO-updates/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (pyqed examples/heom.py spin-boson over a temperature x reorganization grid)",
        "config": {"workload": f"legacy spin-boson sweep{' (RK4 copy, kT 0.5..2)' if rk else ''}: {n} instances, "
                               f"d={d}, nado={nado}, dt={w['dt']}, one observable per step, {args.steps} steps in "
                               "one launch", "instances": n},
        "roofline": {"bound": "mfma",
                     "kernel": (f"legacy_heom_kernel<false, false, true, {2 if d == 2 and nado * d * d <= 256 else 0}>"
                                if rk else f"legacy_heom_kernel<false, {'true' if d * d <= 256 else 'false'}, false, "
                                           f"{2 if d == 2 else 0}>"),
                     "launch_ms": launch_ms,
                     "flops_per_launch": flops, "achieved": achieved, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / FP64_PEAK_TFLOPS, "traffic": None},
    }
    if not args.no_cpu:
        print("[bench] cpu baseline leg", file=sys.stderr, flush=True)
        out["cpu_baseline"] = legacy_cpu_baseline(w, rk=rk)
    print(json.dumps(out), flush=True)


C4_METRIC = ("gradients/s, Lindblad adjoint (3-level Lambda system, 4000 RK4 steps, dJ/d 200 field amplitudes "
             "+ 3 level energies, FP64), 1 GPU")


def c4_cpu_baseline(w, prob, budget_s=10.0):
    """The oracle's FP64 autograd (torch on the CPU, one thread) of the same objective, rho_22(T)."""
    import torch
    from oracle.lindblad_torch import final_state
    nthr = torch.get_num_threads()
    torch.set_num_threads(1)
    rng = np.random.default_rng(45)
    n, t0 = 0, time.perf_counter()
    while time.perf_counter() - t0 < budget_s:
        th = torch.tensor(w["eps"], dtype=torch.float64, requires_grad=True)
        am = torch.tensor(rng.normal(0.0, 0.1, (1, w["nseg"])), dtype=torch.float64, requires_grad=True)
        final_state(th, am, prob.base, prob.G, prob.B, w["rho0"], w["dt"], w["steps_per_seg"])[2, 2].real.backward()
        n += 1
    el = time.perf_counter() - t0
    torch.set_num_threads(nthr)
    return {"value": n / el, "unit": "gradients/s", "cores": 1, "kind": "port",
            "sample": f"{n} gradients (forward + torch.autograd backward over the 4000 steps, FP64, one thread)"}


def bench_c4(args, torch):
    """Batched adjoint gradients of the C4 problem: n_inst seeded field-amplitude sets, one launch of
    pyqed_liouville_adjoint per batch (forward sweep with checkpoints, then the backward sweep)."""
    from pyqed_amd import _lib
    from pyqed_amd import workloads as W
    from pyqed_amd.adjoint import lambda_system
    L = _lib.lib()
    w = W.lambda_c4()
    prob = lambda_system(w["eps"], w["mu"], list(w["c_ops"]), w["dt"], w["steps_per_seg"], w["nseg"],
                         w["checkpoint_every"])
    n, N, nterm, nsteps, ce = args.n_inst, prob.N, 2, prob.nsteps, prob.ckpt_every
    rng = np.random.default_rng(44)
    amps = rng.normal(0.0, 0.1, (n, w["nseg"]))
    amps[0] = w["amps"]
    dev = torch.device("cuda", 0)
    A = torch.as_tensor(prob._gens(w["eps"]), device=dev)
    f = torch.as_tensor(np.ascontiguousarray(np.repeat(amps, prob.sps, axis=1)[:, :, None]), device=dev)
    rho = torch.as_tensor(np.tile(np.asarray(w["rho0"], np.complex128).reshape(1, -1), (n, 1)), device=dev)
    bar_h = np.zeros((n, N), np.complex128)
    bar_h[:, 2 * 3 + 2] = 1.0                      # J = Re rho_22(T)
    bar = torch.as_tensor(bar_h, device=dev)
    gf = torch.zeros((n, nsteps, nterm - 1), dtype=torch.float64, device=dev)
    gA = torch.zeros((n, nterm, N, N), dtype=torch.complex128, device=dev)
    r0b = torch.zeros((n, N), dtype=torch.complex128, device=dev)
    nck = (nsteps + ce - 1) // ce + 1
    ck_bytes = n * (nck + ce) * N * 16
    ck = torch.empty(ck_bytes, dtype=torch.uint8, device=dev)
    stream = torch.cuda.current_stream()

    def launch():
        _lib.check(L.pyqed_liouville_adjoint(N, nterm, A.data_ptr(), n, rho.data_ptr(), f.data_ptr(), nsteps,
                                             w["dt"], ce, bar.data_ptr(), gf.data_ptr(), gA.data_ptr(),
                                             r0b.data_ptr(), ck.data_ptr(), ck_bytes, stream.cuda_stream))

    for _ in range(args.warmup):
        launch()
    torch.cuda.synchronize()
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
    t0 = time.perf_counter()
    ev[0].record(stream)
    for _ in range(args.steps):
        launch()
    ev[1].record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    launch_ms = ev[0].elapsed_time(ev[1]) / args.steps
    # instance 0 against the single-instance path (LindbladAdjoint.gradient)
    gt1, ga1, _ = prob.gradient(w["eps"], w["amps"][None], w["rho0"], bar_h[0].reshape(3, 3))
    gA0 = gA[0, 0].cpu().numpy()
    gt = np.array([np.real(np.sum(np.conj(gA0) * Gj)) for Gj in prob.G])
    ga = gf[0, :, 0].cpu().numpy().reshape(w["nseg"], prob.sps).sum(axis=1)
    check = float(max(np.abs(gt - gt1).max(), np.abs(ga - ga1[0]).max()))
    # forward, checkpoint recompute, adjoint products, field-gradient products and generator-gradient
    # outer products: five sweeps of 4 complex N x N products (8 N^2 flops each) per step
    flops = 5 * 4 * 8 * N * N * nsteps * n
    achieved = flops / (launch_ms / 1e3) / 1e12
    out = {
        "metric": C4_METRIC, "value": n * args.steps / wall, "unit": "gradients/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (seeded field amplitudes N(0, 0.1) per instance)",
        "config": {"workload": f"3-level Lambda system, Lindblad dephasing and decay, 200 segments x 20 steps "
                               f"(dt 0.05), checkpoint every {ce}; {n} gradient problems per launch; a step "
                               "is one batch", "instances": n},
        "roofline": {"bound": "mfma", "kernel": "liouville_adjoint_kernel", "launch_ms": launch_ms,
                     "flops_per_launch": flops, "achieved": achieved, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / FP64_PEAK_TFLOPS, "traffic": None},
        "check_vs_single_instance": check,
    }
    if not args.no_cpu:
        print("[bench] cpu baseline leg", file=sys.stderr, flush=True)
        out["cpu_baseline"] = c4_cpu_baseline(w, prob)
    print(json.dumps(out), flush=True)


LINDBLAD_METRIC = "instance-steps/s, driven Lindblad RK4 (3-level Lambda system, 3 observables per step, FP64), 1 GPU"


def lindblad_cpu_baseline(w, budget_s=10.0):
    """The oracle's driven-Lindblad restatement (numpy, one thread, dense d x d algebra per stage)."""
    from oracle import lindblad as OL
    rng = np.random.default_rng(46)
    H0 = np.diag(w["eps"]).astype(np.complex128)
    e_ops = [np.diag(np.eye(3)[j]).astype(np.complex128) for j in range(3)]
    n, nt, t0 = 0, 400, time.perf_counter()
    while time.perf_counter() - t0 < budget_s:
        amps = rng.normal(0.0, 0.1, w["nseg"])
        f = lambda t, a=amps: a[min(int(t / w["dt"] - 0.5) // w["steps_per_seg"], w["nseg"] - 1)]
        OL.run_driven([H0, [w["mu"], f]], w["rho0"], list(w["c_ops"]), w["dt"], nt, e_ops=e_ops)
        n += 1
    el = time.perf_counter() - t0
    return {"value": n * nt / el, "unit": "instance-steps/s", "cores": 1, "kind": "port",
            "sample": f"{n} instances x {nt} steps of the same driven problem (numpy FP64 restatement, one thread)"}


def bench_lindblad(args, torch):
    """Driven Lindblad propagation (LindbladSolver.run with a field, oqs.py:1699-1806, in batch): the C4
    system forward only, n_inst seeded field-amplitude sets, populations after every step."""
    from pyqed_amd import _lib
    from pyqed_amd import workloads as W
    from pyqed_amd.adjoint import lambda_system
    L = _lib.lib()
    w = W.lambda_c4()
    prob = lambda_system(w["eps"], w["mu"], list(w["c_ops"]), w["dt"], w["steps_per_seg"], w["nseg"],
                         w["checkpoint_every"])
    n, N, nterm, nsteps = args.n_inst, prob.N, 2, prob.nsteps
    rng = np.random.default_rng(44)
    amps = rng.normal(0.0, 0.1, (n, w["nseg"]))
    dev = torch.device("cuda", 0)
    A = torch.as_tensor(prob._gens(w["eps"]), device=dev)
    f = torch.as_tensor(np.ascontiguousarray(np.repeat(amps, prob.sps, axis=1)[:, :, None]), device=dev)
    rho0 = np.tile(np.asarray(w["rho0"], np.complex128).reshape(1, -1), (n, 1))
    rho = torch.as_tensor(rho0, device=dev)
    E = torch.as_tensor(np.array([np.diag(np.eye(3)[j]).astype(np.complex128).T.reshape(-1) for j in range(3)]),
                        device=dev)
    obs = torch.empty((n, nsteps + 1, 3), dtype=torch.complex128, device=dev)
    stream = torch.cuda.current_stream()

    def launch():
        rho.copy_(torch.as_tensor(rho0, device=dev))
        _lib.check(L.pyqed_liouville_run(N, nterm, A.data_ptr(), n, rho.data_ptr(), f.data_ptr(), nsteps, w["dt"],
                                         3, E.data_ptr(), obs.data_ptr(), None, 1, stream.cuda_stream))

    for _ in range(args.warmup):
        launch()
    torch.cuda.synchronize()
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
    t0 = time.perf_counter()
    ev[0].record(stream)
    for _ in range(args.steps):
        launch()
    ev[1].record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    launch_ms = ev[0].elapsed_time(ev[1]) / args.steps
    pops = obs[:, -1, :].real.sum(dim=1)
    # one RK4 step: 4 products of the step's N x N generator (8 N^2 flops each)
    flops = 4 * 8 * N * N * nsteps * n
    achieved = flops / (launch_ms / 1e3) / 1e12
    out = {
        "metric": LINDBLAD_METRIC, "value": n * nsteps * args.steps / wall, "unit": "instance-steps/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (seeded field amplitudes N(0, 0.1) per instance)",
        "config": {"workload": f"3-level Lambda system, Lindblad dephasing and decay, 200 field segments x 20 "
                               f"steps (dt 0.05), populations every step; {n} instances per launch; a step is one "
                               "launch of 4000 RK4 steps", "instances": n},
        "roofline": {"bound": "mfma", "kernel": "liouville_forward_kernel", "launch_ms": launch_ms,
                     "flops_per_launch": flops, "achieved": achieved, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / FP64_PEAK_TFLOPS, "traffic": None},
        "trace_deviation": float((pops - 1.0).abs().max()),
    }
    if not args.no_cpu:
        print("[bench] cpu baseline leg", file=sys.stderr, flush=True)
        out["cpu_baseline"] = lindblad_cpu_baseline(w)
    print(json.dumps(out), flush=True)


LINDBLAD_MATRIX_METRIC = ("instance-steps/s, Lindblad RK4 in matrix form (64-level damped cavity, decay and "
                          "dephasing, FP64), 1 GPU")


def cavity_problem(d=64, n=256, seed=21):
    """A Kerr cavity truncated at d levels, H = w a^dag a + chi (a^dag a)^2, collapse operators sqrt(kappa) a
    and sqrt(gamma) a^dag a; n seeded coherent initial states."""
    a = np.diag(np.sqrt(np.arange(1, d)), 1).astype(complex)
    num = a.conj().T @ a
    H = 1.0 * num + 0.01 * num @ num
    c_ops = [np.sqrt(0.05) * a, np.sqrt(0.02) * num]
    rng = np.random.default_rng(seed)
    rhos = []
    for alpha in rng.uniform(1.0, 3.0, n) * np.exp(2j * np.pi * rng.uniform(size=n)):
        v = np.zeros(d, complex)
        v[0] = 1.0
        for k in range(1, d):
            v[k] = v[k - 1] * alpha / np.sqrt(k)
        v /= np.linalg.norm(v)
        rhos.append(np.outer(v, v.conj()))
    return H, c_ops, np.array(rhos), [num, a + a.conj().T]


def bench_lindblad_matrix(args, torch):
    """LindbladSolver.run for a d = 64 system (beyond the Liouville kernel's d*d <= 1024): the matrix-form
    kernel, n_inst initial states per launch, two observables every step (oqs.py:1596-1696)."""
    from pyqed_amd.liouville import propagate_matrix
    d, n, nsteps, dt = 64, args.n_inst if args.n_inst != 4096 else 1024, 200, 0.01
    H, c_ops, rhos, e_ops = cavity_problem(d, n)
    for _ in range(args.warmup):
        propagate_matrix(H, c_ops, [], rhos, dt, nsteps, None, e_ops, keep_states=False)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        obs, _, final = propagate_matrix(H, c_ops, [], rhos, dt, nsteps, None, e_ops, keep_states=False)
    wall = (time.perf_counter() - t0) / args.steps
    # kernel time alone (the host wrapper also uploads operators and states and downloads results)
    from pyqed_amd import _lib
    L = _lib.lib()
    dev = torch.device("cuda", 0)
    Heff = H - 0.5j * sum(c.conj().T @ c for c in c_ops)
    t = lambda a: torch.as_tensor(np.ascontiguousarray(a), device=dev)  # noqa: E731
    He, HeD = t(Heff), t(Heff.conj().T)
    Cm, CDm = t(np.array(c_ops)), t(np.array([c.conj().T for c in c_ops]))
    rho = t(rhos)
    O = t(np.array([e.T.reshape(-1) for e in e_ops]))
    obs_out = torch.empty((n, nsteps + 1, 2), dtype=torch.complex128, device=dev)
    wb = _lib.C.c_int64()
    _lib.check(L.pyqed_lindblad_work_bytes(d, 2, 0, n, _lib.C.byref(wb)))
    work = torch.empty(wb.value, dtype=torch.uint8, device=dev)
    st = torch.cuda.current_stream()
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]

    def launch():
        _lib.check(L.pyqed_lindblad_run(d, 2, He.data_ptr(), HeD.data_ptr(), Cm.data_ptr(), CDm.data_ptr(), 0, None,
                                        None, n, rho.data_ptr(), None, nsteps, dt, 2, O.data_ptr(), obs_out.data_ptr(),
                                        None, 1, work.data_ptr(), work.numel(), st.cuda_stream))

    launch()
    ev[0].record(st)
    launch()
    ev[1].record(st)
    torch.cuda.synchronize()
    launch_ms = ev[0].elapsed_time(ev[1])
    flops = 4 * (2 + 2 * len(c_ops)) * 8 * d ** 3 * nsteps * n  # complex products of an RK4 step
    achieved = flops / (launch_ms / 1e3) / 1e12
    trace_dev = float(np.abs(np.trace(final, axis1=1, axis2=2) - 1).max())
    out = {
        "metric": LINDBLAD_MATRIX_METRIC, "value": n * nsteps / wall, "unit": "instance-steps/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": wall * 1e3, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (Kerr cavity, seeded coherent initial states)",
        "config": {"workload": f"Kerr cavity truncated at d = {d} (w = 1, chi = 0.01), collapse operators "
                               f"sqrt(0.05) a and sqrt(0.02) a^dag a, {n} initial states per launch, {nsteps} RK4 "
                               f"steps (dt {dt}), <n> and <x> every step; a step is one call (upload, launch, "
                               "download)", "instances": n, "d": d},
        "roofline": {"bound": "mfma", "kernel": "lindblad_matrix_kernel", "launch_ms": launch_ms,
                     "flops_per_launch": flops, "achieved": achieved, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / FP64_PEAK_TFLOPS, "traffic": None,
                     "note": "FP64 VALU FMAs; the MI355X FP64 vector and matrix peaks are equal"},
        "trace_deviation": trace_dev,
    }
    if not args.no_cpu:
        print("[bench] cpu baseline leg", file=sys.stderr, flush=True)
        from oracle import lindblad as OL
        import threadpoolctl
        info = threadpoolctl.threadpool_info()
        nthr = max([i.get("num_threads", 1) for i in info] or [1])
        model, host = cpu_info()
        k = 20
        t0 = time.perf_counter()
        OL.run(H, c_ops, rhos[0], dt, k)
        el = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": k / el, "unit": "instance-steps/s", "cores": nthr, "kind": "port",
                               "cpu_model": model, "host_cores": host,
                               "sample": f"oracle.lindblad.run (numpy FP64 dense products, BLAS threads {nthr}): "
                                         f"{k} RK4 steps of one instance in {el:.2f} s"}
    print(json.dumps(out), flush=True)


SPECTRAL_METRIC = "Liouville-propagator entries/s, eigen-expansion U(t) = U1 e^{lambda t} U1^-1 (FP64), 1 GPU"


def bench_spectral(args, torch):
    """RedfieldSolver.propagator(t, 'SOS') expansion for a 16-level system: N = 256, 1024 times."""
    from pyqed_amd import _lib
    L = _lib.lib()
    N, ns = 256, 1024
    rng = np.random.default_rng(11)
    U1 = rng.normal(size=(N, N)) + 1j * rng.normal(size=(N, N))
    U2 = np.linalg.inv(U1)
    lam = -np.abs(rng.normal(size=N)) + 1j * rng.normal(size=N)
    t = np.arange(ns) * 0.01
    dev = torch.device("cuda", 0)
    A, B = torch.as_tensor(U1, device=dev), torch.as_tensor(U2, device=dev)
    lt, st = torch.as_tensor(lam, device=dev), torch.as_tensor(t, device=dev)
    out = torch.empty((N, N, ns), dtype=torch.complex128, device=dev)
    wb = int(L.pyqed_spectral_work_bytes(N, ns))
    work = torch.empty(wb, dtype=torch.uint8, device=dev)
    stream = torch.cuda.current_stream()

    def launch():
        _lib.check(L.pyqed_spectral_sum(0, N, N, N, ns, A.data_ptr(), lt.data_ptr(), B.data_ptr(), st.data_ptr(),
                                        out.data_ptr(), 1, N * ns, ns, work.data_ptr(), wb, stream.cuda_stream))

    for _ in range(args.warmup):
        launch()
    torch.cuda.synchronize()
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
    t0 = time.perf_counter()
    ev[0].record(stream)
    for _ in range(args.steps):
        launch()
    ev[1].record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    call_ms = ev[0].elapsed_time(ev[1]) / args.steps
    flops = 8.0 * N * N * N * ns
    achieved = flops / (call_ms / 1e3) / 1e12
    res = {
        "metric": SPECTRAL_METRIC, "value": N * N * ns * args.steps / wall, "unit": "entries/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (seeded random eigenvectors, stable eigenvalues)",
        "config": {"workload": f"Redfield propagator expansion: N = {N} (16 levels), {ns} times, [a][b][t] layout"},
        "roofline": {"bound": "mfma", "kernel": "spectral_mfma_kernel (+ spectral_weights_kernel)",
                     "call_ms": call_ms, "flops_per_call": flops, "achieved": achieved, "peak": FP64_PEAK_TFLOPS,
                     "unit": "TFLOP/s", "frac": achieved / FP64_PEAK_TFLOPS, "traffic": None},
    }
    if not args.no_cpu:
        print("[bench] cpu baseline leg", file=sys.stderr, flush=True)
        from threadpoolctl import threadpool_limits
        with threadpool_limits(limits=1):
            nsamp, t0 = 0, time.perf_counter()
            while time.perf_counter() - t0 < 12.0:
                E = np.exp(lam[:, None] * t[None, 32 * nsamp % ns: 32 * nsamp % ns + 32])
                np.einsum("aj,jk,jb->abk", U1, E, U2, optimize=True)
                nsamp += 1
            el = time.perf_counter() - t0
        res["cpu_baseline"] = {"value": N * N * 32 * nsamp / el, "unit": "entries/s", "cores": 1, "kind": "port",
                               "sample": f"{nsamp} blocks of 32 times (numpy einsum 'aj,jk,jb->abk', the reference's "
                                         "contraction, BLAS limited to one thread)"}
    print(json.dumps(res), flush=True)


KRYLOV_METRIC = "two-frequency HEOM correlation grid points/s (Krylov route, FP64), 1 GPU"


def bench_krylov(args, torch):
    """DEOMSolver.correlation_4op_3t(method='krylov') for the 7-site exciton (K = 14, depth 3: 680 ADOs,
    a generator of order 33320, beyond a dense eigen-decomposition), 8 x 8 frequency grid."""
    from pyqed_amd import workloads as W
    from pyqed_amd.heom.deom import Bath, DEOMSolver
    from pyqed_amd.heom.bath import drude
    from pyqed_amd.heom.engine import HeomEngine
    w = W.exciton_c2()
    d = w["H"].shape[0]
    e = drude(w["lam"], w["gam"], w["beta"], w["npsd"])
    bath = Bath.from_arrays(*(np.tile(a, d) for a in e), np.repeat(np.arange(d), len(e[3])))
    s = DEOMSolver(w["H"], None, bath, list(w["Q"]), None, None, None, 3)
    mu = np.zeros((d, d), complex)
    mu[0, 1:] = mu[1:, 0] = 1.0
    wx = np.linspace(-0.002, 0.002, 8)
    wy = np.linspace(-0.002, 0.002, 8)
    s.correlation_4op_3t(mu, mu, mu, mu, w["rho0"], 500.0, wx[:1], wy[:1], lcr="lrcl", method="krylov")  # warm-up
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    C = s.correlation_4op_3t(mu, mu, mu, mu, w["rho0"], 500.0, wx, wy, lcr="lrcl", method="krylov")
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    # the product L x itself: the FP64 right-hand-side kernel on one full-layout vector
    eng = HeomEngine([s._instance(w["rho0"])], packed=False)
    M = int(eng.nmax[0]) * d * d
    x = torch.randn(((M + 15) // 16) * 16, dtype=torch.complex128, device=eng.device)
    for _ in range(3):
        eng.rhs(x)
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
    st = torch.cuda.current_stream()
    ev[0].record(st)
    for _ in range(100):
        eng.rhs(x)
    ev[1].record(st)
    torch.cuda.synchronize()
    mv_ms = ev[0].elapsed_time(ev[1]) / 100
    eng.close()
    byts = 32.0 * M  # read x, write L x, FP64 complex; neighbour re-reads counted once
    res = {
        "metric": KRYLOV_METRIC, "value": C.size / wall, "unit": "grid points/s", "n_gpus": 1, "steps": 1,
        "warmup": 1, "ms_per_step": wall * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic (7-site exciton chain, Drude baths; transition-dipole operators)",
        "config": {"workload": f"7-site exciton, depth 3, generator order {M}, 8 x 8 grid, T = 500 au, tol 1e-10"},
        "roofline": {"bound": "hbm", "kernel": "L x (pyqed_heom_rhs, FP64 generic stage kernel)",
                     "launch_ms": mv_ms, "bytes_per_launch": byts, "achieved": byts / (mv_ms / 1e3) / 1e9,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": byts / (mv_ms / 1e3) / 1e9 / HBM_PEAK_GBS,
                     "traffic": None},
        "cpu_baseline": None,
        "note": "no CPU baseline: the reference's route needs la.eig of the dense generator (order 33320)",
    }
    print(json.dumps(res), flush=True)


HEOM_ADJ_METRIC = "HEOM gradients/s (streamed FP64 adjoint, 7-site exciton depth 4, 1000 steps), 1 GPU"


def heom_adjoint_problem(depth=4, nseg=50, sps=20, ckpt_every=50):
    """7-site exciton (K = 14, Drude baths) with H(theta) = H + sum_j theta_j |j><j| over all sites and a
    transition dipole (nearest-neighbour hopping operator) driven by a piecewise-constant field of
    nseg segments of sps steps (dt = 1 fs); J = population of site 2 at T."""
    from pyqed_amd import workloads as W
    from pyqed_amd.heom.adjoint import HeomControl
    from pyqed_amd.heom.bath import Bath
    w = W.exciton_c2()
    d = w["H"].shape[0]
    b = Bath.drude(w["lam"], w["gam"], w["beta"], 1, nbath=d)
    P = [np.diag(np.eye(d)[j]).astype(complex) for j in range(d)]
    mu = np.zeros((d, d), complex)
    for j in range(d - 1):
        mu[j, j + 1] = mu[j + 1, j] = 1.0
    prob = HeomControl(w["H"], P, mu, w["Q"], (b.etal, b.etar, b.etaa, b.expn, b.mode), depth, w["rho0"], w["dt"],
                       sps, nseg, ckpt_every=ckpt_every)
    theta = np.zeros(d)
    amps = np.random.default_rng(7).normal(0.0, 2e-4, nseg)
    bar = np.zeros((d, d), complex)
    bar[2, 2] = 1.0
    return prob, theta, amps, bar


def bench_heom_adjoint(args, torch):
    """One gradient of a final population with respect to 7 site energies and 50 field amplitudes through
    1000 RK4 steps of a 3060-ADO hierarchy: forward run with checkpoints, then the backward sweep (per
    step 3 forward and 4 adjoint right-hand sides on the FP64 stage kernel and one contraction)."""
    from pyqed_amd import _lib
    from pyqed_amd.heom.adjoint import HeomAdjoint, piecewise_field
    prob, theta, amps, bar = heom_adjoint_problem()
    prob.gradient(theta, amps, bar)  # warm-up (plan build, code load)
    torch.cuda.synchronize()
    reps = max(1, args.steps if args.steps else 1)
    t0 = time.perf_counter()
    for _ in range(reps):
        gt, ga, _, r = prob.gradient(theta, amps, bar)
    torch.cuda.synchronize()
    wall = (time.perf_counter() - t0) / reps
    # the dominant kernel: the FP64 right-hand side (generic stage kernel) on the adjoint plan
    H = prob.H0 + sum(float(t) * p for t, p in zip(theta, prob.P))
    inst = dict(H=H, Q=prob.Q, bath=prob.bath, L=prob.Ldepth, rho0=prob.rho0, sdip=prob.mu)
    ha = HeomAdjoint([inst], prob.dt, prob.nt, ckpt_every=prob.ckpt_every)
    f = torch.as_tensor(piecewise_field(amps, prob.sps), device=ha.fwd.device)
    x = torch.randn(ha.fwd.n_elems, dtype=torch.complex128, device=ha.fwd.device)
    out = torch.empty_like(x)
    st = torch.cuda.current_stream()
    for _ in range(3):
        ha._rhs(ha.adj, x, out, f, 0, 1)
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
    ev[0].record(st)
    for _ in range(200):
        ha._rhs(ha.adj, x, out, f, 0, 1)
    ev[1].record(st)
    torch.cuda.synchronize()
    rhs_ms = ev[0].elapsed_time(ev[1]) / 200
    M = int(ha.fwd.nmax[0]) * 49
    ha.close()
    byts = 32.0 * M  # read X, write L^H X (complex128); neighbour re-reads counted once
    nt = prob.nt
    res = {
        "metric": HEOM_ADJ_METRIC, "value": 1.0 / wall, "unit": "gradients/s", "n_gpus": 1, "steps": reps,
        "warmup": 1, "ms_per_step": wall * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic (7-site exciton chain, seeded energies, Drude baths, seeded field)",
        "config": {"workload": f"7-site exciton HEOM depth 4 (3060 ADOs), {nt} RK4 steps, {prob.nseg} field "
                               f"segments, gradient w.r.t. 7 site energies + {prob.nseg} amplitudes, checkpoints "
                               f"every {prob.ckpt_every} steps",
                   "backward_us_per_step": wall / nt * 1e6},
        "roofline": {"bound": "hbm", "kernel": "heom_stage_kernel<double, double, double, 4, false> (L^H x, "
                                               "adjoint plan)", "launch_ms": rhs_ms, "bytes_per_launch": byts,
                     "achieved": byts / (rhs_ms / 1e3) / 1e9, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": byts / (rhs_ms / 1e3) / 1e9 / HBM_PEAK_GBS, "traffic": None,
                     "note": "state of 2.4 MB: cache-resident and launch-latency bound, not HBM-meaningful"},
        "gradient_norm": float(np.sqrt(np.sum(gt ** 2) + np.sum(ga ** 2))),
    }
    if not args.no_cpu:
        res["cpu_baseline"] = heom_adjoint_cpu_baseline(prob, theta, amps)
    print(json.dumps(res), flush=True)


def heom_adjoint_cpu_baseline(prob, theta, amps, budget_steps=20):
    """FP64 torch autograd of the same propagation (the oracle's restatement) on the host's cores, over
    the first budget_steps steps (one segment), scaled to the full step count."""
    import torch
    from oracle import heom_torch as T
    model, host = cpu_info()
    nthr = max(1, min(16, host))
    torch.set_num_threads(nthr)
    etal, etar, etaa, expn, mode = prob.bath
    th = torch.tensor(theta, dtype=torch.float64, requires_grad=True)
    nseg = max(1, budget_steps // prob.sps)
    am = torch.tensor(amps[:nseg], dtype=torch.float64, requires_grad=True)
    t0 = time.perf_counter()
    rT = T.final_state(th, am, prob.H0, np.array(prob.P), prob.mu, prob.Q, (etal, etar, etaa, expn), mode,
                       prob.Ldepth, prob.rho0, prob.dt, prob.sps)
    rT[2, 2].real.backward()
    el = time.perf_counter() - t0
    steps = nseg * prob.sps
    per_grad = el / steps * prob.nt
    return {"value": 1.0 / per_grad, "unit": "gradients/s", "cores": nthr, "kind": "port", "cpu_model": model,
            "host_cores": host,
            "sample": f"oracle.heom_torch (FP64 torch autograd, {nthr} threads): forward and backward through the "
                      f"first {steps} steps in {el:.1f} s, scaled to {prob.nt} steps"}


# Full-length accuracy check per workload: the step count each configuration is quoted at.
FULL_STEPS = {"c1": 1000, "c2": 1000, "c3": 2000, "c5": 5000}


def error_sample(insts, n=64, seed=11):
    """Indices of a seeded accuracy sample: one instance of every (d, depth) stratum, every instance of
    the largest d at the largest depth (up to 4), then seeded others up to n."""
    strata = _strata(insts)
    idx = sorted(strata.values())
    dmax = max(p["H"].shape[0] for p in insts)
    lmax = max(p["L"] for p in insts if p["H"].shape[0] == dmax)
    big = [i for i, p in enumerate(insts) if p["H"].shape[0] == dmax and p["L"] == lmax]
    for i in big[:4]:
        if i not in idx:
            idx.append(i)
    rest = [i for i in np.random.default_rng(seed).permutation(len(insts)) if i not in idx]
    idx += [int(i) for i in rest[:max(0, n - len(idx))]]
    return sorted(idx)


def full_length_error(insts, dt, nt, torch, n=64, save_every=20):
    """Mixed (packed, FP32 right-hand side) against plain FP64, both from rho0 over nt steps, on the
    seeded sample: max-abs population and coherence difference of rho over every saved time."""
    from pyqed_amd.heom.engine import HeomEngine
    idx = error_sample(insts, n) if len(insts) > 1 else [0]
    sub = [insts[i] for i in idx]
    traj = {}
    t_leg = {}
    for prec in ("mixed", "fp64"):
        eng = HeomEngine(sub)
        t0 = time.perf_counter()
        traj[prec] = eng.run(dt, nt, precision=prec, save_every=save_every)
        t_leg[prec] = time.perf_counter() - t0
        eng.close()
        del eng
        torch.cuda.empty_cache()
    pops, cohs = zip(*[max_errs(a, b) for a, b in zip(traj["mixed"], traj["fp64"])])
    strata = {(p["H"].shape[0], p["L"]) for p in sub}
    worst = int(np.argmax(pops))
    out = {"population": max(pops), "coherence": max(cohs), "steps": nt, "save_every": save_every,
           "instances": len(sub), "strata": len(strata), "d_max": max(d for d, _ in strata),
           "worst_instance": {"index": idx[worst], "d": sub[worst]["H"].shape[0], "depth": sub[worst]["L"],
                              "population": pops[worst]},
           "seconds": t_leg, "reference": "the same kernels in plain FP64 (packed state, FP64 stage inputs and "
                                          "RK4 register), which the golden tests pin to the reference at 1e-12"}
    return out, sub, traj["mixed"]


def oracle_error(sub, mixed, dt, nt, save_every=20, budget_s=30.0):
    """The mixed GPU trajectories against the CPU oracle (oracle.heom.run: the reference's RK4 and
    right-hand side restated in numpy FP64, bit-identical to the reference on the spin-boson
    trajectory) over the full nt steps: the cheapest instance of each (d, depth) stratum first, as many
    strata as the CPU budget allows (an instance that would not finish within it is dropped)."""
    from math import comb
    from oracle import heom as O
    cost = [comb(len(p["bath"][3]) + p["L"], p["L"]) * p["H"].shape[0] ** 2 * len(p["bath"][3]) for p in sub]
    order = sorted(range(len(sub)), key=lambda k: cost[k])
    seen, pops, cohs, used = set(), [], [], []
    rate = None
    t_start = time.perf_counter()
    for k in order:
        p = sub[k]
        key = (p["H"].shape[0], p["L"])
        if key in seen:
            continue
        left = budget_s - (time.perf_counter() - t_start)
        if left <= 0 or (rate is not None and cost[k] * nt * rate > 2 * left):
            continue
        etal, etar, etaa, expn, mode = p["bath"]
        Q = [np.asarray(q, np.complex128) for q in p["Q"]]
        t0 = time.perf_counter()
        if all(np.count_nonzero(q - np.diag(np.diag(q))) == 0 for q in Q):
            # diagonal couplings: the oracle's row/column-scaling form of the same right-hand side
            hier = O.Hierarchy(len(expn), p["L"])
            terms = O.DiagTerms(hier, (etal, etar, etaa, expn), mode, np.array([np.diag(q) for q in Q]))
            H = np.asarray(p["H"], np.complex128)
            y = np.zeros((hier.nmax,) + H.shape, np.complex128)
            y[0] = p["rho0"]
            ref = [y[0].copy()]
            for i in range(nt):
                y = O.rk4_step(y, lambda x, t: O.rhs_vec_diag(x, H, terms), dt, i * dt)
                if (i + 1) % save_every == 0:
                    ref.append(y[0].copy())
                if i % 100 == 99 and time.perf_counter() - t_start > budget_s:
                    break  # over budget: this instance is dropped
            if len(ref) < nt // save_every + 1:
                break
            ref = np.array(ref)
        else:
            _, ref = O.run(p["H"], Q, (etal, etar, etaa, expn), mode, p["L"], p["rho0"], dt, nt,
                           save_every=save_every)
        el = time.perf_counter() - t0
        rate = el / (cost[k] * nt)  # per unit of cost, from the latest (largest) instance
        pop, coh = max_errs(mixed[k], ref)
        pops.append(pop)
        cohs.append(coh)
        used.append(key)
        seen.add(key)
    if not used:
        return None
    return {"population": max(pops), "coherence": max(cohs), "steps": nt, "save_every": save_every,
            "instances": len(used), "strata": len(used), "d_max": max(d for d, _ in used),
            "depth_max": max(L for _, L in used), "seconds": time.perf_counter() - t_start,
            "oracle": "the numpy FP64 restatement of deom.py:641-766 (oracle.heom; its row/column-scaling form for "
                      "diagonal couplings)"}


def fp64_batched(insts, dt, warmup, steps, torch):
    """Plain FP64 with the same batched kernels as the mixed leg (one plan over all instances, packed
    state when Hermitian): time `steps` steps after `warmup`.  When one plan does not fit the device,
    the instances run in chunks and the chunk times add."""
    from pyqed_amd.heom.engine import HeomEngine
    chunks = [insts]
    while True:
        try:
            total, ados, finals = 0.0, 0, []
            for ch in chunks:
                eng = HeomEngine(ch)
                eng.advance(warmup, dt, "fp64")
                torch.cuda.synchronize()
                t0 = time.perf_counter()
                eng.advance(steps, dt, "fp64")
                torch.cuda.synchronize()
                total += time.perf_counter() - t0
                finals += eng.reduced_states()
                ados += int(eng.n_ados)
                eng.close()
                del eng
                torch.cuda.empty_cache()
            return ados * steps / total, total, finals, len(chunks)
        except (torch.cuda.OutOfMemoryError, RuntimeError) as e:
            if "memory" not in str(e).lower() or len(chunks[0]) < 2:
                raise
            torch.cuda.empty_cache()
            chunks = [c for ch in chunks for c in (ch[:len(ch) // 2], ch[len(ch) // 2:])]
            print(f"[bench] fp64 batched leg: out of memory, retrying in {len(chunks)} chunks", file=sys.stderr,
                  flush=True)


def bench_heom(args, torch):
    """ADO-updates/s of DEOMSolver's propagation (pyqed/heom/deom.py:725-766) for one HEOM workload."""
    from pyqed_amd.heom.engine import HeomEngine
    insts, dt, desc = instances(args.workload, args.n_inst)
    print(f"[bench] workload {args.workload}: {len(insts)} instance(s); building plan", file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    eng = HeomEngine(insts, packed=False if args.state == "full" else None)
    if args.tiling:
        eng.autotune_cache = False  # an explicit tiling, e.g. to profile at the tiling another run used
        eng.set_tiling(*(int(x) for x in args.tiling.split(",")))
    torch.cuda.synchronize()
    t_plan = time.perf_counter() - t0
    st = eng.stats()
    print(f"[bench] plan {t_plan:.1f}s: {eng.n_ados} ADOs, {eng.n_elems} elements, {st}", file=sys.stderr,
          flush=True)
    eng.advance(args.warmup, dt, args.precision)
    torch.cuda.synchronize()
    print(f"[bench] warmup done; state at {eng.y.data_ptr():#x}, workspace at {eng._ws.data_ptr():#x}",
          file=sys.stderr, flush=True)
    if args.graph:
        eng.advance_graph(args.graph, dt, args.precision, chunk=args.graph)  # capture (+ one replay) untimed
        torch.cuda.synchronize()
    evs = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    evs[0].record()
    if args.graph:
        for s in range(0, args.steps, args.graph):
            eng.advance_graph(min(args.graph, args.steps - s), dt, args.precision, chunk=args.graph)
        evs[-1].record()
    else:
        # one call for all timed steps, as DEOMSolver.run makes (a packed plan rebuilds the stage-0
        # input from y once per call)
        eng.advance(args.steps, dt, args.precision)
        evs[-1].record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    step_ms = [evs[0].elapsed_time(evs[-1]) / args.steps]
    tiling = eng.tiling()
    tiling_cached = bool(getattr(eng, "_tuned_hit", False))
    value = eng.n_ados * args.steps / wall
    print(f"[bench] timed {args.steps} steps in {wall:.3f}s", file=sys.stderr, flush=True)
    mixed_final = eng.reduced_states()  # the state after warmup + timed steps (accuracy leg)
    # per-launch times: an event pair around every kernel-class launch, on the launch stream, in a
    # separate pass after the timed steps (events would lengthen a launch-bound step)
    timing = not args.graph and not args.no_launch_timing
    n_instr = min(args.steps, 10)
    if timing:
        eng.set_timing(True)
        eng.advance(n_instr, dt, args.precision)
        torch.cuda.synchronize()
    ts = 8 if args.precision == "mixed" else 16
    F = eng.classes()["elems"]
    S = eng.class_state()
    nc = len(F)
    step_bytes = sum(alg_bytes(s, ts, F[c]) for s in range(4) for c in range(nc))
    step_layout = sum(stage_bytes(s, ts, F[c], S[c], eng.packed) for s in range(4) for c in range(nc))
    t_step = float(np.mean(step_ms)) / 1e3
    roofline = {"bound": "hbm", "peak": HBM_PEAK_GBS, "unit": "GB/s", "traffic": None,
                "bytes_model": "classic three-register RK4, %d B per full-matrix element per step"
                               % sum(ALG["mixed" if ts == 8 else "fp64"]),
                "step": {"bytes": step_bytes, "ms": t_step * 1e3, "achieved": step_bytes / t_step / 1e9,
                         "frac": step_bytes / t_step / 1e9 / HBM_PEAK_GBS, "layout_bytes": step_layout}}
    avg = {}
    if timing:
        ms, launches = eng.read_timing()
        eng.set_timing(False)
        avg = {k: ms[k] / launches[k] for k in range(len(ms)) if launches[k]}
    if avg:  # no streamed launches when every instance runs LDS-resident
        k = max(avg, key=avg.get)
        s_dom, c_dom = divmod(k, nc)
        b = alg_bytes(s_dom, ts, F[c_dom])
        bl = stage_bytes(s_dom, ts, F[c_dom], S[c_dom], eng.packed)
        name = kernel_name(s_dom, c_dom, args.precision, eng)
        roofline.update({"kernel": name, "stage": s_dom, "class": c_dom, "launch_ms": avg[k],
                         "bytes_per_launch": b, "achieved": b / (avg[k] / 1e3) / 1e9,
                         "frac": b / (avg[k] / 1e3) / 1e9 / HBM_PEAK_GBS,
                         "layout_bytes_per_launch": bl, "layout_achieved": bl / (avg[k] / 1e3) / 1e9,
                         "launches": {f"s{kk // nc}c{kk % nc}": {
                             "ms": v, "bytes": alg_bytes(kk // nc, ts, F[kk % nc]),
                             "frac": alg_bytes(kk // nc, ts, F[kk % nc]) / (v / 1e3) / 1e9 / HBM_PEAK_GBS}
                             for kk, v in sorted(avg.items())}})
        roofline["traffic"] = measured_traffic(name, args.workload, desc, eng.packed, tiling)
    else:
        roofline.update({"kernel": "all stage kernels of a step" if eng.stats()["stream_items"] else
                         "heom_resident_kernel (every instance LDS-resident)",
                         "achieved": roofline["step"]["achieved"], "frac": roofline["step"]["frac"]})
    if avg:
        roofline["instrumented_steps"] = n_instr
    packed = eng.packed
    eng.close()
    del eng
    torch.cuda.empty_cache()

    mixed = args.precision == "mixed"
    out = {
        "metric": METRIC, "value": value, "unit": "ADO-updates/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": wall / args.steps * 1e3, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None,
        "dtype": ("f32 right-hand side and stage inputs, f32 RK4 increment registers, f64 state" if mixed
                  else "f64"),
        "data": ("synthetic (seeded numpy.random.default_rng; closed-form Drude + Pade baths in physical "
                 "normalisation: the reference's Bath would carry a factor 1/gamma^2 on every eta of an instance "
                 "with gamma > 1)"),
        "config": {"workload": desc, "instances": len(insts), "precision": args.precision},
        "roofline": roofline,
        "plan_build_s": t_plan, "plan": st, "tiling": tiling, "tiling_source": "explicit" if args.tiling else (
            "autotuner cache" if tiling_cached else "automatic"), "graph_chunk": args.graph or None,
        "device": torch.cuda.get_device_name(), 
        "state_layout": "packed Hermitian (upper triangles)" if packed else "full matrices",
    }
    from math import comb
    out["config"]["ados"] = int(sum(comb(len(i["bath"][4]) + i["L"], i["L"]) for i in insts))
    out["config"]["elements"] = int(st["stream_elems"] + st["resident_elems"])
    pre = args.warmup + (args.graph or 0)  # steps the mixed run made before its final state
    if mixed and not args.no_fp64:
        print("[bench] fp64 batched leg", file=sys.stderr, flush=True)
        v64b, t64b, finals_b, nch = fp64_batched(insts, dt, pre, args.steps, torch)
        out["fp64_batched"] = {"value": v64b, "unit": "ADO-updates/s", "seconds": t64b, "chunks": nch,
                               "kernel": "the same batched stage kernels in plain FP64 (one plan over all "
                                         "instances, packed state, FP64 stage inputs and RK4 register)"}
        out["ratio_vs_fp64_batched"] = value / v64b
        pops, cohs = zip(*[max_errs(a, b) for a, b in zip(mixed_final, finals_b)])
        short = {"population": max(pops), "coherence": max(cohs), "steps": pre + args.steps,
                 "instances": len(insts), "reference": "fp64_batched"}
        if not args.no_fp64_plain:
            print("[bench] fp64 plain leg", file=sys.stderr, flush=True)
            v64, t64, _ = fp64_plain(insts, dt, pre, args.steps, torch)
            out["fp64_plain"] = {"value": v64, "unit": "ADO-updates/s", "seconds": t64,
                                 "kernel": "FP64 stage kernels (projector or generic class as the instance allows), "
                                           "one instance per plan, full-matrix state, no LDS residency"}
            out["ratio_vs_fp64_plain"] = value / v64
        if not args.no_full_error and args.workload in FULL_STEPS:
            nt = FULL_STEPS[args.workload]
            print(f"[bench] full-length accuracy leg ({nt} steps)", file=sys.stderr, flush=True)
            out["error_vs_fp64"], sample, mixed_traj = full_length_error(insts, dt, nt, torch)
            out["error_vs_fp64"]["after_timed_steps"] = short
            if not args.no_cpu:
                print("[bench] oracle accuracy leg", file=sys.stderr, flush=True)
                out["error_vs_oracle"] = oracle_error(sample, mixed_traj, dt, nt)
        else:
            out["error_vs_fp64"] = short
    if not args.no_cpu:
        print("[bench] cpu baseline leg", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline(insts, dt, max_steps=FULL_STEPS.get(args.workload, 1))
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None, help="timed steps (default 20; 2000 for --workload legacy)")
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="c3", choices=["c1", "c2", "c3", "c4", "c5", "lindblad", "legacy",
                                                              "legacy-rk4", "spectral", "krylov", "heom-adjoint", "lindblad-matrix"])
    ap.add_argument("--n-inst", type=int, default=None, help="instances (default 4096; 65536 for legacy)")
    ap.add_argument("--precision", default="mixed", choices=["mixed", "fp64"])
    ap.add_argument("--state", default="auto", choices=["auto", "full"],
                    help="RK4 state layout: auto = upper triangles for Hermitian hierarchies, full = full matrices")
    ap.add_argument("--no-launch-timing", action="store_true",
                    help="no per-launch HIP events (roofline from the step time only)")
    ap.add_argument("--no-fp64-plain", action="store_true", help="skip the one-instance-per-plan FP64 leg")
    ap.add_argument("--no-fp64", action="store_true", help="skip every FP64 leg (and the accuracy legs)")
    ap.add_argument("--no-full-error", action="store_true",
                    help="skip the full-length mixed-vs-FP64 accuracy leg on the seeded sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--traffic-from", nargs=5, metavar=("FETCH_DIR", "WRITE_DIR", "COPY_FETCH_DIR", "COPY_WRITE_DIR",
                                                       "BENCH_JSON"),
                    help="only convert rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE runs of a bench command (and of "
                         "--copy-calibration) into the per-kernel traffic summary")
    ap.add_argument("--copy-mib", type=int, default=1024, help="size of the --copy-calibration run (MiB)")
    ap.add_argument("--tiling", default=None, metavar="TARGET,BLOCKS_PER_CU,VARIANT",
                    help="work-queue tiling of the mixed leg instead of the autotuner's cached choice")
    ap.add_argument("--graph", type=int, default=0, metavar="CHUNK",
                    help="time graph replays of CHUNK captured steps instead of eager launches")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run one device-to-device copy of MIB MiB (scale reference for HBM byte counters)")
    ap.add_argument("--stream-probe", action="store_true",
                    help="only measure achievable HBM bandwidth: device copies and a 3-read/2-write stream")
    args = ap.parse_args()
    legacy = args.workload in ("legacy", "legacy-rk4")
    if args.steps is None:
        args.steps = 2000 if legacy else (3 if args.workload in ("c4", "lindblad", "heom-adjoint", "lindblad-matrix") else 20)
    if args.n_inst is None:
        args.n_inst = 65536 if legacy else 4096
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py runs on one GPU only: the propagator has no multi-GPU path (the instances of a "
              "batch are independent and would split over GPUs by a size-balanced partition).", file=sys.stderr)
        sys.exit(2)
    import torch
    from pyqed_amd.heom.engine import HeomEngine

    if args.copy_calibration:
        n = args.copy_calibration * (1 << 20)
        src = torch.ones(n, dtype=torch.uint8, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_bytes": n}), flush=True)
        return
    if args.stream_probe:
        print(json.dumps(stream_probe(torch)), flush=True)
        return
    if legacy:
        bench_legacy(args, torch)
        return
    if args.workload == "spectral":
        bench_spectral(args, torch)
        return
    if args.workload == "c4":
        bench_c4(args, torch)
        return
    if args.workload == "lindblad-matrix":
        return bench_lindblad_matrix(args, torch)
    if args.workload == "lindblad":
        bench_lindblad(args, torch)
        return
    if args.workload == "krylov":
        bench_krylov(args, torch)
        return
    if args.workload == "heom-adjoint":
        bench_heom_adjoint(args, torch)
        return

    if args.traffic_from:
        traffic_from(*args.traffic_from, copy_mib=args.copy_mib, key=args.workload)
        return
    bench_heom(args, torch)


if __name__ == "__main__":
    main()
