"""GPU engine behind DEOMSolver and HEOMBatch: one libpyqed_amd plan over one or
many HEOM instances, with the packed state resident in HBM as a torch tensor.

Each instance is a dict with
    H      (d, d) system Hamiltonian
    Q      (ncoup, d, d) coupling operators Q_m
    bath   (etal, etar, etaa, expn, mode) arrays of length K
    L      hierarchy depth
    rho0   (d, d) initial reduced density matrix
    sdip   optional (d, d) system dipole (H(t) = H + sdip f_s(t))
    cdip   optional (ncoup, d, d) coupling dipoles (Q_m(t) = Q_m + cdip_m f_c(t))

Couplings that are all diagonal are passed as diagonals.  Couplings that are
not diagonal but commute (including a single coupling operator) are
diagonalised together: the instance is propagated in that eigenbasis, which
leaves the hierarchy equations unchanged (every term is a product of
operators), and results are rotated back.  Couplings that do not commute use
the dense-coupling kernel path.
"""
import ctypes as C

import numpy as np

from .. import _lib

PRECISIONS = {"fp64": _lib.PREC_FP64, "mixed": _lib.PREC_MIXED, "fp32": _lib.PREC_MIXED}
NCLASSES = 4  # streamed kernel classes of a plan (pyqed_heom_plan_classes)


def _is_diag(m, tol=0.0):
    off = m - np.diag(np.diag(m))
    return np.abs(off).max(initial=0.0) <= tol


def common_eigenbasis(mats, tol=1e-10):
    """Unitary U with U^H M U diagonal for every Hermitian M in mats, or None."""
    mats = [np.asarray(m, np.complex128) for m in mats]
    if not all(np.allclose(m, m.conj().T, atol=1e-12) for m in mats):
        return None
    rng = np.random.default_rng(12345)
    comb = sum(rng.uniform(0.5, 1.5) * m for m in mats)
    _, U = np.linalg.eigh(comb)
    scale = max(1.0, max(np.abs(m).max() for m in mats))
    for m in mats:
        if not _is_diag(U.conj().T @ m @ U, tol * scale):
            return None
    return U


def prepare_instance(inst):
    """Choose the working basis of one instance; returns a dict of complex128 arrays ready for the plan.
    Non-finite initial states raise ValueError (the library rejects non-finite operators and baths)."""
    if not np.all(np.isfinite(np.asarray(inst["rho0"], np.complex128))):
        raise ValueError("rho0 has a non-finite entry (NaN or Inf)")
    H = np.asarray(inst["H"], np.complex128)
    d = H.shape[0]
    Q = np.asarray(inst["Q"], np.complex128).reshape(-1, d, d)
    sdip = inst.get("sdip")
    cdip = inst.get("cdip")
    sdip = None if sdip is None else np.asarray(sdip, np.complex128)
    cdip = None if cdip is None else np.asarray(cdip, np.complex128).reshape(-1, d, d)
    coup = list(Q) + ([] if cdip is None else list(cdip))
    U = None
    dense = False
    if not all(_is_diag(m) for m in coup):
        U = common_eigenbasis(coup)
        if U is None:
            dense = True
    rot = (lambda m: m) if U is None else (lambda m: U.conj().T @ m @ U)
    out = dict(d=d, U=U, dense=dense, H=rot(H), sdip=None if sdip is None else rot(sdip),
               rho0=rot(np.asarray(inst["rho0"], np.complex128)), L=int(inst["L"]))
    Qw = np.array([rot(q) for q in Q])
    Cw = None if cdip is None else np.array([rot(c) for c in cdip])
    out["Qmat"], out["Cmat"] = Qw, Cw
    out["qdiag"] = np.array([np.diag(q) for q in Qw])
    out["cdiag"] = None if Cw is None else np.array([np.diag(c) for c in Cw])
    etal, etar, etaa, expn, mode = inst["bath"]
    out["bath"] = tuple(np.ascontiguousarray(np.asarray(a, np.complex128)) for a in (etal, etar, etaa, expn))
    out["mode"] = np.ascontiguousarray(np.asarray(mode, np.int32))
    return out


def _herm(m, tol=1e-14):
    return np.abs(m - m.conj().T).max(initial=0.0) <= tol * (1.0 + np.abs(m).max(initial=0.0))


def _hermitian_dynamics(p):
    """True when every ADO stays Hermitian: Hermitian H, dipole and rho0, real diagonal couplings,
    no coupling dipole, real exponents and etar = conj(etal) for every exponential."""
    etal, etar, _, expn = p["bath"]
    return (_herm(p["H"]) and (p["sdip"] is None or _herm(p["sdip"])) and _herm(p["rho0"]) and p["Cmat"] is None
            and np.all(np.imag(p["qdiag"]) == 0) and np.all(np.imag(expn) == 0)
            and np.all(np.abs(etar - np.conj(etal)) <= 1e-14 * np.abs(etal)))


def pack_upper(a):
    """Upper triangles (diagonal included) of matrices [n, d, d], row by row, as one flat array: the
    packed RK4-state layout, element (u, v), u <= v, of an ADO at u(2d-u-1)/2 + v."""
    a = np.asarray(a, np.complex128)
    d = a.shape[-1]
    iu = np.triu_indices(d)
    return a.reshape(-1, d, d)[:, iu[0], iu[1]].reshape(-1)


def unpack_upper(v, d):
    """Full Hermitian matrices [n, d, d] from packed upper triangles (inverse of pack_upper)."""
    iu = np.triu_indices(d)
    v = np.asarray(v, np.complex128).reshape(-1, len(iu[0]))
    up = np.zeros((v.shape[0], d, d), np.complex128)
    up[:, iu[0], iu[1]] = v
    low = np.conj(np.transpose(up, (0, 2, 1)))
    low[:, np.arange(d), np.arange(d)] = 0
    return up + low


class HeomEngine:
    """A plan over one or more HEOM instances plus its device state."""

    def __init__(self, instances, *, resident=True, generic=False, hermitian=None, packed=None, adjoint=False):
        self.torch = torch = _lib.require_gpu()
        self.L = _lib.lib()
        self.prep = [prepare_instance(i) for i in instances]
        self.n = len(self.prep)
        self.dense = any(p["dense"] for p in self.prep)
        keep = []
        arr = (_lib.HeomInstance * self.n)()
        for k, p in enumerate(self.prep):
            d = p["d"]
            bufs = dict(
                mode=p["mode"], expn=p["bath"][3], etal=p["bath"][0], etar=p["bath"][1], etaa=p["bath"][2],
                H=np.ascontiguousarray(p["H"]),
                sdip=None if p["sdip"] is None else np.ascontiguousarray(p["sdip"]),
            )
            if self.dense:
                bufs["qmat"] = np.ascontiguousarray(p["Qmat"])
                bufs["cmat"] = None if p["Cmat"] is None else np.ascontiguousarray(p["Cmat"])
            else:
                bufs["qdiag"] = np.ascontiguousarray(p["qdiag"])
                bufs["cdiag"] = None if p["cdiag"] is None else np.ascontiguousarray(p["cdiag"])
            keep.append(bufs)
            s = arr[k]
            s.d, s.K, s.L, s.ncoup = d, len(p["mode"]), p["L"], p["qdiag"].shape[0]
            for name, b in bufs.items():
                setattr(s, name, None if b is None else b.ctypes.data)
        if hermitian is None:
            hermitian = not adjoint and not self.dense and all(_hermitian_dynamics(p) for p in self.prep)
        self.hermitian = bool(hermitian)
        # Hermitian hierarchies keep only the upper triangle of every ADO in the FP64 state
        self.packed = self.hermitian if packed is None else bool(packed)
        # adjoint: the plan applies L^H (same layout and tables), for the streamed HEOM adjoint
        self.adjoint = bool(adjoint)
        flags = ((0 if resident else _lib.PLAN_NO_RESIDENT) | (_lib.PLAN_DENSE_Q if self.dense else 0) |
                 (_lib.PLAN_GENERIC if generic else 0) | (_lib.PLAN_HERMITIAN if self.hermitian else 0) |
                 (_lib.PLAN_PACKED if self.packed else 0) | (_lib.PLAN_ADJOINT if self.adjoint else 0))
        plan = C.c_void_p()
        _lib.check(self.L.pyqed_heom_plan_create(arr, self.n, flags, C.byref(plan)))
        self.plan = plan
        ne, na, w64, wmx = C.c_int64(), C.c_int64(), C.c_int64(), C.c_int64()
        _lib.check(self.L.pyqed_heom_plan_query(plan, C.byref(ne), C.byref(na), C.byref(w64), C.byref(wmx)))
        self.n_elems, self.n_ados = ne.value, na.value
        self.ws_bytes = {_lib.PREC_FP64: w64.value, _lib.PREC_MIXED: wmx.value}
        so = np.zeros(self.n, np.int64)
        nm = np.zeros(self.n, np.int64)
        to = np.zeros(self.n, np.int64)
        _lib.check(self.L.pyqed_heom_plan_layout(plan, so.ctypes.data, nm.ctypes.data, to.ctypes.data))
        self.state_offset, self.nmax, self.traj_offset = so, nm, to
        self.traj_stride = int(sum(p["d"] ** 2 for p in self.prep))
        self.device = torch.device("cuda", torch.cuda.current_device())
        self.y = torch.zeros(self.n_elems, dtype=torch.complex128, device=self.device)
        self._ws = None
        self._ws_prec = None
        self._tuned = set()
        self.variant = 0
        self.autotune_cache = True
        self.step = 0
        self.reset()

    # ------------------------------------------------------------------ state
    # ------------------------------------------------------------------ state layout
    def state_size(self, k):
        """RK4-state elements per ADO of instance k: d*d, or d(d+1)/2 when packed."""
        d = self.prep[k]["d"]
        return d * (d + 1) // 2 if self.packed else d * d

    def _to_state(self, k, a):
        """[n, d, d] matrices -> [n * state_size] state elements (upper triangles when packed)."""
        d = self.prep[k]["d"]
        a = np.asarray(a, np.complex128).reshape(-1, d, d)
        return pack_upper(a) if self.packed else a.reshape(-1)

    def _from_state(self, k, v):
        """Inverse of _to_state: full [n, d, d] matrices (Hermitian completion when packed)."""
        d = self.prep[k]["d"]
        return unpack_upper(v, d) if self.packed else np.asarray(v).reshape(-1, d, d)

    # ------------------------------------------------------------------ state
    def reset(self):
        """y <- every instance at rho0 in ADO 0, zero elsewhere; time index back to 0."""
        torch = self.torch
        self.y.zero_()
        vals = [self._to_state(k, p["rho0"]) for k, p in enumerate(self.prep)]
        idx = np.concatenate([self.state_offset[k] + np.arange(v.size) for k, v in enumerate(vals)])
        self.y[torch.as_tensor(idx, device=self.device)] = torch.as_tensor(np.concatenate(vals), device=self.device)
        self.step = 0

    def workspace(self, prec):
        torch = self.torch
        need = self.ws_bytes[prec]
        if self._ws is None or self._ws.numel() < need:
            # captured graphs hold the old workspace's address: drop them before it is released
            self._drop_graphs()
            self._ws = torch.empty(max(need, 256), dtype=torch.uint8, device=self.device)
        return self._ws

    def _drop_graphs(self):
        """Forget every captured step graph.  A graph holds raw device addresses of the workspace and of
        the plan's work queue, so it must not outlive either (workspace growth, a tiling or variant
        change, close())."""
        g = getattr(self, "_graphs", None)
        if g:
            self.torch.cuda.synchronize()
            g.clear()

    def set_tiling(self, target_elems, blocks_per_cu, variant=0):
        """Work-item size, persistent blocks per CU and kernel variant of the streamed stage kernels.
        Rebuilding the work queue frees the old one, so captured graphs are dropped first."""
        self._drop_graphs()
        _lib.check(self.L.pyqed_heom_plan_set_tiling(self.plan, int(target_elems), int(blocks_per_cu)))
        _lib.check(self.L.pyqed_heom_plan_set_variant(self.plan, int(variant)))
        self.variant = int(variant)

    def set_state(self, states):
        """Load full hierarchy states (list of [nmax, d, d] arrays in the caller's basis) into y."""
        torch = self.torch
        for k, p in enumerate(self.prep):
            a = np.asarray(states[k], np.complex128)
            if not np.all(np.isfinite(a)):
                raise ValueError(f"set_state: instance {k} has a non-finite entry (NaN or Inf)")
            U = p["U"]
            if U is not None:
                a = U.conj().T @ a @ U
            o = int(self.state_offset[k])
            v = self._to_state(k, a)
            self.y[o:o + v.size] = torch.as_tensor(v, device=self.device)

    def tiling(self):
        t, b = C.c_int64(), C.c_int32()
        _lib.check(self.L.pyqed_heom_plan_get_tiling(self.plan, C.byref(t), C.byref(b)))
        return {"target_elems": t.value, "blocks_per_cu": b.value, "variant": self.variant,
                "classes": [list(x) for x in self.class_tiling()]}

    def set_class_tiling(self, cls, target_elems, blocks_per_cu):
        """Work-item size and blocks per CU of one kernel class (0 generic, 1 projector VALU, 2 MFMA 16 x 16,
        3 MFMA 2 x 2); set_tiling sets every class."""
        self._drop_graphs()
        _lib.check(self.L.pyqed_heom_plan_set_class_tiling(self.plan, int(cls), int(target_elems), int(blocks_per_cu)))

    def class_tiling(self):
        """[(target_elems, blocks_per_cu)] per kernel class."""
        t = np.zeros(NCLASSES, np.int64)
        b = np.zeros(NCLASSES, np.int32)
        _lib.check(self.L.pyqed_heom_plan_get_class_tiling(self.plan, t.ctypes.data, b.ctypes.data))
        return [(int(x), int(y)) for x, y in zip(t, b)]

    def classes(self):
        """Streamed work items and full-matrix elements per kernel class (0 generic, 1 site projector with
        VALU commutator, 2 with MFMA on one 16 x 16 tile, 3 with MFMA on 2 x 2 tiles)."""
        it = np.zeros(NCLASSES, np.int64)
        el = np.zeros(NCLASSES, np.int64)
        _lib.check(self.L.pyqed_heom_plan_classes(self.plan, it.ctypes.data, el.ctypes.data))
        return dict(items=it.tolist(), elems=el.tolist())

    def class_single(self):
        """Per kernel class: for classes 1 and 3 of a packed plan whose streamed work is one instance, its
        index (their FP32 forms then run the single-instance kernel), else -1."""
        out = np.zeros(NCLASSES, np.int32)
        _lib.check(self.L.pyqed_heom_plan_class_single(self.plan, out.ctypes.data))
        return out.tolist()

    def class_state(self):
        """RK4-state elements of the streamed work per kernel class (packed counts when packed)."""
        st = np.zeros(NCLASSES, np.int64)
        _lib.check(self.L.pyqed_heom_plan_class_state(self.plan, st.ctypes.data))
        return st.tolist()

    def set_timing(self, enable):
        """Record an HIP event pair around every kernel-class launch of later advance() calls."""
        _lib.check(self.L.pyqed_heom_plan_set_timing(self.plan, 1 if enable else 0))

    def read_timing(self):
        """(ms[16], launches[16]) per RK4 stage s and kernel class c at index 4s + c, summed over the
        launches recorded since set_timing / the last read."""
        ms = np.zeros(4 * NCLASSES, np.float64)
        n = np.zeros(4 * NCLASSES, np.int64)
        _lib.check(self.L.pyqed_heom_plan_read_timing(self.plan, ms.ctypes.data, n.ctypes.data))
        return ms.tolist(), n.tolist()

    def stats(self):
        s = np.zeros(4, np.int64)
        _lib.check(self.L.pyqed_heom_plan_stats(self.plan, s.ctypes.data))
        return dict(stream_items=int(s[0]), resident_instances=int(s[1]), stream_elems=int(s[2]),
                    resident_elems=int(s[3]))

    # ------------------------------------------------------------------ stepping
    def advance(self, nsteps, dt, precision="fp64", field=None, traj=None, save_every=1, stream=None):
        """Launch nsteps RK4 steps (asynchronous, stream-ordered).  `field` is a device float64 tensor
        [>= step+nsteps, 6]; `traj` a device complex128 tensor [nsteps//save_every, traj_stride]."""
        torch = self.torch
        prec = PRECISIONS[precision]
        if nsteps < 0 or save_every < 1:
            raise ValueError("advance: nsteps must be >= 0 and save_every >= 1")
        if field is not None:
            self._check_tensor("field", field, torch.float64, (self.step + nsteps, 6))
        elif self.driven:
            raise ValueError("advance: an instance has a system or coupling dipole; pass its field table")
        if traj is not None:
            self._check_tensor("traj", traj, torch.complex128, (nsteps // save_every, self.traj_stride))
        if self.autotune_cache and precision not in self._tuned:
            from . import autotune
            self._tuned.add(precision)
            hit = autotune.lookup(self, precision)
            if hit is not None:
                self.set_tiling(*hit)
                for c, tb in enumerate(autotune.lookup_classes(self, precision) or []):
                    if tb:
                        self.set_class_tiling(c, *tb)
                self._tuned_hit = True
        ws = self.workspace(prec)
        _lib.check(self.L.pyqed_heom_run(
            self.plan, prec, self.y.data_ptr(), ws.data_ptr(), ws.numel(),
            None if field is None else field.data_ptr(), self.step, nsteps, float(dt),
            None if traj is None else traj.data_ptr(), save_every, _lib.stream_ptr(torch, stream)))
        self.step += nsteps

    def _check_tensor(self, name, t, dtype, min_shape):
        """t must be a contiguous tensor of `dtype` on this engine's device with at least min_shape[0] rows
        of exactly min_shape[1] columns: the kernels index it by step and column with no bounds of their own."""
        ok = (isinstance(t, self.torch.Tensor) and t.dtype == dtype and t.device == self.device and t.is_contiguous()
              and t.dim() == 2 and t.shape[0] >= min_shape[0] and t.shape[1] == min_shape[1])
        if not ok:
            got = (tuple(t.shape), t.dtype, t.device, t.is_contiguous()) if isinstance(t, self.torch.Tensor) else type(t)
            raise ValueError(f"advance: {name} must be a contiguous {dtype} tensor on {self.device} of shape "
                             f"[>= {min_shape[0]}, {min_shape[1]}]; got {got}")

    @property
    def driven(self):
        """True when some instance has a system or coupling dipole (its kernels read the field table)."""
        return any(p["sdip"] is not None or p["Cmat"] is not None for p in self.prep)

    def advance_graph(self, nsteps, dt, precision="fp64", chunk=25):
        """Advance nsteps steps by replaying a captured HIP graph of `chunk` steps (the stage
        launches of every kernel class, and the LDS-resident launch on its side stream).  Without a
        time-dependent field no kernel reads the step index, so one captured chunk serves every
        replay.  The remainder runs eagerly.  Returns the number of graph replays."""
        torch = self.torch
        if self.driven:
            raise ValueError("advance_graph: a captured chunk replays one step index, so instances with a "
                             "system or coupling dipole (time-dependent fields) must use advance()")
        if getattr(self, "_graphs", None) is None:
            self._graphs = {}
        self.workspace(PRECISIONS[precision])
        if self.autotune_cache and precision not in self._tuned:
            self.advance(0, dt, precision)  # apply the cached tiling before looking up or capturing a graph
        key = (precision, chunk, float(dt))
        g = self._graphs.get(key)
        if g is None:
            step0 = self.step
            y0 = self.y.clone()
            s = torch.cuda.Stream()
            s.wait_stream(torch.cuda.current_stream())
            with torch.cuda.stream(s):
                self.advance(chunk, dt, precision, stream=s)  # warm up on the capture stream
            torch.cuda.current_stream().wait_stream(s)
            self.y.copy_(y0)
            g = torch.cuda.CUDAGraph()
            with torch.cuda.graph(g):
                self.advance(chunk, dt, precision, stream=torch.cuda.current_stream())
            self.y.copy_(y0)  # capture does not execute; restore anyway in case a backend did
            del y0
            self.step = step0
            self._graphs[key] = g
        nrep = nsteps // chunk
        for _ in range(nrep):
            g.replay()
        rest = nsteps - nrep * chunk
        if rest:
            self.advance(rest, dt, precision)
        self.step += nrep * chunk
        return nrep

    def run(self, dt, nt, precision="fp64", field=None, save_every=1):
        """Propagate nt steps from the current state; returns host trajectories of rho per instance,
        each [nt//save_every + 1, d, d] (row 0 = the state before the call), in the caller's basis."""
        torch = self.torch
        nsave = nt // save_every
        traj = torch.empty((max(nsave, 1), self.traj_stride), dtype=torch.complex128, device=self.device)
        first = self.reduced_states()
        if nt > 0:
            self.advance(nt, dt, precision, field, traj, save_every)
        torch.cuda.synchronize()
        host = traj[:nsave].cpu().numpy()
        out = []
        for k, p in enumerate(self.prep):
            d = p["d"]
            t = host[:, self.traj_offset[k]:self.traj_offset[k] + d * d].reshape(nsave, d, d)
            t = np.concatenate([first[k][None], self._unrotate(k, t)], axis=0)
            out.append(t)
        return out

    def _unrotate(self, k, rho):
        U = self.prep[k]["U"]
        if U is None:
            return rho
        return U @ rho @ U.conj().T

    def reduced_states(self):
        """Current rho (ADO 0) of every instance, host complex128, caller's basis."""
        out = []
        for k in range(self.n):
            o = int(self.state_offset[k])
            r = self._from_state(k, self.y[o:o + self.state_size(k)].cpu().numpy())[0]
            out.append(self._unrotate(k, r))
        return out

    def hierarchy(self, k=0):
        """Full ADO state of instance k as host array [nmax, d, d] (caller's basis)."""
        o = int(self.state_offset[k])
        a = self._from_state(k, self.y[o:o + int(self.nmax[k]) * self.state_size(k)].cpu().numpy())
        return self._unrotate(k, a)

    def rhs(self, x, field=None, step=0, col=0):
        """FP64 right-hand side of full hierarchy matrices x (device complex128, every ADO of every
        instance as a full d x d matrix, in the layout of y without packing: instance starts rounded
        up to multiples of 16 elements; also for packed engines).  Padding between instances is zero."""
        torch = self.torch
        out = torch.zeros_like(x)
        _lib.check(self.L.pyqed_heom_rhs(self.plan, x.data_ptr(), out.data_ptr(),
                                         None if field is None else field.data_ptr(), step, col,
                                         _lib.stream_ptr(torch)))
        return out

    def tables(self, k=0):
        K = len(self.prep[k]["mode"])
        nm = int(self.nmax[k])
        keys = np.zeros((nm, K), np.int64)
        plus = np.zeros((nm, K), np.int32)
        minus = np.zeros((nm, K), np.int32)
        _lib.check(self.L.pyqed_heom_tables_export(self.plan, k, keys.ctypes.data, plus.ctypes.data,
                                                   minus.ctypes.data))
        return keys, plus, minus

    def close(self):
        if getattr(self, "plan", None):
            self._drop_graphs()
            self.L.pyqed_heom_plan_destroy(self.plan)
            self.plan = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def field_table(nt, dt, fs=None, fc=None, step0=0):
    """Host table [nt, 6] of f_s and f_c at t, t+dt/2, t+dt for t = (step0+i) dt (generate_time's
    evaluation points, deom.py:730-759)."""
    tab = np.zeros((step0 + nt, 6), np.float64)
    for i in range(step0, step0 + nt):
        t = i * dt
        for c, tt in enumerate((t, t + dt / 2, t + dt)):
            for f, col in ((fs, c), (fc, 3 + c)):
                if f is None:
                    continue
                v = complex(f(tt))
                if v.imag != 0.0:
                    raise NotImplementedError("complex-valued pulse functions are not supported")
                if not np.isfinite(v.real):
                    raise ValueError(f"pulse function returned {v.real} at t = {tt}")
                tab[i, col] = v.real
    return tab
