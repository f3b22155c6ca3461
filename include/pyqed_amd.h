/*
 * pyqed_amd.h -- C ABI of the MI355X (gfx950) open-system propagator.
 *
 * The library (libpyqed_amd.so) advances HEOM hierarchies and Liouville-space
 * master equations (Lindblad, Redfield) with hand-written HIP kernels.  Python
 * callers (pyqed_amd.heom.deom.DEOMSolver, pyqed_amd.oqs.LindbladSolver /
 * RedfieldSolver) bind it with ctypes; any FFI that can pass pointers works.
 *
 * Conventions
 *  - Every entry point returns an int status: PYQED_OK (0) or an error code;
 *    pyqed_last_error() returns a thread-local message for the last failure.
 *  - Complex numbers are interleaved (re, im) pairs of double; matrices are
 *    row-major, element (i, j) of a d x d matrix at index i*d + j.  This is the
 *    index order of the reference's dense generator (pyqed/heom/deom.py:769-771).
 *  - Pointers documented as "device" point to GPU memory owned by the caller
 *    (e.g. torch tensors); "host" pointers are ordinary CPU memory.  The library
 *    owns only opaque plans and the device tables inside them.
 *  - `stream` is a hipStream_t passed as void* (NULL = default stream).  Run
 *    calls are stream-ordered, do not synchronise the host, and are safe to
 *    call concurrently on distinct plans.
 */
#ifndef PYQED_AMD_H
#define PYQED_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PYQED_OK 0
#define PYQED_EINVAL 1       /* bad argument (shape, size, null pointer, unsupported value) */
#define PYQED_ENOMEM 2       /* device or host allocation failed */
#define PYQED_EHIP 3         /* a HIP runtime call failed */
#define PYQED_EUNSUPPORTED 4 /* valid input that this build does not handle */

#define PYQED_PREC_FP64 0  /* plain FP64: right-hand side, stage inputs and state all FP64 */
#define PYQED_PREC_MIXED 1 /* FP32 right-hand side, stage inputs and RK4 increments, FP64 state */

/* Plan flags */
#define PYQED_PLAN_NO_RESIDENT 1 /* never run instances LDS-resident (always stream through HBM) */
#define PYQED_PLAN_DENSE_Q 2     /* coupling operators given as dense d x d matrices (qmat) */
#define PYQED_PLAN_GENERIC 4     /* do not use the site-projector kernels (for cross-checking) */
#define PYQED_PLAN_HERMITIAN 8   /* the caller's initial states are Hermitian; with Hermitian H, real
                                    couplings, real expn and etar = conj(etal) every ADO then stays
                                    Hermitian, and the kernels read column terms as conjugated rows.
                                    Rejected (PYQED_EINVAL) when the coefficients do not allow it. */
#define PYQED_PLAN_PACKED 16     /* with PYQED_PLAN_HERMITIAN: the RK4 state y (and the RK4 registers
                                    in the workspace) hold only the upper triangle of every ADO, row by row,
                                    d(d+1)/2 complex elements per ADO; plan_query's n_elems and
                                    plan_layout's state offsets refer to that layout.  The stage inputs
                                    stay full matrices.  No counterpart in the reference (which stores
                                    full matrices); it halves the FP64 state traffic. */
#define PYQED_PLAN_ADJOINT 32    /* the plan's operator is the adjoint L^H of the HEOM operator (Frobenius
                                    inner product): same tables, H -> -H^H, Q_m -> Q_m^H, and the
                                    coefficients of the minus and plus neighbour terms swapped and
                                    conjugated.  pyqed_heom_rhs then applies L^H; one RK4 step of it with
                                    the stage times reversed is the discrete adjoint of a forward step.
                                    Generic kernels only; excludes PYQED_PLAN_HERMITIAN.  No counterpart in
                                    the reference (which has no gradients). */

/* Thread-local description of the last error (never NULL). */
const char* pyqed_last_error(void);
/* Library version as 10000*major + 100*minor + patch. */
int pyqed_version(void);
/* Number of HIP devices visible; 0 when none (status PYQED_OK either way). */
int pyqed_device_count(int* count);

/* ------------------------------------------------------------------ HEOM
 * One HEOM instance: system Hamiltonian, bath exponents and couplings.
 * Mirrors the state DEOMSolver builds in check_()/init_() and the arguments
 * of rk4()/generate_dot_element() (pyqed/heom/deom.py:641-673, 953-1070):
 *   dRho_n/dt = -(sum_k n_k expn_k) Rho_n - i[H(t), Rho_n]
 *     - i sum_{k: n_k>0} sqrt(n_k)/sqrt(etaa_k) (etal_k Q Rho_{n-e_k} - etar_k Rho_{n-e_k} Q)
 *     - i sum_{k: |n|<L} sqrt(n_k+1) sqrt(etaa_k) [Q, Rho_{n+e_k}],   Q = Q_{mode[k]}(t)
 * with H(t) = H + sdip f_s(t) and Q_m(t) = Q_m + cdip_m f_c(t) (deom.py:676-687).
 * All pointers are host pointers; the plan copies what it needs. */
typedef struct {
  int32_t d;              /* system dimension (nsys), 1..32 */
  int32_t K;              /* number of bath exponentials (nind), 0..255 */
  int32_t L;              /* hierarchy depth (lmax), 0..127 */
  int32_t ncoup;          /* number of coupling operators Q_m */
  const int32_t* mode;    /* [K] coupling index of each exponential (Bath.mode) */
  const double* expn;     /* [K] complex gamma_k (Bath.expn) */
  const double* etal;     /* [K] complex (Bath.etal) */
  const double* etar;     /* [K] complex (Bath.etar) */
  const double* etaa;     /* [K] complex (Bath.etaa) */
  const double* H;        /* [d*d] complex system Hamiltonian */
  const double* sdip;     /* [d*d] complex system dipole, or NULL */
  const double* qdiag;    /* [ncoup*d] complex diagonals of Q_m (when not PYQED_PLAN_DENSE_Q) */
  const double* cdiag;    /* [ncoup*d] complex diagonals of the coupling dipoles, or NULL */
  const double* qmat;     /* [ncoup*d*d] complex dense Q_m (with PYQED_PLAN_DENSE_Q), else NULL */
  const double* cmat;     /* [ncoup*d*d] complex dense coupling dipoles (with PYQED_PLAN_DENSE_Q), or NULL */
} pyqed_heom_instance;

typedef struct pyqed_heom_plan pyqed_heom_plan;

/* Build the hierarchy tables (keys, n+e_k and n-e_k ranks: deom.py:555-638
 * and init_() deom.py:1048-1064) for each instance in host C++, pack every
 * instance into one state layout and upload tables and coefficients.
 * Replaces DEOMSolver.check_()/init_() (deom.py:1035-1064). */
int pyqed_heom_plan_create(const pyqed_heom_instance* inst, int32_t n_inst, int32_t flags,
                           pyqed_heom_plan** plan);
int pyqed_heom_plan_destroy(pyqed_heom_plan* plan);

/* Sizes of the plan's state y (the instances' hierarchies one after another): total complex elements
 * (sum nmax*d*d, or sum nmax*d(d+1)/2 upper-triangle elements with PYQED_PLAN_PACKED), total ADOs,
 * and the workspace bytes run() needs for each precision. */
int pyqed_heom_plan_query(const pyqed_heom_plan* plan, int64_t* n_elems, int64_t* n_ados,
                          int64_t* ws_bytes_fp64, int64_t* ws_bytes_mixed);
/* Per instance: element offset of ADO 0 in the state y (in upper-triangle elements with
 * PYQED_PLAN_PACKED), nmax, and the offset of its d*d block inside one trajectory slot (host arrays
 * [n_inst]). */
int pyqed_heom_plan_layout(const pyqed_heom_plan* plan, int64_t* state_offset, int64_t* nmax,
                           int64_t* traj_offset);
/* Copy instance `inst`'s tables to host arrays of shape [nmax][K]: keys
 * (int64, as DEOMSolver.keys), plus (rank of n+e_k, -1 at tier L) and minus
 * (rank of n-e_k, -1 where n_k = 0) -- what hash_plus / hash_minus return
 * (deom.py:588-605).  Any pointer may be NULL. */
int pyqed_heom_tables_export(const pyqed_heom_plan* plan, int32_t inst, int64_t* keys, int32_t* plus,
                             int32_t* minus);

/* Host-only (no GPU needed): number of ADOs C(K+L, L) of a (K, L) hierarchy,
 * and its tables in the layout of pyqed_heom_tables_export.  Built by the same
 * C++ code the plans use.  Replaces DEOMSolver.init_() + gen_keys
 * (deom.py:609-638, 1048-1064) for inspection. */
int pyqed_heom_tables_size(int32_t K, int32_t L, int64_t* nmax);
int pyqed_heom_tables_build(int32_t K, int32_t L, int64_t* keys, int32_t* plus, int32_t* minus);

/* Advance every instance of the plan by `nsteps` classic RK4 steps of size
 * dt, starting at global step index `step0` (time step0*dt).
 * Replaces the time loop of DEOMSolver.run and rk4() (deom.py:725-766,
 * 1072-1114).
 *   y          device, [n_elems] complex double: the hierarchy state; instance
 *              i starts at plan_layout's state_offset[i] (instance starts are
 *              aligned to 128 bytes), then ADO after ADO (full d x d matrices,
 *              or upper triangles with PYQED_PLAN_PACKED), updated in place
 *              (ADO 0 of each instance is the reduced rho).
 *   workspace  device, >= ws_bytes for `precision` (plan_query); scratch.
 *   field      device [(step0+nsteps)*6] double or NULL: for step s,
 *              field[6s+0..2] = f_s(t), f_s(t+dt/2), f_s(t+dt) and
 *              field[6s+3..5] = f_c at the same times (t = s*dt).
 *   traj       device or NULL: after every `save_every` steps, the rho of
 *              every instance is written to slot j (j = 0, 1, ...) at
 *              traj[j*traj_stride + traj_offset[i]], complex double, where
 *              traj_stride = sum of d*d over instances (complex elements).
 *   stream     hipStream_t. */
int pyqed_heom_run(pyqed_heom_plan* plan, int32_t precision, void* y, void* workspace, int64_t ws_bytes,
                   const double* field, int64_t step0, int64_t nsteps, double dt, void* traj,
                   int64_t save_every, void* stream);

/* Evaluate the right-hand side once: out = RHS(x) at time index `step`
 * stage column `col` (0: t, 1: t+dt/2, 2: t+dt) in FP64 (device pointers,
 * full matrices of every ADO, also for PYQED_PLAN_PACKED plans: instance i
 * starts at element off_i, off_0 = 0, off_{i+1} = off_i + nmax_i*d_i*d_i
 * rounded up to a multiple of 16 -- the layout of y without
 * PYQED_PLAN_PACKED).  Replaces
 * rem_cal (deom.py:667-673); used to check the operator against the dense
 * generator. */
int pyqed_heom_rhs(pyqed_heom_plan* plan, const void* x, void* out, const double* field, int64_t step,
                   int32_t col, void* stream);
/* The right-hand side r = RHS(x) of pyqed_heom_rhs (same layout, step and column), combined in the same
 * launch (no counterpart in the reference; for the stage combinations of the RK4 adjoint and other
 * Krylov-type callers): out = base + alpha r when base is given, else out = r (out may be NULL when
 * acc is given); acc = acc + beta r, or beta r when acc_init is nonzero (acc may be NULL).  out and
 * acc must not alias x, nor out the base. */
int pyqed_heom_rhs_combine(pyqed_heom_plan* plan, const void* x, void* out, const void* base, double alpha, void* acc,
                           double beta, int32_t acc_init, const double* field, int64_t step, int32_t col,
                           void* stream);

/* Kernel-level statistics of the last run (for the benchmark): number of
 * streamed work items, resident instances, streamed elements, resident
 * elements. */
int pyqed_heom_plan_stats(const pyqed_heom_plan* plan, int64_t* stats4);
/* Work-queue tiling: target complex elements per work item (0 = automatic)
 * and persistent-grid blocks per CU of the streamed kernels.  Used by the
 * autotuner (pyqed_amd.heom.autotune); no counterpart in the reference.
 * set_tiling synchronises the device before replacing the queue. */
int pyqed_heom_plan_set_tiling(pyqed_heom_plan* plan, int64_t target_elems, int32_t blocks_per_cu);
int pyqed_heom_plan_get_tiling(const pyqed_heom_plan* plan, int64_t* target_elems, int32_t* blocks_per_cu);
/* Work-item size and blocks per CU of one streamed kernel class (0 generic, 1 projector VALU, 2 projector
 * MFMA 16 x 16, 3 projector MFMA 2 x 2), for the autotuner's per-class choice; set_tiling sets every
 * class (no counterpart in the reference).  get_class_tiling fills host arrays [4]. */
int pyqed_heom_plan_set_class_tiling(pyqed_heom_plan* plan, int32_t cls, int64_t target_elems, int32_t blocks_per_cu);
int pyqed_heom_plan_get_class_tiling(const pyqed_heom_plan* plan, int64_t* target_elems, int32_t* blocks_per_cu);
/* Kernel variants, chosen by the autotuner; results are identical.  Bit 0: the MFMA class (12 <= d <= 16)
 * of Hermitian full-matrix FP32 runs uses the two-deep software-pipelined kernel instead of one ADO per
 * wavefront at a time.  Bit 1: the wide class (17 <= d <= 32) of packed FP32 runs uses the
 * block-cooperative kernel (one ADO per block, one 16 x 16 MFMA tile per wavefront) instead of one ADO
 * per wavefront on 2 x 2 tiles.  variant in 0..3. */
int pyqed_heom_plan_set_variant(pyqed_heom_plan* plan, int32_t variant);
/* Streamed work per kernel class (0 generic, 1 site-projector with VALU
 * commutator, 2 site-projector with MFMA commutator on one 16 x 16 tile (d <= 16),
 * 3 the same on 2 x 2 tiles (17 <= d <= 32)): items and elements
 * (host arrays [4], either may be NULL). */
int pyqed_heom_plan_classes(const pyqed_heom_plan* plan, int64_t* items, int64_t* elems);
/* Per kernel class (host array [4]): for classes 1 and 3 of a packed plan whose streamed work is a
 * single instance, that instance's index (their FP32 forms then run the single-instance kernel:
 * instance record from the kernel arguments, neighbour lists at a fixed stride); -1 otherwise, and
 * always for classes 0 and 2.
 * No counterpart in the reference; used to name the kernels a benchmark's launches run. */
int pyqed_heom_plan_class_single(const pyqed_heom_plan* plan, int32_t* inst);
/* RK4-state elements (size in y) of the streamed work per kernel class (host array [4]): equal to
 * the element counts of plan_classes, or the packed counts with PYQED_PLAN_PACKED.  No counterpart
 * in the reference; used for the benchmark's byte model. */
int pyqed_heom_plan_class_state(const pyqed_heom_plan* plan, int64_t* state_elems);
/* Per-launch timing of the streamed stage kernels with HIP events on the run's stream (for the
 * benchmark's roofline; no counterpart in the reference).  set_timing(1) records an event pair
 * around every kernel-class launch of later pyqed_heom_run calls (launches captured into a graph
 * are not timed); read_timing waits for the recorded launches and returns, for RK4 stage s (0..3)
 * and class c (0..3) at index 4*s + c, the summed milliseconds and the number of launches (host
 * arrays [16]), then clears the record.  set_timing(0) stops recording. */
int pyqed_heom_plan_set_timing(pyqed_heom_plan* plan, int32_t enable);
int pyqed_heom_plan_read_timing(pyqed_heom_plan* plan, double* ms, int64_t* launches);

/* Gradient contraction of the streamed HEOM adjoint (no counterpart in the reference): for every
 * instance and field column c,
 *   G[c] = sum_{s : col(s) = c} w[s] sum_a [x[s]_a, u[s]_a^H],   col(0..3) = (0, 1, 1, 2),
 * over the instance's ADOs a.  x[s], u[s] are host arrays of 4 device pointers to full hierarchy
 * matrices in pyqed_heom_rhs's layout (a NULL pair is skipped): x[s] the forward input of RK4 stage s,
 * u[s] the adjoint variable of that stage's right-hand side divided by w[s].  G is device complex
 * double [3][traj_stride], overwritten.  work: device, >= pyqed_heom_adjoint_work_bytes.  Used with
 * a PYQED_PLAN_ADJOINT plan to differentiate DEOMSolver.run (pyqed/heom/deom.py:725-766, 1072-1114). */
int pyqed_heom_adjoint_work_bytes(const pyqed_heom_plan* plan, int64_t* bytes);
int pyqed_heom_adjoint_contract(pyqed_heom_plan* plan, const void* const* x, const void* const* u, const double* w,
                                void* G, void* work, int64_t work_bytes, void* stream);

/* ------------------------------------------------------------------ Liouville-space RK4
 * Batched propagation of vec(rho) (row-major, length N = d*d) under
 *   d vec(rho)/dt = (A_0 + sum_{i=1..nterm-1} f_i(step) A_i) vec(rho),
 * with the field constant over each RK4 step, using the RK4 of
 * pyqed/phys.py:1051-1063.  Lindblad (oqs.py:697-714, 1596-1806) and
 * Redfield (oqs.py:364-459, 519-570) both map onto it: the host builds the
 * generators A_i (N x N complex, row-major) once.
 *   A        device [nterm][N][N] complex double (shared by all instances)
 *   rho      device [n_inst][N] complex double, advanced in place
 *   field    device [n_inst][nsteps][nterm-1] double, or NULL when nterm == 1
 *   obs      device [n_obs][N] complex double: observable Tr(O rho) is
 *            computed as sum_e obs[o][e] * rho[e]; pass O^T flattened so that
 *            the sum is Tr(O rho) (may be NULL when n_obs == 0)
 *   obs_out  device [n_inst][nsteps+1][n_obs] complex double (row 0 = t0)
 *   rho_out  device [n_inst][nsteps/save_every][N] complex double, or NULL */
int pyqed_liouville_run(int32_t N, int32_t nterm, const void* A, int32_t n_inst, void* rho,
                        const double* field, int64_t nsteps, double dt, int32_t n_obs, const void* obs,
                        void* obs_out, void* rho_out, int64_t save_every, void* stream);

/* ------------------------------------------------------------------ Lindblad, matrix form
 * Batched propagation of rho (d x d, row-major) under the Lindblad equation for any d (the
 * Liouville-space entry point above holds the N x N generator and serves N = d*d <= 1024):
 *   d rho/dt = -i (Hs rho - rho Hs^dag) + sum_k c_k rho c_k^dag,  Hs = Heff - sum_i f_i(step) H_i,
 * Heff = H0 - (i/2) sum_k c_k^dag c_k (formed by the caller), i.e. the reference's liouvillian
 * (pyqed/oqs.py:697-714) with the anticommutators folded into Heff, stepped by the RK4 of
 * pyqed/phys.py:1051-1063 with the drive held over each step.  Replaces the time loops of _lindblad
 * and _lindblad_driven (oqs.py:1596-1806) for large d.
 *   Heff, HeffD  device [d][d] complex double: Heff and its adjoint
 *   C, CD        device [nc][d][d]: the left and right operators of the terms C_k rho CD_k (Lindblad:
 *                c_k and c_k^dag; the Bloch-Redfield tensor of oqs.py:519-570 in operator form: pairs
 *                (A_a, L_a^dag) and (L_a, A_a), with Heff = diag(E) - i sum_a A_a L_a); NULL when nc == 0
 *   Hd, HdD      device [ndrive][d][d]: drive operators H_i and their adjoints (NULL when ndrive == 0)
 *   rho          device [n_inst][d][d], advanced in place
 *   field        device [n_inst][nsteps][ndrive] double (NULL when ndrive == 0)
 *   obs          device [n_obs][d][d], O^T flattened (Tr(O rho) = sum_e obs[e] rho[e]); obs_out
 *                device [n_inst][nsteps+1][n_obs] (row 0 = the initial state)
 *   rho_out      device [n_inst][nsteps/save_every][d][d] or NULL
 *   work         device, >= pyqed_lindblad_work_bytes(d, nc, ndrive, n_inst) bytes */
int pyqed_lindblad_work_bytes(int32_t d, int32_t nc, int32_t ndrive, int32_t n_inst, int64_t* bytes);
int pyqed_lindblad_run(int32_t d, int32_t nc, const void* Heff, const void* HeffD, const void* C, const void* CD,
                       int32_t ndrive, const void* Hd, const void* HdD, int32_t n_inst, void* rho,
                       const double* field, int64_t nsteps, double dt, int32_t n_obs, const void* obs,
                       void* obs_out, void* rho_out, int64_t save_every, void* work, int64_t work_bytes,
                       void* stream);

/* Adjoint of pyqed_liouville_run for a real objective J(rho(T)) (no
 * reference counterpart: pyqed has no gradients).  Forward states are
 * checkpointed every `ckpt_every` steps and each segment is recomputed
 * before it is swept backwards (discrete adjoint of the RK4 step).
 *   rhoT_bar  device [n_inst][N] complex double: dJ/dRe rho(T) + i dJ/dIm rho(T)
 *   grad_field device [n_inst][nsteps][nterm-1] double: dJ/df_i(step)
 *   grad_A     device [n_inst][nterm][N][N] complex double or NULL:
 *              sum over steps of the adjoint outer products, so that
 *              dJ/dtheta = Re sum conj(grad_A) * dA/dtheta for any generator
 *              parameter theta (Hamiltonian energies, rates)
 *   rho0_bar   device [n_inst][N] complex double or NULL: gradient w.r.t. rho(0)
 *   ckpt       device workspace, >= n_inst*(nsteps/ckpt_every+1)*N*16 bytes */
int pyqed_liouville_adjoint(int32_t N, int32_t nterm, const void* A, int32_t n_inst, const void* rho0,
                            const double* field, int64_t nsteps, double dt, int32_t ckpt_every,
                            const void* rhoT_bar, double* grad_field, void* grad_A, void* rho0_bar,
                            void* ckpt, int64_t ckpt_bytes, void* stream);

/* Legacy single-exponential HEOM of oqs.HEOMSolver (modes 0, 1: forward Euler
 * with the reference's in-place sweep over the hierarchy, ADO n seeing ADO n-1
 * of the same step and ADO n+1 of the previous one) and of its second copy
 * HEOM.heom.HEOMSolver (modes 2, 3).  One workgroup per instance; the whole
 * hierarchy is held in LDS, so (nado + 4) * M^2 * 16 bytes must fit (modes 0, 1, 3).
 *   mode 0 replaces _heom (pyqed/oqs.py:1808-1874, called by HEOMSolver.run,
 *     oqs.py:1360-1364): M = d, A = H, S = c_ops[0], Sp unused,
 *     params[i] = (gamma, Re D0, Im D0), ado [n_inst][nado][d][d] starts with
 *     rho0 in ADO 0; obs_out [n_inst][nt][n_obs] receives Tr(e_ops[o] rho(t))
 *     after every step (e_ops shared, [n_obs][d][d]).
 *   mode 1 replaces _heom_propagator (oqs.py:1876-1935, HEOMSolver.propagator,
 *     oqs.py:1366-1370): M = d^2, A = L0, S = Sm, Sp = the commutator and
 *     anticommutator superoperators of S, params[i] = (gamma, a, b), ado the
 *     propagator blocks u [n_inst][nado][M][M] (u[0] = identity); n_obs = 0.
 *   mode 2 replaces _heom of the second copy, pyqed/HEOM/heom.py:275-346
 *     (HEOM.heom.HEOMSolver.run, HEOM/heom.py:195-199): arguments as mode 0 with
 *     D0 = reorg (2T - i gamma), classic RK4 over the whole hierarchy instead of
 *     the sweep; (4 nado + 2) * M^2 * 16 bytes of LDS per instance.
 *   mode 3 replaces _heom_propagator of that copy (HEOM/heom.py:349-413,
 *     HEOMSolver.propagator, HEOM/heom.py:201-205): arguments as mode 1, u[0]
 *     updated once per step.
 * A, S, Sp are device [n_inst][M][M] complex double; params is a device
 * [n_inst][3] double array; ado is advanced in place by nt steps of dt. */
int pyqed_legacy_heom_run(int32_t mode, int32_t M, int32_t nado, int32_t n_inst, const void* A, const void* S,
                          const void* Sp, const double* params, void* ado, int64_t nt, double dt, int32_t n_obs,
                          const void* e_ops, void* obs_out, void* stream);

/* Spectral sums of a diagonalised generator L = U1 diag(lambda) U2:
 *   out[t][a][b] = sum_j A[a][j] f(lambda_j, s[t]) B[j][b]
 *   kind 0: f = exp(lambda s)  (time domain); kind 1: f = -1/(lambda + i s) (frequency domain).
 * Replaces the expansion loops of the eigen-decomposition solvers:
 * superoperator.Lindblad_solver.evolve / correlation_2op_1t / 2op_1w / 3op_1t / 3op_1w /
 * 3op_2t (pyqed/superoperator.py:535-739) and RedfieldSolver.propagator's
 * 'SOS' branch, contract('aj, jk, jb -> abk', U1, E, U2) (pyqed/oqs.py:197-207).
 *   A   device [nrow][k] complex double      lam device [k] complex double
 *   B   device [k][ncol] complex double      s   device [ns] double
 *   out device complex double, element (t, a, b) at t*so_t + a*so_a + b*so_b
 *   work device workspace of pyqed_spectral_work_bytes(k, ns) bytes (the weights)
 * nrow <= 65535.  The eigen-decomposition itself runs on the host (LAPACK). */
int64_t pyqed_spectral_work_bytes(int32_t k, int32_t ns);
int pyqed_spectral_sum(int32_t kind, int32_t nrow, int32_t k, int32_t ncol, int32_t ns, const void* A,
                       const void* lam, const void* B, const double* s, void* out, int64_t so_t, int64_t so_a,
                       int64_t so_b, void* work, int64_t work_bytes, void* stream);

/* pyqed_spectral_sum without a workspace argument and without any device allocation: no counterpart in
 * the reference.  It is for callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  The weights are staged in a static device
 * array of PYQED_SPECTRAL_STATIC_MAX complex doubles (16 MiB), one per GPU, held until the process
 * ends; k * ns above that limit returns PYQED_EINVAL.  Launches go to `stream`, which must belong to
 * the current device.  Every call on one GPU shares that GPU's array: calls that overlap in time on
 * one GPU (different streams, or host threads) race and must be serialised by the caller.  Results
 * are bitwise those of pyqed_spectral_sum (same kernels). */
#define PYQED_SPECTRAL_STATIC_MAX (1 << 20)
int pyqed_spectral_sum_static(int32_t kind, int32_t nrow, int32_t k, int32_t ncol, int32_t ns, const void* A,
                              const void* lam, const void* B, const double* s, void* out, int64_t so_t,
                              int64_t so_a, int64_t so_b, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* PYQED_AMD_H */
