// Single-exponential (high-temperature) HEOM of pyqed's legacy solver oqs.HEOMSolver
// (pyqed/oqs.py:1332-1372, 1808-1953): forward Euler with the reference's in-place sweep.
//
// run (mode 0): ADOs X_n, d x d, n = 0..nado-1.  Per step (oqs.py:1849-1862):
//   X_0 <- X_0 - i [H, X_0] dt - [S, X_1] dt
//   for n = 1..nado-2:
//     X_n <- X_n + (-i [H, X_n] - [S, X_{n+1}] - n gamma X_n
//                   + n (Re D0 [S, X_{n-1}] + i Im D0 {S, X_{n-1}})) dt
//   with X_{n-1} already updated in this step and X_{n+1} not yet; X_{nado-1} is never updated
//   (the truncation, zero when started from zero).  Observables Tr(E_o X_0) after every step.
// propagator (mode 1): blocks U_n, N x N (N = d^2), left products with the superoperators
//   L0 = H x I - I x H^T, Sm (commutator) and Sp (anticommutator) of S.  Per step
//   (oqs.py:1920-1932):
//   U_0 <- U_0 - i L0 U_0 dt - Sm U_1 dt
//   for n = 0..nado-2:  (n = 0 updates U_0 a second time, as the reference's loop does)
//     U_n <- U_n + (-i L0 U_n - Sm U_{n+1} - n gamma U_n + n (a Sm + i b Sp) U_{n-1}) dt
// Both are one update formula, u(n), with the products two-sided (mode 0) or left-only (mode 1).
// pyqed.HEOM.heom keeps a second copy of this solver (pyqed/HEOM/heom.py:161-414) that differs in
// two places, served by two more modes:
// run, classic RK4 (mode 2; HEOM/heom.py:275-346): the right-hand side F_n of the update formula
//   for n = 0..nado-2 (F_{nado-1} = 0), evaluated for all n from one stage input (no sweep), and
//   y <- y + (k1 + 2 k2 + 2 k3 + k4) / 6 dt with k1 = F(y), k2 = F(y + k1 dt/2),
//   k3 = F(y + k2 dt/2), k4 = F(y + k3 dt) (pyqed/phys.py:1051-1064).
// propagator with U_0 updated once (mode 3; HEOM/heom.py:349-413): the sweep of mode 1 with the
//   loop over n starting at 1.
// One workgroup owns one instance and keeps every ADO in LDS; the sweep over n is sequential and
// the elements of one ADO are updated in parallel.
#include <hip/hip_runtime.h>

#include <cstring>
#include <string>

#include "common.h"

using namespace pqa;

namespace {

constexpr int kMaxThreads = 256;
#ifndef PQA_LEGACY_MT
#define PQA_LEGACY_MT 1  // 0: always the runtime-M kernels
#endif

struct LgParams {
  int M, nado, n_obs, n_inst;
  int n_first;        // first n of the sweep after the update of ADO 0 (Euler modes)
  int gsz;            // lanes per instance group: a power of two; gsz < blockDim.x packs several
                      // instances (blockDim.x / gsz) into one workgroup
  const double2* A;   // run: H; propagator: L0              [n_inst][M][M]
  const double2* S;   // run: S; propagator: Sm
  const double2* P;   // propagator: Sp
  const double* par;  // [n_inst][3]: gamma, cr, ci (run: Re D0, Im D0; propagator: a, b)
  double2* ado;       // [n_inst][nado][M][M], advanced in place
  int64_t nt;
  double dt;
  const double2* eops;  // [n_obs][M][M] (run)
  double2* obs_out;     // [n_inst][nt][n_obs]
};

// LDS blocks of M^2 complex doubles per instance group.
__host__ __device__ constexpr int lds_blocks(int nado, bool prop, bool one, bool rk) {
  return rk ? 4 * nado + 2 : nado + (one ? 0 : 1) + 2 + (prop ? 1 : 0);
}

// PROP = false: two-sided products (commutators of d x d ADOs); true: left products of N x N blocks.
// ONE = true: every lane owns at most one element of its instance (M^2 <= gsz), the new value waits
// in a register; otherwise lanes stride over the elements and new values go through LDS.
// RK = true (with PROP = false): classic RK4 over the whole hierarchy instead of the Euler sweep.
// MT > 0: M = MT at compile time, and each lane keeps the row and column of the operators its
// element needs in registers (the small-d instances, where LDS reads and address arithmetic
// dominate a step).  Needs one element per lane: ONE for the sweep, nado M^2 <= gsz for RK.
// Instances are packed gsz lanes apart, so a group never straddles a wavefront when gsz <= 64.
template <bool PROP, bool ONE, bool RK, int MT>
__global__ __launch_bounds__(kMaxThreads, MT > 0 ? 4 : 1) void legacy_heom_kernel(LgParams p) {
  constexpr bool REG = MT > 0;
  static_assert(!REG || ONE || RK, "register operators need one element per lane");
  extern __shared__ __align__(16) char smem[];
  const int tid = threadIdx.x, nth = blockDim.x, M = REG ? MT : p.M, M2 = M * M, nado = p.nado, gsz = p.gsz;
  const int g = tid / gsz, li = tid - g * gsz;
  const int b = blockIdx.x * (nth / gsz) + g;
  const bool live = b < p.n_inst;  // lanes of a missing instance still take every barrier
  // per-group LDS: X [nado][M2], Tn [M2] (!ONE), sA, sS [M2], sP [M2] (PROP);
  // RK: X (the state y), Ks (k1 + 2 k2 + 2 k3), Sa, Sb (stage inputs) [nado][M2] each, sA, sS
  const int per = lds_blocks(nado, PROP, ONE, RK) * M2;
  cpx<double>* X = (cpx<double>*)smem + (size_t)g * per;
  cpx<double>* Tn = X + (size_t)nado * M2;
  cpx<double>* Ks = Tn;
  cpx<double>* Sa = Ks + (size_t)nado * M2;
  cpx<double>* Sb = Sa + (size_t)nado * M2;
  cpx<double>* sA = RK ? Sb + (size_t)nado * M2 : Tn + (ONE ? 0 : M2);
  cpx<double>* sS = sA + M2;
  cpx<double>* sP = sS + M2;
  double* red = (double*)((cpx<double>*)smem + (size_t)(nth / gsz) * per);  // [2 * nth]
  double gamma = 0.0, cr = 0.0, ci = 0.0;
  double2* gado = p.ado + (size_t)(live ? b : 0) * nado * M2;
  if (live) {
    const double2* Ag = p.A + (size_t)b * M2;
    const double2* Sg = p.S + (size_t)b * M2;
    for (int e = li; e < M2; e += gsz) {
      sA[e] = mk(Ag[e].x, Ag[e].y);
      sS[e] = mk(Sg[e].x, Sg[e].y);
      if (PROP) sP[e] = mk(p.P[(size_t)b * M2 + e].x, p.P[(size_t)b * M2 + e].y);
    }
    for (int e = li; e < nado * M2; e += gsz) X[e] = mk(gado[e].x, gado[e].y);
    gamma = p.par[3 * b];
    cr = p.par[3 * b + 1];
    ci = p.par[3 * b + 2];
  }
  const double dt = p.dt;
  __syncthreads();
  // REG: the lane's element (i, j) = (li / M, li % M) needs row i and column j of A and S (row i of P)
  constexpr int NR = REG ? MT : 1;
  cpx<double> rA[NR], cA[NR], rS[NR], cS[NR], rP[NR];
  if constexpr (REG) {
    const int e0 = RK ? li % (MT * MT) : li;  // RK: lane li holds pair (li / M^2, li % M^2)
    const int i = e0 / MT, j = e0 - (e0 / MT) * MT;
#pragma unroll
    for (int k = 0; k < MT; ++k) {
      const bool ok = live && li < (RK ? nado * M2 : M2);
      rA[k] = ok ? sA[i * MT + k] : mk(0.0, 0.0);
      cA[k] = ok ? sA[k * MT + j] : mk(0.0, 0.0);
      rS[k] = ok ? sS[i * MT + k] : mk(0.0, 0.0);
      cS[k] = ok ? sS[k * MT + j] : mk(0.0, 0.0);
      rP[k] = ok && PROP ? sP[i * MT + k] : mk(0.0, 0.0);
    }
  }

  // right-hand side of element e of ADO n (n <= nado - 2) from the hierarchy src
  auto rhs_el = [&](const cpx<double>* src, int n, int e) {
    const cpx<double>* Xn = src + (size_t)n * M2;
    const cpx<double>* Y = src + (size_t)(n + 1) * M2;
    const cpx<double>* Z = src + (size_t)(n > 0 ? n - 1 : 0) * M2;
    const int i = e / M, j = e - (e / M) * M;
    cpx<double> h = mk(0.0, 0.0), s = mk(0.0, 0.0), zc = mk(0.0, 0.0), za = mk(0.0, 0.0);
    if constexpr (REG) {  // e is the lane's element: the operators' entries are its registers
#pragma unroll
      for (int k = 0; k < MT; ++k) {
        cmac(h, rA[k], Xn[k * MT + j]);
        cmac(s, rS[k], Y[k * MT + j]);
        if (!PROP) {
          cmsub(h, Xn[i * MT + k], cA[k]);
          cmsub(s, Y[i * MT + k], cS[k]);
        }
        if (n > 0) {
          if (!PROP) {
            const cpx<double> l = rS[k] * Z[k * MT + j], r = Z[i * MT + k] * cS[k];
            zc = zc + (l - r);
            za = za + (l + r);
          } else {
            cmac(zc, rS[k], Z[k * MT + j]);
            cmac(za, rP[k], Z[k * MT + j]);
          }
        }
      }
    }
    for (int k = 0; !REG && k < M; ++k) {
      cmac(h, sA[i * M + k], Xn[k * M + j]);
      cmac(s, sS[i * M + k], Y[k * M + j]);
      if (!PROP) {
        cmsub(h, Xn[i * M + k], sA[k * M + j]);
        cmsub(s, Y[i * M + k], sS[k * M + j]);
      }
      if (n > 0) {
        if (!PROP) {
          const cpx<double> l = sS[i * M + k] * Z[k * M + j], r = Z[i * M + k] * sS[k * M + j];
          zc = zc + (l - r);
          za = za + (l + r);
        } else {
          cmac(zc, sS[i * M + k], Z[k * M + j]);
          cmac(za, sP[i * M + k], Z[k * M + j]);
        }
      }
    }
    // -i h - s - n gamma X_n + n (cr zc + i ci za)
    cpx<double> rhs = mul_mi(h) - s - Xn[e] * ((double)n * gamma);
    if (n > 0) rhs = rhs + (zc * cr + mk(-ci * za.y, ci * za.x)) * (double)n;
    return rhs;
  };

  // one update of ADO n from the current LDS contents (n-1 already new, n+1 still old)
  auto update = [&](int n) {
    cpx<double>* Xn = X + (size_t)n * M2;
    cpx<double> nv = mk(0.0, 0.0);
    for (int e = li; live && e < M2; e += gsz) {
      nv = Xn[e] + rhs_el(X, n, e) * dt;
      if (!ONE) Tn[e] = nv;
    }
    __syncthreads();
    if (ONE) {
      if (live && li < M2) Xn[li] = nv;
    } else {
      for (int e = li; live && e < M2; e += gsz) Xn[e] = Tn[e];
    }
    __syncthreads();
  };

  // one RK4 stage: k = F(src) for every element of every ADO (k = 0 for ADO nado - 1), then
  // Ks and the next stage input dst (or y, at the last stage) from k.  The lanes of a group
  // stride over (ADO, element) pairs: all ADOs of a stage are independent.
  auto rk_stage = [&](int stage, const cpx<double>* src, cpx<double>* dst) {
    const double dt2 = dt / 2.0;
    for (int q = li; live && q < nado * M2; q += gsz) {
      const int n = q / M2, e = q - n * M2;
      const cpx<double> k = n <= nado - 2 ? rhs_el(src, n, e) : mk(0.0, 0.0);
      if (stage == 0) {
        Ks[q] = k;
        dst[q] = X[q] + k * dt2;
      } else if (stage < 3) {
        Ks[q] = Ks[q] + k * 2.0;
        dst[q] = X[q] + k * (stage == 1 ? dt2 : dt);
      } else {
        const cpx<double> t = Ks[q] + k;
        X[q] = X[q] + mk(t.x / 6.0, t.y / 6.0) * dt;
      }
    }
    __syncthreads();
  };

  // lanes li >= M^2 hold no element of ADO 0: the observable sums reduce over the first red_w
  // lanes only (a power of two >= min(M^2, gsz))
  int red_w = 1;
  while (red_w < M2 && red_w < gsz) red_w <<= 1;

  for (int64_t step = 0; step < p.nt; ++step) {
    if (RK) {
      rk_stage(0, X, Sa);
      rk_stage(1, Sa, Sb);
      rk_stage(2, Sb, Sa);
      rk_stage(3, Sa, nullptr);
    } else {
      update(0);
      for (int n = PROP ? p.n_first : 1; n <= nado - 2; ++n) update(n);
    }
    for (int o = 0; o < p.n_obs; ++o) {
      // Tr(E_o X_0) = sum_ij E_o[j][i] X_0[i][j], summed over the group's lanes
      const double2* E = p.eops + (size_t)o * M2;
      double re = 0.0, im = 0.0;
      for (int e = li; live && e < M2; e += gsz) {
        const int i = e / M, j = e - (e / M) * M;
        const cpx<double> v = mk(E[j * M + i].x, E[j * M + i].y) * X[e];
        re += v.x;
        im += v.y;
      }
      red[tid] = re;
      red[nth + tid] = im;
      __syncthreads();
      for (int w = red_w / 2; w > 0; w >>= 1) {
        if (li < w) {
          red[tid] += red[tid + w];
          red[nth + tid] += red[nth + tid + w];
        }
        __syncthreads();
      }
      if (live && li == 0) p.obs_out[((size_t)b * p.nt + step) * p.n_obs + o] = make_double2(red[tid], red[nth + tid]);
      __syncthreads();
    }
  }
  if (live)
    for (int e = li; e < nado * M2; e += gsz) gado[e] = make_double2(X[e].x, X[e].y);
}

template <bool PROP, bool ONE, bool RK, int MT = 0>
int launch_legacy(const LgParams& p, int blocks, int nth, size_t lds, hipStream_t stream) {
  if (lds > 65536)
    PQA_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(&legacy_heom_kernel<PROP, ONE, RK, MT>),
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  hipLaunchKernelGGL((legacy_heom_kernel<PROP, ONE, RK, MT>), dim3(blocks), dim3(nth), lds, stream, p);
  PQA_HIP(hipGetLastError());
  return PYQED_OK;
}

}  // namespace

extern "C" int pyqed_legacy_heom_run(int32_t mode, int32_t M, int32_t nado, int32_t n_inst, const void* A,
                                     const void* S, const void* Sp, const double* params, void* ado, int64_t nt,
                                     double dt, int32_t n_obs, const void* e_ops, void* obs_out, void* stream) {
  if (mode < 0 || mode > 3)
    return fail(PYQED_EINVAL, "legacy_heom_run: mode must be 0 (run, Euler sweep), 1 (propagator), 2 (run, RK4) "
                              "or 3 (propagator, U_0 updated once)");
  const bool prop = mode == 1 || mode == 3, rk = mode == 2;
  if (M < 1 || M > 4096 || nado < 2 || n_inst < 1 || !A || !S || !params || !ado || (prop && !Sp))
    return fail(PYQED_EINVAL, "legacy_heom_run: need 1 <= M <= 4096, nado >= 2, n_inst >= 1 and every operator");
  if (nt < 0) return fail(PYQED_EINVAL, "legacy_heom_run: negative step count");
  if (n_obs < 0 || (n_obs > 0 && (prop || !e_ops || !obs_out)))
    return fail(PYQED_EINVAL, "legacy_heom_run: observables need a run mode (0 or 2), e_ops and obs_out");
  const int M2 = M * M;
  // span <= 64: one wavefront holds 64 / gsz instances, gsz = span rounded up to a power of two;
  // larger spans: one instance per workgroup of 128 or 256 lanes.  The span is one ADO (the sweep
  // updates one ADO at a time) or, for RK4, the whole hierarchy.
  const int64_t span = rk ? (int64_t)nado * M2 : M2;
  int gsz = 1;
  while (gsz < span && gsz < 256) gsz *= 2;
  const int nth = gsz <= 64 ? 64 : (gsz <= 128 ? 128 : 256);
  if (gsz > nth) gsz = nth;
  const int groups = nth / gsz;
  const bool one = M2 <= gsz;
  const size_t per = (size_t)lds_blocks(nado, prop, one, rk) * M2;
  const size_t lds = sizeof(double2) * per * groups + 2 * sizeof(double) * nth;
  int dev = 0, maxlds = 65536;
  PQA_HIP(hipGetDevice(&dev));
  {
    int v = 0;
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) == hipSuccess && v > 0) maxlds = v;
  }
  if (lds > (size_t)maxlds)
    return fail(PYQED_EINVAL, "legacy_heom_run: the hierarchy (nado + 2 to 4 blocks of M^2 complex doubles per "
                              "instance, 4 nado + 2 for RK4) exceeds one workgroup's LDS (" + std::to_string(lds) + " > " +
                              std::to_string(maxlds) + " bytes)");
  LgParams p;
  std::memset(&p, 0, sizeof(p));
  p.M = M;
  p.nado = nado;
  p.n_obs = n_obs;
  p.n_inst = n_inst;
  p.n_first = mode == 1 ? 0 : 1;
  p.gsz = gsz;
  p.A = (const double2*)A;
  p.S = (const double2*)S;
  p.P = (const double2*)Sp;
  p.par = params;
  p.ado = (double2*)ado;
  p.nt = nt;
  p.dt = dt;
  p.eops = (const double2*)e_ops;
  p.obs_out = (double2*)obs_out;
  const int blocks = (n_inst + groups - 1) / groups;
  hipStream_t st = (hipStream_t)stream;
  if (rk) {
    // one (ADO, element) pair per lane; larger M spill at the 4-wave register budget
    if (PQA_LEGACY_MT && span <= gsz && M == 2) return launch_legacy<false, false, true, 2>(p, blocks, nth, lds, st);
    return launch_legacy<false, false, true>(p, blocks, nth, lds, st);
  }
  if (!prop && PQA_LEGACY_MT && one && M == 2) return launch_legacy<false, true, false, 2>(p, blocks, nth, lds, st);
  if (!prop)
    return one ? launch_legacy<false, true, false>(p, blocks, nth, lds, st)
               : launch_legacy<false, false, false>(p, blocks, nth, lds, st);
  if (PQA_LEGACY_MT && one && M == 4) return launch_legacy<true, true, false, 4>(p, blocks, nth, lds, st);
  return one ? launch_legacy<true, true, false>(p, blocks, nth, lds, st)
             : launch_legacy<true, false, false>(p, blocks, nth, lds, st);
}
