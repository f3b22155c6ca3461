// Liouville-space RK4 propagation (Lindblad, Redfield) and its discrete adjoint.
//
// Forward:  d v/dt = A(step) v,  A(step) = A_0 + sum_{i>=1} f_i(step) A_i,  v = vec(rho) (row-major).
// One workgroup owns one instance; the state and RK4 registers live in LDS and
// the generators are read from global memory (L2-resident: they are shared by
// every instance).  The step follows pyqed/phys.py:1051-1063:
//   k1 = A v; k2 = A(v + k1 dt/2); k3 = A(v + k2 dt/2); k4 = A(v + k3 dt)
//   v += (k1 + 2 k2 + 2 k3 + k4)/6 * dt
// with the field held fixed over the step, as _lindblad_driven does
// (pyqed/oqs.py:1792-1796).
//
// Adjoint (no counterpart in the reference): for a real objective J(v(T)) with
// vbar = dJ/dRe v + i dJ/dIm v, each RK4 step is reversed stage by stage
// (X_s are the stage inputs, k_s = A X_s):
//   kbar_4 = dt/6 vbar, kbar_3 = kbar_2 = dt/3 vbar, kbar_1 = dt/6 vbar, ubar = vbar
//   for s = 4..1:  Xbar = A^H kbar_s;  ubar += Xbar;  kbar_{s-1} += c_s dt Xbar
//                  (c_4 = 1, c_3 = c_2 = 1/2) and  fbar_i += Re <kbar_s, A_i X_s>
// Forward states are checkpointed every `ckpt_every` steps; each segment is
// recomputed from its checkpoint into a segment buffer and swept backwards.
#include <hip/hip_runtime.h>

#include <cstring>
#include <string>

#include "common.h"

using namespace pqa;

namespace {

constexpr int kT = 256;

struct LvParams {
  int N, nterm, n_inst;
  const double2* A;
  double2* rho;
  const double* field;
  int64_t nsteps;
  double dt;
  int n_obs;
  const double2* obs;
  double2* obs_out;
  double2* rho_out;
  int64_t save_every;
  // adjoint
  int ckpt_every;
  const double2* rhoT_bar;
  double* grad_field;
  double2* grad_A;
  double2* rho0_bar;
  double2* ckpt;  // [n_inst][nck][N] then segment buffers [n_inst][ckpt_every][N]
  int64_t nck;
};

__device__ __forceinline__ double coeff(const LvParams& p, int64_t step, int i) {
  return i == 0 ? 1.0 : p.field[((int64_t)blockIdx.x * p.nsteps + step) * (p.nterm - 1) + (i - 1)];
}

// out = A(step) x   (row r handled by thread r, r += blockDim)
__device__ void matvec(const LvParams& p, int64_t step, const cpx<double>* x, cpx<double>* out) {
  const int N = p.N;
  for (int r = threadIdx.x; r < N; r += kT) {
    cpx<double> s = mk(0.0, 0.0);
    for (int i = 0; i < p.nterm; ++i) {
      const double c = coeff(p, step, i);
      if (c == 0.0) continue;
      const double2* Ai = p.A + (int64_t)i * N * N + (int64_t)r * N;
      cpx<double> si = mk(0.0, 0.0);
      for (int q = 0; q < N; ++q) cmac(si, mk(Ai[q].x, Ai[q].y), x[q]);
      s = s + si * c;
    }
    out[r] = s;
  }
}

// out = A_i x for one generator
__device__ void matvec_term(const LvParams& p, int i, const cpx<double>* x, cpx<double>* out) {
  const int N = p.N;
  for (int r = threadIdx.x; r < N; r += kT) {
    const double2* Ai = p.A + (int64_t)i * N * N + (int64_t)r * N;
    cpx<double> s = mk(0.0, 0.0);
    for (int q = 0; q < N; ++q) cmac(s, mk(Ai[q].x, Ai[q].y), x[q]);
    out[r] = s;
  }
}

// out = A(step)^H v
__device__ void matvec_h(const LvParams& p, int64_t step, const cpx<double>* v, cpx<double>* out) {
  const int N = p.N;
  for (int c = threadIdx.x; c < N; c += kT) {
    cpx<double> s = mk(0.0, 0.0);
    for (int i = 0; i < p.nterm; ++i) {
      const double f = coeff(p, step, i);
      if (f == 0.0) continue;
      const double2* Ai = p.A + (int64_t)i * N * N + c;
      cpx<double> si = mk(0.0, 0.0);
      for (int r = 0; r < N; ++r) cmac(si, mk(Ai[(int64_t)r * N].x, -Ai[(int64_t)r * N].y), v[r]);
      s = s + si * f;
    }
    out[c] = s;
  }
}

__device__ double block_sum(double v, double* red) {
  red[threadIdx.x] = v;
  __syncthreads();
  for (int s = kT / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  const double r = red[0];
  __syncthreads();
  return r;
}

__device__ void observe(const LvParams& p, const cpx<double>* v, int64_t row, double* red) {
  const int N = p.N;
  for (int o = 0; o < p.n_obs; ++o) {
    cpx<double> s = mk(0.0, 0.0);
    for (int e = threadIdx.x; e < N; e += kT) {
      const double2 w = p.obs[(int64_t)o * N + e];
      cmac(s, mk(w.x, w.y), v[e]);
    }
    const double re = block_sum(s.x, red);
    const double im = block_sum(s.y, red);
    if (threadIdx.x == 0)
      p.obs_out[((int64_t)blockIdx.x * (p.nsteps + 1) + row) * p.n_obs + o] = make_double2(re, im);
  }
}

// One RK4 step in place on v (LDS).  Scratch: X, k, acc (each N).
__device__ void rk4_step(const LvParams& p, int64_t step, cpx<double>* v, cpx<double>* X, cpx<double>* k,
                         cpx<double>* acc) {
  const int N = p.N;
  const double dt = p.dt, dt2 = p.dt / 2.0;
  matvec(p, step, v, k);  // k1
  __syncthreads();
  for (int e = threadIdx.x; e < N; e += kT) {
    acc[e] = k[e];
    X[e] = v[e] + k[e] * dt2;
  }
  __syncthreads();
  matvec(p, step, X, k);  // k2
  __syncthreads();
  for (int e = threadIdx.x; e < N; e += kT) {
    acc[e] = acc[e] + k[e] * 2.0;
    X[e] = v[e] + k[e] * dt2;
  }
  __syncthreads();
  matvec(p, step, X, k);  // k3
  __syncthreads();
  for (int e = threadIdx.x; e < N; e += kT) {
    acc[e] = acc[e] + k[e] * 2.0;
    X[e] = v[e] + k[e] * dt;
  }
  __syncthreads();
  matvec(p, step, X, k);  // k4
  __syncthreads();
  for (int e = threadIdx.x; e < N; e += kT) {
    const cpx<double> a = acc[e] + k[e];
    v[e] = v[e] + mk(a.x / 6.0 * dt, a.y / 6.0 * dt);
  }
  __syncthreads();
}

__global__ __launch_bounds__(kT) void liouville_forward_kernel(LvParams p) {
  extern __shared__ __align__(16) char smem[];
  const int N = p.N;
  cpx<double>* v = (cpx<double>*)smem;
  cpx<double>* X = v + N;
  cpx<double>* k = X + N;
  cpx<double>* acc = k + N;
  double* red = (double*)(acc + N);
  double2* rho = p.rho + (int64_t)blockIdx.x * N;
  for (int e = threadIdx.x; e < N; e += kT) v[e] = mk(rho[e].x, rho[e].y);
  __syncthreads();
  if (p.n_obs) observe(p, v, 0, red);
  for (int64_t s = 0; s < p.nsteps; ++s) {
    rk4_step(p, s, v, X, k, acc);
    if (p.n_obs) observe(p, v, s + 1, red);
    if (p.rho_out && ((s + 1) % p.save_every) == 0) {
      const int64_t slot = (s + 1) / p.save_every - 1;
      const int64_t nslot = p.nsteps / p.save_every;
      for (int e = threadIdx.x; e < N; e += kT)
        p.rho_out[((int64_t)blockIdx.x * nslot + slot) * N + e] = make_double2(v[e].x, v[e].y);
    }
  }
  for (int e = threadIdx.x; e < N; e += kT) rho[e] = make_double2(v[e].x, v[e].y);
}

// Re <a, b> = Re sum conj(a) b over the block
__device__ double redot(const cpx<double>* a, const cpx<double>* b, int N, double* red) {
  double s = 0.0;
  for (int e = threadIdx.x; e < N; e += kT) s += a[e].x * b[e].x + a[e].y * b[e].y;
  return block_sum(s, red);
}

__global__ __launch_bounds__(kT) void liouville_adjoint_kernel(LvParams p) {
  extern __shared__ __align__(16) char smem[];
  const int N = p.N;
  cpx<double>* v = (cpx<double>*)smem;  // current forward state
  cpx<double>* X1 = v + N;              // stage inputs (X1 = y)
  cpx<double>* X2 = X1 + N;
  cpx<double>* X3 = X2 + N;
  cpx<double>* X4 = X3 + N;
  cpx<double>* k = X4 + N;
  cpx<double>* kb = k + N;   // kbar of the current stage
  cpx<double>* lam = kb + N; // adjoint of the step's output
  cpx<double>* ub = lam + N; // accumulating adjoint of the step's input
  cpx<double>* Xb = ub + N;
  cpx<double>* acc = Xb + N;
  double* red = (double*)(acc + N);
  const int64_t b = blockIdx.x;
  const double dt = p.dt;
  const int ce = p.ckpt_every;
  double2* ck = p.ckpt + b * p.nck * N;
  double2* seg = p.ckpt + (int64_t)p.n_inst * p.nck * N + b * (int64_t)ce * N;
  if (p.grad_field && p.nterm > 1)
    for (int64_t e = threadIdx.x; e < p.nsteps * (p.nterm - 1); e += kT) p.grad_field[b * p.nsteps * (p.nterm - 1) + e] = 0.0;
  // ---- forward sweep with checkpoints
  for (int e = threadIdx.x; e < N; e += kT) {
    const double2 r = p.rho[b * N + e];
    v[e] = mk(r.x, r.y);
    ck[e] = r;
  }
  __syncthreads();
  for (int64_t s = 0; s < p.nsteps; ++s) {
    rk4_step(p, s, v, X1, k, acc);
    if ((s + 1) % ce == 0 && (s + 1) / ce < p.nck)
      for (int e = threadIdx.x; e < N; e += kT) ck[((s + 1) / ce) * N + e] = make_double2(v[e].x, v[e].y);
  }
  __syncthreads();
  // ---- backward sweep
  for (int e = threadIdx.x; e < N; e += kT) {
    const double2 r = p.rhoT_bar[b * N + e];
    lam[e] = mk(r.x, r.y);
  }
  if (p.grad_A)
    for (int64_t e = threadIdx.x; e < (int64_t)p.nterm * N * N; e += kT) p.grad_A[b * p.nterm * N * N + e] = make_double2(0, 0);
  __syncthreads();
  const int64_t nseg = (p.nsteps + ce - 1) / ce;
  for (int64_t c = nseg - 1; c >= 0; --c) {
    const int64_t s0 = c * ce, s1 = (s0 + ce < p.nsteps) ? s0 + ce : p.nsteps;
    // recompute the segment's input states y_s0 .. y_{s1-1}
    for (int e = threadIdx.x; e < N; e += kT) {
      const double2 r = ck[c * N + e];
      v[e] = mk(r.x, r.y);
    }
    __syncthreads();
    for (int64_t s = s0; s < s1; ++s) {
      for (int e = threadIdx.x; e < N; e += kT) seg[(s - s0) * N + e] = make_double2(v[e].x, v[e].y);
      __syncthreads();
      if (s + 1 < s1) rk4_step(p, s, v, X1, k, acc);
    }
    __syncthreads();
    for (int64_t s = s1 - 1; s >= s0; --s) {
      // stage inputs of step s
      for (int e = threadIdx.x; e < N; e += kT) {
        const double2 r = seg[(s - s0) * N + e];
        X1[e] = mk(r.x, r.y);
      }
      __syncthreads();
      matvec(p, s, X1, k);
      __syncthreads();
      for (int e = threadIdx.x; e < N; e += kT) X2[e] = X1[e] + k[e] * (dt / 2.0);
      __syncthreads();
      matvec(p, s, X2, k);
      __syncthreads();
      for (int e = threadIdx.x; e < N; e += kT) X3[e] = X1[e] + k[e] * (dt / 2.0);
      __syncthreads();
      matvec(p, s, X3, k);
      __syncthreads();
      for (int e = threadIdx.x; e < N; e += kT) {
        X4[e] = X1[e] + k[e] * dt;
        ub[e] = lam[e];
        kb[e] = lam[e] * (dt / 6.0);
      }
      __syncthreads();
      cpx<double>* Xs[4] = {X1, X2, X3, X4};
      const double carry[4] = {0.0, dt / 2.0, dt / 2.0, dt};  // X_s = y + carry_s k_{s-1}
      const double wk[4] = {dt / 6.0, dt / 3.0, dt / 3.0, dt / 6.0};
      for (int st = 3; st >= 0; --st) {
        cpx<double>* Xst = Xs[st];
        // field gradients: Re <kbar_s, A_i X_s>
        for (int i = 1; i < p.nterm; ++i) {
          matvec_term(p, i, Xst, k);
          __syncthreads();
          const double g = redot(kb, k, N, red);
          if (threadIdx.x == 0) p.grad_field[(b * p.nsteps + s) * (p.nterm - 1) + (i - 1)] += g;
        }
        if (p.grad_A) {
          for (int64_t e = threadIdx.x; e < (int64_t)N * N; e += kT) {
            const int r = (int)(e / N), q = (int)(e % N);
            const cpx<double> o = kb[r] * conjg(Xst[q]);
            for (int i = 0; i < p.nterm; ++i) {
              const double f = coeff(p, s, i);
              double2& gA = p.grad_A[(b * p.nterm + i) * (int64_t)N * N + e];
              gA.x += o.x * f;
              gA.y += o.y * f;
            }
          }
        }
        matvec_h(p, s, kb, Xb);
        __syncthreads();
        for (int e = threadIdx.x; e < N; e += kT) {
          ub[e] = ub[e] + Xb[e];
          if (st > 0) kb[e] = lam[e] * wk[st - 1] + Xb[e] * carry[st];
        }
        __syncthreads();
      }
      for (int e = threadIdx.x; e < N; e += kT) lam[e] = ub[e];
      __syncthreads();
    }
  }
  if (p.rho0_bar)
    for (int e = threadIdx.x; e < N; e += kT) p.rho0_bar[b * N + e] = make_double2(lam[e].x, lam[e].y);
}

}  // namespace

extern "C" int pyqed_liouville_run(int32_t N, int32_t nterm, const void* A, int32_t n_inst, void* rho,
                                   const double* field, int64_t nsteps, double dt, int32_t n_obs, const void* obs,
                                   void* obs_out, void* rho_out, int64_t save_every, void* stream) {
  if (N < 1 || N > 1024 || nterm < 1 || n_inst < 1 || !A || !rho)
    return fail(PYQED_EINVAL, "liouville_run: need 1 <= N <= 1024, nterm >= 1, n_inst >= 1 and A, rho");
  if (nterm > 1 && !field) return fail(PYQED_EINVAL, "liouville_run: nterm > 1 needs a field table");
  if (n_obs < 0 || (n_obs > 0 && (!obs || !obs_out))) return fail(PYQED_EINVAL, "liouville_run: bad observables");
  if (nsteps < 0) return fail(PYQED_EINVAL, "liouville_run: negative step count");
  LvParams p;
  std::memset(&p, 0, sizeof(p));
  p.N = N;
  p.nterm = nterm;
  p.n_inst = n_inst;
  p.A = (const double2*)A;
  p.rho = (double2*)rho;
  p.field = field;
  p.nsteps = nsteps;
  p.dt = dt;
  p.n_obs = n_obs;
  p.obs = (const double2*)obs;
  p.obs_out = (double2*)obs_out;
  p.rho_out = (double2*)rho_out;
  p.save_every = save_every > 0 ? save_every : 1;
  const size_t lds = 4 * sizeof(double2) * (size_t)N + sizeof(double) * kT;
  if (lds > 65536)
    PQA_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(&liouville_forward_kernel),
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  hipLaunchKernelGGL(liouville_forward_kernel, dim3(n_inst), dim3(kT), lds, (hipStream_t)stream, p);
  PQA_HIP(hipGetLastError());
  return PYQED_OK;
}

extern "C" int pyqed_liouville_adjoint(int32_t N, int32_t nterm, const void* A, int32_t n_inst, const void* rho0,
                                       const double* field, int64_t nsteps, double dt, int32_t ckpt_every,
                                       const void* rhoT_bar, double* grad_field, void* grad_A, void* rho0_bar,
                                       void* ckpt, int64_t ckpt_bytes, void* stream) {
  if (N < 1 || N > 512 || nterm < 1 || n_inst < 1 || !A || !rho0 || !rhoT_bar || !ckpt)
    return fail(PYQED_EINVAL, "liouville_adjoint: need 1 <= N <= 512 and A, rho0, rhoT_bar, ckpt");
  if (nterm > 1 && (!field || !grad_field))
    return fail(PYQED_EINVAL, "liouville_adjoint: nterm > 1 needs field and grad_field");
  if (ckpt_every < 1 || nsteps < 1) return fail(PYQED_EINVAL, "liouville_adjoint: ckpt_every >= 1, nsteps >= 1");
  const int64_t nck = (nsteps + ckpt_every - 1) / ckpt_every + 1;
  const int64_t need = ((int64_t)n_inst * nck * N + (int64_t)n_inst * ckpt_every * N) * 16;
  if (ckpt_bytes < need)
    return fail(PYQED_EINVAL, "liouville_adjoint: checkpoint workspace too small (" + std::to_string(ckpt_bytes) +
                                  " < " + std::to_string(need) + ")");
  LvParams p;
  std::memset(&p, 0, sizeof(p));
  p.N = N;
  p.nterm = nterm;
  p.n_inst = n_inst;
  p.A = (const double2*)A;
  p.rho = (double2*)const_cast<void*>(rho0);
  p.field = field;
  p.nsteps = nsteps;
  p.dt = dt;
  p.ckpt_every = ckpt_every;
  p.rhoT_bar = (const double2*)rhoT_bar;
  p.grad_field = grad_field;
  p.grad_A = (double2*)grad_A;
  p.rho0_bar = (double2*)rho0_bar;
  p.ckpt = (double2*)ckpt;
  p.nck = nck;
  const size_t lds = 11 * sizeof(double2) * (size_t)N + sizeof(double) * kT;
  if (lds > 65536)
    PQA_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(&liouville_adjoint_kernel),
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  hipLaunchKernelGGL(liouville_adjoint_kernel, dim3(n_inst), dim3(kT), lds, (hipStream_t)stream, p);
  PQA_HIP(hipGetLastError());
  return PYQED_OK;
}
