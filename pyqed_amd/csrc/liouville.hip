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

constexpr int kT = 256;       // largest workgroup
constexpr int kLdsGen = 65536;  // generators (and gradient accumulators) go to LDS up to this many bytes

struct LvParams {
  int N, nterm, n_inst;
  int gsz;  // lanes per instance: a power of two >= min(N, 256); 64 / gsz instances share a wavefront
  int ch;   // steps per LDS window of the field table (and of the field gradients)
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

// The lanes of one instance: instance b, lane li of gsz; dead groups (b >= n_inst, the tail of the
// last workgroup) skip every element loop and only take the barriers.
struct Grp {
  int64_t b;
  int li, gsz, tid;
  bool live;
};

__device__ __forceinline__ Grp group_of(const LvParams& p) {
  Grp q;
  q.tid = threadIdx.x;
  q.gsz = p.gsz;
  const int g = q.tid / p.gsz;
  q.li = q.tid - g * p.gsz;
  q.b = (int64_t)blockIdx.x * (blockDim.x / p.gsz) + g;
  q.live = q.b < p.n_inst;
  return q;
}

// A window of rows [lo, hi) of the instance's field table (or of its field gradients) in the group's
// LDS, ch rows of nterm - 1 values.  Every group of a workgroup walks the same steps, so moving a
// window is uniform across the workgroup and may take barriers.
struct Win {
  double* w;
  int64_t lo, hi;
};

// make step s resident in the coefficient window, loading ch rows upward (or downward) from s
__device__ void cf_ensure(const LvParams& p, const Grp& q, Win& c, int64_t s, bool down) {
  if (s >= c.lo && s < c.hi) return;
  __syncthreads();
  const int64_t lo = down ? (s - p.ch + 1 > 0 ? s - p.ch + 1 : 0) : s;
  const int64_t hi = lo + p.ch < p.nsteps ? lo + p.ch : p.nsteps;
  const int nt1 = p.nterm - 1;
  const double* src = p.field + (q.b * p.nsteps + lo) * nt1;
  for (int64_t e = q.li; q.live && e < (hi - lo) * nt1; e += q.gsz) c.w[e] = src[e];
  c.lo = lo;
  c.hi = hi;
  __syncthreads();
}

// write the gradient window [lo, hi) to grad_field and open the window of ch rows ending at s
// (the backward sweep visits steps in decreasing order); s < 0 only flushes
__device__ void gf_move(const LvParams& p, const Grp& q, Win& g, int64_t s) {
  __syncthreads();
  const int nt1 = p.nterm - 1;
  if (g.hi > g.lo) {
    double* dst = p.grad_field + (q.b * p.nsteps + g.lo) * nt1;
    for (int64_t e = q.li; q.live && e < (g.hi - g.lo) * nt1; e += q.gsz) dst[e] = g.w[e];
  }
  if (s >= 0) {
    g.lo = s - p.ch + 1 > 0 ? s - p.ch + 1 : 0;
    g.hi = s + 1;
    for (int64_t e = q.li; q.live && e < (g.hi - g.lo) * nt1; e += q.gsz) g.w[e] = 0.0;
  }
  __syncthreads();
}

// coefficient i of the step whose field row is cf (cf = nullptr when nterm == 1)
__device__ __forceinline__ double coeff(const double* cf, int i) { return i == 0 ? 1.0 : cf[i - 1]; }

// generator element A_i[r][c]: from the workgroup's LDS copy (ALDS) or global memory
template <bool ALDS, int NT>
__device__ __forceinline__ cpx<double> gen(const LvParams& p, const cpx<double>* sA, int i, int r, int c) {
  const int N = NT > 0 ? NT : p.N;
  const int64_t e = ((int64_t)i * N + r) * N + c;
  if (ALDS) return sA[e];
  const double2 a = p.A[e];
  return mk(a.x, a.y);
}

// sum_k A_i[r][k] x[k] (HT = false) or sum_k conj(A_i[k][r]) x[k] (HT = true), with four
// independent partial sums: a single wave per SIMD is common here, and one accumulator would make
// every complex multiply-add wait for the previous one
template <bool ALDS, bool HT, int NT>
__device__ __forceinline__ cpx<double> row_dot(const LvParams& p, const cpx<double>* sA, int i, int r,
                                               const cpx<double>* x) {
  const int N = NT > 0 ? NT : p.N;
  cpx<double> a0 = mk(0.0, 0.0), a1 = a0, a2 = a0, a3 = a0;
  int k = 0;
  for (; k + 3 < N; k += 4) {
    cmac(a0, HT ? conjg(gen<ALDS, NT>(p, sA, i, k, r)) : gen<ALDS, NT>(p, sA, i, r, k), x[k]);
    cmac(a1, HT ? conjg(gen<ALDS, NT>(p, sA, i, k + 1, r)) : gen<ALDS, NT>(p, sA, i, r, k + 1), x[k + 1]);
    cmac(a2, HT ? conjg(gen<ALDS, NT>(p, sA, i, k + 2, r)) : gen<ALDS, NT>(p, sA, i, r, k + 2), x[k + 2]);
    cmac(a3, HT ? conjg(gen<ALDS, NT>(p, sA, i, k + 3, r)) : gen<ALDS, NT>(p, sA, i, r, k + 3), x[k + 3]);
  }
  for (; k < N; ++k) cmac(a0, HT ? conjg(gen<ALDS, NT>(p, sA, i, k, r)) : gen<ALDS, NT>(p, sA, i, r, k), x[k]);
  return (a0 + a1) + (a2 + a3);
}

// Compile-time N (with at most two generators, which the host guarantees for those instantiations):
// the lane's row r = li of A_0 and A_1 lives in registers, so a product reads only x from LDS.
template <int NT>
struct Rows {
  cpx<double> a[2][NT > 0 ? NT : 1];
};

template <bool ALDS, int NT>
__device__ void load_rows(const LvParams& p, const Grp& q, const cpx<double>* sA, Rows<NT>& R) {
  if constexpr (NT > 0) {
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int k = 0; k < NT; ++k)
        R.a[i][k] = (q.live && q.li < NT && i < p.nterm) ? gen<ALDS, NT>(p, sA, i, q.li, k) : mk(0.0, 0.0);
  }
}

template <int NT>
__device__ __forceinline__ cpx<double> reg_dot(const cpx<double> (&row)[NT > 0 ? NT : 1], const cpx<double>* x) {
  cpx<double> a0 = mk(0.0, 0.0), a1 = a0, a2 = a0, a3 = a0;
#pragma unroll
  for (int k = 0; k < NT; ++k) {
    if (k % 4 == 0) cmac(a0, row[k], x[k]);
    if (k % 4 == 1) cmac(a1, row[k], x[k]);
    if (k % 4 == 2) cmac(a2, row[k], x[k]);
    if (k % 4 == 3) cmac(a3, row[k], x[k]);
  }
  return (a0 + a1) + (a2 + a3);
}

// out = A(step) x   (row r by lane r of the group, r += gsz)
template <bool ALDS, int NT>
__device__ void matvec(const LvParams& p, const Grp& q, const cpx<double>* sA, const double* cf,
                       const cpx<double>* x, cpx<double>* out, const Rows<NT>& R) {
  const int N = NT > 0 ? NT : p.N;
  if constexpr (NT > 0) {
    if (q.live && q.li < NT) {
      cpx<double> s = reg_dot<NT>(R.a[0], x);
      if (p.nterm > 1) {
        const double c = coeff(cf, 1);
        if (c != 0.0) s = s + reg_dot<NT>(R.a[1], x) * c;
      }
      out[q.li] = s;
    }
    return;
  }
  for (int r = q.li; q.live && r < N; r += q.gsz) {
    cpx<double> s = mk(0.0, 0.0);
    for (int i = 0; i < p.nterm; ++i) {
      const double c = coeff(cf, i);
      if (c == 0.0) continue;
      s = s + row_dot<ALDS, false, NT>(p, sA, i, r, x) * c;
    }
    out[r] = s;
  }
}

// out = A_i x for one generator
template <bool ALDS, int NT>
__device__ void matvec_term(const LvParams& p, const Grp& q, const cpx<double>* sA, int i, const cpx<double>* x,
                            cpx<double>* out, const Rows<NT>& R) {
  const int N = NT > 0 ? NT : p.N;
  if constexpr (NT > 0) {
    if (q.live && q.li < NT) out[q.li] = i == 0 ? reg_dot<NT>(R.a[0], x) : reg_dot<NT>(R.a[1], x);
    return;
  }
  for (int r = q.li; q.live && r < N; r += q.gsz) out[r] = row_dot<ALDS, false, NT>(p, sA, i, r, x);
}

// out = A(step)^H v
template <bool ALDS, int NT>
__device__ void matvec_h(const LvParams& p, const Grp& q, const cpx<double>* sA, const double* cf,
                         const cpx<double>* v, cpx<double>* out) {
  const int N = NT > 0 ? NT : p.N;
  for (int c = q.li; q.live && c < N; c += q.gsz) {
    cpx<double> s = mk(0.0, 0.0);
    for (int i = 0; i < p.nterm; ++i) {
      const double f = coeff(cf, i);
      if (f == 0.0) continue;
      s = s + row_dot<ALDS, true, NT>(p, sA, i, c, v) * f;
    }
    out[c] = s;
  }
}

// sum of v over the lanes of the group (every thread of the workgroup must call it).  A group of at
// most 64 lanes lies inside one wavefront (groups are aligned to gsz), so lane exchanges replace
// the LDS tree and its barriers.
__device__ double group_sum(const Grp& q, double v, double* red) {
  if (q.gsz <= 64) {
    for (int w = q.gsz / 2; w > 0; w >>= 1) v += __shfl_xor(v, w, q.gsz);
    return v;
  }
  red[q.tid] = v;
  __syncthreads();
  for (int w = q.gsz / 2; w > 0; w >>= 1) {
    if (q.li < w) red[q.tid] += red[q.tid + w];
    __syncthreads();
  }
  const double r = red[q.tid - q.li];
  __syncthreads();
  return r;
}

template <bool ALDS, int NT>
__device__ void observe(const LvParams& p, const Grp& q, const cpx<double>* v, int64_t row, double* red) {
  const int N = NT > 0 ? NT : p.N;
  for (int o = 0; o < p.n_obs; ++o) {
    cpx<double> s = mk(0.0, 0.0);
    for (int e = q.li; q.live && e < N; e += q.gsz) {
      const double2 w = p.obs[(int64_t)o * N + e];
      cmac(s, mk(w.x, w.y), v[e]);
    }
    const double re = group_sum(q, s.x, red);
    const double im = group_sum(q, s.y, red);
    if (q.live && q.li == 0) p.obs_out[(q.b * (p.nsteps + 1) + row) * p.n_obs + o] = make_double2(re, im);
  }
}

// One RK4 step in place on v (LDS), field row cf.  Scratch: X, k, acc (each N).
template <bool ALDS, int NT>
__device__ void rk4_step(const LvParams& p, const Grp& q, const cpx<double>* sA, const double* cf, cpx<double>* v,
                         cpx<double>* X, cpx<double>* k, cpx<double>* acc, const Rows<NT>& R) {
  const int N = NT > 0 ? NT : p.N;
  const double dt = p.dt, dt2 = p.dt / 2.0;
  matvec<ALDS, NT>(p, q, sA, cf, v, k, R);  // k1
  __syncthreads();
  for (int e = q.li; q.live && e < N; e += q.gsz) {
    acc[e] = k[e];
    X[e] = v[e] + k[e] * dt2;
  }
  __syncthreads();
  matvec<ALDS, NT>(p, q, sA, cf, X, k, R);  // k2
  __syncthreads();
  for (int e = q.li; q.live && e < N; e += q.gsz) {
    acc[e] = acc[e] + k[e] * 2.0;
    X[e] = v[e] + k[e] * dt2;
  }
  __syncthreads();
  matvec<ALDS, NT>(p, q, sA, cf, X, k, R);  // k3
  __syncthreads();
  for (int e = q.li; q.live && e < N; e += q.gsz) {
    acc[e] = acc[e] + k[e] * 2.0;
    X[e] = v[e] + k[e] * dt;
  }
  __syncthreads();
  matvec<ALDS, NT>(p, q, sA, cf, X, k, R);  // k4
  __syncthreads();
  for (int e = q.li; q.live && e < N; e += q.gsz) {
    const cpx<double> a = acc[e] + k[e];
    v[e] = v[e] + mk(a.x / 6.0 * dt, a.y / 6.0 * dt);
  }
  __syncthreads();
}

// LDS copy of the generators (shared by every instance of the workgroup)
template <bool ALDS, int NT>
__device__ void stage_generators(const LvParams& p, cpx<double>* sA) {
  if (!ALDS) return;
  const int N = NT > 0 ? NT : p.N;
  const int64_t n = (int64_t)p.nterm * N * N;
  for (int64_t e = threadIdx.x; e < n; e += blockDim.x) sA[e] = mk(p.A[e].x, p.A[e].y);
}

// LDS layout: [generators (ALDS)] [per group: 4 N state] [red: blockDim doubles]
// [per group: field window, ch (nterm - 1) doubles]
template <bool ALDS, int NT>
__global__ __launch_bounds__(kT) void liouville_forward_kernel(LvParams p) {
  extern __shared__ __align__(16) char smem[];
  const int N = NT > 0 ? NT : p.N;
  const Grp q = group_of(p);
  const int groups = blockDim.x / p.gsz;
  const int g = q.tid / p.gsz, nt1 = p.nterm - 1;
  cpx<double>* sA = (cpx<double>*)smem;
  cpx<double>* base = sA + (ALDS ? (int64_t)p.nterm * N * N : 0);
  cpx<double>* v = base + (int64_t)g * 4 * N;
  cpx<double>* X = v + N;
  cpx<double>* k = X + N;
  cpx<double>* acc = k + N;
  double* red = (double*)(base + (int64_t)groups * 4 * N);
  Win cw{red + blockDim.x + (int64_t)g * p.ch * nt1, 0, 0};
  stage_generators<ALDS, NT>(p, sA);
  double2* rho = p.rho + q.b * N;
  for (int e = q.li; q.live && e < N; e += q.gsz) v[e] = mk(rho[e].x, rho[e].y);
  __syncthreads();
  Rows<NT> R;
  load_rows<ALDS, NT>(p, q, sA, R);
  if (p.n_obs) observe<ALDS, NT>(p, q, v, 0, red);
  for (int64_t s = 0; s < p.nsteps; ++s) {
    if (nt1) cf_ensure(p, q, cw, s, false);
    rk4_step<ALDS, NT>(p, q, sA, nt1 ? cw.w + (s - cw.lo) * nt1 : nullptr, v, X, k, acc, R);
    if (p.n_obs) observe<ALDS, NT>(p, q, v, s + 1, red);
    if (p.rho_out && ((s + 1) % p.save_every) == 0) {
      const int64_t slot = (s + 1) / p.save_every - 1;
      const int64_t nslot = p.nsteps / p.save_every;
      for (int e = q.li; q.live && e < N; e += q.gsz)
        p.rho_out[(q.b * nslot + slot) * N + e] = make_double2(v[e].x, v[e].y);
    }
  }
  for (int e = q.li; q.live && e < N; e += q.gsz) rho[e] = make_double2(v[e].x, v[e].y);
}

// Re <a, b> = Re sum conj(a) b over the group
__device__ double redot(const Grp& q, const cpx<double>* a, const cpx<double>* b, int N, double* red) {
  double s = 0.0;
  for (int e = q.li; q.live && e < N; e += q.gsz) s += a[e].x * b[e].x + a[e].y * b[e].y;
  return group_sum(q, s, red);
}

// LDS layout: [generators (ALDS)] [per group: 11 N state, then nterm N^2 gradient accumulators (GALDS)]
// [red: blockDim doubles] [per group: field window and field-gradient window, ch (nterm - 1) doubles each]
template <bool ALDS, bool GALDS, int NT>
__global__ __launch_bounds__(kT) void liouville_adjoint_kernel(LvParams p) {
  extern __shared__ __align__(16) char smem[];
  const int N = NT > 0 ? NT : p.N;
  const Grp q = group_of(p);
  const int groups = blockDim.x / p.gsz;
  const int64_t NN = (int64_t)N * N;
  const int64_t per = 11 * (int64_t)N + (GALDS ? p.nterm * NN : 0);
  cpx<double>* sA = (cpx<double>*)smem;
  cpx<double>* base = sA + (ALDS ? p.nterm * NN : 0);
  cpx<double>* v = base + (int64_t)(q.tid / p.gsz) * per;  // current forward state
  cpx<double>* X1 = v + N;                                 // stage inputs (X1 = y)
  cpx<double>* X2 = X1 + N;
  cpx<double>* X3 = X2 + N;
  cpx<double>* X4 = X3 + N;
  cpx<double>* k = X4 + N;
  cpx<double>* kb = k + N;    // kbar of the current stage
  cpx<double>* lam = kb + N;  // adjoint of the step's output
  cpx<double>* ub = lam + N;  // accumulating adjoint of the step's input
  cpx<double>* Xb = ub + N;
  cpx<double>* acc = Xb + N;
  cpx<double>* sG = acc + N;  // [nterm][N][N] gradient accumulators (GALDS)
  double* red = (double*)(base + (int64_t)groups * per);
  const int g = q.tid / p.gsz, nt1 = p.nterm - 1;
  Win cw{red + blockDim.x + (int64_t)g * 2 * p.ch * nt1, 0, 0};
  Win gw{cw.w + (int64_t)p.ch * nt1, p.nsteps, p.nsteps};
  stage_generators<ALDS, NT>(p, sA);
  const int64_t b = q.b;
  const double dt = p.dt;
  const int ce = p.ckpt_every;
  double2* ck = p.ckpt + b * p.nck * N;
  double2* seg = p.ckpt + (int64_t)p.n_inst * p.nck * N + b * (int64_t)ce * N;
  // ---- forward sweep with checkpoints
  for (int e = q.li; q.live && e < N; e += q.gsz) {
    const double2 r = p.rho[b * N + e];
    v[e] = mk(r.x, r.y);
    ck[e] = r;
  }
  __syncthreads();
  Rows<NT> R;
  load_rows<ALDS, NT>(p, q, sA, R);
  for (int64_t s = 0; s < p.nsteps; ++s) {
    if (nt1) cf_ensure(p, q, cw, s, false);
    rk4_step<ALDS, NT>(p, q, sA, nt1 ? cw.w + (s - cw.lo) * nt1 : nullptr, v, X1, k, acc, R);
    if ((s + 1) % ce == 0 && (s + 1) / ce < p.nck)
      for (int e = q.li; q.live && e < N; e += q.gsz) ck[((s + 1) / ce) * N + e] = make_double2(v[e].x, v[e].y);
  }
  __syncthreads();
  // ---- backward sweep
  for (int e = q.li; q.live && e < N; e += q.gsz) {
    const double2 r = p.rhoT_bar[b * N + e];
    lam[e] = mk(r.x, r.y);
  }
  if (p.grad_A && q.live) {
    if (GALDS)
      for (int64_t e = q.li; e < p.nterm * NN; e += q.gsz) sG[e] = mk(0.0, 0.0);
    else
      for (int64_t e = q.li; e < p.nterm * NN; e += q.gsz) p.grad_A[b * p.nterm * NN + e] = make_double2(0, 0);
  }
  __syncthreads();
  const int64_t nseg = (p.nsteps + ce - 1) / ce;
  for (int64_t c = nseg - 1; c >= 0; --c) {
    const int64_t s0 = c * ce, s1 = (s0 + ce < p.nsteps) ? s0 + ce : p.nsteps;
    // recompute the segment's input states y_s0 .. y_{s1-1}
    for (int e = q.li; q.live && e < N; e += q.gsz) {
      const double2 r = ck[c * N + e];
      v[e] = mk(r.x, r.y);
    }
    __syncthreads();
    for (int64_t s = s0; s < s1; ++s) {
      for (int e = q.li; q.live && e < N; e += q.gsz) seg[(s - s0) * N + e] = make_double2(v[e].x, v[e].y);
      __syncthreads();
      if (s + 1 < s1) {
        if (nt1) cf_ensure(p, q, cw, s, false);
        rk4_step<ALDS, NT>(p, q, sA, nt1 ? cw.w + (s - cw.lo) * nt1 : nullptr, v, X1, k, acc, R);
      }
    }
    __syncthreads();
    for (int64_t s = s1 - 1; s >= s0; --s) {
      const double* cf = nullptr;
      if (nt1) {
        cf_ensure(p, q, cw, s, true);
        if (s < gw.lo) gf_move(p, q, gw, s);
        cf = cw.w + (s - cw.lo) * nt1;
      }
      // stage inputs of step s
      for (int e = q.li; q.live && e < N; e += q.gsz) {
        const double2 r = seg[(s - s0) * N + e];
        X1[e] = mk(r.x, r.y);
      }
      __syncthreads();
      matvec<ALDS, NT>(p, q, sA, cf, X1, k, R);
      __syncthreads();
      for (int e = q.li; q.live && e < N; e += q.gsz) X2[e] = X1[e] + k[e] * (dt / 2.0);
      __syncthreads();
      matvec<ALDS, NT>(p, q, sA, cf, X2, k, R);
      __syncthreads();
      for (int e = q.li; q.live && e < N; e += q.gsz) X3[e] = X1[e] + k[e] * (dt / 2.0);
      __syncthreads();
      matvec<ALDS, NT>(p, q, sA, cf, X3, k, R);
      __syncthreads();
      for (int e = q.li; q.live && e < N; e += q.gsz) {
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
          matvec_term<ALDS, NT>(p, q, sA, i, Xst, k, R);
          __syncthreads();
          const double gi = redot(q, kb, k, N, red);
          if (q.live && q.li == 0) gw.w[(s - gw.lo) * nt1 + (i - 1)] += gi;
        }
        if (p.grad_A) {
          for (int64_t e = q.li; q.live && e < NN; e += q.gsz) {
            const int r = (int)(e / N), cc = (int)(e % N);
            const cpx<double> o = kb[r] * conjg(Xst[cc]);
            for (int i = 0; i < p.nterm; ++i) {
              const double f = coeff(cf, i);
              if (GALDS) {
                sG[i * NN + e] = sG[i * NN + e] + o * f;
              } else {
                double2& gA = p.grad_A[(b * p.nterm + i) * NN + e];
                gA.x += o.x * f;
                gA.y += o.y * f;
              }
            }
          }
        }
        matvec_h<ALDS, NT>(p, q, sA, cf, kb, Xb);
        __syncthreads();
        for (int e = q.li; q.live && e < N; e += q.gsz) {
          ub[e] = ub[e] + Xb[e];
          if (st > 0) kb[e] = lam[e] * wk[st - 1] + Xb[e] * carry[st];
        }
        __syncthreads();
      }
      for (int e = q.li; q.live && e < N; e += q.gsz) lam[e] = ub[e];
      __syncthreads();
    }
  }
  if (nt1) gf_move(p, q, gw, -1);
  if (GALDS && p.grad_A && q.live)
    for (int64_t e = q.li; e < p.nterm * NN; e += q.gsz) p.grad_A[b * p.nterm * NN + e] = make_double2(sG[e].x, sG[e].y);
  if (p.rho0_bar)
    for (int e = q.li; q.live && e < N; e += q.gsz) p.rho0_bar[b * N + e] = make_double2(lam[e].x, lam[e].y);
}

// Launch shape: gsz lanes per instance (a power of two >= N, at most kT), workgroups of
// max(64, gsz) lanes; returns the workgroup size.
int lv_shape(int N, int* gsz) {
  int g = 1;
  while (g < N && g < kT) g *= 2;
  *gsz = g;
  return g < 64 ? 64 : g;
}

// rows per field window: about 2 KB of windows (n_win per group) per workgroup, 1..256 rows
int field_window(int nterm, int groups, int n_win) {
  if (nterm <= 1) return 1;
  const int rows = 256 / (groups * n_win * (nterm - 1));
  return rows < 1 ? 1 : (rows > 256 ? 256 : rows);
}

int max_lds_bytes() {
  int dev = 0, v = 0;
  if (hipGetDevice(&dev) != hipSuccess) return 65536;
  if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess || v <= 0) return 65536;
  return v;
}

template <typename K>
int launch_lv(K kernel, int blocks, int nth, size_t lds, hipStream_t st, const LvParams& p) {
  if (lds > 65536)
    PQA_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(kernel), hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)lds));
  hipLaunchKernelGGL(kernel, dim3(blocks), dim3(nth), lds, st, p);
  PQA_HIP(hipGetLastError());
  return PYQED_OK;
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
  int gsz = 1;
  const int nth = lv_shape(N, &gsz);
  p.gsz = gsz;
  const int groups = nth / gsz;
  const int blocks = (n_inst + groups - 1) / groups;
  const size_t maxl = (size_t)max_lds_bytes();
  const size_t a_bytes = (size_t)nterm * N * N * sizeof(double2);
  p.ch = field_window(nterm, groups, 1);
  const size_t st_bytes = (size_t)groups * 4 * N * sizeof(double2) + sizeof(double) * nth +
                          (size_t)groups * p.ch * (nterm - 1) * sizeof(double);
  const bool alds = a_bytes <= (size_t)kLdsGen && a_bytes + st_bytes <= maxl;
  const size_t lds = (alds ? a_bytes : 0) + st_bytes;
  if (lds > maxl) return fail(PYQED_EINVAL, "liouville_run: the state does not fit one workgroup's LDS");
  hipStream_t st = (hipStream_t)stream;
  // N = 4 (a qubit) and N = 9 (three levels) at compile time, at most two generators: the lane's
  // generator rows in registers, the row products unrolled
  if (alds && nterm <= 2 && N == 4) return launch_lv(liouville_forward_kernel<true, 4>, blocks, nth, lds, st, p);
  if (alds && nterm <= 2 && N == 9) return launch_lv(liouville_forward_kernel<true, 9>, blocks, nth, lds, st, p);
  return alds ? launch_lv(liouville_forward_kernel<true, 0>, blocks, nth, lds, st, p)
              : launch_lv(liouville_forward_kernel<false, 0>, blocks, nth, lds, st, p);
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
  int gsz = 1;
  const int nth = lv_shape(N, &gsz);
  p.gsz = gsz;
  const int groups = nth / gsz;
  const int blocks = (n_inst + groups - 1) / groups;
  const size_t maxl = (size_t)max_lds_bytes();
  const size_t a_bytes = (size_t)nterm * N * N * sizeof(double2);
  const size_t g_bytes = (size_t)groups * nterm * N * N * sizeof(double2);
  p.ch = field_window(nterm, groups, 2);
  const size_t st_bytes = (size_t)groups * 11 * N * sizeof(double2) + sizeof(double) * nth +
                          (size_t)groups * 2 * p.ch * (nterm - 1) * sizeof(double);
  // gradient accumulators first (they replace a read-modify-write of global memory per stage), then
  // the generators
  const bool galds = grad_A && g_bytes <= (size_t)kLdsGen && g_bytes + st_bytes <= maxl;
  const size_t st2 = st_bytes + (galds ? g_bytes : 0);
  const bool alds = a_bytes <= (size_t)kLdsGen && a_bytes + st2 <= maxl;
  const size_t lds = (alds ? a_bytes : 0) + st2;
  if (lds > maxl) return fail(PYQED_EINVAL, "liouville_adjoint: the state does not fit one workgroup's LDS");
  hipStream_t st = (hipStream_t)stream;
  if (alds && galds && nterm <= 2 && N == 4)
    return launch_lv(liouville_adjoint_kernel<true, true, 4>, blocks, nth, lds, st, p);
  if (alds && galds && nterm <= 2 && N == 9)
    return launch_lv(liouville_adjoint_kernel<true, true, 9>, blocks, nth, lds, st, p);
  if (alds && galds) return launch_lv(liouville_adjoint_kernel<true, true, 0>, blocks, nth, lds, st, p);
  if (alds) return launch_lv(liouville_adjoint_kernel<true, false, 0>, blocks, nth, lds, st, p);
  if (galds) return launch_lv(liouville_adjoint_kernel<false, true, 0>, blocks, nth, lds, st, p);
  return launch_lv(liouville_adjoint_kernel<false, false, 0>, blocks, nth, lds, st, p);
}
