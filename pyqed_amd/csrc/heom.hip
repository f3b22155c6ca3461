// HEOM propagator for gfx950: host-side plan (hierarchy tables, packed
// layout, work queue) and the fused RK4-stage kernels.
//
// Operator (pyqed/heom/deom.py:641-673, generate_dot_element / rem_cal):
//   dRho_n/dt = -(sum_k n_k g_k) Rho_n - i[H(t), Rho_n]
//     + sum_{k: n_k>0}  sqrt(n_k)   (cLm_k Q Rho_{n-e_k} + cRm_k Rho_{n-e_k} Q)
//     + sum_{k: |n|<L}  sqrt(n_k+1) (cLp_k Q Rho_{n+e_k} + cRp_k Rho_{n+e_k} Q)
// with cLm = -i etal/sqrt(etaa), cRm = i etar/sqrt(etaa), cLp = -i sqrt(etaa),
// cRp = i sqrt(etaa) and Q = Q_{mode[k]}(t).
//
// Time stepping (deom.py:725-766) is classic RK4.  Each stage is one kernel
// launch that evaluates the right-hand side of every ADO and applies that
// stage's update, so the state crosses HBM once per stage.  The accumulator
// register holds the running new state (ynew = y + dt/6 k1 + dt/3 k2 + ...),
// which lets the last stage write y without re-reading it:
//   stage 0: k1 = F(y);  ynew  = y + k1 dt/6;  X = y + k1 dt/2
//   stage 1: k2 = F(X);  ynew += k2 dt/3;      X = y + k2 dt/2
//   stage 2: k3 = F(X);  ynew += k3 dt/3;      X = y + k3 dt
//   stage 3: k4 = F(X);  y     = ynew + k4 dt/6
// Right-hand-side arithmetic runs in T (double for the plain FP64 path,
// float for the mixed path); y is always double.  The FP64 path keeps ynew
// in double.  The mixed path keeps FP32 increments instead, in the ynew
// buffer, since an increment is O(dt) and FP32 loses nothing FP64 y keeps:
//   full-matrix state, one register inc = ynew - y:
//     stage 0: inc = k1 dt/6;  stages 1-2: inc += k dt/3;  stage 3: y += inc + k4 dt/6
//   packed Hermitian state, two registers A (first n_state elements of the
//   ynew buffer) and B (the next n_state), with y for stages 0-2 read from
//   the FP32 stage input (stage 0's input is y itself; X = y + c k, so
//   y = X - c k for the k that produced X, which the previous stage wrote):
//     stage 0: A = k1 dt/6;  X = y + k1 dt/2
//     stage 1: B = k2 dt/3;  y = X - 3A;  X = y + k2 dt/2
//     stage 2: B = A + B + k3 dt/3;  y = X - 1.5 B_old;  X = y + k3 dt
//     stage 3: y = y + B + k4 dt/6   (y read from and written to FP64)
//
// Layout in HBM: instances packed one after another; inside an instance ADO
// a (rank in the reference's combinadic order) occupies elements
// [a*d*d, (a+1)*d*d), row-major (deom.py:769-771).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <tuple>
#include <type_traits>
#include <string>
#include <vector>

#include "common.h"

using namespace pqa;

namespace {

constexpr int kThreads = 256;
constexpr int kMaxD = 32;
// Streamed work classes: 0 generic, 1 site projector with VALU commutator, 2 site projector with MFMA
// commutator on one 16 x 16 tile (d <= 16), 3 the same on 2 x 2 tiles (17 <= d <= 32)
constexpr int kClasses = 4;
constexpr int kMaxK = 255;
constexpr int kMaxL = 126;

// ------------------------------------------------------------------ device structures
struct InstDev {
  int32_t d, d2, K, L, ncoup, nmax;
  int32_t flags;  // bit0: sdip present, bit1: coupling dipole present, bit2: dense Q
  int32_t grp;    // lanes cooperating on one ADO in the streamed kernel
  int64_t off;    // element offset of ADO 0 in the packed state
  int64_t traj_off;
  int32_t rowptr_off, ent_off, coef_off, mode_off;
  int32_t h_off, q_off;  // offsets (complex elements) into the matrix pool
  // site-projector path: site-ordered tables and coefficients scaled by the projector weight
  int32_t path;          // work class: 0 generic, 1 projector/VALU, 2 projector/MFMA 16, 3 projector/MFMA 2x2x16
  int32_t coefp_off;     // coefficients scaled by the projector weights
  int32_t site_off;      // site of every exponential, [K]
  int32_t dp;            // RK4 state elements per ADO: d*d, or d(d+1)/2 in packed Hermitian mode
  int64_t poff;          // offset of ADO 0 in the RK4 state (y, ynew)
};

struct WorkItem {
  int32_t inst, ado0, nado, pad;
};

// F_SITESORT: site(k) nondecreasing; F_TINY: at most kThreads elements, d <= kTinyD, diagonal couplings and
// at most kTinyEnt neighbour entries per ADO (the resident kernel keeps one element per thread in registers);
// F_TINY4: at most 4 neighbour entries per ADO
// F_REALH: H and the system dipole are real, so H(t) is real and the MFMA commutator skips the products
// with Im H (exact zeros)
enum { F_SDIP = 1, F_CDIP = 2, F_DENSE = 4, F_HERM = 8, F_SITESORT = 16, F_TINY = 32, F_TINY4 = 64, F_REALH = 128 };
constexpr int kTinyD = 8;
constexpr int kTinyEnt = 8;

// Plan tables (instances, work items, row pointers) are never written while a kernel runs. Read
// through the constant address space at a wave-uniform index they become scalar loads, which
// are counted apart from vector loads: waiting for one does not drain the vector loads in flight.
#define PQA_CONST __attribute__((address_space(4)))

template <typename V>
__device__ __forceinline__ V ro_load(const V* q) {
  static_assert(sizeof(V) % 4 == 0, "word-sized records only");
  const PQA_CONST uint32_t* s = (const PQA_CONST uint32_t*)q;
  V v;
  uint32_t* w = (uint32_t*)&v;
#pragma unroll
  for (int i = 0; i < (int)(sizeof(V) / 4); ++i) w[i] = s[i];
  return v;
}

__device__ __forceinline__ int32_t ro_int(const int32_t* q, int64_t i) { return ((const PQA_CONST int32_t*)q)[i]; }

struct StageParams {
  const InstDev* insts;
  const WorkItem* items;
  const int32_t* rowptr;
  const int2* ent;
  const double2* coef;  // 5 per mode: g, cLm, cRm, cLp, cRp
  const int32_t* mode;
  const double2* mat;   // H, sdip, qdiag/qmat, cdiag/cmat
  const int32_t* resident;  // resident instance ids (resident kernel)
  const int32_t* site;      // projector path: site of each exponential
  const double2* coefp;
  const double* field;
  int64_t step;
  double dt;
  const void* xin;
  void* xout;
  double2* y;
  double2* ynew;
  int64_t nstate;     // RK4 state elements (packed mixed runs: offset of the second FP32 register)
  double2* traj;
  int64_t traj_stride;
  int64_t save_slot;  // streamed: slot for this step (-1 = none)
  int64_t nsteps;     // resident kernel: steps to run
  int64_t save_every; // resident kernel
  int32_t item_begin, item_end;
  int32_t xr[9];  // XCD ranges of the item queue (item_begin-relative), xr[8] = count
  int32_t use_xcd;
  int32_t col_override;  // >= 0: field column used instead of the stage's own (RHS-only evaluation)
  int32_t rowbuf, colbuf;  // projector path: per-wave neighbour row / column buffer (complex elements)
  int32_t d2max, qmax, Kmax, Lmax;
  int32_t d2pad;  // projector path: max d*(d+1), matrices in LDS with row stride d+1
  // right-hand-side evaluation (STAGE 4) with combinations, r = L x: xout = cbase + calpha r when
  // cbase is set (else xout = r, when xout is set); cacc = (cinit ? 0 : cacc) + cbeta r when cacc is set
  const double2* cbase;
  double2* cacc;
  double calpha, cbeta;
  int32_t cinit;
  // projector classes with a single streamed instance (one_per > 0): its record is inst0 and item k of
  // the class holds ADOs [k one_per, min((k + 1) one_per, nmax)), so neither table is read; its
  // neighbour lists are read from ent_fix at fix_stride entries per ADO (the row pointers give only
  // their lengths).  Each saves a dependent round trip per work item.
  InstDev inst0;
  int32_t one_per, fix_stride;
  const int2* ent_fix;
};

// LDS carve-up shared by both kernels.
template <typename T>
struct Lds {
  cpx<T>* H;
  cpx<T>* Q;
  cpx<T>* C;
  int32_t* M;
  T* S;
  char* rest;
};

template <typename T>
__device__ __forceinline__ Lds<T> carve(char* base, const StageParams& p) {
  Lds<T> l;
  size_t o = 0;
  l.H = (cpx<T>*)(base + o);
  o += sizeof(cpx<T>) * p.d2max;
  l.Q = (cpx<T>*)(base + o);
  o += sizeof(cpx<T>) * p.qmax;
  l.C = (cpx<T>*)(base + o);
  o += sizeof(cpx<T>) * 5 * p.Kmax;
  l.S = (T*)(base + o);  // 8-byte data before the int array keeps every field aligned
  o += sizeof(T) * (p.Lmax + 2);
  l.M = (int32_t*)(base + o);
  o += sizeof(int32_t) * p.Kmax;
  o = (o + 15) & ~size_t(15);
  l.rest = base + o;
  return l;
}

__host__ __device__ inline size_t lds_header_bytes(int d2max, int qmax, int Kmax, int Lmax, size_t tsize) {
  size_t o = 2 * tsize * (size_t)d2max + 2 * tsize * (size_t)qmax + 2 * tsize * 5 * (size_t)Kmax + 4 * (size_t)Kmax +
             tsize * (size_t)(Lmax + 2);
  return (o + 15) & ~size_t(15);
}

__device__ __forceinline__ int stage_col(int stage) { return stage == 0 ? 0 : (stage == 3 ? 2 : 1); }

// Load H(t), Q(t) diagonals, per-mode coefficients and sqrt table of one instance into LDS.
template <typename T>
__device__ __forceinline__ void load_instance(const StageParams& p, const InstDev& I, const Lds<T>& l, int c,
                                              int64_t step) {
  const int tid = threadIdx.x;
  double fs = 0.0, fc = 0.0;
  if (p.field) {
    fs = p.field[6 * step + c];
    fc = p.field[6 * step + 3 + c];
  }
  const double2* H = p.mat + I.h_off;
  const bool sd = (I.flags & F_SDIP) != 0;
  for (int e = tid; e < I.d2; e += kThreads) {
    double2 h = H[e];
    if (sd) {
      double2 s = H[I.d2 + e];
      h.x += s.x * fs;
      h.y += s.y * fs;
    }
    l.H[e] = mk<T>((T)h.x, (T)h.y);
  }
  if (!(I.flags & F_DENSE)) {
    const int nq = I.ncoup * I.d;
    const double2* q = p.mat + I.q_off;
    const bool cd = (I.flags & F_CDIP) != 0;
    for (int e = tid; e < nq; e += kThreads) {
      double2 v = q[e];
      if (cd) {
        double2 s = q[nq + e];
        v.x += s.x * fc;
        v.y += s.y * fc;
      }
      l.Q[e] = mk<T>((T)v.x, (T)v.y);
    }
  }
  for (int e = tid; e < 5 * I.K; e += kThreads) {
    double2 v = p.coef[I.coef_off + e];
    l.C[e] = mk<T>((T)v.x, (T)v.y);
  }
  for (int e = tid; e < I.K; e += kThreads) l.M[e] = p.mode[I.mode_off + e];
  for (int e = tid; e < I.L + 2; e += kThreads) l.S[e] = (T)sqrt((double)e);
}

template <typename T, typename S>
__device__ __forceinline__ cpx<T> ldc(const S* base, int64_t idx);

template <>
__device__ __forceinline__ cpx<double> ldc<double, cpx_store<double>>(const cpx_store<double>* b, int64_t i) {
  return ld(b + i);
}
template <>
__device__ __forceinline__ cpx<float> ldc<float, cpx_store<float>>(const cpx_store<float>* b, int64_t i) {
  return ld(b + i);
}
template <>
__device__ __forceinline__ cpx<float> ldc<float, cpx_store<double>>(const cpx_store<double>* b, int64_t i) {
  return cvt<float>(ld(b + i));
}
template <>
__device__ __forceinline__ cpx<double> ldc<double, cpx<double>>(const cpx<double>* b, int64_t i) {
  return b[i];
}
template <>
__device__ __forceinline__ cpx<float> ldc<float, cpx<float>>(const cpx<float>* b, int64_t i) {
  return b[i];
}
template <>
__device__ __forceinline__ cpx<float> ldc<float, cpx<double>>(const cpx<double>* b, int64_t i) {
  return cvt<float>(b[i]);
}

// Right-hand side of element (i, j) of ADO `a`.
//   own : the ADO's own d x d matrix (LDS), in T
//   xin : base of the instance's stage input (ADO 0), indexable by ldc
//   rowp, entp: the instance's row pointers and neighbour entries (global memory, or LDS copies)
template <typename T, typename OS, typename XS>
__device__ __forceinline__ cpx<T> element_rhs(const StageParams& p, const InstDev& I, const Lds<T>& l,
                                              const OS* own, const XS* xin, int a, int i, int j,
                                              double fc, const int32_t* rowp, const int2* ent) {
  const int d = I.d, d2 = I.d2;
  cpx<T> c = mk<T>(0, 0);
  for (int k = 0; k < d; ++k) {
    cmac(c, l.H[i * d + k], ldc<T>(own, k * d + j));
    cmsub(c, ldc<T>(own, i * d + k), l.H[k * d + j]);
  }
  cpx<T> r = mul_mi(c);
  cpx<T> damp = mk<T>(0, 0);
  const int r0 = rowp[a], r1 = rowp[a + 1];
  const bool dense = (I.flags & F_DENSE) != 0;
  for (int t = r0; t < r1; ++t) {
    const int2 en = ent[t];
    const int k = en.y & 0xffff;
    const int nf = (en.y >> 16) & 0x7f;
    const bool plus = en.y < 0;
    const cpx<T>* C = l.C + 5 * k;
    if (!plus) damp = damp + C[0] * (T)nf;
    const cpx<T> cL = plus ? C[3] : C[1];
    const cpx<T> cR = plus ? C[4] : C[2];
    const T sq = l.S[nf];
    const int m = l.M[k];
    const int64_t nb = (int64_t)en.x * d2;
    if (!dense) {
      cpx<T> coef = (cL * l.Q[m * d + i] + cR * l.Q[m * d + j]) * sq;
      if (coef.x != (T)0 || coef.y != (T)0) cmac(r, coef, ldc<T>(xin, nb + i * d + j));
    } else {
      // dense coupling: cL (Q X)_{ij} + cR (X Q)_{ij}, Q read from the FP64 matrix pool
      const double2* Qm = p.mat + I.q_off + (int64_t)m * d2;
      const double2* Cm = (I.flags & F_CDIP) ? Qm + (int64_t)I.ncoup * d2 : nullptr;
      cpx<T> s1 = mk<T>(0, 0), s2 = mk<T>(0, 0);
      for (int q = 0; q < d; ++q) {
        double2 qa = Qm[i * d + q], qb = Qm[q * d + j];
        if (Cm) {
          double2 ca = Cm[i * d + q], cb = Cm[q * d + j];
          qa.x += ca.x * fc;
          qa.y += ca.y * fc;
          qb.x += cb.x * fc;
          qb.y += cb.y * fc;
        }
        cmac(s1, mk<T>((T)qa.x, (T)qa.y), ldc<T>(xin, nb + q * d + j));
        cmac(s2, ldc<T>(xin, nb + i * d + q), mk<T>((T)qb.x, (T)qb.y));
      }
      cpx<T> v = cL * s1 + cR * s2;
      r = r + v * sq;
    }
  }
  cmsub(r, damp, ldc<T>(own, i * d + j));
  return r;
}

// Address of element (i, j) of ADO a in the RK4 state (y, ynew).  The state holds either the full
// matrix of every ADO (d*d elements from I.off) or, in packed Hermitian mode, its upper triangle row
// by row (d(d+1)/2 elements from I.poff).  A lower element (i > j) of a packed ADO reads the
// conjugate of (j, i) and writes no state; the stage inputs X are always full matrices.
struct StateRef {
  int64_t g;
  bool lower;
};

template <bool PACK>
__device__ __forceinline__ StateRef state_ref(const InstDev& I, int64_t a, int i, int j) {
  StateRef r;
  if (!PACK) {
    r.g = I.off + a * I.d2 + (i * I.d + j);
    r.lower = false;
  } else {
    r.lower = i > j;
    const int u = r.lower ? j : i, v = r.lower ? i : j;
    r.g = I.poff + a * I.dp + ((u * (2 * I.d - u - 1)) >> 1) + v;  // row u starts at u(2d-u+1)/2
  }
  return r;
}

__device__ __forceinline__ double2 ld_state(const double2* q, const StateRef& s) {
  double2 v = q[s.g];
  if (s.lower) v.y = -v.y;
  return v;
}

// The RK4 register.  FP64 runs keep the running new state in it (ynew = y + dt/6 k1 + dt/3 k2 + ...,
// complex double).  Mixed-precision runs (DELTA) keep only the increment ynew - y as complex float:
// stages 0-2 accumulate dt/6 k1 + dt/3 k2 + dt/3 k3 into it and stage 3 forms y + (inc + dt/6 k4).
// The increments come from FP32 right-hand sides, so FP32 storage loses nothing they carry, and the
// register moves 8 instead of 16 bytes per access; stage 3 then reads y (16 bytes) as well.
template <bool DELTA>
struct Pre {
  double2 y;
  typename std::conditional<DELTA, float2, double2>::type an;
};

__device__ __forceinline__ float2 ld_delta(const double2* q, const StateRef& s) {
  float2 v = ((const float2*)q)[s.g];
  if (s.lower) v.y = -v.y;
  return v;
}

template <int STAGE, bool DELTA>
__device__ __forceinline__ Pre<DELTA> rk4_prefetch(const StageParams& p, const StateRef& s) {
  Pre<DELTA> v;
  v.y = (STAGE <= 2 || (DELTA && STAGE == 3)) ? ld_state(p.y, s) : make_double2(0, 0);
  if constexpr (DELTA) {
    v.an = (STAGE >= 1 && STAGE <= 3) ? ld_delta(p.ynew, s) : make_float2(0.f, 0.f);
  } else {
    v.an = (STAGE >= 1 && STAGE <= 3) ? ld_state(p.ynew, s) : make_double2(0, 0);
  }
  return v;
}

// Packed mixed-precision RK4.  Each upper element has two FP32 registers, A (the low half of the
// register buffer) and B (the high half, from element nstate on), and stages 0-2 take y from their
// FP32 stage input, which is in LDS, instead of the FP64 state:
//   stage 0: A = dt/6 k1;              X   = y + dt/2 k1   (input: y itself)
//   stage 1: B = dt/3 k2;              X'  = y + dt/2 k2   (input X: y = X - 3 A)
//   stage 2: B = A + B + dt/3 k3;      X'' = y + dt k3     (input X': y = X' - 3/2 B)
//   stage 3: y = y + (B + dt/6 k4)     (the FP64 state: read and written)
// The recovered y carries FP32 accuracy, all that the FP32 stage outputs hold; the increments come
// from FP32 right-hand sides, and only stage 3 adds them to the FP64 state.
struct PreM {
  double2 y;
  float2 a, b;
};

template <int STAGE>
__device__ __forceinline__ PreM rk4_prefetch_m(const StageParams& p, int64_t g) {
  PreM v;
  const float2* A = (const float2*)p.ynew;
  if (STAGE == 1 || STAGE == 2) v.a = A[g];
  if (STAGE == 2 || STAGE == 3) v.b = A[p.nstate + g];
  if (STAGE == 3) v.y = p.y[g];
  return v;
}

// y of an upper element from its FP32 stage input x (stages 0-2).
template <int STAGE, typename TX>
__device__ __forceinline__ double2 y_from_input_m(const cpx<TX>& x, const PreM& v) {
  if (STAGE == 1) return make_double2((double)x.x - 3.0 * (double)v.a.x, (double)x.y - 3.0 * (double)v.a.y);
  if (STAGE == 2) return make_double2((double)x.x - 1.5 * (double)v.b.x, (double)x.y - 1.5 * (double)v.b.y);
  return make_double2((double)x.x, (double)x.y);
}

// The update of upper element g (v.y: y recovered from the input for stages 0-2, the FP64 state for
// stage 3); returns the element of the stage output.
template <int STAGE>
__device__ __forceinline__ double2 rk4_upper_m(const StageParams& p, int64_t g, double rx, double ry, const PreM& v) {
  const double dt = p.dt;
  float2* A = (float2*)p.ynew;
  if (STAGE == 0) {
    A[g] = make_float2((float)(rx * (dt / 6)), (float)(ry * (dt / 6)));
    return make_double2(v.y.x + rx * (dt / 2), v.y.y + ry * (dt / 2));
  } else if (STAGE == 1) {
    A[p.nstate + g] = make_float2((float)(rx * (dt / 3)), (float)(ry * (dt / 3)));
    return make_double2(v.y.x + rx * (dt / 2), v.y.y + ry * (dt / 2));
  } else if (STAGE == 2) {
    A[p.nstate + g] = make_float2((float)((double)v.a.x + (double)v.b.x + rx * (dt / 3)),
                                  (float)((double)v.a.y + (double)v.b.y + ry * (dt / 3)));
    return make_double2(v.y.x + rx * dt, v.y.y + ry * dt);
  } else {
    const double2 yn = make_double2(v.y.x + ((double)v.b.x + rx * (dt / 6)), v.y.y + ((double)v.b.y + ry * (dt / 6)));
    p.y[g] = yn;
    return yn;
  }
}

// Writes the RK4 register of state element g: value (x, y) as complex double, or as complex float
// in DELTA mode.
template <bool DELTA>
__device__ __forceinline__ void st_reg(const StageParams& p, int64_t g, double x, double y) {
  if constexpr (DELTA) {
    ((float2*)p.ynew)[g] = make_float2((float)x, (float)y);
  } else {
    p.ynew[g] = make_double2(x, y);
  }
}

// RK4 stage update of one element (i, j) from its right-hand side (rx, ry): gx is its index in the
// stage buffers (full matrices) and gxT that of (j, i), s its RK4 state address.  In packed mode
// stage 3 also writes the new state to xout (the stage-0 input of the next step) as the element
// and its conjugate at (j, i), from the upper element only, so that this copy is exactly the one
// the unpack kernel makes from y.  `save` marks ADO 0 of an instance on a trajectory step; tpos is
// the element's trajectory slot and tposT that of (j, i).
template <typename TO, int STAGE, bool PACK, bool DELTA>
__device__ __forceinline__ void rk4_store(const StageParams& p, int64_t gx, int64_t gxT, const StateRef& s, double rx,
                                          double ry, const Pre<DELTA>& v, cpx_store<TO>* xout, bool save, int64_t tpos,
                                          int64_t tposT) {
  const double dt = p.dt;
  if (STAGE == 4) {
    if (p.cacc) {
      const double2 a = p.cinit ? make_double2(0, 0) : p.cacc[gx];
      p.cacc[gx] = make_double2(a.x + p.cbeta * rx, a.y + p.cbeta * ry);
    }
    if (p.cbase) {
      const double2 b = p.cbase[gx];
      ((double2*)p.xout)[gx] = make_double2(b.x + p.calpha * rx, b.y + p.calpha * ry);
    } else if (p.xout) {
      ((double2*)p.xout)[gx] = make_double2(rx, ry);
    }
  } else if (STAGE == 0) {
    if (!s.lower) {
      if (DELTA) st_reg<DELTA>(p, s.g, rx * (dt / 6), ry * (dt / 6));
      else st_reg<DELTA>(p, s.g, v.y.x + rx * (dt / 6), v.y.y + ry * (dt / 6));
    }
    st(xout + gx, mk<TO>((TO)(v.y.x + rx * (dt / 2)), (TO)(v.y.y + ry * (dt / 2))));
  } else if (STAGE == 1 || STAGE == 2) {
    if (!s.lower) st_reg<DELTA>(p, s.g, (double)v.an.x + rx * (dt / 3), (double)v.an.y + ry * (dt / 3));
    const double h = STAGE == 1 ? dt / 2 : dt;
    st(xout + gx, mk<TO>((TO)(v.y.x + rx * h), (TO)(v.y.y + ry * h)));
  } else {
    const double2 yn = DELTA ? make_double2(v.y.x + ((double)v.an.x + rx * (dt / 6)),
                                            v.y.y + ((double)v.an.y + ry * (dt / 6)))
                             : make_double2((double)v.an.x + rx * (dt / 6), (double)v.an.y + ry * (dt / 6));
    if (!s.lower) {
      p.y[s.g] = yn;
      // packed (Hermitian) mode: (j, i) is written as the conjugate of (i, j), except on the diagonal
      if (save) {
        p.traj[tpos] = yn;
        if (PACK && tposT != tpos) p.traj[tposT] = make_double2(yn.x, -yn.y);
      }
      if (PACK && xout) {
        st(xout + gx, mk<TO>((TO)yn.x, (TO)yn.y));
        if (gxT != gx) st(xout + gxT, mk<TO>((TO)yn.x, (TO)(-yn.y)));
      }
    }
  }
}

// Packed (Hermitian) mode, upper element (i <= j) at state index g: writes ynew (stages 0-2) or y
// (stage 3) and returns the element of the stage output (X' for stages 0-2, the new state for
// stage 3, whose full-matrix copy is the next step's stage-0 input).  The lower element (j, i) of
// the output is its conjugate.
template <int STAGE, bool DELTA>
__device__ __forceinline__ double2 rk4_upper(const StageParams& p, int64_t g, double rx, double ry, const Pre<DELTA>& v) {
  const double dt = p.dt;
  if (STAGE == 0) {
    if (DELTA) st_reg<DELTA>(p, g, rx * (dt / 6), ry * (dt / 6));
    else st_reg<DELTA>(p, g, v.y.x + rx * (dt / 6), v.y.y + ry * (dt / 6));
    return make_double2(v.y.x + rx * (dt / 2), v.y.y + ry * (dt / 2));
  } else if (STAGE == 1 || STAGE == 2) {
    st_reg<DELTA>(p, g, (double)v.an.x + rx * (dt / 3), (double)v.an.y + ry * (dt / 3));
    const double h = STAGE == 1 ? dt / 2 : dt;
    return make_double2(v.y.x + rx * h, v.y.y + ry * h);
  } else {
    const double2 yn = DELTA ? make_double2(v.y.x + ((double)v.an.x + rx * (dt / 6)),
                                            v.y.y + ((double)v.an.y + ry * (dt / 6)))
                             : make_double2((double)v.an.x + rx * (dt / 6), (double)v.an.y + ry * (dt / 6));
    p.y[g] = yn;
    return yn;
  }
}

// Packed index of upper element (i, j), i <= j, within its ADO (rows of the upper triangle in order).
__device__ __forceinline__ int packed_index(int i, int j, int d) { return ((i * (2 * d - i - 1)) >> 1) + j; }

template <typename TO, int STAGE, bool PACK>
__device__ __forceinline__ void rk4_apply(const StageParams& p, int64_t gx, int64_t gxT, const StateRef& s, double rx,
                                          double ry, cpx_store<TO>* xout, bool save, int64_t tpos, int64_t tposT) {
  constexpr bool DELTA = std::is_same<TO, float>::value;
  const Pre<DELTA> v = rk4_prefetch<STAGE, DELTA>(p, s);
  rk4_store<TO, STAGE, PACK, DELTA>(p, gx, gxT, s, rx, ry, v, xout, save, tpos, tposT);
}

__device__ __forceinline__ void item_range(const StageParams& p, int& it0, int& itend, int& itstep) {
  if (p.use_xcd) {
    // blocks are dispatched round-robin over the 8 XCDs: block b runs on XCD b % 8, and the blocks of
    // one XCD sweep one contiguous eighth of the queue so an instance's tiles share that XCD's L2
    const int x = blockIdx.x & 7, j = blockIdx.x >> 3, nbx = gridDim.x >> 3;
    it0 = p.item_begin + p.xr[x] + j;
    itend = p.item_begin + p.xr[x + 1];
    itstep = nbx;
  } else {
    it0 = p.item_begin + blockIdx.x;
    itend = p.item_end;
    itstep = gridDim.x;
  }
}

// work item `it` of a single-instance projector class (StageParams::one_per)
__device__ __forceinline__ WorkItem one_item(const StageParams& p, int it) {
  WorkItem w;
  w.inst = 0;
  w.ado0 = (it - p.item_begin) * p.one_per;
  w.nado = min(p.one_per, p.inst0.nmax - w.ado0);
  w.pad = 0;
  return w;
}

// ------------------------------------------------------------------ streamed stage kernel
// STAGE 0..3: RK4 stage; STAGE 4: out = RHS(xin) only (FP64 operator check).
template <typename T, typename TI, typename TO, int STAGE, bool PACK>
__global__ __launch_bounds__(kThreads) void heom_stage_kernel(StageParams p) {
  extern __shared__ __align__(16) char smem[];
  const Lds<T> l = carve<T>(smem, p);
  cpx<T>* sX = (cpx<T>*)l.rest;
  const int tid = threadIdx.x;
  const cpx_store<TI>* xin = (const cpx_store<TI>*)p.xin;
  cpx_store<TO>* xout = (cpx_store<TO>*)p.xout;
  const int col = p.col_override >= 0 ? p.col_override : stage_col(STAGE & 3);
  double fc = 0.0;
  if (p.field) fc = p.field[6 * p.step + 3 + col];

  int it0, itstep, itend;
  item_range(p, it0, itend, itstep);
  for (int it = it0; it < itend; it += itstep) {
    const WorkItem w = p.items[it];
    const InstDev I = p.insts[w.inst];
    __syncthreads();
    load_instance<T>(p, I, l, col, p.step);
    const int d2 = I.d2, d = I.d, G = I.grp;
    const int nslot = kThreads / G;
    const int slot = tid / G, le = tid % G;
    cpx<T>* own = sX + slot * d2;
    const cpx_store<TI>* xi = xin + I.off;
    for (int a0 = w.ado0; a0 < w.ado0 + w.nado; a0 += nslot) {
      const int a = a0 + slot;
      const bool act = a < w.ado0 + w.nado;
      __syncthreads();
      if (act)
        for (int e = le; e < d2; e += G) own[e] = ldc<T>(xi, (int64_t)a * d2 + e);
      __syncthreads();
      if (act) {
        for (int e = le; e < d2; e += G) {
          const int i = e / d, j = e - (e / d) * d;
          const cpx<T> r = element_rhs<T>(p, I, l, own, xi, a, i, j, fc, p.rowptr + I.rowptr_off, p.ent + I.ent_off);
          const int64_t g = I.off + (int64_t)a * d2 + e;
          const bool save = STAGE == 3 && a == 0 && p.traj && p.save_slot >= 0;
          const int64_t tb = p.save_slot * p.traj_stride + I.traj_off;
          rk4_apply<TO, STAGE, PACK>(p, g, I.off + (int64_t)a * d2 + j * d + i, state_ref<PACK>(I, a, i, j), (double)r.x,
                               (double)r.y, xout, save, tb + e, tb + j * d + i);
        }
      }
    }
  }
}

// ------------------------------------------------------------------ site-projector stage kernel
// For couplings Q_m = c_m |p_m><p_m| the coupling term of a neighbour X' entering through
// exponential k (site s = p_{mode[k]}) touches only row s and column s of the output:
//   out(s, j) += c cL_k X'(s, j),   out(i, s) += c cR_k X'(i, s).
// One wavefront owns one ADO and reaches memory in two round trips:
//   1. its neighbour list, its own matrix and the RK4 operands (y, ynew) it will update;
//   2. row s of every neighbour, lanes spread over (entry, column), staged in an LDS row buffer
//      (for non-Hermitian ADOs column s as well).
// The neighbour list is sorted by site in LDS (ballot counting sort), so element (i, j) reads only
// its two groups from the row buffer: sum_{site(t)=i} cL_t R_t(j) + sum_{site(t)=j} cR_t R_t(i)^*
// (Hermitian ADOs: X'(i, s) = conj(X'(s, i))).  The commutator -i[H, X] runs on FP32 MFMA 16x16x4
// for PQA_MFMA_DMIN <= d <= 16 and on VALU otherwise, while the round-trip-2 loads are in flight.
typedef float floatx4 __attribute__((ext_vector_type(4)));

constexpr int kMaxEnt = 64;  // neighbour entries per ADO handled by the projector path
#ifndef PQA_PROJ_WAVES
#define PQA_PROJ_WAVES 3  // minimum waves per SIMD the projector kernel is compiled for
#endif
#ifndef PQA_VALU_WAVES
#define PQA_VALU_WAVES 4  // the same for its FP32 VALU-commutator form (small d, fewer registers)
#endif
#ifndef PQA_MFMA_DMIN
#define PQA_MFMA_DMIN 12  // smallest d whose projector kernel forms the commutator on MFMA (8: same step time)
#endif
#ifndef PQA_PACK_WAVES
#define PQA_PACK_WAVES 4  // the same for the packed-state MFMA form
#endif
#ifndef PQA_PACK_WAVES_LAST
#define PQA_PACK_WAVES_LAST PQA_PACK_WAVES  // the same for its last RK4 stage
#endif
#ifndef PQA_CHUNK
#define PQA_CHUNK 2
#endif
constexpr int kChunk = PQA_CHUNK;  // neighbour loads per lane issued together
#ifndef PQA_CHUNK_WIDE
#define PQA_CHUNK_WIDE 8  // the same for the wide MFMA class (17 <= d <= 32: about 50 neighbour rows of d)
#endif
#ifndef PQA_WIDE_WAVES
#define PQA_WIDE_WAVES 2  // minimum waves per SIMD of the wide MFMA class (LDS allows two blocks per CU)
#endif
#ifndef PQA_CHUNK_S1
#define PQA_CHUNK_S1 3  // the same for stage 1 of the MFMA projector class
#endif
#ifndef PQA_RK4_HOIST
// MFMA path, RK4 operands of a lane's four elements: 0 one request per element at its update,
// 1 requested together after the neighbour rows have arrived, 2 requested together right behind
// the neighbour-row requests (both round trips overlap the commutator)
#define PQA_RK4_HOIST 1
#endif
#ifndef PQA_SIDE_CLASSES
// kernel classes [0, PQA_SIDE_CLASSES) run their stage launches on a second stream of the plan,
// concurrently with the others (the classes share no elements): 1 the generic class, 2 with the VALU
// projector class (measured slower on the 4096-instance batch: the two projector classes slow each
// other down more than the overlap gains)
#define PQA_SIDE_CLASSES 1
#endif
#ifndef PQA_HALF_COMMUTATOR
// packed Hermitian MFMA forms: 1 form only P = H X on MFMA and complete the commutator from the
// output tile; 0 form H X - X H on MFMA
#define PQA_HALF_COMMUTATOR 1
#endif
#ifndef PQA_ROWS_DMA
// wide projector class: neighbour rows by LDS DMA (global_load_lds_dword) instead of register chunks
#define PQA_ROWS_DMA 1
#endif
#ifndef PQA_ROWS_DMA_NARROW
// the same for the packed 16 x 16 MFMA class (0: register chunks; 1 measured the same stage times)
#define PQA_ROWS_DMA_NARROW 0
#endif
#ifndef PQA_PACK_HOIST
// packed MFMA forms: 0 request the RK4 operands of an ADO's upper elements once its neighbour rows
// have arrived; 1 request them right behind the first chunk of neighbour rows (they then arrive during
// the commutator, at the cost of the registers that hold them meanwhile)
#define PQA_PACK_HOIST 0
#endif

__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <typename T>
__device__ __forceinline__ cpx<T> wave_sum(cpx<T> v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    v.x += __shfl_xor(v.x, o, 64);
    v.y += __shfl_xor(v.y, o, 64);
  }
  return v;
}

// Site order of a wave's neighbour list (one entry per lane, my_site = INT_MAX past the end):
// returns the entry's position in site order and, in lanes 0..d, the start of each site's group.
// The tables list entries by exponential k (minus, then plus), so when site(k) is nondecreasing
// (F_SITESORT) the list is already in site order and only the group starts are counted; otherwise
// a stable counting sort over ballots orders it.  Both keep the table order within a site.
__device__ __forceinline__ int site_order(int my_site, int d, int lane, bool presorted, int& gsv) {
  gsv = 0;
  if (presorted) {
    for (int s2 = 0; s2 <= d; ++s2) {
      const int c = __popcll(__ballot(my_site < s2));
      if (lane == s2) gsv = c;
    }
    return lane;
  }
  int pos = 0, run = 0;
  const uint64_t below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  for (int s2 = 0; s2 < d; ++s2) {
    const uint64_t m = __ballot(my_site == s2);
    if (lane == s2) gsv = run;
    if (my_site == s2) pos = run + __popcll(m & below);
    run += __popcll(m);
  }
  if (lane == d) gsv = run;
  return pos;
}

template <typename T>
struct WaveLds {
  cpx<T>* X;    // own ADO, d*d
  cpx<T>* R;    // neighbour rows R[t][0..d), t in site order
  cpx<T>* Cc;   // neighbour columns (non-Hermitian only)
  cpx<T>* cl;   // per entry: coefficient of the row term
  cpx<T>* cr;   // per entry: coefficient of the column term
  int32_t* nb;  // per entry: neighbour ADO
  int32_t* st;  // per entry: site
  int32_t* gs;  // site-group starts, d+1
  cpx<T>* Tt;   // packed mode: the upper elements of the stage output, d*d (row stride d+1)
};

// Per-wave LDS of the projector kernel; `tile` adds the stage-output tile of packed mode.
__host__ __device__ inline size_t proj_wave_bytes(int xsize, int rowbuf, int colbuf, size_t ts, bool tile) {
  size_t b = 2 * ts * ((size_t)xsize + rowbuf + colbuf) + 2 * 2 * ts * kMaxEnt + 2 * 4 * kMaxEnt + 4 * 36;
  b = (b + 15) & ~size_t(15);
  return tile ? b + ((2 * ts * (size_t)xsize + 15) & ~size_t(15)) : b;
}
// Per-block LDS of the projector kernel: H, coefficients, sqrt table, sites and, in packed mode, the
// (i, j) of every packed index (at most xsize / 2 of them).
__host__ __device__ inline size_t proj_block_bytes(int xsize, int Kmax, int Lmax, size_t ts, bool pack) {
  size_t b = 2 * ts * (size_t)xsize + 2 * ts * 5 * (size_t)Kmax + ts * (size_t)(Lmax + 2) + 4 * (size_t)Kmax;
  if (pack) b += 4 * ((size_t)xsize / 2 + 1);
  return (b + 15) & ~size_t(15);
}

// Per-item tables of the projector kernels in LDS: H (+ field x system dipole) with row stride d+1,
// the projector-scaled coefficients, the site of every exponential and sqrt(0..L+1).  One pass
// issues all of its loads before the first LDS store waits for them.
template <typename T>
__device__ __forceinline__ void proj_header(const StageParams& p, const InstDev& I, int col, uint32_t dmag,
                                            cpx<T>* sH, cpx<T>* sC, int32_t* sSite, T* sS, int32_t* sIJ) {
  const double fs = p.field ? ro_load(p.field + 6 * p.step + col) : 0.0;
  const double2* H = p.mat + I.h_off;
  const bool sd = (I.flags & F_SDIP) != 0;
  const int d2 = I.d2, nc = 5 * I.K, n = max(max(d2, nc), I.L + 2);
  for (int e = threadIdx.x; e < n; e += kThreads) {
    double2 h = make_double2(0, 0), s2 = make_double2(0, 0), c = make_double2(0, 0);
    int32_t st = 0;
    if (e < d2) h = H[e];
    if (sd && e < d2) s2 = H[d2 + e];
    if (e < nc) c = p.coefp[I.coefp_off + e];
    if (e < I.K) st = p.site[I.site_off + e];
    if (e < d2) {
      if (sd) {
        h.x += s2.x * fs;
        h.y += s2.y * fs;
      }
      sH[e + (int)(((uint32_t)e * dmag) >> 20)] = mk<T>((T)h.x, (T)h.y);
    }
    if (e < nc) sC[e] = mk<T>((T)c.x, (T)c.y);
    if (e < I.K) sSite[e] = st;
    if (e < I.L + 2) sS[e] = (T)sqrt((double)e);
  }
  if (sIJ) {
    const int d = I.d, np = (d * (d + 1)) >> 1;
    for (int q = threadIdx.x; q < np; q += kThreads) {
      int i = 0;
      while (((i + 1) * (2 * d - i)) / 2 <= q) ++i;  // row i starts at i(2d-i+1)/2
      sIJ[q] = (i << 8) | (q - ((i * (2 * d - i + 1)) >> 1) + i);
    }
  }
}

// Minimum waves per SIMD each form of the projector kernel is compiled for.
constexpr int proj_waves(bool mfma, bool f32, bool pack, int stage, int nt) {
  return (mfma && nt == 2) ? PQA_WIDE_WAVES
         : mfma ? (pack ? (stage == 3 ? PQA_PACK_WAVES_LAST : PQA_PACK_WAVES) : PQA_PROJ_WAVES)
                : (f32 ? PQA_VALU_WAVES : PQA_PROJ_WAVES);
}

// NT: MFMA tiles per dimension, 1 for d <= 16 (one 16 x 16 tile) or 2 for d <= 32 (2 x 2 tiles of 16).
// ONE: the class holds a single instance (StageParams::one_per, inst0, ent_fix); a compile-time form so
// that the table-driven forms keep their registers.
template <typename T, typename TI, typename TO, int STAGE, bool MFMA, bool HERM, bool PACK, int NT, bool ONE = false>
__global__ __launch_bounds__(kThreads, proj_waves(MFMA, std::is_same<T, float>::value, PACK, STAGE, NT))
void heom_proj_kernel(StageParams p) {
  constexpr bool kDelta = std::is_same<TO, float>::value;  // mixed precision: FP32 RK4 increment
  constexpr bool kPM = kDelta && PACK;  // packed mixed run: two FP32 registers (rk4_upper_m)
  static_assert(NT == 1 || (NT == 2 && MFMA), "2 x 2 tiles are an MFMA form");
#ifndef PQA_WIDE_PF
#define PQA_WIDE_PF 9  // own-matrix elements per lane the wide wave form prefetches (the rest load in place)
#endif
  constexpr int kPF = NT == 1 ? 4 : PQA_WIDE_PF;  // own-matrix elements per lane prefetched (64 kPF)
  constexpr int kQP = NT == 1 ? 3 : 9;  // packed upper elements per lane, d(d+1)/2 <= 64 kQP
  extern __shared__ __align__(16) char smem[];
  const int tid = threadIdx.x, lane = tid & 63;
  const int wv = __builtin_amdgcn_readfirstlane(tid >> 6);  // wave-uniform: ADO indices stay scalar
  cpx<T>* sH = (cpx<T>*)smem;
  cpx<T>* sC = sH + p.d2pad;
  T* sS = (T*)(sC + 5 * p.Kmax);
  int32_t* sSite = (int32_t*)(sS + p.Lmax + 2);
  int32_t* sIJ = sSite + p.Kmax;  // packed mode: (i << 8) | j of packed index p
  char* wbase = smem + proj_block_bytes(p.d2pad, p.Kmax, p.Lmax, sizeof(T), PACK);
  WaveLds<T> w;
  {
    // packed mode: the stage-output tile; the MFMA form reuses the own-matrix buffer for it (the
    // own matrix is dead once the commutator and damping are formed)
    char* b = wbase + wv * proj_wave_bytes(p.d2pad, p.rowbuf, p.colbuf, sizeof(T), PACK && !MFMA);
    w.X = (cpx<T>*)b;
    w.Tt = MFMA ? w.X : (cpx<T>*)(b + proj_wave_bytes(p.d2pad, p.rowbuf, p.colbuf, sizeof(T), false));
    w.R = w.X + p.d2pad;
    w.Cc = w.R + p.rowbuf;
    w.cl = w.Cc + p.colbuf;
    w.cr = w.cl + kMaxEnt;
    w.nb = (int32_t*)(w.cr + kMaxEnt);
    w.st = w.nb + kMaxEnt;
    w.gs = w.st + kMaxEnt;
  }
  const cpx_store<TI>* xin = (const cpx_store<TI>*)p.xin;
  cpx_store<TO>* xout = (cpx_store<TO>*)p.xout;
  const int col = p.col_override >= 0 ? p.col_override : stage_col(STAGE & 3);
  // neighbour rows by LDS DMA in the wide class only: it removes that class's register spills and
  // shortens the 24-site chain's step by 6%; the 16 x 16 MFMA and VALU classes ran 1-10% slower with it.
  // Hermitian DMA plans keep, per entry, the row's index in the instance (neighbour * d + site) in
  // place of the site.
  constexpr bool kDMA = PQA_ROWS_DMA && (NT == 2 || (PQA_ROWS_DMA_NARROW && MFMA && PACK)) && std::is_same<TI, T>::value;
  // packed Hermitian MFMA forms: with H and the stage input X Hermitian, X H = (H X)^H, so only
  // P = H X is formed on MFMA and the commutator is completed from the output tile
  constexpr bool kHalf = PQA_HALF_COMMUTATOR && MFMA && PACK && HERM;
  int it0, itstep, itend;
  item_range(p, it0, itend, itstep);
  for (int it = it0; it < itend; it += itstep) {
    const WorkItem wi = ONE ? one_item(p, it) : ro_load(p.items + it);
    const InstDev I = ONE ? p.inst0 : ro_load(p.insts + wi.inst);
    const int d = I.d, d2 = I.d2;
    // floor(idx / d) = (idx * dmag) >> 20 for idx < 2^11, d <= 32 (dmag*d - 2^20 < d < 2^20 / 2^11)
    const uint32_t dmag = ((1u << 20) + d - 1) / d;
    // LDS row stride of the own matrix, H and the neighbour rows: d+1 (odd), so that column reads of
    // the A operand (stride rs) and the rows of the row buffer fall on different banks
    const int rs = d + 1;
    const cpx_store<TI>* xi = xin + I.off;
    // Software pipeline over this wave's ADOs: the neighbour list (and, on the MFMA path, the own
    // matrix) of the next ADO are loaded while the current one is processed; they are issued after
    // the current ADO's neighbour loads, so waiting for those does not wait for the prefetch.
    const int aend = wi.ado0 + wi.nado;
    int nent_n = 0;
    int2 en_n = make_int2(0, 0);
    cpx_store<TI> xt_n[kPF];
    auto prefetch = [&](int an) {
      if constexpr (ONE) {
        // the list's address does not depend on the row pointers, which only give its length
        nent_n = ro_int(p.rowptr, I.rowptr_off + an + 1) - ro_int(p.rowptr, I.rowptr_off + an);
        en_n = lane < p.fix_stride ? p.ent_fix[(int64_t)an * p.fix_stride + lane] : make_int2(0, 0);
      } else {
        const int r0 = ro_int(p.rowptr, I.rowptr_off + an);
        nent_n = ro_int(p.rowptr, I.rowptr_off + an + 1) - r0;
        en_n = make_int2(0, 0);
        if (lane < nent_n) en_n = p.ent[I.ent_off + r0 + lane];
      }
      // own matrix: its first 64 kPF elements (all of it for d <= 16 NT)
#pragma unroll
      for (int q = 0; q < kPF; ++q) {
        const int e = lane + 64 * q;
        if (e < d2) xt_n[q] = xi[(int64_t)an * d2 + e];
      }
    };
    // the first ADO's list and matrix are requested before the instance header, so the two round
    // trips overlap (the prefetch touches registers only)
    if (wi.ado0 + wv < aend) prefetch(wi.ado0 + wv);
    __syncthreads();
    proj_header<T>(p, I, col, dmag, sH, sC, sSite, sS, PACK ? sIJ : nullptr);
    __syncthreads();
    // MFMA lane coordinates: lane = (gk, gi), gk = lane / 16, gi = lane % 16
    const int gi = lane & 15, gk = lane >> 4;
    for (int a = wi.ado0 + wv; a < aend; a += kThreads / 64) {
      const int64_t base = (int64_t)a * d2;
      // ---- neighbour list and own matrix (prefetched)
      const int nent = nent_n;
      const int2 en = en_n;
#pragma unroll
      for (int q = 0; q < kPF; ++q) {
        const int e = lane + 64 * q;
        if (e < d2) w.X[e + (int)(((uint32_t)e * dmag) >> 20)] = cvt<T>(mk(xt_n[q].x, xt_n[q].y));
      }
      for (int e = 64 * kPF + lane; e < d2; e += 64) w.X[e + (int)(((uint32_t)e * dmag) >> 20)] = ldc<T>(xi, base + e);
      // ---- neighbour list in site order, with coefficients
      const int k = en.y & 0xffff, nf = (en.y >> 16) & 0x7f;
      const bool plus = en.y < 0;
      const int my_site = lane < nent ? sSite[k] : 0x7fffffff;
      int gsv;
      const int pos = site_order(my_site, d, lane, (I.flags & F_SITESORT) != 0, gsv);
      if (lane <= d) w.gs[lane] = gsv;
      cpx<T> dp = mk<T>(0, 0);
      if (lane < nent) {
        const cpx<T>* C = sC + 5 * k;
        const T sq = sS[nf];
        w.nb[pos] = en.x;
        w.st[pos] = (kDMA && HERM) ? en.x * d + my_site : my_site;
        w.cl[pos] = (plus ? C[3] : C[1]) * sq;
        w.cr[pos] = (plus ? C[4] : C[2]) * sq;
        if (!plus) dp = C[0] * (T)nf;
      }
      const cpx<T> damp = wave_sum(dp);
      wave_sync();
      // ---- round trip 2: row s (pass 0) and, unless Hermitian, column s (pass 1) of every
      // neighbour into LDS; value idx = t*d + jj of a pass is element jj of entry t's row/column
      const int nval = nent * d;
      // the MFMA class's stage 1 keeps one more load per lane in flight: it has the registers to
      // spare (no FP64 y, one FP32 register read), and its stage time drops by about 3%
      constexpr int kCh = NT == 2 ? PQA_CHUNK_WIDE : (MFMA && STAGE == 1) ? PQA_CHUNK_S1 : kChunk;
      cpx<T> v[kCh];
      auto issue = [&](int pass, int q0) {
#pragma unroll
        for (int q = 0; q < kCh; ++q) {
          const int idx = q0 + lane + 64 * q;
          if (idx < nval) {
            const int t = (int)(((uint32_t)idx * dmag) >> 20), jj = idx - t * d, s2 = w.st[t];
            v[q] = ldc<T>(xi, (int64_t)w.nb[t] * d2 + (pass == 0 ? s2 * d + jj : jj * d + s2));
          }
        }
      };
      auto stash = [&](int pass, int q0) {
        cpx<T>* dst = pass == 0 ? w.R : w.Cc;
#pragma unroll
        for (int q = 0; q < kCh; ++q) {
          const int idx = q0 + lane + 64 * q;
          if (idx < nval) {
            const int t = (int)(((uint32_t)idx * dmag) >> 20);
            dst[idx + t] = v[q];  // row t starts at t * (d + 1)
          }
        }
      };
      // LDS-DMA form (kDMA): every 64-dword group of the row (and column) buffer is one
      // global_load_lds_dword (lane l fills dword l of the group; lanes on padding dwords read a harmless
      // in-bounds dword), the last partial group goes through one register per lane.  All rows of the
      // ADO are then in flight at once with no registers held, instead of kCh values per lane per round
      // trip.
      constexpr int kW = (int)(sizeof(cpx<T>) / 4);  // dwords per complex value
      float tail[2] = {0.f, 0.f};
      auto dma_rows = [&]() {
        const int m = kW * rs;  // dwords per row (or column) in LDS, padding included
        // floor(p / m) = (p * mmag) >> 24 for p < 2^24 / m (p < 64 * 4 * 33 here)
        const uint64_t mmag = ((1ull << 24) + m - 1) / m;
        const int total = nent * m;
        const int J = total >> 6;
        const float* src = (const float*)xi;
        if constexpr (HERM) {
          // rows only: lane position (t, k) in the buffer advanced by 64 dwords per group (wide class:
          // m >= 36, so at most two wraps; 16 x 16 class: m >= 18, up to four), the row start read once
          // per group from the per-entry row index
          float* dst = (float*)w.R;
          int t = lane / m, k = lane - (lane / m) * m;
          const int rw = kW * d;  // dwords of a row
#pragma unroll 4
          for (int j = 0; j < J; ++j) {
            const int64_t off = k < rw ? (int64_t)w.st[t] * rw + k : 0;
            __builtin_amdgcn_global_load_lds(src + off, (__attribute__((address_space(3))) void*)(dst + 64 * j), 4, 0, 0);
            k += 64;
            if constexpr (NT == 2) {
              if (k >= m) {
                k -= m;
                ++t;
              }
              if (k >= m) {
                k -= m;
                ++t;
              }
            } else {
              while (k >= m) {  // at most 64 / m + 1 <= 4 times (m = 2 (d + 1) >= 18)
                k -= m;
                ++t;
              }
            }
          }
          if (64 * J + lane < total) tail[0] = src[k < rw ? (int64_t)w.st[t] * rw + k : 0];
          return;
        }
        auto source = [&](int pass, int pdw) -> int64_t {
          const int t = (int)(((uint64_t)pdw * mmag) >> 24), k = pdw - t * m;
          const int jj = k / kW;
          if (jj >= d) return 0;
          const int s2 = w.st[t];
          const int64_t e = (int64_t)w.nb[t] * d2 + (pass == 0 ? s2 * d + jj : jj * d + s2);
          return e * kW + (k - jj * kW);
        };
#pragma unroll
        for (int pass = 0; pass < (HERM ? 1 : 2); ++pass) {
          float* dst = (float*)(pass == 0 ? w.R : w.Cc);
#pragma unroll 4
          for (int j = 0; j < J; ++j)
            __builtin_amdgcn_global_load_lds(src + source(pass, 64 * j + lane),
                                             (__attribute__((address_space(3))) void*)(dst + 64 * j), 4, 0, 0);
          const int pdw = 64 * J + lane;
          if (pdw < total) tail[pass] = src[source(pass, pdw)];
        }
      };
      auto dma_finish = [&]() {
        // the DMA writes are complete (vmcnt counts them); the asm's memory clobber keeps every LDS read
        // of the buffers below it
        asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
        const int total = nent * kW * rs;
        const int pdw = 64 * (total >> 6) + lane;
        if (pdw < total) {
          ((float*)w.R)[pdw] = tail[0];
          if (!HERM) ((float*)w.Cc)[pdw] = tail[1];
        }
      };
      if constexpr (kDMA) dma_rows();
      else issue(0, 0);
      cpx<T> acc[4 * NT * NT];  // MFMA layout: element (16 ti + 4 gk + r, 16 tj + gi) at (ti NT + tj) 4 + r
      Pre<kDelta> pr[NT == 1 ? 4 : kQP];
      PreM pm[NT == 1 ? 4 : kQP];
      // packed mode: the RK4 operands of this lane's upper elements (packed indices lane + 64q, at most
      // 64 kQP of them), read in packed order, either here or once the neighbour rows are in
      auto request_packed = [&]() {
#pragma unroll
        for (int q = 0; q < kQP; ++q) {
          const int pk = lane + 64 * q;
          if (pk < ((d * (d + 1)) >> 1)) {
            const int64_t g = I.poff + (int64_t)a * I.dp + pk;
            if constexpr (kPM) pm[q] = rk4_prefetch_m<STAGE>(p, g);
            else pr[q] = rk4_prefetch<STAGE, kDelta>(p, StateRef{g, false});
          }
        }
      };
      if (MFMA && PACK && !(kPM && STAGE == 0) && (PQA_PACK_HOIST == 1 || kDMA)) request_packed();
      if (MFMA && NT == 1 && !PACK && PQA_RK4_HOIST == 2) {
#pragma unroll
        for (int r = 0; r < 4; ++r)
          if (4 * gk + r < d && gi < d) pr[r] = rk4_prefetch<STAGE, kDelta>(p, state_ref<PACK>(I, a, 4 * gk + r, gi));
      }
      if (a + kThreads / 64 < aend) prefetch(a + kThreads / 64);
      // ---- commutator and damping (while the neighbour loads are in flight)
      if (MFMA) {
        floatx4 aR[NT][NT], aI[NT][NT];
#pragma unroll
        for (int ti = 0; ti < NT; ++ti)
#pragma unroll
          for (int tj = 0; tj < NT; ++tj) {
            aR[ti][tj] = floatx4{0.f, 0.f, 0.f, 0.f};
            aI[ti][tj] = floatx4{0.f, 0.f, 0.f, 0.f};
          }
        // H is read from LDS for every ADO rather than kept in registers across the item's ADOs
        // (an opaque zero offset stops the compiler from hoisting the loads out of the ADO loop)
        int hz = 0;
        asm volatile("" : "+v"(hz));
        const cpx<T>* sHa = sH + hz;
        // 2 x 2 tiles: one k-chunk at a time keeps 16 operands and 32 accumulators live
        constexpr int kUnrollKC = NT == 1 ? 4 : 1;
        // 2 x 2 tiles: k-chunks past d hold only zero padding and are skipped (6 of 8 at d = 24)
        const int nkc = NT == 1 ? 4 : (d + 3) >> 2;
        // real H (F_REALH): the four products with Im H are exact zeros and are skipped.  Packed
        // Hermitian plans form only P = H X (kHalf)
        auto kloop = [&](auto RH) {
          constexpr bool kRealH = decltype(RH)::value;
#pragma unroll kUnrollKC
          for (int kc = 0; kc < nkc; ++kc) {
            const int kk = 4 * kc + gk;
            float xar[NT], xai[NT], har[NT], hai[NT], xbr[NT], xbi[NT], hbr[NT], hbi[NT];
#pragma unroll
            for (int t = 0; t < NT; ++t) {
              const int g = 16 * t + gi;
              const bool ok = g < d && kk < d;
              const cpx<T> xb = ok ? w.X[kk * rs + g] : mk<T>(0, 0);  // B-layout X[kk][g]
              const cpx<T> xa = ok ? w.X[g * rs + kk] : mk<T>(0, 0);  // A-layout X[g][kk]
              const cpx<T> ha = ok ? sHa[g * rs + kk] : mk<T>(0, 0);  // A-layout H[g][kk]
              const cpx<T> hb = ok ? sHa[kk * rs + g] : mk<T>(0, 0);  // B-layout H[kk][g]
              // kHalf: P is formed from the Hermitian part of X, whose diagonal is real.  The stored
              // diagonal carries rounding-level imaginary parts, which H X - X H cancels exactly and
              // H X alone would turn into a trace drift of 2 Im(H_ii X_ii)
              xbr[t] = (float)xb.x, xbi[t] = (kHalf && kk == g) ? 0.f : (float)xb.y;
              xar[t] = (float)xa.x, xai[t] = (float)xa.y;
              har[t] = (float)ha.x, hai[t] = (float)ha.y, hbr[t] = (float)hb.x, hbi[t] = (float)hb.y;
            }
#pragma unroll
            for (int ti = 0; ti < NT; ++ti)
#pragma unroll
              for (int tj = 0; tj < NT; ++tj) {
                floatx4 r = aR[ti][tj], m = aI[ti][tj];
                // + H X: rows of tile ti from H, columns of tile tj from X
                r = __builtin_amdgcn_mfma_f32_16x16x4f32(har[ti], xbr[tj], r, 0, 0, 0);
                if constexpr (!kRealH) r = __builtin_amdgcn_mfma_f32_16x16x4f32(-hai[ti], xbi[tj], r, 0, 0, 0);
                m = __builtin_amdgcn_mfma_f32_16x16x4f32(har[ti], xbi[tj], m, 0, 0, 0);
                if constexpr (!kRealH) m = __builtin_amdgcn_mfma_f32_16x16x4f32(hai[ti], xbr[tj], m, 0, 0, 0);
                // - X H
                if constexpr (!kHalf) {
                  r = __builtin_amdgcn_mfma_f32_16x16x4f32(xar[ti], -hbr[tj], r, 0, 0, 0);
                  if constexpr (!kRealH) r = __builtin_amdgcn_mfma_f32_16x16x4f32(xai[ti], hbi[tj], r, 0, 0, 0);
                  if constexpr (!kRealH) m = __builtin_amdgcn_mfma_f32_16x16x4f32(xar[ti], -hbi[tj], m, 0, 0, 0);
                  m = __builtin_amdgcn_mfma_f32_16x16x4f32(xai[ti], -hbr[tj], m, 0, 0, 0);
                }
                aR[ti][tj] = r;
                aI[ti][tj] = m;
              }
          }
        };
        if (I.flags & F_REALH)
          kloop(std::true_type());
        else
          kloop(std::false_type());
#pragma unroll
        for (int ti = 0; ti < NT; ++ti)
#pragma unroll
          for (int tj = 0; tj < NT; ++tj)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
              const int i = 16 * ti + 4 * gk + r, j = 16 * tj + gi;
              cpx<T>& a2 = acc[(ti * NT + tj) * 4 + r];
              a2 = mul_mi(mk<T>((T)aR[ti][tj][r], (T)aI[ti][tj][r]));
              // kHalf: -i P - (damp / 2) X here; the reader adds the conjugate of the (j, i) entry
              // (damp is real in Hermitian plans and X Hermitian, so the two halves sum to damp X)
              if (i < d && j < d) cmsub(a2, kHalf ? mk<T>(damp.x * (T)0.5, damp.y * (T)0.5) : damp, w.X[i * rs + j]);
            }
      }
      if constexpr (kDMA) {
        dma_finish();
      } else {
        stash(0, 0);
        for (int pass = 0; pass < (HERM ? 1 : 2); ++pass)
          for (int q0 = pass == 0 ? 64 * kCh : 0; q0 < nval; q0 += 64 * kCh) {
            issue(pass, q0);
            stash(pass, q0);
          }
      }
      // RK4 operands of the lane's four elements, requested together (one round trip instead of
      // one per element); the gather below overlaps their latency
      if (MFMA && PACK && !(kPM && STAGE == 0)) {
        if (PQA_PACK_HOIST == 0 && !kDMA) request_packed();
      } else if (MFMA && NT == 1 && PQA_RK4_HOIST == 1) {
#pragma unroll
        for (int r = 0; r < 4; ++r)
          if (4 * gk + r < d && gi < d) pr[r] = rk4_prefetch<STAGE, kDelta>(p, state_ref<PACK>(I, a, 4 * gk + r, gi));
      }
      wave_sync();
      // ---- gather the neighbour terms from the row buffer, RK4 update
      const bool save_ado = STAGE == 3 && a == 0 && p.traj && p.save_slot >= 0;
      const int64_t tbase = p.save_slot * p.traj_stride + I.traj_off;
      if constexpr (PACK) {
        // Packed (Hermitian) mode: lane takes the upper elements of packed index lane + 64q; their
        // stage outputs and the conjugates at (j, i) go to the tile, which is then stored as full
        // matrices.  Neighbour terms of upper element (i, j): row term of site i and column term of
        // site j read as conjugated rows (the diagonal keeps the row value itself).
        const int np = (d * (d + 1)) >> 1;
        auto neighbours = [&](cpx<T>& o, int i, int j) {
          for (int t = w.gs[i]; t < w.gs[i + 1]; ++t) cmac(o, w.cl[t], w.R[t * rs + j]);
          for (int t = w.gs[j]; t < w.gs[j + 1]; ++t)
            cmac(o, w.cr[t], i != j ? conjg(w.R[t * rs + i]) : w.R[t * rs + i]);
        };
        auto output = [&](int i, int j, const double2& o) {
          w.Tt[i * rs + j] = mk<T>((T)o.x, (T)o.y);
          if (i != j) w.Tt[j * rs + i] = mk<T>((T)o.x, (T)(-o.y));
          if (save_ado) {
            p.traj[tbase + i * d + j] = o;
            if (i != j) p.traj[tbase + j * d + i] = make_double2(o.x, -o.y);
          }
        };
        if (MFMA) {
          // stages 0-2 of a packed mixed run: this lane's upper elements of the FP32 stage input (the
          // own matrix), kept in FP32 and read before the tile takes its place
          cpx<T> yo[kQP];
          if constexpr (kPM && STAGE <= 2) {
#pragma unroll
            for (int q = 0; q < kQP; ++q) {
              const int pk = lane + 64 * q;
              if (pk < np) {
                const int ij = sIJ[pk];
                yo[q] = w.X[(ij >> 8) * rs + (ij & 255)];
              }
            }
          }
          // commutator and damping of every element (MFMA lane layout) through the tile
#pragma unroll
          for (int ti = 0; ti < NT; ++ti)
#pragma unroll
            for (int tj = 0; tj < NT; ++tj)
#pragma unroll
              for (int r = 0; r < 4; ++r) {
                const int i = 16 * ti + 4 * gk + r, j = 16 * tj + gi;
                if (i < d && j < d) w.Tt[i * rs + j] = acc[(ti * NT + tj) * 4 + r];
              }
          wave_sync();
          cpx<T> au[kQP];
#pragma unroll
          for (int q = 0; q < kQP; ++q) {
            const int pk = lane + 64 * q;
            if (pk < np) {
              const int ij = sIJ[pk], i = ij >> 8, j = ij & 255;
              au[q] = w.Tt[i * rs + j];
              if constexpr (kHalf) {  // -i (P - P^H) - damp X = T(i, j) + conj(T(j, i))
                const cpx<T> tt = w.Tt[j * rs + i];
                au[q].x += tt.x;
                au[q].y -= tt.y;
              }
              neighbours(au[q], i, j);
            }
          }
          wave_sync();  // every tile read is done: the tile now takes the stage output
#pragma unroll
          for (int q = 0; q < kQP; ++q) {
            const int pk = lane + 64 * q;
            if (pk < np) {
              const int ij = sIJ[pk];
              const int64_t g = I.poff + (int64_t)a * I.dp + pk;
              if constexpr (kPM) {
                PreM v = pm[q];
                if (STAGE <= 2) v.y = y_from_input_m<STAGE>(yo[q], v);
                output(ij >> 8, ij & 255, rk4_upper_m<STAGE>(p, g, (double)au[q].x, (double)au[q].y, v));
              } else {
                output(ij >> 8, ij & 255, rk4_upper<STAGE, kDelta>(p, g, (double)au[q].x, (double)au[q].y, pr[q]));
              }
            }
          }
        } else {
          for (int p0 = 0; p0 < np; p0 += 256) {
#pragma unroll
            for (int q = 0; q < 4; ++q) {
              const int pk = p0 + lane + 64 * q;
              if (pk < np) {
                if constexpr (kPM) {
                  pm[q] = rk4_prefetch_m<STAGE>(p, I.poff + (int64_t)a * I.dp + pk);
                } else {
                  pr[q] = rk4_prefetch<STAGE, kDelta>(p, StateRef{I.poff + (int64_t)a * I.dp + pk, false});
                }
              }
            }
#pragma unroll
            for (int q = 0; q < 4; ++q) {
              const int pk = p0 + lane + 64 * q;
              if (pk < np) {
                const int ij = sIJ[pk], i = ij >> 8, j = ij & 255;
                cpx<T> c = mk<T>(0, 0);
                // real H (H(t) too): the products with Im H are zeros, two FMAs instead of four (not in
                // stage 0, whose form the second loop would take from 5 to 4 waves per SIMD)
                if (STAGE >= 1 && (I.flags & F_REALH)) {
                  for (int kk = 0; kk < d; ++kk) {
                    const T hik = sH[i * rs + kk].x, hkj = sH[kk * rs + j].x;
                    const cpx<T> xkj = w.X[kk * rs + j], xik = w.X[i * rs + kk];
                    c.x += hik * xkj.x - xik.x * hkj;
                    c.y += hik * xkj.y - xik.y * hkj;
                  }
                } else {
                  for (int kk = 0; kk < d; ++kk) {
                    cmac(c, sH[i * rs + kk], w.X[kk * rs + j]);
                    cmsub(c, w.X[i * rs + kk], sH[kk * rs + j]);
                  }
                }
                cpx<T> o = mul_mi(c);
                cmsub(o, damp, w.X[i * rs + j]);
                neighbours(o, i, j);
                const int64_t g = I.poff + (int64_t)a * I.dp + pk;
                if constexpr (kPM) {
                  if (STAGE <= 2) pm[q].y = y_from_input_m<STAGE>(w.X[i * rs + j], pm[q]);
                  output(i, j, rk4_upper_m<STAGE>(p, g, (double)o.x, (double)o.y, pm[q]));
                } else {
                  output(i, j, rk4_upper<STAGE, kDelta>(p, g, (double)o.x, (double)o.y, pr[q]));
                }
              }
            }
          }
        }
        wave_sync();
        for (int e = lane; e < d2; e += 64) {
          const int i = (int)(((uint32_t)e * dmag) >> 20), j = e - i * d;
          const cpx<T> u = w.Tt[i * rs + j];
          st(xout + I.off + base + e, mk<TO>((TO)u.x, (TO)u.y));
        }
      } else if (MFMA) {
        // tile by tile (one for d <= 16): lane (gk, gi) holds elements (16 ti + 4 gk + r, 16 tj + gi)
#pragma unroll
        for (int ti = 0; ti < NT; ++ti)
#pragma unroll
          for (int tj = 0; tj < NT; ++tj) {
            const int j = 16 * tj + gi, i0 = 16 * ti + 4 * gk;
            cpx<T>* at = acc + (ti * NT + tj) * 4;
            if (j < d) {
              if (NT == 2) {
#pragma unroll
                for (int r = 0; r < 4; ++r)
                  if (i0 + r < d) pr[r] = rk4_prefetch<STAGE, kDelta>(p, state_ref<PACK>(I, a, i0 + r, j));
              }
              // column term, site j: element (i0+r, j) <- cR_t * conj(R_t[i0+r])
              for (int t = w.gs[j]; t < w.gs[j + 1]; ++t) {
                const cpx<T> c = w.cr[t];
                const cpx<T>* row = (HERM ? w.R : w.Cc) + t * rs + i0;
#pragma unroll
                for (int r = 0; r < 4; ++r)
                  // the diagonal element keeps X'(s, s) itself, so that (cL + cR) X'(s, s) is formed exactly
                  // as without the Hermitian shortcut (cL = -cR for the n + e_k terms: no trace drift)
                  if (i0 + r < d) cmac(at[r], c, (HERM && i0 + r != j) ? conjg(row[r]) : row[r]);
              }
#pragma unroll
              for (int r = 0; r < 4; ++r) {
                const int i = i0 + r;
                // row term, site i: element (i, j) <- cL_t * R_t[j]
                if (i < d)
                  for (int t = w.gs[i]; t < w.gs[i + 1]; ++t) cmac(at[r], w.cl[t], w.R[t * rs + j]);
              }
#pragma unroll
              for (int r = 0; r < 4; ++r) {
                const int i = i0 + r;
                if (i < d) {
                  const int e = i * d + j;
                  const StateRef sr = state_ref<PACK>(I, a, i, j);
                  if (NT == 1 && !PQA_RK4_HOIST) pr[r] = rk4_prefetch<STAGE, kDelta>(p, sr);
                  rk4_store<TO, STAGE, PACK, kDelta>(p, I.off + base + e, I.off + base + j * d + i, sr, (double)at[r].x,
                                             (double)at[r].y, pr[r], xout, save_ado, tbase + e, tbase + j * d + i);
                }
              }
            }
          }
      } else {
        // four elements per lane at a time: their RK4 operands are requested together
        for (int e0 = 0; e0 < d2; e0 += 256) {
#pragma unroll
          for (int q = 0; q < 4; ++q) {
            const int e = e0 + lane + 64 * q;
            const int i = (int)(((uint32_t)e * dmag) >> 20);
            if (e < d2) pr[q] = rk4_prefetch<STAGE, kDelta>(p, state_ref<PACK>(I, a, i, e - i * d));
          }
#pragma unroll
          for (int q = 0; q < 4; ++q) {
            const int e = e0 + lane + 64 * q;
            if (e < d2) {
              const int i = (int)(((uint32_t)e * dmag) >> 20), j = e - i * d;
              cpx<T> c = mk<T>(0, 0);
              for (int kk = 0; kk < d; ++kk) {
                cmac(c, sH[i * rs + kk], w.X[kk * rs + j]);
                cmsub(c, w.X[i * rs + kk], sH[kk * rs + j]);
              }
              cpx<T> o = mul_mi(c);
              cmsub(o, damp, w.X[i * rs + j]);
              for (int t = w.gs[i]; t < w.gs[i + 1]; ++t) cmac(o, w.cl[t], w.R[t * rs + j]);
              for (int t = w.gs[j]; t < w.gs[j + 1]; ++t)
                cmac(o, w.cr[t], HERM ? (i != j ? conjg(w.R[t * rs + i]) : w.R[t * rs + i]) : w.Cc[t * rs + i]);
              rk4_store<TO, STAGE, PACK, kDelta>(p, I.off + base + e, I.off + base + j * d + i, state_ref<PACK>(I, a, i, j),
                                         (double)o.x, (double)o.y, pr[q], xout, save_ado, tbase + e, tbase + j * d + i);
            }
          }
        }
      }
      wave_sync();
    }
  }
}

// ------------------------------------------------------------------ block-cooperative wide kernel
// The wide class (17 <= d <= 32, site projectors) in packed Hermitian mixed runs, one ADO per block at
// a time: the four waves share its neighbour-row, own-matrix and RK4-operand requests (one round trip
// for all of them, a few values per lane), and each wave forms one 16 x 16 tile of the 2 x 2 MFMA
// commutator.  Wave 0 builds the site-ordered neighbour list (as the projector kernel does per wave).
#ifndef PQA_WIDE_ROWS
#define PQA_WIDE_ROWS 6  // neighbour-row values per lane requested together (a further pass beyond)
#endif
#ifndef PQA_WIDE_BLOCK_WAVES
#define PQA_WIDE_BLOCK_WAVES 3  // minimum waves per SIMD the block kernel is compiled for (no spills)
#endif

// Per-block ADO buffers of the wide block kernel: own matrix / output tile (d2pad), row buffer, and two
// copies (this ADO's and the next one's) of the neighbour list with its coefficients, site-group starts
// and damping.
constexpr int kWideList = 2 * 8 * kMaxEnt + 2 * 4 * kMaxEnt + 4 * 36 + 16;  // bytes of one list copy
__host__ __device__ inline size_t wide_ado_bytes(int d2pad, int rowbuf) {
  size_t b = 8 * ((size_t)d2pad + rowbuf) + 2 * (size_t)kWideList;
  return (b + 15) & ~size_t(15);
}

// One neighbour-list copy of the wide block kernel.
struct WideList {
  cpx<float>* cl;   // per entry: coefficient of the row term
  cpx<float>* cr;   // per entry: coefficient of the column term
  int32_t* nb;      // per entry: neighbour ADO
  int32_t* ste;     // per entry: site
  int32_t* gs;      // site-group starts (d + 1 <= 33); gs[35]: entries of the ADO
  cpx<float>* damp; // sum_k n_k gamma_k
};

__device__ __forceinline__ WideList wide_list(char* b) {
  WideList w;
  w.cl = (cpx<float>*)b;
  w.cr = w.cl + kMaxEnt;
  w.nb = (int32_t*)(w.cr + kMaxEnt);
  w.ste = w.nb + kMaxEnt;
  w.gs = w.ste + kMaxEnt;
  w.damp = (cpx<float>*)(w.gs + 36);
  return w;
}

// Wave 0 of the block: the neighbour list of one ADO (entry en in lane < nent) in site order, with its
// coefficients and damping, into list copy L.
__device__ __forceinline__ void wide_build_list(const WideList& L, int2 en, int nent, int lane, int d, bool sorted,
                                                const cpx<float>* sC, const float* sS, const int32_t* sSite) {
  const int k = en.y & 0xffff, nf = (en.y >> 16) & 0x7f;
  const bool plus = en.y < 0;
  const int my_site = lane < nent ? sSite[k] : 0x7fffffff;
  int gsv;
  const int pos = site_order(my_site, d, lane, sorted, gsv);
  if (lane <= d) L.gs[lane] = gsv;
  cpx<float> dp = mk<float>(0, 0);
  if (lane < nent) {
    const cpx<float>* C = sC + 5 * k;
    const float sq = sS[nf];
    L.nb[pos] = en.x;
    L.ste[pos] = my_site;
    L.cl[pos] = (plus ? C[3] : C[1]) * sq;
    L.cr[pos] = (plus ? C[4] : C[2]) * sq;
    if (!plus) dp = C[0] * (float)nf;
  }
  const cpx<float> damp = wave_sum(dp);
  if (lane == 0) {
    L.damp[0] = damp;
    L.gs[35] = nent;
  }
}

// The ADO loop is software-pipelined: while ADO a computes, wave 0 builds the neighbour list of a + 1
// in the other list copy and every thread holds a + 1's own matrix in registers, so an ADO waits only
// for its neighbour rows and RK4 operands, requested together.
template <int STAGE>
__global__ __launch_bounds__(kThreads, PQA_WIDE_BLOCK_WAVES) void heom_wide_kernel(StageParams p) {
  using T = float;
  extern __shared__ __align__(16) char smem[];
  const int tid = threadIdx.x, lane = tid & 63;
  const int wv = __builtin_amdgcn_readfirstlane(tid >> 6);
  cpx<T>* sH = (cpx<T>*)smem;
  cpx<T>* sC = sH + p.d2pad;
  T* sS = (T*)(sC + 5 * p.Kmax);
  int32_t* sSite = (int32_t*)(sS + p.Lmax + 2);
  int32_t* sIJ = sSite + p.Kmax;
  cpx<T>* X = (cpx<T>*)(smem + proj_block_bytes(p.d2pad, p.Kmax, p.Lmax, sizeof(T), true));
  cpx<T>* R = X + p.d2pad;  // neighbour rows in site order, row t at t (d + 1)
  char* lists = (char*)(R + p.rowbuf);  // two list copies, kWideList bytes each
  const cpx_store<T>* xin = (const cpx_store<T>*)p.xin;
  cpx_store<T>* xout = (cpx_store<T>*)p.xout;
  const int col = stage_col(STAGE);
  const int ti = wv >> 1, tj = wv & 1, gi = lane & 15, gk = lane >> 4;
  constexpr int kOwn = (32 * 32 + kThreads - 1) / kThreads;  // own-matrix elements per thread
  int it0, itstep, itend;
  item_range(p, it0, itend, itstep);
  for (int it = it0; it < itend; it += itstep) {
    const WorkItem wi = ro_load(p.items + it);
    const InstDev I = ro_load(p.insts + wi.inst);
    const int d = I.d, d2 = I.d2, rs = d + 1, np = (d * (d + 1)) >> 1;
    const uint32_t dmag = ((1u << 20) + d - 1) / d;
    const bool sorted = (I.flags & F_SITESORT) != 0;
    const int a0 = wi.ado0, aend = wi.ado0 + wi.nado;
    const cpx_store<T>* xi = xin + I.off;
    __syncthreads();  // the previous item's buffers are free
    proj_header<T>(p, I, col, dmag, sH, sC, sSite, sS, sIJ);
    // prologue: own matrix of a0 in registers; neighbour list of a0 in list copy 0
    cpx_store<T> xo[kOwn];
#pragma unroll
    for (int q = 0; q < kOwn; ++q) {
      const int e = tid + kThreads * q;
      if (e < d2) xo[q] = xi[(int64_t)a0 * d2 + e];
    }
    if (wv == 0) {
      const int r0 = ro_int(p.rowptr, I.rowptr_off + a0);
      const int nent = ro_int(p.rowptr, I.rowptr_off + a0 + 1) - r0;
      int2 en = make_int2(0, 0);
      if (lane < nent) en = p.ent[I.ent_off + r0 + lane];
      __syncthreads();  // header in place (sSite, sC, sS)
      wide_build_list(wide_list(lists), en, nent, lane, d, sorted, sC, sS, sSite);
    } else {
      __syncthreads();
    }
    for (int a = a0; a < aend; ++a) {
      const int par = (a - a0) & 1;
      const WideList L = wide_list(lists + par * kWideList);
      const WideList Ln = wide_list(lists + (par ^ 1) * kWideList);
      const int64_t base = (int64_t)a * d2;
      const bool has_next = a + 1 < aend;
      __syncthreads();  // list(a) is built; the previous ADO's tile and rows are no longer read
      // ---- own matrix of a from registers into LDS, then a + 1's requested; wave 0 requests list(a + 1)
#pragma unroll
      for (int q = 0; q < kOwn; ++q) {
        const int e = tid + kThreads * q;
        if (e < d2) X[e + (int)(((uint32_t)e * dmag) >> 20)] = cvt<T>(mk(xo[q].x, xo[q].y));
      }
      int2 en_n = make_int2(0, 0);
      int nent_n = 0;
      if (wv == 0 && has_next) {
        const int r0 = ro_int(p.rowptr, I.rowptr_off + a + 1);
        nent_n = ro_int(p.rowptr, I.rowptr_off + a + 2) - r0;
        if (lane < nent_n) en_n = p.ent[I.ent_off + r0 + lane];
      }
      const int nval = L.gs[35] * d;
      // ---- neighbour rows and this thread's RK4 operands (one round trip)
      cpx<T> v[PQA_WIDE_ROWS];
#pragma unroll
      for (int q = 0; q < PQA_WIDE_ROWS; ++q) {
        const int idx = tid + kThreads * q;
        if (idx < nval) {
          const int t = (int)(((uint32_t)idx * dmag) >> 20), jj = idx - t * d;
          v[q] = ldc<T>(xi, (int64_t)L.nb[t] * d2 + L.ste[t] * d + jj);
        }
      }
      PreM pm[3];
      if (STAGE != 0) {
#pragma unroll
        for (int q = 0; q < 3; ++q) {
          const int pk = tid + kThreads * q;
          if (pk < np) pm[q] = rk4_prefetch_m<STAGE>(p, I.poff + (int64_t)a * I.dp + pk);
        }
      }
      if (has_next) {
#pragma unroll
        for (int q = 0; q < kOwn; ++q) {
          const int e = tid + kThreads * q;
          if (e < d2) xo[q] = xi[base + d2 + e];
        }
      }
      __syncthreads();  // own matrix of a in place
      // ---- commutator -i[H, X] on MFMA: wave (ti, tj) forms tile (ti, tj) of the 2 x 2 grid
      cpx<T> acc[4];
      {
        floatx4 aR = {0.f, 0.f, 0.f, 0.f}, aI = {0.f, 0.f, 0.f, 0.f};
        const int ga = 16 * ti + gi, gb = 16 * tj + gi;
        const int nkc = (d + 3) >> 2;  // k-chunks past d hold only zero padding
#pragma unroll 1
        for (int kc = 0; kc < nkc; ++kc) {
          const int kk = 4 * kc + gk;
          const bool oka = ga < d && kk < d, okb = gb < d && kk < d;
          const cpx<T> xa = oka ? X[ga * rs + kk] : mk<T>(0, 0);   // A-layout X[ga][kk]
          const cpx<T> ha = oka ? sH[ga * rs + kk] : mk<T>(0, 0);  // A-layout H[ga][kk]
          const cpx<T> xb = okb ? X[kk * rs + gb] : mk<T>(0, 0);   // B-layout X[kk][gb]
          const cpx<T> hb = okb ? sH[kk * rs + gb] : mk<T>(0, 0);  // B-layout H[kk][gb]
          aR = __builtin_amdgcn_mfma_f32_16x16x4f32(ha.x, xb.x, aR, 0, 0, 0);
          aR = __builtin_amdgcn_mfma_f32_16x16x4f32(-ha.y, xb.y, aR, 0, 0, 0);
          aI = __builtin_amdgcn_mfma_f32_16x16x4f32(ha.x, xb.y, aI, 0, 0, 0);
          aI = __builtin_amdgcn_mfma_f32_16x16x4f32(ha.y, xb.x, aI, 0, 0, 0);
          aR = __builtin_amdgcn_mfma_f32_16x16x4f32(xa.x, -hb.x, aR, 0, 0, 0);
          aR = __builtin_amdgcn_mfma_f32_16x16x4f32(xa.y, hb.y, aR, 0, 0, 0);
          aI = __builtin_amdgcn_mfma_f32_16x16x4f32(xa.x, -hb.y, aI, 0, 0, 0);
          aI = __builtin_amdgcn_mfma_f32_16x16x4f32(xa.y, -hb.x, aI, 0, 0, 0);
        }
        const cpx<T> damp = L.damp[0];
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int i = 16 * ti + 4 * gk + r, j = 16 * tj + gi;
          acc[r] = mul_mi(mk<T>(aR[r], aI[r]));
          if (i < d && j < d) cmsub(acc[r], damp, X[i * rs + j]);
        }
      }
      // ---- wave 0: list of a + 1 into the other copy (nobody reads that copy during this ADO)
      if (wv == 0 && has_next) wide_build_list(Ln, en_n, nent_n, lane, d, sorted, sC, sS, sSite);
      // ---- rows into the row buffer (a further pass for lists longer than the registers hold)
#pragma unroll
      for (int q = 0; q < PQA_WIDE_ROWS; ++q) {
        const int idx = tid + kThreads * q;
        if (idx < nval) R[idx + (int)(((uint32_t)idx * dmag) >> 20)] = v[q];
      }
      for (int idx = tid + kThreads * PQA_WIDE_ROWS; idx < nval; idx += kThreads) {
        const int t = (int)(((uint32_t)idx * dmag) >> 20), jj = idx - t * d;
        R[idx + t] = ldc<T>(xi, (int64_t)L.nb[t] * d2 + L.ste[t] * d + jj);
      }
      // y of this thread's upper elements from the FP32 stage input (stages 0-2), before the tile
      // takes the own matrix's place
      cpx<T> yo[3];
      if (STAGE <= 2) {
#pragma unroll
        for (int q = 0; q < 3; ++q) {
          const int pk = tid + kThreads * q;
          if (pk < np) {
            const int ij = sIJ[pk];
            yo[q] = X[(ij >> 8) * rs + (ij & 255)];
          }
        }
      }
      __syncthreads();  // every read of the own matrix is done; the row buffer is complete
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int i = 16 * ti + 4 * gk + r, j = 16 * tj + gi;
        if (i < d && j < d) X[i * rs + j] = acc[r];
      }
      __syncthreads();
      // ---- neighbour terms of the upper elements: row term of site i, column term of site j read as
      // conjugated rows (the diagonal keeps the row value itself)
      cpx<T> au[3];
#pragma unroll
      for (int q = 0; q < 3; ++q) {
        const int pk = tid + kThreads * q;
        if (pk < np) {
          const int ij = sIJ[pk], i = ij >> 8, j = ij & 255;
          cpx<T> o = X[i * rs + j];
          for (int t = L.gs[i]; t < L.gs[i + 1]; ++t) cmac(o, L.cl[t], R[t * rs + j]);
          for (int t = L.gs[j]; t < L.gs[j + 1]; ++t) cmac(o, L.cr[t], i != j ? conjg(R[t * rs + i]) : R[t * rs + i]);
          au[q] = o;
        }
      }
      __syncthreads();  // every tile read is done: the tile now takes the stage output
      const bool save_ado = STAGE == 3 && a == 0 && p.traj && p.save_slot >= 0;
      const int64_t tbase = p.save_slot * p.traj_stride + I.traj_off;
#pragma unroll
      for (int q = 0; q < 3; ++q) {
        const int pk = tid + kThreads * q;
        if (pk < np) {
          const int ij = sIJ[pk], i = ij >> 8, j = ij & 255;
          PreM w = pm[q];
          if (STAGE <= 2) w.y = y_from_input_m<STAGE>(yo[q], w);
          const double2 o = rk4_upper_m<STAGE>(p, I.poff + (int64_t)a * I.dp + pk, (double)au[q].x, (double)au[q].y, w);
          X[i * rs + j] = mk<T>((T)o.x, (T)o.y);
          if (i != j) X[j * rs + i] = mk<T>((T)o.x, (T)(-o.y));
          if (save_ado) {
            p.traj[tbase + i * d + j] = o;
            if (i != j) p.traj[tbase + j * d + i] = make_double2(o.x, -o.y);
          }
        }
      }
      __syncthreads();
      // ---- the full-matrix stage output (the next stage's input; after stage 3 the next step's)
      for (int e = tid; e < d2; e += kThreads) {
        const cpx<T> u = X[e + (int)(((uint32_t)e * dmag) >> 20)];
        st(xout + I.off + base + e, mk<T>(u.x, u.y));
      }
    }
  }
}

// ------------------------------------------------------------------ pipelined site-projector kernel
// The dominant class (PQA_MFMA_DMIN <= d <= 16, Hermitian hierarchy, FP32 right-hand side) in a two-deep
// software pipeline: while ADO a is gathered and written, the neighbour rows of ADO a+1 are in
// flight and the neighbour list and matrix of ADO a+2 are being fetched.  Per-wave LDS buffers
// (own matrix, neighbour list, row buffer) are doubled.  Loads are issued in the order
//   rows(a+1) ... y/ynew(a) ... rows(a+2) ... list/matrix(a+3)
// relative to where they are waited for, so the in-order vector-memory counter never makes a wait
// for one ADO's data wait for a later ADO's.
#ifndef PQA_PIPE_WAVES
#define PQA_PIPE_WAVES 2
#endif

__host__ __device__ inline size_t pipe_wave_bytes(int xsize, int rowbuf) {
  size_t b = 2 * (8 * ((size_t)xsize + rowbuf) + 2 * 8 * kMaxEnt + 2 * 4 * kMaxEnt + 4 * 36);
  return (b + 15) & ~size_t(15);
}

template <int STAGE>
__global__ __launch_bounds__(kThreads, PQA_PIPE_WAVES) void heom_pipe_kernel(StageParams p) {
  using T = float;
  using TI = typename std::conditional<STAGE == 0, double, float>::type;
  extern __shared__ __align__(16) char smem[];
  const int tid = threadIdx.x, lane = tid & 63;
  const int wv = __builtin_amdgcn_readfirstlane(tid >> 6);  // wave-uniform: ADO indices stay scalar
  cpx<T>* sH = (cpx<T>*)smem;
  cpx<T>* sC = sH + p.d2pad;
  T* sS = (T*)(sC + 5 * p.Kmax);
  int32_t* sSite = (int32_t*)(sS + p.Lmax + 2);
  char* wb = smem + proj_block_bytes(p.d2pad, p.Kmax, p.Lmax, sizeof(T), false) + wv * pipe_wave_bytes(p.d2pad, p.rowbuf);
  // two buffers of everything per wave
  cpx<T>* Xb[2];
  cpx<T>* Rb[2];
  cpx<T>* clb[2];
  cpx<T>* crb[2];
  int32_t* nbb[2];
  int32_t* stb[2];
  int32_t* gsb[2];
  {
    char* q = wb;
    for (int u = 0; u < 2; ++u) {
      Xb[u] = (cpx<T>*)q;
      q += 8 * (size_t)p.d2pad;
      Rb[u] = (cpx<T>*)q;
      q += 8 * (size_t)p.rowbuf;
      clb[u] = (cpx<T>*)q;
      q += 8 * kMaxEnt;
      crb[u] = (cpx<T>*)q;
      q += 8 * kMaxEnt;
      nbb[u] = (int32_t*)q;
      q += 4 * kMaxEnt;
      stb[u] = (int32_t*)q;
      q += 4 * kMaxEnt;
      gsb[u] = (int32_t*)q;
      q += 4 * 36;
    }
  }
  const cpx_store<TI>* xin = (const cpx_store<TI>*)p.xin;
  cpx_store<T>* xout = (cpx_store<T>*)p.xout;
  const int col = p.col_override >= 0 ? p.col_override : stage_col(STAGE & 3);
  int it0, itstep, itend;
  item_range(p, it0, itend, itstep);
  for (int it = it0; it < itend; it += itstep) {
    const WorkItem wi = ro_load(p.items + it);
    const InstDev I = ro_load(p.insts + wi.inst);
    const int d = I.d, d2 = I.d2, rs = d + 1;
    const uint32_t dmag = ((1u << 20) + d - 1) / d;
    __syncthreads();
    proj_header<T>(p, I, col, dmag, sH, sC, sSite, sS, nullptr);
    __syncthreads();
    const int gi = lane & 15, gk = lane >> 4;
    const cpx_store<TI>* xi = xin + I.off;
    const int aend = wi.ado0 + wi.nado, astep = kThreads / 64;
    const int afirst = wi.ado0 + wv;
    if (afirst >= aend) continue;
    // ---- pipeline state in registers
    int nent_n = 0;
    int2 en_n = make_int2(0, 0);
    cpx_store<TI> xt_n[4];
    cpx_store<TI> v[kChunk];
    int nval_n = 0;  // number of row values in flight for the next ADO
    auto fetch_list = [&](int an) {
      const int r0 = ro_int(p.rowptr, I.rowptr_off + an);
      nent_n = ro_int(p.rowptr, I.rowptr_off + an + 1) - r0;
      en_n = make_int2(0, 0);
      if (lane < nent_n) en_n = p.ent[I.ent_off + r0 + lane];
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int e = lane + 64 * q;
        if (e < d2) xt_n[q] = xi[(int64_t)an * d2 + e];
      }
    };
    // sort the fetched list into buffer u, copy the fetched matrix into X[u]; returns damping
    auto prep = [&](int u) -> cpx<T> {
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int e = lane + 64 * q;
        if (e < d2) Xb[u][e + (int)(((uint32_t)e * dmag) >> 20)] = cvt<T>(mk(xt_n[q].x, xt_n[q].y));
      }
      const int nent = nent_n;
      const int k = en_n.y & 0xffff, nf = (en_n.y >> 16) & 0x7f;
      const bool plus = en_n.y < 0;
      const int my_site = lane < nent ? sSite[k] : 0x7fffffff;
      int gsv;
      const int pos = site_order(my_site, d, lane, (I.flags & F_SITESORT) != 0, gsv);
      if (lane <= d) gsb[u][lane] = gsv;
      cpx<T> dp = mk<T>(0, 0);
      if (lane < nent) {
        const cpx<T>* C = sC + 5 * k;
        const T sq = sS[nf];
        nbb[u][pos] = en_n.x;
        stb[u][pos] = my_site;
        clb[u][pos] = (plus ? C[3] : C[1]) * sq;
        crb[u][pos] = (plus ? C[4] : C[2]) * sq;
        if (!plus) dp = C[0] * (T)nf;
      }
      nval_n = nent * d;
      return wave_sum(dp);
    };
    auto issue_rows = [&](int u) {
#pragma unroll
      for (int q = 0; q < kChunk; ++q) {
        const int idx = lane + 64 * q;
        if (idx < nval_n) {
          const int t = (int)(((uint32_t)idx * dmag) >> 20), jj = idx - t * d;
          v[q] = xi[(int64_t)nbb[u][t] * d2 + stb[u][t] * d + jj];
        }
      }
    };
    auto stash_rows = [&](int u, int nval) {
#pragma unroll
      for (int q = 0; q < kChunk; ++q) {
        const int idx = lane + 64 * q;
        if (idx < nval) {
          const int t = (int)(((uint32_t)idx * dmag) >> 20);
          Rb[u][idx + t] = cvt<T>(mk(v[q].x, v[q].y));
        }
      }
      // values beyond the first chunk (more than 64*kChunk) are fetched and stored directly
      for (int q0 = 64 * kChunk; q0 < nval; q0 += 64) {
        const int idx = q0 + lane;
        if (idx < nval) {
          const int t = (int)(((uint32_t)idx * dmag) >> 20), jj = idx - t * d;
          const cpx_store<TI> w2 = xi[(int64_t)nbb[u][t] * d2 + stb[u][t] * d + jj];
          Rb[u][idx + t] = cvt<T>(mk(w2.x, w2.y));
        }
      }
    };
    // ---- prologue: ADO afirst
    fetch_list(afirst);
    cpx<T> damp = prep(0);
    int nval = nval_n;
    wave_sync();
    issue_rows(0);
    if (afirst + astep < aend) fetch_list(afirst + astep);
    int u = 0;
    for (int a = afirst; a < aend; a += astep, u ^= 1) {
      const int64_t base = (int64_t)a * d2;
      // commutator of ADO a (its matrix is in X[u])
      floatx4 aR = {0.f, 0.f, 0.f, 0.f}, aI = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int kc = 0; kc < 4; ++kc) {
        const int kk = 4 * kc + gk;
        const bool ok = gi < d && kk < d;
        const cpx<T> xb = ok ? Xb[u][kk * rs + gi] : mk<T>(0, 0);
        const cpx<T> xa = ok ? Xb[u][gi * rs + kk] : mk<T>(0, 0);
        const cpx<T> ha = ok ? sH[gi * rs + kk] : mk<T>(0, 0);
        const cpx<T> hb = ok ? sH[kk * rs + gi] : mk<T>(0, 0);
        aR = __builtin_amdgcn_mfma_f32_16x16x4f32(ha.x, xb.x, aR, 0, 0, 0);
        aR = __builtin_amdgcn_mfma_f32_16x16x4f32(-ha.y, xb.y, aR, 0, 0, 0);
        aI = __builtin_amdgcn_mfma_f32_16x16x4f32(ha.x, xb.y, aI, 0, 0, 0);
        aI = __builtin_amdgcn_mfma_f32_16x16x4f32(ha.y, xb.x, aI, 0, 0, 0);
        aR = __builtin_amdgcn_mfma_f32_16x16x4f32(xa.x, -hb.x, aR, 0, 0, 0);
        aR = __builtin_amdgcn_mfma_f32_16x16x4f32(xa.y, hb.y, aR, 0, 0, 0);
        aI = __builtin_amdgcn_mfma_f32_16x16x4f32(xa.x, -hb.y, aI, 0, 0, 0);
        aI = __builtin_amdgcn_mfma_f32_16x16x4f32(xa.y, -hb.x, aI, 0, 0, 0);
      }
      cpx<T> acc[4];
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int i = 4 * gk + r;
        acc[r] = mul_mi(mk<T>(aR[r], aI[r]));
        if (i < d && gi < d) cmsub(acc[r], damp, Xb[u][i * rs + gi]);
      }
      // rows of ADO a arrive; then its RK4 operands are requested
      stash_rows(u, nval);
      Pre<true> pre[4];
#pragma unroll
      for (int r = 0; r < 4; ++r)
        if (4 * gk + r < d && gi < d) pre[r] = rk4_prefetch<STAGE, true>(p, state_ref<false>(I, a, 4 * gk + r, gi));
      wave_sync();
      // next ADO: sort its list, request its rows, fetch the list after it
      cpx<T> damp_n = mk<T>(0, 0);
      const bool has_next = a + astep < aend;
      if (has_next) {
        damp_n = prep(u ^ 1);
        wave_sync();
        issue_rows(u ^ 1);
        if (a + 2 * astep < aend) fetch_list(a + 2 * astep);
      }
      // gather ADO a's neighbour terms from its row buffer and write its update
      const bool save_ado = STAGE == 3 && a == 0 && p.traj && p.save_slot >= 0;
      const int64_t tbase = p.save_slot * p.traj_stride + I.traj_off;
      if (gi < d) {
        for (int t = gsb[u][gi]; t < gsb[u][gi + 1]; ++t) {
          const cpx<T> c = crb[u][t];
          const cpx<T>* row = Rb[u] + t * rs + 4 * gk;
#pragma unroll
          for (int r = 0; r < 4; ++r)
            if (4 * gk + r < d) cmac(acc[r], c, 4 * gk + r != gi ? conjg(row[r]) : row[r]);
        }
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int i = 4 * gk + r;
          if (i < d) {
            for (int t = gsb[u][i]; t < gsb[u][i + 1]; ++t) cmac(acc[r], clb[u][t], Rb[u][t * rs + gi]);
            const int e = i * d + gi;
            rk4_store<T, STAGE, false, true>(p, I.off + base + e, I.off + base + gi * d + i, state_ref<false>(I, a, i, gi),
                                (double)acc[r].x, (double)acc[r].y, pre[r], xout, save_ado, tbase + e, tbase + gi * d + i);
          }
        }
      }
      damp = damp_n;
      nval = nval_n;
      wave_sync();
    }
  }
}

// ------------------------------------------------------------------ LDS-resident kernel
// One workgroup owns one small instance for the whole run: y, ynew and both
// stage buffers live in LDS, so HBM is touched only to load the initial state,
// write trajectory samples and store the final state.
template <typename T, bool PACK>
__global__ __launch_bounds__(kThreads) void heom_resident_kernel(StageParams p) {
  extern __shared__ __align__(16) char smem[];
  const Lds<T> l = carve<T>(smem, p);
  const InstDev I = p.insts[p.resident[blockIdx.x]];
  const int n = I.nmax * I.d2, d2 = I.d2, d = I.d;
  const int tid = threadIdx.x, nth = blockDim.x;
  const double dt = p.dt;
  // ---- LDS: base operators (FP64), the instance's neighbour records, then the state and stage buffers
  const bool dense = (I.flags & F_DENSE) != 0, sd = (I.flags & F_SDIP) != 0, cd = (I.flags & F_CDIP) != 0;
  const int nq = dense ? 0 : I.ncoup * d;
  const int nent = p.rowptr[I.rowptr_off + I.nmax];
  double2* Hb = (double2*)l.rest;  // H, then the system dipole
  double2* Qb = Hb + 2 * d2;       // coupling diagonals, then the coupling dipoles
  cpx<T>* sCL = (cpx<T>*)(Qb + 2 * nq);  // per entry: sqrt-scaled coefficient of Q X' (row term)
  cpx<T>* sCR = sCL + nent;               // per entry: the same of X' Q (column term)
  cpx<T>* sDamp = sCR + nent;             // per ADO: sum_k n_k gamma_k
  int2* sRec = (int2*)(sDamp + I.nmax);   // per entry: neighbour ADO, offset of its coupling (m d, or m)
  int32_t* sRow = (int32_t*)(sRec + nent);
  cpx<double>* Y = (cpx<double>*)(((uintptr_t)(sRow + I.nmax + 1) + 15) & ~(uintptr_t)15);
  cpx<double>* Yn = Y + n;
  cpx<T>* X0 = (cpx<T>*)(Yn + n);
  cpx<T>* X1 = X0 + n;
  const double2* mat = p.mat;
  for (int e = tid; e < d2; e += nth) {
    Hb[e] = mat[I.h_off + e];
    Hb[d2 + e] = sd ? mat[I.h_off + d2 + e] : make_double2(0, 0);
  }
  for (int e = tid; e < nq; e += nth) {
    Qb[e] = mat[I.q_off + e];
    Qb[nq + e] = cd ? mat[I.q_off + nq + e] : make_double2(0, 0);
  }
  for (int e = tid; e <= I.nmax; e += nth) sRow[e] = p.rowptr[I.rowptr_off + e];
  // neighbour records: the coefficient, sqrt(n) and coupling of every entry resolved once
  for (int t = tid; t < nent; t += nth) {
    const int2 en = p.ent[I.ent_off + t];
    const int k = en.y & 0xffff, nf = (en.y >> 16) & 0x7f;
    const bool plus = en.y < 0;
    const double2* C = p.coef + I.coef_off + 5 * k;
    const double sq = sqrt((double)nf);
    const double2 cl = plus ? C[3] : C[1], cr = plus ? C[4] : C[2];
    sCL[t] = mk<T>((T)(cl.x * sq), (T)(cl.y * sq));
    sCR[t] = mk<T>((T)(cr.x * sq), (T)(cr.y * sq));
    const int m = p.mode[I.mode_off + k];
    sRec[t] = make_int2(en.x, dense ? m : m * d);
  }
  __syncthreads();
  for (int a = tid; a < I.nmax; a += nth) {
    double gx = 0.0, gy = 0.0;
    for (int t = sRow[a]; t < sRow[a + 1]; ++t) {
      const int2 en = p.ent[I.ent_off + t];
      if (en.y >= 0) {  // minus entries list every k with n_k > 0, with nf = n_k
        const double2 g = p.coef[I.coef_off + 5 * (en.y & 0xffff)];
        const double nf = (double)((en.y >> 16) & 0x7f);
        gx += g.x * nf;
        gy += g.y * nf;
      }
    }
    sDamp[a] = mk<T>((T)gx, (T)gy);
  }
  for (int e = tid; e < n; e += nth) {
    const int a = e / d2, r = e - a * d2, i = r / d;
    const double2 v = ld_state(p.y, state_ref<PACK>(I, a, i, r - i * d));
    Y[e] = mk(v.x, v.y);
  }
  // H(t) = H + f_s(t) mu and Q(t) = Q + f_c(t) C for field values f_s, f_c, from the LDS copies
  auto set_ops = [&](double fs, double fc) {
    for (int e = tid; e < d2; e += nth) {
      const double2 h = Hb[e], m = Hb[d2 + e];
      l.H[e] = mk<T>((T)(h.x + m.x * fs), (T)(h.y + m.y * fs));
    }
    for (int e = tid; e < nq; e += nth) {
      const double2 q = Qb[e], c = Qb[nq + e];
      l.Q[e] = mk<T>((T)(q.x + c.x * fc), (T)(q.y + c.y * fc));
    }
  };
  // right-hand side of element (i, j) of ADO a from input xs (ADO a's matrix at own)
  auto rhs = [&](const auto* own, const auto* xs, int a, int i, int j, double fc) {
    cpx<T> c = mk<T>(0, 0);
    for (int k = 0; k < d; ++k) {
      cmac(c, l.H[i * d + k], ldc<T>(own, k * d + j));
      cmsub(c, ldc<T>(own, i * d + k), l.H[k * d + j]);
    }
    cpx<T> r = mul_mi(c);
    cmsub(r, sDamp[a], ldc<T>(own, i * d + j));
    const int r1 = sRow[a + 1];
    if (!dense) {
#pragma unroll 4
      for (int t = sRow[a]; t < r1; ++t) {
        const int2 rc = sRec[t];
        const cpx<T> coef = sCL[t] * l.Q[rc.y + i] + sCR[t] * l.Q[rc.y + j];
        cmac(r, coef, ldc<T>(xs, (int64_t)rc.x * d2 + i * d + j));
      }
    } else {
      for (int t = sRow[a]; t < r1; ++t) {
        const int2 rc = sRec[t];
        // dense coupling: cL (Q X')_ij + cR (X' Q)_ij, Q (+ fc C) read from the FP64 matrix pool
        const double2* Qm = mat + I.q_off + (int64_t)rc.y * d2;
        const double2* Cm = cd ? Qm + (int64_t)I.ncoup * d2 : nullptr;
        const int64_t nb = (int64_t)rc.x * d2;
        cpx<T> s1 = mk<T>(0, 0), s2 = mk<T>(0, 0);
        for (int q = 0; q < d; ++q) {
          double2 qa = Qm[i * d + q], qb = Qm[q * d + j];
          if (Cm) {
            const double2 ca = Cm[i * d + q], cb = Cm[q * d + j];
            qa.x += ca.x * fc;
            qa.y += ca.y * fc;
            qb.x += cb.x * fc;
            qb.y += cb.y * fc;
          }
          cmac(s1, mk<T>((T)qa.x, (T)qa.y), ldc<T>(xs, nb + q * d + j));
          cmac(s2, ldc<T>(xs, nb + i * d + q), mk<T>((T)qb.x, (T)qb.y));
        }
        r = r + sCL[t] * s1 + sCR[t] * s2;
      }
    }
    return r;
  };
  const bool driven = p.field != nullptr;
  __syncthreads();
  if (!driven) set_ops(0.0, 0.0);
  // the field values of a step (f_s, f_c at t, t + dt/2, t + dt), read one step ahead
  double fnext[6] = {0, 0, 0, 0, 0, 0};
  if (driven && p.nsteps > 0)
    for (int c = 0; c < 6; ++c) fnext[c] = p.field[6 * p.step + c];
  auto next_field = [&](int64_t s, double* f) {
    for (int c = 0; c < 6; ++c) f[c] = fnext[c];
    if (driven && s + 1 < p.nsteps)
      for (int c = 0; c < 6; ++c) fnext[c] = p.field[6 * (p.step + s + 1) + c];
  };
  int64_t since_save = 0;
  auto save_sample = [&](int64_t s) {  // after step s, once every save_every steps
    if (p.traj && ++since_save == p.save_every) {
      since_save = 0;
      const int64_t slot = (s + 1) / p.save_every - 1;
      for (int e = tid; e < d2; e += nth) p.traj[slot * p.traj_stride + I.traj_off + e] = make_double2(Y[e].x, Y[e].y);
    }
  };
  if (I.flags & F_TINY) {
    // Tiny instances: thread e owns element e for the whole run.  Its indices, damping, the element
    // offsets of its neighbour entries and their coefficients sit in registers; d and the entry count
    // are bounded at compile time (D, E) and the four stages are unrolled, so a stage is one round of
    // independent LDS loads (row i and column j of the input, the neighbours' (i, j) elements) and
    // its arithmetic.  Coefficients are formed once Q is in LDS, and again per stage only when a
    // coupling dipole makes Q time dependent.
    const int te = tid < n ? tid : 0;
    const int ta = te / d2, tr = te - ta * d2, ti = tr / d, tj = tr - ti * d;
    const bool tact = tid < n && !(PACK && ti > tj);
    const int teT = ta * d2 + tj * d + ti;
    const int tt0 = sRow[ta], tcnt = tact ? sRow[ta + 1] - tt0 : 0;
    const cpx<T> tdamp = sDamp[ta];
    const bool qvar = driven && cd;
    auto tiny_run = [&](auto DC, auto EC) {
      constexpr int D = decltype(DC)::value, E = decltype(EC)::value;
      int nb[E];
      cpx<T> cf[E];
#pragma unroll
      for (int u = 0; u < E; ++u) {
        nb[u] = te;
        cf[u] = mk<T>(0, 0);
      }
      auto coefs = [&]() {
#pragma unroll
        for (int u = 0; u < E; ++u) {
          const int t = u < tcnt ? tt0 + u : 0;
          const int2 rc = sRec[t];
          const cpx<T> c = sCL[t] * l.Q[rc.y + ti] + sCR[t] * l.Q[rc.y + tj];
          cf[u] = u < tcnt ? c : mk<T>(0, 0);
          nb[u] = u < tcnt ? rc.x * d2 + ti * d + tj : te;
        }
      };
      if (nent > 0 && !driven) coefs();
      // right-hand side of this thread's element from input xs (every index in bounds; unused terms
      // are dropped by selects, not branches, so the loads issue together)
      auto rhs_t = [&](const auto* xs) {
        const auto* own = xs + ta * d2;
        cpx<T> c = mk<T>(0, 0);
#pragma unroll
        for (int k = 0; k < D; ++k) {
          const int kk = k < d ? k : 0;
          const cpx<T> h1 = l.H[ti * d + kk], x1 = ldc<T>(own, kk * d + tj);
          const cpx<T> x2 = ldc<T>(own, ti * d + kk), h2 = l.H[kk * d + tj];
          if (k < d) {
            cmac(c, h1, x1);
            cmsub(c, x2, h2);
          }
        }
        cpx<T> r = mul_mi(c);
        cmsub(r, tdamp, ldc<T>(own, tr));
#pragma unroll
        for (int u = 0; u < E; ++u) {
          const cpx<T> v = ldc<T>(xs, nb[u]);
          if (u < tcnt) cmac(r, cf[u], v);
        }
        return r;
      };
      auto stage_t = [&](auto SC, const double* f, bool first) {
        constexpr int STG = decltype(SC)::value;
        constexpr int col = STG == 0 ? 0 : (STG == 3 ? 2 : 1);
        if (driven && STG != 2) {  // stages 1 and 2 share the time t + dt/2
          __syncthreads();          // every read of the previous operators is done
          set_ops(f[col], f[3 + col]);
        }
        __syncthreads();  // operators and the previous stage's output are in place
        if (driven && nent > 0 && (qvar ? STG != 2 : (first && STG == 0))) coefs();
        if (!tact) return;
        cpx<T> rr;
        if constexpr (STG == 0)
          rr = rhs_t(Y);  // RHS input is y (double), converted to T as it is read
        else
          rr = rhs_t(STG == 2 ? X1 : X0);
        const double rx = (double)rr.x, ry = (double)rr.y;
        const cpx<double> yv = Y[te];
        if constexpr (STG < 3) {
          const cpx<double> an = STG == 0 ? yv : Yn[te];
          constexpr double wa = STG == 0 ? 1.0 / 6 : 1.0 / 3;
          Yn[te] = mk(an.x + rx * (dt * wa), an.y + ry * (dt * wa));
          const double h = STG == 2 ? dt : dt / 2;
          const cpx<T> o = mk<T>((T)(yv.x + rx * h), (T)(yv.y + ry * h));
          cpx<T>* xd = STG == 1 ? X1 : X0;
          xd[te] = o;
          if (PACK && teT != te) xd[teT] = conjg(o);
        } else {
          // Y is not read by other threads in stage 3 (its input is X0), so it is updated in place
          const cpx<double> an = Yn[te];
          const cpx<double> yn = mk(an.x + rx * (dt / 6), an.y + ry * (dt / 6));
          Y[te] = yn;
          if (PACK && teT != te) Y[teT] = conjg(yn);
        }
      };
      for (int64_t s = 0; s < p.nsteps; ++s) {
        double f[6];
        next_field(s, f);
        stage_t(std::integral_constant<int, 0>(), f, s == 0);
        stage_t(std::integral_constant<int, 1>(), f, s == 0);
        stage_t(std::integral_constant<int, 2>(), f, s == 0);
        stage_t(std::integral_constant<int, 3>(), f, s == 0);
        __syncthreads();
        save_sample(s);
      }
    };
    using I4 = std::integral_constant<int, 4>;
    using I8 = std::integral_constant<int, 8>;
    const bool e4 = (I.flags & F_TINY4) != 0;
    if (d <= 2) {
      if (e4) tiny_run(std::integral_constant<int, 2>(), I4());
      else tiny_run(std::integral_constant<int, 2>(), I8());
    } else if (d <= 4) {
      if (e4) tiny_run(I4(), I4());
      else tiny_run(I4(), I8());
    } else {
      if (e4) tiny_run(I8(), I4());
      else tiny_run(I8(), I8());
    }
  } else {
    for (int64_t s = 0; s < p.nsteps; ++s) {
      double f[6];
      next_field(s, f);
#pragma unroll 1
      for (int stage = 0; stage < 4; ++stage) {
        const int col = stage_col(stage);
        if (driven && stage != 2) {  // stages 1 and 2 share the time t + dt/2
          __syncthreads();            // every read of the previous operators is done
          set_ops(f[col], f[3 + col]);
        }
        __syncthreads();  // operators and the previous stage's output are in place
        const double fc = f[3 + col];
        const cpx<T>* xs = stage == 1 ? X0 : (stage == 2 ? X1 : X0);
        cpx<T>* xd = stage == 1 ? X1 : X0;
        for (int e = tid; e < n; e += nth) {
          const int a = e / d2, r = e - a * d2;
          const int i = r / d, j = r - (r / d) * d;
          // Hermitian (packed) plans: every ADO stays Hermitian, so only upper elements are computed
          // and the lower ones are written as their conjugates
          if (PACK && i > j) continue;
          const int eT = a * d2 + j * d + i;
          cpx<T> rr;
          if (stage == 0) {
            // RHS input is y (double), converted to T as it is read
            rr = rhs(Y + a * d2, Y, a, i, j, fc);
          } else {
            rr = rhs(xs + a * d2, xs, a, i, j, fc);
          }
          const double rx = (double)rr.x, ry = (double)rr.y;
          const cpx<double> yv = Y[e];
          if (stage == 0) {
            Yn[e] = mk(yv.x + rx * (dt / 6), yv.y + ry * (dt / 6));
            const cpx<T> o = mk<T>((T)(yv.x + rx * (dt / 2)), (T)(yv.y + ry * (dt / 2)));
            xd[e] = o;
            if (PACK && eT != e) xd[eT] = conjg(o);
          } else if (stage < 3) {
            const cpx<double> an = Yn[e];
            Yn[e] = mk(an.x + rx * (dt / 3), an.y + ry * (dt / 3));
            const double h = stage == 1 ? dt / 2 : dt;
            const cpx<T> o = mk<T>((T)(yv.x + rx * h), (T)(yv.y + ry * h));
            xd[e] = o;
            if (PACK && eT != e) xd[eT] = conjg(o);
          }
          if (stage == 3) {
            const cpx<double> an = Yn[e];
            // Y is not read by other threads in stage 3 (its input is X0), so it can be updated in place
            const cpx<double> yn = mk(an.x + rx * (dt / 6), an.y + ry * (dt / 6));
            Y[e] = yn;
            if (PACK && eT != e) Y[eT] = conjg(yn);
          }
        }
      }
      __syncthreads();
      save_sample(s);
    }
  }
  __syncthreads();
  for (int e = tid; e < n; e += nth) {
    const int a = e / d2, r = e - a * d2, i = r / d;
    const StateRef sr = state_ref<PACK>(I, a, i, r - i * d);
    if (!sr.lower) p.y[sr.g] = make_double2(Y[e].x, Y[e].y);
  }
}

}  // namespace

// ====================================================================== host side
namespace {

// Binomial table C(n, r) for n <= nmaxn, r <= rmax, saturating at 2^62.
struct Binom {
  int rows = 0, cols = 0;
  std::vector<int64_t> v;
  Binom(int n, int r) : rows(n + 1), cols(r + 1), v((size_t)(n + 1) * (r + 1), 0) {
    const int64_t cap = int64_t(1) << 62;
    for (int i = 0; i <= n; ++i) {
      at(i, 0) = 1;
      for (int j = 1; j <= r && j <= i; ++j) {
        int64_t s = at(i - 1, j - 1) + (j <= i - 1 ? at(i - 1, j) : 0);
        at(i, j) = s > cap ? cap : s;
      }
    }
  }
  int64_t& at(int n, int r) { return v[(size_t)n * cols + r]; }
  int64_t get(int n, int r) const { return (n < 0 || r < 0 || r > n || r >= cols) ? 0 : v[(size_t)n * cols + r]; }
};

// Tables of one (K, L) hierarchy.  Ranks follow the combinatorial number
// system the reference uses (deom.py:555-565): rank(n) = sum_i C(s_i + i, i+1)
// with s_i = n_0 + ... + n_i.  Here every rank is unranked directly (greedy
// decomposition), not generated by the reference's tier-by-tier search, and
// neighbour ranks use rank(n + e_k) - rank(n) = sum_{i>=k} C(s_i + i, i).
struct Tables {
  int K = 0, L = 0, nmax = 0;
  std::vector<int64_t> keys;    // [nmax][K]
  std::vector<int32_t> rowptr;  // [nmax+1]
  std::vector<int2> ent;        // reference order: for k: minus (n_k>0), plus (|n|<L)
};

int build_tables(int K, int L, Tables& t) {
  Binom C(K + L + 1, K + 1);
  const int64_t nm = C.get(K + L, K);
  if (nm <= 0 || nm > (int64_t)INT32_MAX) return fail(PYQED_EINVAL, "hierarchy too large: C(K+L, L) exceeds 2^31-1");
  t.K = K;
  t.L = L;
  t.nmax = (int)nm;
  t.keys.assign((size_t)t.nmax * K, 0);
  t.rowptr.assign((size_t)t.nmax + 1, 0);
  t.ent.clear();
  std::vector<int64_t> s(K), cvec(K + 1), up(K + 1, 0), dn(K + 1, 0);
  t.ent.reserve((size_t)t.nmax * (size_t)std::max(1, std::min(K, L)) * 2);
  for (int r = 0; r < t.nmax; ++r) {
    // unrank: N = sum_{j=1..K} C(c_j, j), c_1 < c_2 < ... < c_K
    int64_t N = r;
    int hi = K + L;  // c_K <= K + L - 1 and c_j < c_{j+1}
    for (int j = K; j >= 1; --j) {
      // largest c in [j-1, hi] with C(c, j) <= N (C(c, j) is nondecreasing in c)
      int lo = j - 1, top = hi;
      while (lo < top) {
        const int mid = (lo + top + 1) / 2;
        if (C.get(mid, j) <= N) lo = mid; else top = mid - 1;
      }
      N -= C.get(lo, j);
      cvec[j] = lo;
      hi = lo - 1;
    }
    int tier = 0;
    for (int i = 0; i < K; ++i) {
      s[i] = cvec[i + 1] - i;  // s_i = c_{i+1} - i
      const int64_t n = i == 0 ? s[0] : s[i] - s[i - 1];
      t.keys[(size_t)r * K + i] = n;
    }
    tier = K ? (int)s[K - 1] : 0;
    // suffix sums for neighbour ranks
    up[K] = dn[K] = 0;
    for (int i = K - 1; i >= 0; --i) {
      up[i] = up[i + 1] + C.get((int)s[i] + i, i);
      dn[i] = dn[i + 1] + C.get((int)s[i] + i - 1, i);
    }
    for (int k = 0; k < K; ++k) {
      const int64_t nk = t.keys[(size_t)r * K + k];
      if (nk > 0) {
        int2 e;
        e.x = (int)(r - dn[k]);
        e.y = (k & 0xffff) | ((int)(nk & 0x7f) << 16);
        t.ent.push_back(e);
      }
      if (tier < L) {
        int2 e;
        e.x = (int)(r + up[k]);
        e.y = (int)((uint32_t)(k & 0xffff) | ((uint32_t)((nk + 1) & 0x7f) << 16) | 0x80000000u);
        t.ent.push_back(e);
      }
    }
    t.rowptr[r + 1] = (int32_t)t.ent.size();
    if (t.ent.size() > (size_t)INT32_MAX / 2) return fail(PYQED_EINVAL, "neighbour table too large");
  }
  return PYQED_OK;
}

// Process-wide cache of built tables keyed by (K, L): many plans share few shapes.
int cached_tables(int K, int L, std::shared_ptr<const Tables>& out) {
  static std::mutex mu;
  static std::map<std::pair<int, int>, std::shared_ptr<const Tables>> cache;
  {
    std::lock_guard<std::mutex> g(mu);
    auto it = cache.find({K, L});
    if (it != cache.end()) {
      out = it->second;
      return PYQED_OK;
    }
  }
  auto t = std::make_shared<Tables>();
  int rc = build_tables(K, L, *t);
  if (rc) return rc;
  std::lock_guard<std::mutex> g(mu);
  auto& slot = cache[{K, L}];
  if (!slot) slot = t;
  out = slot;
  return PYQED_OK;
}

template <typename K>
void* kernel_ptr(K k) {
  return reinterpret_cast<void*>(k);
}

}  // namespace

struct pyqed_heom_plan {
  int n_inst = 0, flags = 0, device = 0;
  std::vector<InstDev> inst;
  std::vector<int64_t> nmax;
  std::map<std::pair<int, int>, std::shared_ptr<const Tables>> tabs;
  std::vector<int> inst_K, inst_L;
  int64_t n_elems = 0, n_ados = 0, traj_stride = 0;
  int d2max = 1, qmax = 1, Kmax = 1, Lmax = 0;
  int d2pad = 2;
  std::vector<WorkItem> items;  // streamed items by class (generic, projector/VALU, projector/MFMA), then resident
  int cls_begin[kClasses + 1] = {0};  // class c occupies [cls_begin[c], cls_begin[c+1])
  int n_items_stream = 0;
  int xr[kClasses][9] = {{0}};
  std::vector<int32_t> resident;
  int64_t res_elems = 0, stream_elems = 0;
  int num_cu = 256, max_lds = 65536;
  // device copies
  InstDev* d_inst = nullptr;
  WorkItem* d_items = nullptr;
  int32_t* d_rowptr = nullptr;
  int2* d_ent = nullptr;
  // projector classes with a single streamed instance: that instance, and its neighbour lists at a fixed
  // stride (StageParams::one_per, ent_fix)
  int one_inst[kClasses] = {-1, -1, -1, -1};
  int fix_stride[kClasses] = {0, 0, 0, 0};
  int2* d_entfix[kClasses] = {nullptr, nullptr, nullptr, nullptr};
  double2* d_coef = nullptr;
  int32_t* d_mode = nullptr;
  double2* d_mat = nullptr;
  int32_t* d_res = nullptr;
  int32_t* d_site = nullptr;
  int2* d_inst_items = nullptr;  // per instance, its work items [first, last) (the adjoint reduction)
  double2* d_coefp = nullptr;
  int n_proj = 0;
  std::vector<char> is_res;
  std::vector<size_t> res_extra;  // per instance: resident_extra_bytes
  int64_t tile_target = 0;
  int blocks_per_cu = 8;
  int64_t cls_target[kClasses] = {0, 0, 0, 0};  // per streamed class: work-item size and blocks per CU
  int cls_bpc[kClasses] = {8, 8, 8, 8};
  int variant = 0;  // bit 0: pipelined kernel for the MFMA class (Hermitian full-matrix FP32); bit 1: block
                    // kernel for the wide class (packed FP32)
  int rowbuf = 1, colbuf = 0;  // projector path: max entries*d over instances (columns: non-Hermitian)
  // LDS dimensions of the projector kernels per item class (1: VALU, 2: MFMA commutator), so that
  // a class of small instances is not sized by the largest
  struct ProjDims {
    int d2pad = 2, rowbuf = 1, colbuf = 0, Kmax = 1, Lmax = 0;
  } pd[kClasses];
  bool hermitian = false;
  bool packed = false;     // y, ynew hold upper triangles (PYQED_PLAN_PACKED)
  bool adjoint = false;    // the plan's operator is the adjoint L^H (PYQED_PLAN_ADJOINT)
  int64_t n_state = 0;     // RK4 state elements (size of y): n_elems, or the packed count
  int64_t cls_state[kClasses] = {0};  // RK4 state elements of the streamed items per class
  hipStream_t side = nullptr;
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  // the generic class (small d, latency-bound launches) runs its stages on a stream of its own,
  // concurrently with the projector classes: the classes share no elements
  hipStream_t side0 = nullptr;
  hipEvent_t ev_fork0 = nullptr, ev_join0 = nullptr;
  // per-launch timing (pyqed_heom_plan_set_timing): event pairs around every class launch
  bool timing = false;
  std::vector<hipEvent_t> ev_pool;
  std::vector<int> ev_tag;  // stage * kClasses + class of pair k (events 2k, 2k+1)
};

namespace {

template <typename V>
int upload(const std::vector<V>& h, V** d) {
  if (h.empty()) {
    *d = nullptr;
    return PYQED_OK;
  }
  PQA_HIP(hipMalloc((void**)d, sizeof(V) * h.size()));
  PQA_HIP(hipMemcpy(*d, h.data(), sizeof(V) * h.size(), hipMemcpyHostToDevice));
  return PYQED_OK;
}

// Workspace of a run: ynew in the RK4 state layout (complex double), then the two stage buffers
// (full matrices, complex of xsize bytes).
int64_t ws_state_bytes(const pyqed_heom_plan* P) { return ((P->n_state * 16 + 255) / 256) * 256; }
int64_t workspace_bytes(const pyqed_heom_plan* P, size_t xsize) { return ws_state_bytes(P) + 2 * P->n_elems * (int64_t)xsize; }

size_t stream_lds_bytes(const pyqed_heom_plan* P, size_t tsize) {
  size_t h = lds_header_bytes(P->d2max, P->qmax, P->Kmax, P->Lmax, tsize);
  size_t x = 2 * tsize * (size_t)std::max(256, 4 * P->d2max);
  return h + x;
}

size_t proj_lds_bytes(const pyqed_heom_plan* P, int c, size_t tsize) {
  const auto& q = P->pd[c];
  const bool mfma = c >= 2 && tsize == sizeof(float);  // the MFMA forms keep their tile in the own-matrix buffer
  return proj_block_bytes(q.d2pad, q.Kmax, q.Lmax, tsize, P->packed) +
         (kThreads / 64) * proj_wave_bytes(q.d2pad, q.rowbuf, q.colbuf, tsize, P->packed && !mfma);
}

size_t pipe_lds_bytes(const pyqed_heom_plan* P, int c) {
  const auto& q = P->pd[c];
  return proj_block_bytes(q.d2pad, q.Kmax, q.Lmax, sizeof(float), false) + (kThreads / 64) * pipe_wave_bytes(q.d2pad, q.rowbuf);
}

// LDS of the resident kernel for an instance of n_el elements whose base operators and tables take
// `extra` bytes (resident_extra_bytes): header, base operators, tables, state and stage buffers.
size_t resident_lds_bytes(const pyqed_heom_plan* P, int64_t n_el, size_t extra, size_t tsize) {
  size_t h = lds_header_bytes(P->d2max, P->qmax, P->Kmax, P->Lmax, tsize);
  return h + extra + 16 + (size_t)n_el * (32 + 2 * 2 * tsize);
}

// Base operators (H and system dipole; coupling diagonals and coupling dipoles, FP64 complex), the
// neighbour records (two complex coefficients and an int2 per entry), the damping per ADO and the row
// pointers of one instance, as the resident kernel lays them out (sized for FP64 records).
size_t resident_extra_bytes(int d, int ncoup, bool dense, int nmax, int64_t nent) {
  const size_t nq = dense ? 0 : (size_t)ncoup * d;
  return 16 * (2 * (size_t)d * d + 2 * nq) + (2 * 16 + 8) * (size_t)nent + 16 * (size_t)nmax + 4 * (size_t)(nmax + 1);
}

// The work queue and, per instance, the range [first, last) of its items: tile() lists the items of one
// instance consecutively.
int upload_items(pyqed_heom_plan* P) {
  (void)hipFree(P->d_items);
  (void)hipFree(P->d_inst_items);
  P->d_items = nullptr;
  P->d_inst_items = nullptr;
  std::vector<int2> rng(P->n_inst, make_int2(0, 0));
  std::vector<char> seen(P->n_inst, 0);
  for (int it = 0; it < (int)P->items.size(); ++it) {
    const int i = P->items[it].inst;
    if (!seen[i]) rng[i].x = it;
    seen[i] = 1;
    rng[i].y = it + 1;
  }
  int rc = upload(P->items, &P->d_items);
  if (rc) return rc;
  return upload(rng, &P->d_inst_items);
}

int group_size(int d2) {
  int g = 4;
  while (g < d2 && g < 64) g <<= 1;
  return g;
}

void destroy_plan(pyqed_heom_plan* P) {
  if (!P) return;
  (void)hipFree(P->d_inst);
  (void)hipFree(P->d_items);
  (void)hipFree(P->d_rowptr);
  (void)hipFree(P->d_ent);
  for (int c = 0; c < kClasses; ++c) (void)hipFree(P->d_entfix[c]);
  (void)hipFree(P->d_coef);
  (void)hipFree(P->d_mode);
  (void)hipFree(P->d_mat);
  (void)hipFree(P->d_res);
  (void)hipFree(P->d_site);
  (void)hipFree(P->d_inst_items);
  (void)hipFree(P->d_coefp);
  if (P->side) (void)hipStreamDestroy(P->side);
  if (P->ev_fork) (void)hipEventDestroy(P->ev_fork);
  if (P->ev_join) (void)hipEventDestroy(P->ev_join);
  if (P->side0) (void)hipStreamDestroy(P->side0);
  if (P->ev_fork0) (void)hipEventDestroy(P->ev_fork0);
  if (P->ev_join0) (void)hipEventDestroy(P->ev_join0);
  for (hipEvent_t e : P->ev_pool) (void)hipEventDestroy(e);
  delete P;
}

std::complex<double> cz(const double* p, int i) { return std::complex<double>(p[2 * i], p[2 * i + 1]); }

// Cut every instance into work items of about `target` complex elements (0: automatic), streamed
// classes first (generic, projector/VALU, projector/MFMA), resident instances last, and split each
// class into eight contiguous ranges of equal element count (one per XCD).  `blocks_per_cu` sets the
// persistent grid of the streamed kernels.
void build_items(pyqed_heom_plan* P);

void retile(pyqed_heom_plan* P, int64_t target, int blocks_per_cu) {
  if (target <= 0)
    target = std::min<int64_t>(8192, std::max<int64_t>(64, P->stream_elems / std::max<int64_t>(1, (int64_t)P->num_cu * 8)));
  P->tile_target = target;
  P->blocks_per_cu = std::max(1, std::min(blocks_per_cu, 32));
  for (int c = 0; c < kClasses; ++c) {
    P->cls_target[c] = P->tile_target;
    P->cls_bpc[c] = P->blocks_per_cu;
  }
  build_items(P);
}

// The work items of every class at that class's item size; resident instances at the plan's.
void build_items(pyqed_heom_plan* P) {
  P->items.clear();
  auto tile = [&](int i) {
    const InstDev& I = P->inst[i];
    const int nslot = (I.path == 0 || P->is_res[i]) ? kThreads / I.grp : kThreads / 64;
    const int64_t target = P->is_res[i] ? P->tile_target : P->cls_target[I.path];
    int64_t per = (target + I.d2 - 1) / I.d2;
    per = ((per + nslot - 1) / nslot) * nslot;
    per = std::max<int64_t>(per, nslot);
    for (int a = 0; a < I.nmax; a += (int)per) {
      WorkItem w;
      w.inst = i;
      w.ado0 = a;
      w.nado = (int)std::min<int64_t>(per, I.nmax - a);
      w.pad = 0;
      P->items.push_back(w);
    }
  };
  for (int c = 0; c < kClasses; ++c) {
    P->cls_begin[c] = (int)P->items.size();
    for (int i = 0; i < P->n_inst; ++i)
      if (!P->is_res[i] && P->inst[i].path == c) tile(i);
  }
  P->cls_begin[kClasses] = P->n_items_stream = (int)P->items.size();
  for (int i : P->resident) tile(i);
  for (int c = 0; c < kClasses; ++c) {
    const int b0 = P->cls_begin[c], n = P->cls_begin[c + 1] - b0;
    std::vector<int64_t> pre(n + 1, 0);
    for (int k = 0; k < n; ++k) pre[k + 1] = pre[k] + (int64_t)P->items[b0 + k].nado * P->inst[P->items[b0 + k].inst].d2;
    P->xr[c][0] = 0;
    for (int x = 1; x < 8; ++x) {
      const int64_t goal = pre.back() * x / 8;
      P->xr[c][x] = (int)(std::lower_bound(pre.begin(), pre.end(), goal) - pre.begin());
      P->xr[c][x] = std::max(P->xr[c][x], P->xr[c][x - 1]);
    }
    P->xr[c][8] = n;
  }
}

}  // namespace

extern "C" int pyqed_heom_plan_create(const pyqed_heom_instance* in, int32_t n_inst, int32_t flags,
                                      pyqed_heom_plan** out) {
  if (!in || !out || n_inst <= 0) return fail(PYQED_EINVAL, "plan_create: null argument or n_inst <= 0");
  *out = nullptr;
  pyqed_heom_plan* P = new pyqed_heom_plan();
  P->n_inst = n_inst;
  P->flags = flags;
  P->hermitian = (flags & PYQED_PLAN_HERMITIAN) != 0;
  P->packed = (flags & PYQED_PLAN_PACKED) != 0;
  P->adjoint = (flags & PYQED_PLAN_ADJOINT) != 0;
  const bool dense = (flags & PYQED_PLAN_DENSE_Q) != 0;
  if (P->packed && !P->hermitian) {
    destroy_plan(P);
    return fail(PYQED_EINVAL, "plan_create: PYQED_PLAN_PACKED needs PYQED_PLAN_HERMITIAN");
  }
  if (P->adjoint && P->hermitian) {
    destroy_plan(P);
    return fail(PYQED_EINVAL, "plan_create: PYQED_PLAN_ADJOINT excludes PYQED_PLAN_HERMITIAN and PYQED_PLAN_PACKED");
  }
  // ---- validate and build tables
  for (int i = 0; i < n_inst; ++i) {
    const pyqed_heom_instance& s = in[i];
    if (s.d < 1 || s.d > kMaxD) {
      destroy_plan(P);
      return fail(PYQED_EINVAL, "instance " + std::to_string(i) + ": d must be in 1..32");
    }
    if (s.K < 0 || s.K > kMaxK || s.L < 0 || s.L > kMaxL || s.ncoup < 1) {
      destroy_plan(P);
      return fail(PYQED_EINVAL, "instance " + std::to_string(i) + ": K in 0..255, L in 0..126, ncoup >= 1 required");
    }
    if (!s.H || (s.K > 0 && (!s.mode || !s.expn || !s.etal || !s.etar || !s.etaa)) || (!dense && !s.qdiag) ||
        (dense && !s.qmat)) {
      destroy_plan(P);
      return fail(PYQED_EINVAL, "instance " + std::to_string(i) + ": missing H, bath or coupling arrays");
    }
    // non-finite inputs would propagate silently through every ADO: reject them here
    auto finite = [](const double* v, int64_t n) {
      for (int64_t e = 0; e < n; ++e)
        if (!std::isfinite(v[e])) return false;
      return true;
    };
    const int64_t nqv = 2 * (int64_t)s.ncoup * (dense ? s.d * s.d : s.d);
    const char* bad = !finite(s.H, 2 * (int64_t)s.d * s.d) ? "H"
                      : (s.sdip && !finite(s.sdip, 2 * (int64_t)s.d * s.d)) ? "sdip"
                      : !finite(dense ? s.qmat : s.qdiag, nqv) ? "coupling"
                      : ((dense ? s.cmat : s.cdiag) && !finite(dense ? s.cmat : s.cdiag, nqv)) ? "coupling dipole"
                      : (s.K > 0 && !finite(s.expn, 2 * (int64_t)s.K)) ? "expn"
                      : (s.K > 0 && !finite(s.etal, 2 * (int64_t)s.K)) ? "etal"
                      : (s.K > 0 && !finite(s.etar, 2 * (int64_t)s.K)) ? "etar"
                      : (s.K > 0 && !finite(s.etaa, 2 * (int64_t)s.K)) ? "etaa" : nullptr;
    if (bad) {
      destroy_plan(P);
      return fail(PYQED_EINVAL, "instance " + std::to_string(i) + ": non-finite value (NaN or Inf) in " + bad);
    }
    for (int k = 0; k < s.K; ++k) {
      if (s.mode[k] < 0 || s.mode[k] >= s.ncoup) {
        destroy_plan(P);
        return fail(PYQED_EINVAL, "instance " + std::to_string(i) + ": mode[" + std::to_string(k) +
                                      "] out of range of the coupling operators");
      }
      if (std::abs(cz(s.etaa, k)) == 0.0) {
        destroy_plan(P);
        return fail(PYQED_EINVAL, "instance " + std::to_string(i) + ": etaa[" + std::to_string(k) +
                                      "] == 0 (the operator divides by sqrt(etaa))");
      }
    }
    auto key = std::make_pair((int)s.K, (int)s.L);
    if (!P->tabs.count(key)) {
      std::shared_ptr<const Tables> t;
      int rc = cached_tables(s.K, s.L, t);
      if (rc) {
        destroy_plan(P);
        return rc;
      }
      P->tabs.emplace(key, t);
    }
  }
  // ---- pack pools
  std::vector<int32_t> rowptr;
  std::vector<int2> ent;
  std::map<std::pair<int, int>, std::pair<int, int>> tab_off;
  for (auto& kv : P->tabs) {
    tab_off[kv.first] = {(int)rowptr.size(), (int)ent.size()};
    rowptr.insert(rowptr.end(), kv.second->rowptr.begin(), kv.second->rowptr.end());
    ent.insert(ent.end(), kv.second->ent.begin(), kv.second->ent.end());
  }
  std::vector<double2> coef, mat, coefp;
  std::vector<int32_t> mode, sites;
  int64_t off = 0, toff = 0, ados = 0, soff = 0;
  for (int i = 0; i < n_inst; ++i) {
    const pyqed_heom_instance& s = in[i];
    const Tables& t = *P->tabs.at({s.K, s.L});
    InstDev I;
    std::memset(&I, 0, sizeof(I));
    I.d = s.d;
    I.d2 = s.d * s.d;
    I.K = s.K;
    I.L = s.L;
    I.ncoup = s.ncoup;
    I.nmax = t.nmax;
    I.flags = (s.sdip ? F_SDIP : 0) | (dense ? (s.cmat ? F_CDIP : 0) : (s.cdiag ? F_CDIP : 0)) | (dense ? F_DENSE : 0);
    I.grp = group_size(I.d2);
    // every instance starts on a 128-byte boundary of the stage buffers (16 FP32 complex elements)
    // and of the packed FP64 state (8 elements), so that its rows fall on whole cache lines as often
    // as d allows
    off = (off + 15) & ~int64_t(15);
    soff = (soff + 7) & ~int64_t(7);
    I.off = off;
    I.dp = P->packed ? s.d * (s.d + 1) / 2 : I.d2;
    I.poff = P->packed ? soff : off;
    I.traj_off = toff;
    I.rowptr_off = tab_off[{s.K, s.L}].first;
    I.ent_off = tab_off[{s.K, s.L}].second;
    I.coef_off = (int)coef.size();
    I.mode_off = (int)mode.size();
    for (int k = 0; k < s.K; ++k) {
      const std::complex<double> I1(0.0, 1.0);
      const std::complex<double> sq = std::sqrt(cz(s.etaa, k));
      std::complex<double> v[5] = {cz(s.expn, k), -I1 * cz(s.etal, k) / sq, I1 * cz(s.etar, k) / sq, -I1 * sq,
                                   I1 * sq};
      if (P->adjoint) {
        // block (n, n - e_k) of L^H is the conjugate of block (n - e_k, n) of L, a plus term, and
        // block (n, n + e_k) that of the minus term of n + e_k: swap the two pairs and conjugate
        const std::complex<double> a[5] = {std::conj(v[0]), std::conj(v[3]), std::conj(v[4]), std::conj(v[1]),
                                           std::conj(v[2])};
        for (int q = 0; q < 5; ++q) v[q] = a[q];
      }
      for (int q = 0; q < 5; ++q) coef.push_back(make_double2(v[q].real(), v[q].imag()));
      mode.push_back(s.mode[k]);
    }
    // the adjoint operator: -i[H, .]^H = -i[-H^H, .] and (A . B)^H = A^H . B^H for the couplings
    auto push_mat = [&](const double* src, int dd, double sign, bool adj) {
      for (int e = 0; e < dd * dd; ++e) {
        const int a2 = e / dd, b2 = e - a2 * dd;
        const int f = adj ? b2 * dd + a2 : e;
        mat.push_back(make_double2(sign * src[2 * f], (adj ? -sign : sign) * src[2 * f + 1]));
      }
    };
    auto push_diag = [&](const double* src, int n, bool adj) {
      for (int e = 0; e < n; ++e) mat.push_back(make_double2(src[2 * e], adj ? -src[2 * e + 1] : src[2 * e + 1]));
    };
    const bool adj = P->adjoint;
    I.h_off = (int)mat.size();
    push_mat(s.H, s.d, adj ? -1.0 : 1.0, adj);
    if (s.sdip) push_mat(s.sdip, s.d, adj ? -1.0 : 1.0, adj);
    I.q_off = (int)mat.size();
    const double* qsrc = dense ? s.qmat : s.qdiag;
    const double* csrc = dense ? s.cmat : s.cdiag;
    for (int m = 0; m < s.ncoup; ++m) {
      if (dense) push_mat(qsrc + 2 * (int64_t)m * I.d2, s.d, 1.0, adj);
      else push_diag(qsrc + 2 * (int64_t)m * s.d, s.d, adj);
    }
    if (csrc)
      for (int m = 0; m < s.ncoup; ++m) {
        if (dense) push_mat(csrc + 2 * (int64_t)m * I.d2, s.d, 1.0, adj);
        else push_diag(csrc + 2 * (int64_t)m * s.d, s.d, adj);
      }
    if (flags & PYQED_PLAN_HERMITIAN) {
      // the caller asserts Hermitian initial states; check what makes the dynamics preserve it
      bool ok = !dense && !s.cdiag;
      for (int a2 = 0; a2 < s.d && ok; ++a2)
        for (int b2 = 0; b2 < s.d && ok; ++b2) {
          ok = std::abs(cz(s.H, a2 * s.d + b2) - std::conj(cz(s.H, b2 * s.d + a2))) <= 1e-14 * (1 + std::abs(cz(s.H, a2 * s.d + b2)));
          if (ok && s.sdip)
            ok = std::abs(cz(s.sdip, a2 * s.d + b2) - std::conj(cz(s.sdip, b2 * s.d + a2))) <= 1e-14 * (1 + std::abs(cz(s.sdip, a2 * s.d + b2)));
        }
      for (int e = 0; e < s.ncoup * s.d && ok; ++e) ok = cz(s.qdiag, e).imag() == 0.0;
      for (int k = 0; k < s.K && ok; ++k)
        ok = cz(s.expn, k).imag() == 0.0 && std::abs(cz(s.etar, k) - std::conj(cz(s.etal, k))) <= 1e-14 * std::abs(cz(s.etal, k));
      if (!ok) {
        destroy_plan(P);
        return fail(PYQED_EINVAL, "instance " + std::to_string(i) +
                                      ": PYQED_PLAN_HERMITIAN needs Hermitian H and dipole, real diagonal couplings, "
                                      "real expn and etar = conj(etal)");
      }
      I.flags |= F_HERM;
    }
    // ---- site-projector path: every coupling is c_m |p_m><p_m| (diagonal with one nonzero entry)
    I.path = 0;
    if (!dense && !s.cdiag && s.d >= 6 && !(flags & PYQED_PLAN_GENERIC)) {
      std::vector<int> psite(s.ncoup, -1);
      std::vector<std::complex<double>> pw(s.ncoup);
      bool proj = true;
      for (int m = 0; m < s.ncoup && proj; ++m) {
        int nz = 0;
        for (int e = 0; e < s.d; ++e) {
          const std::complex<double> q = cz(s.qdiag, m * s.d + e);
          if (q != 0.0) {
            ++nz;
            psite[m] = e;
            pw[m] = P->adjoint ? std::conj(q) : q;  // the adjoint plan's couplings are Q^H
          }
        }
        proj = nz == 1;
      }
      // projector path: at most kMaxEnt neighbour entries per ADO (K plus + min(K, L) minus)
      if (proj && s.K + std::min(s.K, s.L) <= kMaxEnt) {
        // MFMA commutator for PQA_MFMA_DMIN <= d <= 16 (one 16 x 16 tile, d padded to 16) and for
        // 17 <= d <= 32 (2 x 2 tiles, d padded to 32); VALU below
        I.path = (s.d >= PQA_MFMA_DMIN && s.d <= 16) ? 2 : (s.d > 16 ? 3 : 1);
        I.site_off = (int)sites.size();
        bool sorted_sites = true;
        for (int k = 0; k < s.K; ++k) {
          sites.push_back(psite[s.mode[k]]);
          if (k > 0 && sites.back() < sites[sites.size() - 2]) sorted_sites = false;
        }
        // neighbour entries are listed by k, so a nondecreasing site map lists them in site order
        if (sorted_sites) I.flags |= F_SITESORT;
        I.coefp_off = (int)coefp.size();
        for (int k = 0; k < s.K; ++k) {
          const std::complex<double> wgt = pw[s.mode[k]];
          for (int q = 0; q < 5; ++q) {
            const double2 c = coef[I.coef_off + 5 * k + q];
            std::complex<double> v(c.x, c.y);
            if (q > 0) v *= wgt;
            coefp.push_back(make_double2(v.real(), v.imag()));
          }
        }
        P->n_proj++;
        const int nb = (s.K + std::min(s.K, s.L)) * (s.d + 1);
        P->rowbuf = std::max(P->rowbuf, nb);
        if (!(flags & PYQED_PLAN_HERMITIAN)) P->colbuf = std::max(P->colbuf, nb);
        auto& q = P->pd[I.path];
        q.d2pad = std::max(q.d2pad, I.d2 + I.d);
        q.rowbuf = std::max(q.rowbuf, nb);
        if (!(flags & PYQED_PLAN_HERMITIAN)) q.colbuf = std::max(q.colbuf, nb);
        q.Kmax = std::max(q.Kmax, s.K);
        q.Lmax = std::max(q.Lmax, s.L);
      }
    }
    if (mat.size() > (size_t)INT32_MAX || coef.size() > (size_t)INT32_MAX || coefp.size() > (size_t)INT32_MAX) {
      destroy_plan(P);
      return fail(PYQED_EINVAL, "coefficient pools exceed 2^31 entries");
    }
    {
      bool realh = true;
      for (int e = 0; e < I.d2 && realh; ++e) realh = s.H[2 * e + 1] == 0.0 && (!s.sdip || s.sdip[2 * e + 1] == 0.0);
      if (realh) I.flags |= F_REALH;
      int maxrow = 0;
      for (int a = 0; a < t.nmax; ++a) maxrow = std::max(maxrow, t.rowptr[a + 1] - t.rowptr[a]);
      if (!dense && s.d <= kTinyD && (int64_t)t.nmax * I.d2 <= kThreads && maxrow <= kTinyEnt) {
        I.flags |= F_TINY;
        if (maxrow <= 4) I.flags |= F_TINY4;
      }
    }
    P->inst.push_back(I);
    P->nmax.push_back(t.nmax);
    P->res_extra.push_back(resident_extra_bytes(s.d, s.ncoup, dense, t.nmax, (int64_t)t.ent.size()));
    P->inst_K.push_back(s.K);
    P->inst_L.push_back(s.L);
    off += (int64_t)t.nmax * I.d2;
    soff += (int64_t)t.nmax * I.dp;
    toff += I.d2;
    ados += t.nmax;
    P->d2max = std::max(P->d2max, I.d2);
    P->d2pad = std::max(P->d2pad, I.d2 + I.d);
    P->qmax = std::max(P->qmax, dense ? 1 : s.ncoup * s.d);
    P->Kmax = std::max(P->Kmax, s.K);
    P->Lmax = std::max(P->Lmax, s.L);
  }
  P->n_elems = off;
  P->n_state = P->packed ? soff : off;
  P->n_ados = ados;
  P->traj_stride = toff;
  // ---- device properties
  if (hipGetDevice(&P->device) != hipSuccess) {
    destroy_plan(P);
    return fail(PYQED_EHIP, "no HIP device");
  }
  int v = 0;
  if (hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, P->device) == hipSuccess && v > 0)
    P->num_cu = v;
  if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, P->device) == hipSuccess && v > 0)
    P->max_lds = v;
  // ---- residency: the whole instance (FP64 sizes) must fit one workgroup's LDS
  std::vector<char> is_res(n_inst, 0);
  if (!(flags & PYQED_PLAN_NO_RESIDENT)) {
    for (int i = 0; i < n_inst; ++i) {
      const int64_t n = (int64_t)P->inst[i].nmax * P->inst[i].d2;
      if (resident_lds_bytes(P, n, P->res_extra[i], sizeof(double)) <= (size_t)P->max_lds) is_res[i] = 1;
    }
  }
  // ---- work items: tiles of ADOs, streamed instances first
  for (int i = 0; i < n_inst; ++i) (is_res[i] ? P->res_elems : P->stream_elems) += (int64_t)P->inst[i].nmax * P->inst[i].d2;
  P->is_res = is_res;
  for (int i = 0; i < n_inst; ++i)
    if (is_res[i]) P->resident.push_back(i);
  retile(P, 0, 8);
  for (int i = 0; i < n_inst; ++i)
    if (!is_res[i]) P->cls_state[P->inst[i].path] += (int64_t)P->inst[i].nmax * P->inst[i].dp;
  if (stream_lds_bytes(P, sizeof(double)) > (size_t)P->max_lds) {
    destroy_plan(P);
    return fail(PYQED_EUNSUPPORTED, "instance too large for the streamed kernel's LDS budget");
  }
  // ---- projector classes with a single streamed instance: neighbour lists at a fixed stride, for the
  // classes whose packed forms have a single-instance kernel (the VALU class and the wide class)
  std::vector<std::vector<int2>> entfix(kClasses);
  for (int c = 1; c < kClasses && P->packed; c += 2) {
    int only = -1, count = 0;
    for (int i = 0; i < n_inst; ++i)
      if (!is_res[i] && P->inst[i].path == c) only = i, ++count;
    if (count != 1) continue;
    const InstDev& I = P->inst[only];
    const int S = I.K + std::min(I.K, I.L);  // entries per ADO at most (K plus, min(K, L) minus)
    if ((int64_t)I.nmax * S * (int64_t)sizeof(int2) > (int64_t(1) << 28)) continue;
    std::vector<int2>& f = entfix[c];
    f.assign((size_t)I.nmax * S, make_int2(-1, 0));
    bool fits = true;
    for (int a = 0; a < I.nmax && fits; ++a) {
      const int r0 = rowptr[I.rowptr_off + a], r1 = rowptr[I.rowptr_off + a + 1];
      fits = r1 - r0 <= S;
      for (int r = r0; r < r1 && fits; ++r) f[(size_t)a * S + (r - r0)] = ent[I.ent_off + r];
    }
    if (!fits) {
      f.clear();
      continue;
    }
    P->one_inst[c] = only;
    P->fix_stride[c] = S;
  }
  // ---- upload
  int rc = 0;
  for (int c = 1; c < kClasses && rc == 0; ++c)
    if (!entfix[c].empty()) rc = upload(entfix[c], &P->d_entfix[c]);
  if (rc || (rc = upload(P->inst, &P->d_inst)) || (rc = upload_items(P)) ||
      (rc = upload(rowptr, &P->d_rowptr)) || (rc = upload(ent, &P->d_ent)) || (rc = upload(coef, &P->d_coef)) ||
      (rc = upload(mode, &P->d_mode)) || (rc = upload(mat, &P->d_mat)) || (rc = upload(P->resident, &P->d_res)) ||
      (rc = upload(sites, &P->d_site)) ||
      (rc = upload(coefp, &P->d_coefp))) {
    destroy_plan(P);
    return rc;
  }
  if (hipStreamCreateWithFlags(&P->side, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreateWithFlags(&P->ev_fork, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&P->ev_join, hipEventDisableTiming) != hipSuccess ||
      hipStreamCreateWithFlags(&P->side0, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreateWithFlags(&P->ev_fork0, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&P->ev_join0, hipEventDisableTiming) != hipSuccess) {
    destroy_plan(P);
    return fail(PYQED_EHIP, "stream/event creation failed");
  }
  *out = P;
  return PYQED_OK;
}

extern "C" int pyqed_heom_plan_destroy(pyqed_heom_plan* plan) {
  destroy_plan(plan);
  return PYQED_OK;
}

extern "C" int pyqed_heom_plan_query(const pyqed_heom_plan* P, int64_t* n_elems, int64_t* n_ados,
                                     int64_t* ws64, int64_t* wsmix) {
  if (!P) return fail(PYQED_EINVAL, "plan_query: null plan");
  if (n_elems) *n_elems = P->n_state;
  if (n_ados) *n_ados = P->n_ados;
  if (ws64) *ws64 = workspace_bytes(P, sizeof(double2));
  if (wsmix) *wsmix = workspace_bytes(P, sizeof(float2));
  return PYQED_OK;
}

extern "C" int pyqed_heom_plan_layout(const pyqed_heom_plan* P, int64_t* so, int64_t* nm, int64_t* to) {
  if (!P) return fail(PYQED_EINVAL, "plan_layout: null plan");
  for (int i = 0; i < P->n_inst; ++i) {
    if (so) so[i] = P->inst[i].poff;
    if (nm) nm[i] = P->inst[i].nmax;
    if (to) to[i] = P->inst[i].traj_off;
  }
  return PYQED_OK;
}

extern "C" int pyqed_heom_plan_class_state(const pyqed_heom_plan* P, int64_t* state_elems) {
  if (!P || !state_elems) return fail(PYQED_EINVAL, "plan_class_state: null argument");
  for (int c = 0; c < kClasses; ++c) state_elems[c] = P->cls_state[c];
  return PYQED_OK;
}

extern "C" int pyqed_heom_plan_set_timing(pyqed_heom_plan* P, int32_t enable) {
  if (!P) return fail(PYQED_EINVAL, "plan_set_timing: null plan");
  P->timing = enable != 0;
  P->ev_tag.clear();
  return PYQED_OK;
}

extern "C" int pyqed_heom_plan_read_timing(pyqed_heom_plan* P, double* ms, int64_t* launches) {
  if (!P || !ms || !launches) return fail(PYQED_EINVAL, "plan_read_timing: null argument");
  for (int k = 0; k < 4 * kClasses; ++k) {
    ms[k] = 0.0;
    launches[k] = 0;
  }
  for (size_t k = 0; k < P->ev_tag.size(); ++k) {
    PQA_HIP(hipEventSynchronize(P->ev_pool[2 * k + 1]));
    float t = 0.f;
    PQA_HIP(hipEventElapsedTime(&t, P->ev_pool[2 * k], P->ev_pool[2 * k + 1]));
    ms[P->ev_tag[k]] += t;
    launches[P->ev_tag[k]] += 1;
  }
  P->ev_tag.clear();
  return PYQED_OK;
}

extern "C" int pyqed_heom_plan_stats(const pyqed_heom_plan* P, int64_t* s) {
  if (!P || !s) return fail(PYQED_EINVAL, "plan_stats: null argument");
  s[0] = P->n_items_stream;
  s[1] = (int64_t)P->resident.size();
  s[2] = P->stream_elems;
  s[3] = P->res_elems;
  return PYQED_OK;
}

extern "C" int pyqed_heom_plan_set_tiling(pyqed_heom_plan* P, int64_t target_elems, int32_t blocks_per_cu) {
  if (!P || blocks_per_cu < 1 || target_elems < 0) return fail(PYQED_EINVAL, "plan_set_tiling: bad argument");
  PQA_HIP(hipDeviceSynchronize());  // the old queue may still be in use by queued launches
  retile(P, target_elems, blocks_per_cu);
  return upload_items(P);
}

extern "C" int pyqed_heom_plan_set_class_tiling(pyqed_heom_plan* P, int32_t cls, int64_t target_elems,
                                               int32_t blocks_per_cu) {
  if (!P || cls < 0 || cls >= kClasses || target_elems < 1 || blocks_per_cu < 1 || blocks_per_cu > 32)
    return fail(PYQED_EINVAL, "plan_set_class_tiling: need 0 <= cls < 4, target_elems >= 1, 1 <= blocks_per_cu <= 32");
  PQA_HIP(hipDeviceSynchronize());  // the old queue may still be in use by queued launches
  P->cls_target[cls] = target_elems;
  P->cls_bpc[cls] = blocks_per_cu;
  build_items(P);
  return upload_items(P);
}

extern "C" int pyqed_heom_plan_get_class_tiling(const pyqed_heom_plan* P, int64_t* target_elems,
                                               int32_t* blocks_per_cu) {
  if (!P || !target_elems || !blocks_per_cu) return fail(PYQED_EINVAL, "plan_get_class_tiling: null argument");
  for (int c = 0; c < kClasses; ++c) {
    target_elems[c] = P->cls_target[c];
    blocks_per_cu[c] = P->cls_bpc[c];
  }
  return PYQED_OK;
}

extern "C" int pyqed_heom_plan_set_variant(pyqed_heom_plan* P, int32_t variant) {
  if (!P || variant < 0 || variant > 3) return fail(PYQED_EINVAL, "plan_set_variant: variant must be in 0..3");
  P->variant = variant;
  return PYQED_OK;
}

extern "C" int pyqed_heom_plan_get_tiling(const pyqed_heom_plan* P, int64_t* target_elems, int32_t* blocks_per_cu) {
  if (!P) return fail(PYQED_EINVAL, "plan_get_tiling: null plan");
  if (target_elems) *target_elems = P->tile_target;
  if (blocks_per_cu) *blocks_per_cu = P->blocks_per_cu;
  return PYQED_OK;
}

extern "C" int pyqed_heom_plan_class_single(const pyqed_heom_plan* P, int32_t* inst) {
  if (!P || !inst) return fail(PYQED_EINVAL, "plan_class_single: null argument");
  for (int c = 0; c < kClasses; ++c) inst[c] = P->one_inst[c];
  return PYQED_OK;
}

extern "C" int pyqed_heom_plan_classes(const pyqed_heom_plan* P, int64_t* items, int64_t* elems) {
  if (!P) return fail(PYQED_EINVAL, "plan_classes: null plan");
  for (int c = 0; c < kClasses; ++c) {
    if (items) items[c] = P->cls_begin[c + 1] - P->cls_begin[c];
    if (elems) {
      elems[c] = 0;
      for (int k = P->cls_begin[c]; k < P->cls_begin[c + 1]; ++k)
        elems[c] += (int64_t)P->items[k].nado * P->inst[P->items[k].inst].d2;
    }
  }
  return PYQED_OK;
}

namespace {

void export_tables(const Tables& t, int64_t* keys, int32_t* plus, int32_t* minus) {
  const int K = t.K;
  if (keys) std::memcpy(keys, t.keys.data(), sizeof(int64_t) * t.keys.size());
  if (plus) std::fill(plus, plus + (size_t)t.nmax * K, -1);
  if (minus) std::fill(minus, minus + (size_t)t.nmax * K, -1);
  for (int a = 0; a < t.nmax; ++a)
    for (int q = t.rowptr[a]; q < t.rowptr[a + 1]; ++q) {
      const int2 e = t.ent[q];
      const int k = e.y & 0xffff;
      if (e.y < 0) {
        if (plus) plus[(size_t)a * K + k] = e.x;
      } else if (minus) {
        minus[(size_t)a * K + k] = e.x;
      }
    }
}

}  // namespace

extern "C" int pyqed_heom_tables_export(const pyqed_heom_plan* P, int32_t inst, int64_t* keys, int32_t* plus,
                                        int32_t* minus) {
  if (!P || inst < 0 || inst >= P->n_inst) return fail(PYQED_EINVAL, "tables_export: bad plan or instance");
  export_tables(*P->tabs.at({P->inst_K[inst], P->inst_L[inst]}), keys, plus, minus);
  return PYQED_OK;
}


extern "C" int pyqed_heom_tables_size(int32_t K, int32_t L, int64_t* nmax) {
  if (!nmax || K < 0 || K > kMaxK || L < 0 || L > kMaxL) return fail(PYQED_EINVAL, "tables_size: bad K or L");
  Binom C(K + L + 1, K + 1);
  const int64_t n = C.get(K + L, K);
  if (n <= 0 || n > (int64_t)INT32_MAX) return fail(PYQED_EINVAL, "hierarchy too large: C(K+L, L) exceeds 2^31-1");
  *nmax = n;
  return PYQED_OK;
}

extern "C" int pyqed_heom_tables_build(int32_t K, int32_t L, int64_t* keys, int32_t* plus, int32_t* minus) {
  if (K < 0 || K > kMaxK || L < 0 || L > kMaxL) return fail(PYQED_EINVAL, "tables_build: bad K or L");
  std::shared_ptr<const Tables> t;
  int rc = cached_tables(K, L, t);
  if (rc) return rc;
  export_tables(*t, keys, plus, minus);
  return PYQED_OK;
}

namespace {

StageParams base_params(const pyqed_heom_plan* P) {
  StageParams p;
  std::memset(&p, 0, sizeof(p));
  p.insts = P->d_inst;
  p.items = P->d_items;
  p.rowptr = P->d_rowptr;
  p.ent = P->d_ent;
  p.coef = P->d_coef;
  p.mode = P->d_mode;
  p.mat = P->d_mat;
  p.resident = P->d_res;
  p.site = P->d_site;
  p.rowbuf = P->rowbuf;
  p.colbuf = P->colbuf;
  p.coefp = P->d_coefp;
  p.traj_stride = P->traj_stride;
  p.d2max = P->d2max;
  p.d2pad = P->d2pad;
  p.qmax = P->qmax;
  p.Kmax = P->Kmax;
  p.Lmax = P->Lmax;
  p.save_slot = -1;
  p.col_override = -1;
  return p;
}

template <typename Kern>
int launch(Kern k, int grid, size_t lds, hipStream_t s, const StageParams& p, int threads = kThreads) {
  // raise the kernel's dynamic-LDS limit only when a larger size is first needed, so launches that
  // are being captured into a graph make no attribute calls
  static std::mutex mu;
  static std::map<void*, size_t> limit;
  if (lds > 65536) {
    std::lock_guard<std::mutex> g(mu);
    size_t& cur = limit[kernel_ptr(k)];
    if (lds > cur) {
      PQA_HIP(hipFuncSetAttribute(kernel_ptr(k), hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
      cur = lds;
    }
  }
  hipLaunchKernelGGL(k, dim3(grid), dim3(threads), lds, s, p);
  PQA_HIP(hipGetLastError());
  return PYQED_OK;
}

int stream_grid(const pyqed_heom_plan* P, int n_items, int c) {
  int g = std::min(n_items, P->num_cu * P->cls_bpc[c]);
  g = ((g + 7) / 8) * 8;
  return std::max(g, 8);
}

// One RK4 stage over every streamed item class: generic kernel, projector kernel with VALU
// commutator, projector kernel with MFMA commutator (FP32 only; FP64 uses VALU).
// Event pair around one class launch when the plan records per-launch timing (never while the
// stream is being captured into a graph).
struct LaunchTimer {
  pyqed_heom_plan* P;
  hipStream_t st;
  int k = -1;
  LaunchTimer(pyqed_heom_plan* P_, hipStream_t st_, int tag) : P(P_), st(st_) {
    if (!P->timing || tag < 0) return;
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(st, &cs) != hipSuccess || cs != hipStreamCaptureStatusNone) return;
    k = (int)P->ev_tag.size();
    while ((int)P->ev_pool.size() < 2 * k + 2) {
      hipEvent_t e;
      if (hipEventCreate(&e) != hipSuccess) {
        k = -1;
        return;
      }
      P->ev_pool.push_back(e);
    }
    P->ev_tag.push_back(tag);
    (void)hipEventRecord(P->ev_pool[2 * k], st);
  }
  ~LaunchTimer() {
    if (k >= 0) (void)hipEventRecord(P->ev_pool[2 * k + 1], st);
  }
};

// One RK4 stage over the streamed item classes c_lo <= c < c_hi.  TI is the stage-input precision:
// FP64 for stage 0 of an unpacked plan (it reads y), the stage precision otherwise.
template <typename T, int S, typename TI, bool PACK>
int launch_stage_ti(pyqed_heom_plan* P, const StageParams& p0, hipStream_t st, int c_lo, int c_hi) {
  for (int c = c_lo; c < c_hi; ++c) {
    const int n = P->cls_begin[c + 1] - P->cls_begin[c];
    if (n == 0) continue;
    StageParams p = p0;
    bool one = false;  // a single streamed instance in this projector class
    if (c > 0) {  // projector kernels: LDS dimensions of this class
      p.d2pad = P->pd[c].d2pad;
      p.rowbuf = P->pd[c].rowbuf;
      p.colbuf = P->pd[c].colbuf;
      p.Kmax = P->pd[c].Kmax;
      p.Lmax = P->pd[c].Lmax;
      if (P->one_inst[c] >= 0) {  // the class's items are consecutive ADO ranges of one instance
        one = true;
        p.inst0 = P->inst[P->one_inst[c]];
        p.one_per = P->items[P->cls_begin[c]].nado;
        p.ent_fix = P->d_entfix[c];
        p.fix_stride = P->fix_stride[c];
      }
    }
    p.item_begin = P->cls_begin[c];
    p.item_end = P->cls_begin[c + 1];
    p.use_xcd = 1;
    std::memcpy(p.xr, P->xr[c], sizeof(p.xr));
    const int grid = stream_grid(P, n, c);
    int rc;
    // the projector kernel keeps neighbour rows (and columns) in LDS; when they do not fit, the
    // generic kernel (which any instance can use) takes the class
    const size_t plds = c == 0 ? 0 : proj_lds_bytes(P, c, sizeof(T));
    LaunchTimer timer(P, st, S < 4 ? kClasses * S + c : -1);  // right-hand-side launches are not timed
    constexpr bool F32 = std::is_same<T, float>::value;
    if (c == 0 || plds > (size_t)P->max_lds) {
      rc = launch(heom_stage_kernel<T, TI, T, S, PACK>, grid, stream_lds_bytes(P, sizeof(T)), st, p);
    } else if constexpr (PACK) {  // packed plans are Hermitian
      const size_t wlds = c == 3 ? proj_block_bytes(P->pd[3].d2pad, P->pd[3].Kmax, P->pd[3].Lmax, sizeof(float), true) +
                                       wide_ado_bytes(P->pd[3].d2pad, P->pd[3].rowbuf)
                                 : 0;
      if (c == 2 && F32)
        rc = launch(heom_proj_kernel<T, TI, T, S, F32, true, true, 1>, grid, plds, st, p);
      else if (c == 3 && F32 && (P->variant & 2) && wlds <= (size_t)P->max_lds)
        rc = launch(heom_wide_kernel<S>, grid, wlds, st, p);
      else if (c == 3 && F32 && one)
        rc = launch(heom_proj_kernel<T, TI, T, S, F32, true, true, F32 ? 2 : 1, F32>, grid, plds, st, p);
      else if (c == 3 && F32)
        rc = launch(heom_proj_kernel<T, TI, T, S, F32, true, true, F32 ? 2 : 1>, grid, plds, st, p);
      else if (F32 && one)
        rc = launch(heom_proj_kernel<T, TI, T, S, false, true, true, 1, F32>, grid, plds, st, p);
      else
        rc = launch(heom_proj_kernel<T, TI, T, S, false, true, true, 1>, grid, plds, st, p);
    } else if (P->hermitian && c == 2 && F32 && (P->variant & 1) && pipe_lds_bytes(P, c) <= (size_t)P->max_lds) {
      rc = launch(heom_pipe_kernel<S>, grid, pipe_lds_bytes(P, c), st, p);
    } else if (P->hermitian) {
      if (c == 2 && F32)
        rc = launch(heom_proj_kernel<T, TI, T, S, F32, true, false, 1>, grid, plds, st, p);
      else if (c == 3 && F32)
        rc = launch(heom_proj_kernel<T, TI, T, S, F32, true, false, F32 ? 2 : 1>, grid, plds, st, p);
      else
        rc = launch(heom_proj_kernel<T, TI, T, S, false, true, false, 1>, grid, plds, st, p);
    } else {
      if (c == 2 && F32)
        rc = launch(heom_proj_kernel<T, TI, T, S, F32, false, false, 1>, grid, plds, st, p);
      else if (c == 3 && F32)
        rc = launch(heom_proj_kernel<T, TI, T, S, F32, false, false, F32 ? 2 : 1>, grid, plds, st, p);
      else
        rc = launch(heom_proj_kernel<T, TI, T, S, false, false, false, 1>, grid, plds, st, p);
    }
    if (rc) return rc;
  }
  return PYQED_OK;
}

template <typename T, int S>
int launch_stage(pyqed_heom_plan* P, const StageParams& p0, hipStream_t st, int c_lo = 0, int c_hi = kClasses) {
  if (P->packed) return launch_stage_ti<T, S, T, true>(P, p0, st, c_lo, c_hi);
  // unpacked stage 0 reads y itself (FP64)
  return launch_stage_ti<T, S, typename std::conditional<S == 0, double, T>::type, false>(P, p0, st, c_lo, c_hi);
}

// The stages of one RK4 step over the streamed classes c_lo <= c < c_hi, on stream st.
template <typename T>
int launch_step(pyqed_heom_plan* P, StageParams p, const void* x1, void* x0, bool save, int64_t slot,
                hipStream_t st, int c_lo, int c_hi) {
  int rc;
  p.xin = P->packed ? x1 : (const void*)p.y;
  p.xout = x0;
  if ((rc = launch_stage<T, 0>(P, p, st, c_lo, c_hi))) return rc;
  p.xin = x0;
  p.xout = (void*)x1;
  if ((rc = launch_stage<T, 1>(P, p, st, c_lo, c_hi))) return rc;
  p.xin = x1;
  p.xout = x0;
  if ((rc = launch_stage<T, 2>(P, p, st, c_lo, c_hi))) return rc;
  p.xin = x0;
  p.xout = P->packed ? (void*)x1 : nullptr;
  p.save_slot = save ? slot : -1;
  return launch_stage<T, 3>(P, p, st, c_lo, c_hi);
}

// Packed mode: the stage-0 input of the first step, X = y as full matrices in the stage precision
// (stage 3 writes it for every later step).
template <typename T>
__global__ __launch_bounds__(kThreads) void heom_unpack_kernel(StageParams p) {
  cpx_store<T>* xout = (cpx_store<T>*)p.xout;
  for (int it = p.item_begin + blockIdx.x; it < p.item_end; it += gridDim.x) {
    const WorkItem w = ro_load(p.items + it);
    const InstDev I = ro_load(p.insts + w.inst);
    const int d = I.d, d2 = I.d2;
    const int n = w.nado * d2;
    for (int q = threadIdx.x; q < n; q += kThreads) {
      const int a = w.ado0 + q / d2, e = q - (q / d2) * d2, i = e / d;
      const double2 v = ld_state(p.y, state_ref<true>(I, a, i, e - i * d));
      st(xout + I.off + (int64_t)a * d2 + e, mk<T>((T)v.x, (T)v.y));
    }
  }
}

template <typename T>
int run_typed(pyqed_heom_plan* P, double2* y, char* ws, const double* field, int64_t step0, int64_t nsteps,
              double dt, double2* traj, int64_t save_every, hipStream_t st) {
  const int64_t n = P->n_elems;
  double2* ynew = (double2*)ws;
  cpx_store<T>* x0 = (cpx_store<T>*)(ws + ws_state_bytes(P));
  cpx_store<T>* x1 = x0 + n;
  StageParams p = base_params(P);
  p.dt = dt;
  p.y = y;
  p.ynew = ynew;
  p.nstate = P->n_state;
  p.traj = traj;
  p.field = field;
  if (!P->resident.empty()) {
    PQA_HIP(hipEventRecord(P->ev_fork, st));
    PQA_HIP(hipStreamWaitEvent(P->side, P->ev_fork, 0));
    StageParams r = p;
    r.step = step0;
    r.nsteps = nsteps;
    r.save_every = save_every;
    size_t lds = 0;
    for (int i : P->resident)
      lds = std::max(lds, resident_lds_bytes(P, (int64_t)P->inst[i].nmax * P->inst[i].d2, P->res_extra[i], sizeof(T)));
    // four waves per instance: an element's right-hand side is a chain of dependent LDS reads, so
    // even small instances run faster with one element per lane than with one wave and cheaper barriers
    int rc = P->packed ? launch(heom_resident_kernel<T, true>, (int)P->resident.size(), lds, P->side, r)
                       : launch(heom_resident_kernel<T, false>, (int)P->resident.size(), lds, P->side, r);
    if (rc) return rc;
    PQA_HIP(hipEventRecord(P->ev_join, P->side));
  }
  if (P->n_items_stream > 0) {
    // packed mode: stage 0 reads x1, the full-matrix copy of y that stage 3 leaves there
    if (P->packed) {
      StageParams u = p;
      u.item_begin = 0;
      u.item_end = P->n_items_stream;
      u.xout = x1;
      const int rc = launch(heom_unpack_kernel<T>, std::max(1, std::min(P->n_items_stream, P->num_cu * 8)), 0, st, u);
      if (rc) return rc;
    }
    // the small classes on their own stream when the other classes have work too
    constexpr int cs = PQA_SIDE_CLASSES;
    const bool split = P->cls_begin[cs] > P->cls_begin[0] && P->cls_begin[kClasses] > P->cls_begin[cs];
    if (split) {
      PQA_HIP(hipEventRecord(P->ev_fork0, st));
      PQA_HIP(hipStreamWaitEvent(P->side0, P->ev_fork0, 0));
    }
    for (int64_t s = 0; s < nsteps; ++s) {
      p.step = step0 + s;
      const bool save = traj && ((s + 1) % save_every) == 0;
      const int64_t slot = save ? (s + 1) / save_every - 1 : -1;
      int rc;
      if (split) {
        if ((rc = launch_step<T>(P, p, x1, x0, save, slot, P->side0, 0, cs))) return rc;
        if ((rc = launch_step<T>(P, p, x1, x0, save, slot, st, cs, kClasses))) return rc;
      } else if ((rc = launch_step<T>(P, p, x1, x0, save, slot, st, 0, kClasses))) {
        return rc;
      }
    }
    if (split) {
      PQA_HIP(hipEventRecord(P->ev_join0, P->side0));
      PQA_HIP(hipStreamWaitEvent(st, P->ev_join0, 0));
    }
  }
  if (!P->resident.empty()) PQA_HIP(hipStreamWaitEvent(st, P->ev_join, 0));
  return PYQED_OK;
}

}  // namespace

extern "C" int pyqed_heom_run(pyqed_heom_plan* P, int32_t precision, void* y, void* workspace, int64_t ws_bytes,
                              const double* field, int64_t step0, int64_t nsteps, double dt, void* traj,
                              int64_t save_every, void* stream) {
  if (!P || !y || !workspace) return fail(PYQED_EINVAL, "heom_run: null plan, state or workspace");
  if (nsteps < 0 || step0 < 0) return fail(PYQED_EINVAL, "heom_run: negative step count");
  if (save_every <= 0) save_every = 1;
  const int64_t need = workspace_bytes(P, precision == PYQED_PREC_FP64 ? sizeof(double2) : sizeof(float2));
  if (precision != PYQED_PREC_FP64 && precision != PYQED_PREC_MIXED)
    return fail(PYQED_EINVAL, "heom_run: precision must be PYQED_PREC_FP64 or PYQED_PREC_MIXED");
  if (ws_bytes < need)
    return fail(PYQED_EINVAL, "heom_run: workspace too small (" + std::to_string(ws_bytes) + " < " +
                                  std::to_string(need) + " bytes)");
  if (nsteps == 0) return PYQED_OK;
  hipStream_t st = (hipStream_t)stream;
  if (precision == PYQED_PREC_FP64)
    return run_typed<double>(P, (double2*)y, (char*)workspace, field, step0, nsteps, dt, (double2*)traj, save_every,
                             st);
  return run_typed<float>(P, (double2*)y, (char*)workspace, field, step0, nsteps, dt, (double2*)traj, save_every, st);
}

namespace {

// ------------------------------------------------------------------ adjoint contraction
// For the gradients of the discrete adjoint: per instance and field column c,
//   G[c] = sum_{s : col(s) = c} w_s sum_a [X_s,a, U_s,a^H],   col = (0, 1, 1, 2) for s = 0..3,
// where X_s is the forward input of RK4 stage s, U_s the matching adjoint stage input and a runs over
// the instance's ADOs.  Then dJ/dH (as a cotangent) is i sum G^H and dJ/df = Re(-i Tr(mu G)) for a
// field f entering H(t) = H + mu f.
struct AdjParams {
  const InstDev* insts;
  const WorkItem* items;
  const int2* inst_items;
  const double2* x[4];
  const double2* u[4];
  double w[4];
  double2* part;  // [nitems][3][d2max]
  double2* G;     // [3][traj_stride]
  int nitems, d2max;
  int64_t traj_stride;
};

constexpr int kAdjMaxEl = 4;  // elements per lane: I.grp >= d2 lanes per ADO for d2 <= 64, else 256 lanes

__global__ __launch_bounds__(kThreads) void heom_adjoint_contract_kernel(AdjParams q) {
  extern __shared__ __align__(16) char smem[];
  const int tid = threadIdx.x;
  for (int it = blockIdx.x; it < q.nitems; it += gridDim.x) {
    const WorkItem wi = q.items[it];
    const InstDev I = q.insts[wi.inst];
    const int d = I.d, d2 = I.d2, G = d2 <= 64 ? I.grp : kThreads;
    const int nslot = kThreads / G, slot = tid / G, le = tid - slot * G;
    double2* sx = (double2*)smem + (int64_t)slot * 2 * d2;
    double2* su = sx + d2;
    cpx<double> acc[3][kAdjMaxEl];
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
      for (int t = 0; t < kAdjMaxEl; ++t) acc[c][t] = mk(0.0, 0.0);
    for (int a0 = wi.ado0; a0 < wi.ado0 + wi.nado; a0 += nslot) {
      const int a = a0 + slot;
      const bool act = a < wi.ado0 + wi.nado;
#pragma unroll
      for (int st = 0; st < 4; ++st) {
        if (!q.x[st] || !q.u[st]) continue;  // block-uniform
        const int c = st == 0 ? 0 : (st == 3 ? 2 : 1);
        __syncthreads();
        if (act) {
          const int64_t base = I.off + (int64_t)a * d2;
          for (int e = le; e < d2; e += G) {
            sx[e] = q.x[st][base + e];
            su[e] = q.u[st][base + e];
          }
        }
        __syncthreads();
        if (act) {
#pragma unroll
          for (int t = 0; t < kAdjMaxEl; ++t) {
            const int e = le + t * G;
            if (e >= d2) break;
            const int i = e / d, j = e - i * d;
            cpx<double> r = mk(0.0, 0.0);
            for (int k = 0; k < d; ++k) {
              const double2 xa = sx[i * d + k], ua = su[j * d + k];  // x_ik conj(u_jk)
              const double2 ub = su[k * d + i], xb = sx[k * d + j];  // conj(u_ki) x_kj
              r.x += xa.x * ua.x + xa.y * ua.y - (ub.x * xb.x + ub.y * xb.y);
              r.y += xa.y * ua.x - xa.x * ua.y - (ub.x * xb.y - ub.y * xb.x);
            }
            acc[c][t] = acc[c][t] + r * q.w[st];
          }
        }
      }
    }
    // sum over the slots of the block, then one partial per item
    __syncthreads();
    double2* red = (double2*)smem;  // [nslot][3][d2]
#pragma unroll
    for (int t = 0; t < kAdjMaxEl; ++t) {
      const int e = le + t * G;
      if (e >= d2) break;
#pragma unroll
      for (int c = 0; c < 3; ++c) red[((int64_t)slot * 3 + c) * d2 + e] = make_double2(acc[c][t].x, acc[c][t].y);
    }
    __syncthreads();
    for (int e = tid; e < 3 * d2; e += kThreads) {
      double2 sum = make_double2(0.0, 0.0);
      for (int sl = 0; sl < nslot; ++sl) {
        const double2 v = red[(int64_t)sl * 3 * d2 + e];
        sum.x += v.x;
        sum.y += v.y;
      }
      const int c = e / d2, f = e - c * d2;
      q.part[((int64_t)it * 3 + c) * q.d2max + f] = sum;
    }
  }
}

// G[c][traj_off + f] = sum over the instance's items of their partials.  Block (instance, x) covers 4
// of the instance's 3 d^2 values; its 256 threads are 4 values x 64 item lanes.  Lane ln sums items
// ln, ln + 64, ... into four accumulators (item k into accumulator (k / 64) mod 4, so four loads are
// in flight), and the 64 lane sums are added by a fixed-shape tree: the order never depends on
// timing, so results are deterministic.
constexpr int kRedVals = 4, kRedLanes = kThreads / kRedVals;

__global__ __launch_bounds__(kThreads) void heom_adjoint_reduce_kernel(AdjParams q) {
  __shared__ double2 red[kRedLanes][kRedVals];
  const int i = blockIdx.y;
  const InstDev I = q.insts[i];
  const int2 r = q.inst_items[i];
  const int v = threadIdx.x % kRedVals, ln = threadIdx.x / kRedVals;
  const int e = blockIdx.x * kRedVals + v;
  const bool act = e < 3 * I.d2;
  const int c = act ? e / I.d2 : 0, f = act ? e - c * I.d2 : 0;
  double2 s4[4];
#pragma unroll
  for (int u = 0; u < 4; ++u) s4[u] = make_double2(0.0, 0.0);
  if (act) {
    int it = r.x + ln;
    for (; it + 3 * kRedLanes < r.y; it += 4 * kRedLanes) {
      double2 w[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) w[u] = q.part[((int64_t)(it + u * kRedLanes) * 3 + c) * q.d2max + f];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        s4[u].x += w[u].x;
        s4[u].y += w[u].y;
      }
    }
    for (int u = 0; it < r.y; it += kRedLanes, ++u) {
      const double2 w = q.part[((int64_t)it * 3 + c) * q.d2max + f];
      s4[u].x += w.x;
      s4[u].y += w.y;
    }
  }
  red[ln][v] = make_double2((s4[0].x + s4[1].x) + (s4[2].x + s4[3].x), (s4[0].y + s4[1].y) + (s4[2].y + s4[3].y));
  __syncthreads();
  for (int h = kRedLanes / 2; h > 0; h >>= 1) {
    if (ln < h) {
      red[ln][v].x += red[ln + h][v].x;
      red[ln][v].y += red[ln + h][v].y;
    }
    __syncthreads();
  }
  if (ln == 0 && act) q.G[c * q.traj_stride + I.traj_off + f] = red[0][v];
}

size_t adjoint_lds_bytes(const pyqed_heom_plan* P) {
  // staging: nslot * 2 * d2 <= 2 * max(256, d2max) complex; reduction: nslot * 3 * d2 <= 3 * max(256, d2max)
  return (size_t)3 * std::max(256, P->d2max) * sizeof(double2);
}

}  // namespace

extern "C" int pyqed_heom_adjoint_work_bytes(const pyqed_heom_plan* P, int64_t* bytes) {
  if (!P || !bytes) return fail(PYQED_EINVAL, "heom_adjoint_work_bytes: null argument");
  *bytes = (int64_t)P->items.size() * 3 * P->d2max * (int64_t)sizeof(double2);
  return PYQED_OK;
}

extern "C" int pyqed_heom_adjoint_contract(pyqed_heom_plan* P, const void* const* x, const void* const* u,
                                           const double* w, void* G, void* work, int64_t work_bytes, void* stream) {
  if (!P || !x || !u || !w || !G || !work) return fail(PYQED_EINVAL, "heom_adjoint_contract: null argument");
  int64_t need = 0;
  pyqed_heom_adjoint_work_bytes(P, &need);
  if (work_bytes < need)
    return fail(PYQED_EINVAL, "heom_adjoint_contract: workspace of " + std::to_string(work_bytes) + " bytes < " +
                                  std::to_string(need) + " needed");
  if (P->items.empty()) return PYQED_OK;
  AdjParams q;
  std::memset(&q, 0, sizeof(q));
  q.insts = P->d_inst;
  q.items = P->d_items;
  q.inst_items = P->d_inst_items;
  for (int s = 0; s < 4; ++s) {
    q.x[s] = (const double2*)x[s];
    q.u[s] = (const double2*)u[s];
    q.w[s] = w[s];
  }
  q.part = (double2*)work;
  q.G = (double2*)G;
  q.nitems = (int)P->items.size();
  q.d2max = P->d2max;
  q.traj_stride = P->traj_stride;
  const hipStream_t st = (hipStream_t)stream;
  const int grid = std::max(1, std::min(q.nitems, P->num_cu * 8));
  hipLaunchKernelGGL(heom_adjoint_contract_kernel, dim3(grid), dim3(kThreads), adjoint_lds_bytes(P), st, q);
  PQA_HIP(hipGetLastError());
  if (P->n_inst > 65535) return fail(PYQED_EINVAL, "heom_adjoint_contract: at most 65535 instances per plan");
  const unsigned nx = (unsigned)((3 * P->d2max + kRedVals - 1) / kRedVals);
  hipLaunchKernelGGL(heom_adjoint_reduce_kernel, dim3(nx, P->n_inst), dim3(kThreads), 0, st, q);
  PQA_HIP(hipGetLastError());
  return PYQED_OK;
}

namespace {
int rhs_launch(pyqed_heom_plan* P, StageParams p, int32_t col, void* stream) {
  p.item_begin = 0;
  p.item_end = (int)P->items.size();
  p.use_xcd = 0;
  p.col_override = col;
  const size_t lds = stream_lds_bytes(P, sizeof(double));
  if (!P->hermitian && P->n_proj > 0) {
    // streamed items through their class kernels (the projector classes on the FP64 projector
    // kernel, rows and columns of the neighbours from the input), then the LDS-resident instances'
    // items on the generic kernel.  Hermitian plans keep the generic kernel for every item: their
    // projector kernels read columns as conjugated rows, which holds only for Hermitian inputs.
    const int rc = launch_stage_ti<double, 4, double, false>(P, p, (hipStream_t)stream, 0, kClasses);
    if (rc || P->cls_begin[kClasses] == (int)P->items.size()) return rc;
    p.item_begin = P->cls_begin[kClasses];
    p.use_xcd = 0;
    const int nres = (int)P->items.size() - p.item_begin;
    return launch(heom_stage_kernel<double, double, double, 4, false>, std::max(1, std::min(nres, P->num_cu * 8)), lds,
                  (hipStream_t)stream, p);
  }
  const int grid = std::max(1, std::min((int)P->items.size(), P->num_cu * 8));
  return launch(heom_stage_kernel<double, double, double, 4, false>, grid, lds, (hipStream_t)stream, p);
}
}  // namespace

extern "C" int pyqed_heom_rhs(pyqed_heom_plan* P, const void* x, void* out, const double* field, int64_t step,
                              int32_t col, void* stream) {
  if (!P || !x || !out) return fail(PYQED_EINVAL, "heom_rhs: null argument");
  if (col < 0 || col > 2) return fail(PYQED_EINVAL, "heom_rhs: col must be 0, 1 or 2");
  StageParams p = base_params(P);
  p.field = field;
  p.step = step;
  p.xin = x;
  p.xout = out;
  return rhs_launch(P, p, col, stream);
}

extern "C" int pyqed_heom_rhs_combine(pyqed_heom_plan* P, const void* x, void* out, const void* base, double alpha,
                                      void* acc, double beta, int32_t acc_init, const double* field, int64_t step,
                                      int32_t col, void* stream) {
  if (!P || !x) return fail(PYQED_EINVAL, "heom_rhs_combine: null plan or input");
  if (!out && !acc) return fail(PYQED_EINVAL, "heom_rhs_combine: neither out nor acc given");
  if (base && !out) return fail(PYQED_EINVAL, "heom_rhs_combine: base given without out");
  if (out == x || acc == x || (base && base == out))
    return fail(PYQED_EINVAL, "heom_rhs_combine: out and acc must not alias the input, nor out the base");
  if (col < 0 || col > 2) return fail(PYQED_EINVAL, "heom_rhs_combine: col must be 0, 1 or 2");
  StageParams p = base_params(P);
  p.field = field;
  p.step = step;
  p.xin = x;
  p.xout = out;
  p.cbase = (const double2*)base;
  p.calpha = alpha;
  p.cacc = (double2*)acc;
  p.cbeta = beta;
  p.cinit = acc_init ? 1 : 0;
  return rhs_launch(P, p, col, stream);
}
