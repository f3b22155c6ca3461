// Lindblad master equation in matrix form, for any d (the Liouville-space kernel of liouville.hip
// holds vec(rho) and its N x N generator and serves N = d*d <= 1024):
//
//   d rho/dt = -i (Hs rho - rho Hs^dag) + sum_k c_k (rho c_k^dag),
//   Hs = Heff - sum_i f_i(step) H_i,   Heff = H0 - (i/2) sum_k c_k^dag c_k  (formed on the host),
//
// which is the reference's liouvillian (pyqed/oqs.py:697-714: -i[H, rho] + sum_k c rho c^dag
// - 1/2 {c^dag c, rho}) with the anticommutators folded into one effective Hamiltonian, stepped by
// the RK4 of pyqed/phys.py:1051-1063 in its summation order, with the drive held constant over a
// step as _lindblad_driven (oqs.py:1699-1806) evaluates it.
//
// One workgroup owns one instance for the whole run.  Its stage inputs, RK4 accumulator and the
// products rho c_k^dag live in a per-instance workspace in global memory (L2-resident at these
// sizes); the right-hand side is a sum of complex FP64 products computed in 64 x 64 output tiles,
// each thread accumulating a 4 x 4 block from k-panels of 16 staged in LDS, and the RK4 update is
// applied as the tile's epilogue.  FP64 VALU: on this part its FMA rate equals the FP64 MFMA rate.

#include <cstring>

#include "common.h"

namespace pqa {
namespace {

constexpr int kLT = 64;        // output tile edge
constexpr int kLK = 16;        // k-panel depth
constexpr int kLThreads = 256;  // 16 x 16 threads, each a 4 x 4 block of the tile

struct LbParams {
  int d, nc, ndrive, n_obs;
  int64_t nsteps, save_every, ws_stride;
  double dt;
  const double2* Heff;   // [d*d]
  const double2* HeffD;  // [d*d] Heff^dag
  const double2* C;      // [nc][d*d]
  const double2* CD;     // [nc][d*d] c_k^dag
  const double2* Hd;     // [ndrive][d*d] drive operators
  const double2* HdD;    // [ndrive][d*d] their adjoints
  const double* field;   // [n_inst][nsteps][ndrive]
  double2* rho;          // [n_inst][d*d], advanced in place
  const double2* obs;    // [n_obs][d*d], O^T flattened: Tr(O rho) = sum_e obs[e] rho[e]
  double2* obs_out;      // [n_inst][nsteps+1][n_obs]
  double2* rho_out;      // [n_inst][nsteps/save_every][d*d] or NULL
  double2* ws;           // per instance (ws_stride elements): X0, X1, acc, T[nc], Hs, HsD
};

struct LbLds {
  double2 a[kLK][kLT + 1];  // A panel, transposed (a[k][row]); the pad keeps the staging stores apart
  double2 b[kLK][kLT];      // B panel (b[k][col])
  double2 red[kLThreads];
};

__device__ __forceinline__ double2 cmul2(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// acc (this thread's block: rows r0 + 4 ty + r, columns c0 + 4 tx + c) += (s A)(rows, :) B(:, cols),
// A and B row-major d x d in global memory; s scales A as it is staged.
__device__ void gemm_acc(const double2* __restrict__ A, const double2* __restrict__ B, double2 s, int d, int r0, int c0,
                         double2 (&acc)[4][4], LbLds& L) {
  const int tid = threadIdx.x, tx = tid & 15, ty = tid >> 4;
  for (int k0 = 0; k0 < d; k0 += kLK) {
    __syncthreads();  // every read of the previous panel is done
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int idx = tid + kLThreads * q, row = idx >> 4, k = idx & 15;
      const int gr = r0 + row, gk = k0 + k;
      const double2 v = (gr < d && gk < d) ? A[(int64_t)gr * d + gk] : make_double2(0.0, 0.0);
      L.a[k][row] = cmul2(s, v);
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int idx = tid + kLThreads * q, k = idx >> 6, col = idx & 63;
      const int gk = k0 + k, gc = c0 + col;
      L.b[k][col] = (gk < d && gc < d) ? B[(int64_t)gk * d + gc] : make_double2(0.0, 0.0);
    }
    __syncthreads();
#pragma unroll 4
    for (int k = 0; k < kLK; ++k) {
      double2 a[4], b[4];
#pragma unroll
      for (int r = 0; r < 4; ++r) a[r] = L.a[k][4 * ty + r];
#pragma unroll
      for (int c = 0; c < 4; ++c) b[c] = L.b[k][4 * tx + c];
#pragma unroll
      for (int r = 0; r < 4; ++r)
#pragma unroll
        for (int c = 0; c < 4; ++c) {
          acc[r][c].x = fma(a[r].x, b[c].x, fma(-a[r].y, b[c].y, acc[r][c].x));
          acc[r][c].y = fma(a[r].x, b[c].y, fma(a[r].y, b[c].x, acc[r][c].y));
        }
    }
  }
}

__device__ __forceinline__ void zero_acc(double2 (&acc)[4][4]) {
#pragma unroll
  for (int r = 0; r < 4; ++r)
#pragma unroll
    for (int c = 0; c < 4; ++c) acc[r][c] = make_double2(0.0, 0.0);
}

// sum over the block of one value per thread (result valid in thread 0)
__device__ double2 block_sum(double2 v, LbLds& L) {
  __syncthreads();
  L.red[threadIdx.x] = v;
  __syncthreads();
  for (int h = kLThreads / 2; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) {
      L.red[threadIdx.x].x += L.red[threadIdx.x + h].x;
      L.red[threadIdx.x].y += L.red[threadIdx.x + h].y;
    }
    __syncthreads();
  }
  return L.red[0];
}

__global__ __launch_bounds__(kLThreads) void lindblad_matrix_kernel(LbParams p) {
  __shared__ LbLds L;
  const int inst = blockIdx.x, tid = threadIdx.x, tx = tid & 15, ty = tid >> 4;
  const int d = p.d, nc = p.nc;
  const int64_t d2 = (int64_t)d * d;
  double2* y = p.rho + inst * d2;
  double2* ws = p.ws + inst * p.ws_stride;
  double2* X0 = ws;
  double2* X1 = X0 + d2;
  double2* acR = X1 + d2;
  double2* T = acR + d2;  // [nc][d2]
  double2* Hs = T + nc * d2;
  double2* HsD = Hs + d2;
  const bool driven = p.ndrive > 0;
  const double2* H = driven ? Hs : p.Heff;
  const double2* HD = driven ? HsD : p.HeffD;
  const double dt = p.dt, dt2 = dt / 2.0;
  const int ntile = (d + kLT - 1) / kLT;
  const double2 mi = make_double2(0.0, -1.0), pi = make_double2(0.0, 1.0), one = make_double2(1.0, 0.0);

  auto observe = [&](int64_t row) {
    for (int o = 0; o < p.n_obs; ++o) {
      double2 s = make_double2(0.0, 0.0);
      const double2* O = p.obs + o * d2;
      for (int64_t e = tid; e < d2; e += kLThreads) {
        const double2 v = cmul2(O[e], y[e]);
        s.x += v.x;
        s.y += v.y;
      }
      const double2 t = block_sum(s, L);
      if (tid == 0) p.obs_out[((int64_t)inst * (p.nsteps + 1) + row) * p.n_obs + o] = t;
    }
  };
  if (p.n_obs) observe(0);
  int64_t since = 0;
  for (int64_t s = 0; s < p.nsteps; ++s) {
    if (driven) {
      const double* f = p.field + ((int64_t)inst * p.nsteps + s) * p.ndrive;
      for (int64_t e = tid; e < d2; e += kLThreads) {
        double2 h = p.Heff[e], hd = p.HeffD[e];
        for (int i = 0; i < p.ndrive; ++i) {
          const double2 a = p.Hd[i * d2 + e], b = p.HdD[i * d2 + e];
          h.x -= f[i] * a.x;
          h.y -= f[i] * a.y;
          hd.x -= f[i] * b.x;
          hd.y -= f[i] * b.y;
        }
        Hs[e] = h;
        HsD[e] = hd;
      }
    }
#pragma unroll 1
    for (int stage = 0; stage < 4; ++stage) {
      const double2* in = stage == 0 ? y : (stage == 2 ? X1 : X0);
      double2* out = stage == 1 ? X1 : X0;
      // phase 1: T_k = in c_k^dag
      for (int k = 0; k < nc; ++k)
        for (int tr = 0; tr < ntile; ++tr)
          for (int tc = 0; tc < ntile; ++tc) {
            double2 acc[4][4];
            zero_acc(acc);
            gemm_acc(in, p.CD + k * d2, one, d, tr * kLT, tc * kLT, acc, L);
#pragma unroll
            for (int r = 0; r < 4; ++r)
#pragma unroll
              for (int c = 0; c < 4; ++c) {
                const int i = tr * kLT + 4 * ty + r, j = tc * kLT + 4 * tx + c;
                if (i < d && j < d) T[k * d2 + (int64_t)i * d + j] = acc[r][c];
              }
          }
      // phase 2: the right-hand side -i H in + i in H^dag + sum_k c_k T_k, and the RK4 update
      for (int tr = 0; tr < ntile; ++tr)
        for (int tc = 0; tc < ntile; ++tc) {
          double2 acc[4][4];
          zero_acc(acc);
          gemm_acc(H, in, mi, d, tr * kLT, tc * kLT, acc, L);
          gemm_acc(in, HD, pi, d, tr * kLT, tc * kLT, acc, L);
          for (int k = 0; k < nc; ++k) gemm_acc(p.C + k * d2, T + k * d2, one, d, tr * kLT, tc * kLT, acc, L);
#pragma unroll
          for (int r = 0; r < 4; ++r)
#pragma unroll
            for (int c = 0; c < 4; ++c) {
              const int i = tr * kLT + 4 * ty + r, j = tc * kLT + 4 * tx + c;
              if (i >= d || j >= d) continue;
              const int64_t e = (int64_t)i * d + j;
              const double2 k = acc[r][c], yv = y[e];
              if (stage == 0) {
                acR[e] = k;
                out[e] = make_double2(yv.x + k.x * dt2, yv.y + k.y * dt2);
              } else if (stage < 3) {
                const double2 a = acR[e];
                acR[e] = make_double2(a.x + 2.0 * k.x, a.y + 2.0 * k.y);
                const double h = stage == 1 ? dt2 : dt;
                out[e] = make_double2(yv.x + k.x * h, yv.y + k.y * h);
              } else {
                // y is not a product input in stage 3 (its input is X0): update in place
                const double2 a = acR[e];
                y[e] = make_double2(yv.x + (a.x + k.x) / 6.0 * dt, yv.y + (a.y + k.y) / 6.0 * dt);
              }
            }
        }
      __syncthreads();  // this stage's outputs are in place for the next stage's products
    }
    if (p.n_obs) observe(s + 1);
    if (p.rho_out && ++since == p.save_every) {
      since = 0;
      const int64_t slot = (s + 1) / p.save_every - 1;
      double2* dst = p.rho_out + ((int64_t)inst * (p.nsteps / p.save_every) + slot) * d2;
      for (int64_t e = tid; e < d2; e += kLThreads) dst[e] = y[e];
    }
  }
}

}  // namespace
}  // namespace pqa

using namespace pqa;

extern "C" int pyqed_lindblad_work_bytes(int32_t d, int32_t nc, int32_t ndrive, int32_t n_inst, int64_t* bytes) {
  if (!bytes || d < 1 || nc < 0 || ndrive < 0 || n_inst < 1)
    return fail(PYQED_EINVAL, "lindblad_work_bytes: need d >= 1, nc >= 0, ndrive >= 0, n_inst >= 1");
  const int64_t d2 = (int64_t)d * d;
  *bytes = (int64_t)n_inst * (3 + nc + (ndrive > 0 ? 2 : 0)) * d2 * (int64_t)sizeof(double2);
  return PYQED_OK;
}

extern "C" int pyqed_lindblad_run(int32_t d, int32_t nc, const void* Heff, const void* HeffD, const void* C,
                                  const void* CD, int32_t ndrive, const void* Hd, const void* HdD, int32_t n_inst,
                                  void* rho, const double* field, int64_t nsteps, double dt, int32_t n_obs,
                                  const void* obs, void* obs_out, void* rho_out, int64_t save_every, void* work,
                                  int64_t work_bytes, void* stream) {
  if (d < 1 || d > 46340 || nc < 0 || ndrive < 0 || n_inst < 1 || !Heff || !HeffD || !rho || !work)
    return fail(PYQED_EINVAL, "lindblad_run: need 1 <= d <= 46340, nc >= 0, ndrive >= 0, n_inst >= 1 and "
                              "Heff, HeffD, rho, work");
  if (nc > 0 && (!C || !CD)) return fail(PYQED_EINVAL, "lindblad_run: nc > 0 needs C and CD");
  if (ndrive > 0 && (!Hd || !HdD || !field)) return fail(PYQED_EINVAL, "lindblad_run: drives need Hd, HdD and a field");
  if (n_obs < 0 || (n_obs > 0 && (!obs || !obs_out))) return fail(PYQED_EINVAL, "lindblad_run: bad observables");
  if (nsteps < 0) return fail(PYQED_EINVAL, "lindblad_run: negative step count");
  int64_t need = 0;
  pyqed_lindblad_work_bytes(d, nc, ndrive, n_inst, &need);
  if (work_bytes < need)
    return fail(PYQED_EINVAL, "lindblad_run: workspace of " + std::to_string(work_bytes) + " bytes < " +
                                  std::to_string(need));
  LbParams p;
  std::memset(&p, 0, sizeof(p));
  p.d = d;
  p.nc = nc;
  p.ndrive = ndrive;
  p.n_obs = n_obs;
  p.nsteps = nsteps;
  p.save_every = save_every > 0 ? save_every : 1;
  p.ws_stride = (int64_t)(3 + nc + (ndrive > 0 ? 2 : 0)) * d * d;
  p.dt = dt;
  p.Heff = (const double2*)Heff;
  p.HeffD = (const double2*)HeffD;
  p.C = (const double2*)C;
  p.CD = (const double2*)CD;
  p.Hd = (const double2*)Hd;
  p.HdD = (const double2*)HdD;
  p.field = field;
  p.rho = (double2*)rho;
  p.obs = (const double2*)obs;
  p.obs_out = (double2*)obs_out;
  p.rho_out = (double2*)rho_out;
  p.ws = (double2*)work;
  hipLaunchKernelGGL(lindblad_matrix_kernel, dim3(n_inst), dim3(kLThreads), 0, (hipStream_t)stream, p);
  PQA_HIP(hipGetLastError());
  return PYQED_OK;
}
