// Spectral sums of a diagonalised generator, for the eigen-decomposition propagators and
// correlation functions of pyqed (superoperator.Lindblad_solver, superoperator.py:455-772;
// RedfieldSolver.propagator, oqs.py:160-215):
//
//   out[t][a][b] = sum_j A[a][j] f(lambda_j, s_t) B[j][b]
//   f = exp(lambda s)          (kind 0, time domain: e^{Lt} = U1 diag(e^{lambda t}) U2)
//   f = -1 / (lambda + i s)    (kind 1, frequency domain: the Laplace transform of e^{lambda t})
//
// Two launches: the weights W[t][j] = f(lambda_j, s_t), then a batched complex GEMM per row a,
// out_a = W (ns x k) . C_a (k x ncol) with C_a[j][b] = A[a][j] B[j][b] formed as the tile is staged.
// The output is written through element strides (so_t, so_a, so_b), which gives both the
// [t][a][b] and the reference's [a][b][t] layouts.
#include <hip/hip_runtime.h>

#include <cstring>
#include <mutex>
#include <string>

#include "common.h"

using namespace pqa;

namespace {

constexpr int kTT = 64;   // t rows per tile
constexpr int kTB = 16;   // b columns per tile
constexpr int kTJ = 16;   // j per staged chunk
constexpr int kRT = 4;    // t rows per thread (register block)
constexpr int kThreads = (kTT / kRT) * kTB;  // 256

__global__ __launch_bounds__(256) void spectral_weights_kernel(int kind, int ns, int k, const double2* lam,
                                                               const double* s, double2* W) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (int64_t)ns * k) return;
  const int t = (int)(idx / k), j = (int)(idx - (int64_t)t * k);
  const double lr = lam[j].x, li = lam[j].y, st = s[t];
  double2 w;
  if (kind == 0) {
    const double m = exp(lr * st);
    double sn, cs;
    sincos(li * st, &sn, &cs);
    w = make_double2(m * cs, m * sn);
  } else {
    // -1 / (lr + i (li + s))
    const double re = lr, im = li + st, den = re * re + im * im;
    w = make_double2(-re / den, im / den);
  }
  W[idx] = w;
}

// RB columns per thread: RB = 1 gives a 64 x 16 tile (narrow outputs, e.g. the correlation
// functions with ncol = 1); RB = 4 a 64 x 64 tile with a 4 x 4 register block per thread, which
// halves the LDS reads per complex FMA for wide outputs (propagators).
// TFAST: consecutive lanes take consecutive t (rows tx + 16 r, columns ty + 16 q), so stores are
// coalesced when t is the contiguous output axis; otherwise consecutive lanes take consecutive b
// (rows 4 ty + r, columns tx + 16 q).
template <int RB, bool TFAST>
__global__ __launch_bounds__(kThreads) void spectral_contract_kernel(int ns, int k, int ncol, const double2* W,
                                                                      const double2* A, const double2* B,
                                                                      double2* out, int64_t so_t, int64_t so_a,
                                                                      int64_t so_b) {
  constexpr int TBW = kTB * RB;              // tile width in b
  __shared__ cpx<double> sW[kTJ][kTT + 1];   // W[t0 + r][j0 + c] stored as [c][r]
  __shared__ cpx<double> sC[kTJ][TBW + 1];   // A[a][j0 + c] * B[j0 + c][b0 + q]
  const int tid = threadIdx.x, tx = tid % kTB, ty = tid / kTB;
  const int t0 = blockIdx.x * kTT, b0 = blockIdx.y * TBW, a = blockIdx.z;
  cpx<double> acc[kRT][RB];
#pragma unroll
  for (int r = 0; r < kRT; ++r)
#pragma unroll
    for (int q = 0; q < RB; ++q) acc[r][q] = mk(0.0, 0.0);
  const double2* Arow = A + (int64_t)a * k;
  for (int j0 = 0; j0 < k; j0 += kTJ) {
    // stage W: 64 x 16 values, 4 per thread
    for (int e = tid; e < kTT * kTJ; e += kThreads) {
      const int r = e / kTJ, c = e - (e / kTJ) * kTJ;
      const int t = t0 + r, j = j0 + c;
      double2 v = make_double2(0.0, 0.0);
      if (t < ns && j < k) v = W[(int64_t)t * k + j];
      sW[c][r] = mk(v.x, v.y);
    }
    // stage C: 16 x TBW values, RB per thread
    for (int e = tid; e < kTJ * TBW; e += kThreads) {
      const int c = e / TBW, q = e - (e / TBW) * TBW;
      const int j = j0 + c, b = b0 + q;
      cpx<double> v = mk(0.0, 0.0);
      if (j < k && b < ncol) {
        const double2 x = Arow[j], y = B[(int64_t)j * ncol + b];
        v = mk(x.x, x.y) * mk(y.x, y.y);
      }
      sC[c][q] = v;
    }
    __syncthreads();
#pragma unroll 4
    for (int c = 0; c < kTJ; ++c) {
      cpx<double> cv[RB], wv[kRT];
#pragma unroll
      for (int q = 0; q < RB; ++q) cv[q] = sC[c][(TFAST ? ty : tx) + kTB * q];
#pragma unroll
      for (int r = 0; r < kRT; ++r) wv[r] = sW[c][TFAST ? tx + kTB * r : ty * kRT + r];
#pragma unroll
      for (int r = 0; r < kRT; ++r)
#pragma unroll
        for (int q = 0; q < RB; ++q) cmac(acc[r][q], wv[r], cv[q]);
    }
    __syncthreads();
  }
#pragma unroll
  for (int q = 0; q < RB; ++q) {
    const int b = b0 + (TFAST ? ty : tx) + kTB * q;
    if (b >= ncol) continue;
#pragma unroll
    for (int r = 0; r < kRT; ++r) {
      const int t = t0 + (TFAST ? tx + kTB * r : ty * kRT + r);
      if (t < ns)
        out[(int64_t)t * so_t + (int64_t)a * so_a + (int64_t)b * so_b] = make_double2(acc[r][q].x, acc[r][q].y);
    }
  }
}

typedef double f64x4 __attribute__((ext_vector_type(4)));

// Wide outputs on FP64 MFMA (v_mfma_f64_16x16x4_f64: D[16][16] += A[16][4] B[4][16]; lane l holds
// A[l % 16][l / 16], B[l / 16][l % 16] and D[l / 16 + 4 r][l % 16], r = 0..3 -- the output rows of
// the FP64 form interleave by 4, unlike the FP32 16x16x4 form's D[4 (l / 16) + r][l % 16]).  A block of four
// wavefronts computes a 64 (t) x 64 (b) tile of one row a, each wavefront a 32 x 32 quadrant as 2 x 2
// MFMA tiles; the complex product takes four real MFMAs (re += Wr Cr - Wi Ci, im += Wr Ci + Wi Cr).
// TFAST: the MFMA rows run along b and its columns along t (D = C^T W^T), so a lane's consecutive
// neighbours store consecutive t; otherwise rows along t and columns along b.
template <bool TFAST>
__global__ __launch_bounds__(256) void spectral_mfma_kernel(int ns, int k, int ncol, const double2* W,
                                                            const double2* A, const double2* B, double2* out,
                                                            int64_t so_t, int64_t so_a, int64_t so_b) {
  constexpr int TB = 64;                     // tile width in b
  __shared__ cpx<double> sW[kTJ][kTT + 1];   // W[t0 + r][j0 + c] stored as [c][r]
  __shared__ cpx<double> sC[kTJ][TB + 1];    // A[a][j0 + c] * B[j0 + c][b0 + q]
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int t0 = blockIdx.x * kTT, b0 = blockIdx.y * TB, a = blockIdx.z;
  const int qt = 32 * (wv >> 1), qb = 32 * (wv & 1);  // the wavefront's quadrant
  const int lr = lane & 15, lk = lane >> 4;
  f64x4 re[2][2], im[2][2];
#pragma unroll
  for (int u = 0; u < 2; ++u)
#pragma unroll
    for (int v = 0; v < 2; ++v) {
      re[u][v] = f64x4{0.0, 0.0, 0.0, 0.0};
      im[u][v] = f64x4{0.0, 0.0, 0.0, 0.0};
    }
  const double2* Arow = A + (int64_t)a * k;
  for (int j0 = 0; j0 < k; j0 += kTJ) {
    for (int e = tid; e < kTT * kTJ; e += 256) {
      const int r = e / kTJ, c = e - (e / kTJ) * kTJ;
      const int t = t0 + r, j = j0 + c;
      double2 v = make_double2(0.0, 0.0);
      if (t < ns && j < k) v = W[(int64_t)t * k + j];
      sW[c][r] = mk(v.x, v.y);
    }
    for (int e = tid; e < kTJ * TB; e += 256) {
      const int c = e / TB, q = e - (e / TB) * TB;
      const int j = j0 + c, b = b0 + q;
      cpx<double> v = mk(0.0, 0.0);
      if (j < k && b < ncol) {
        const double2 x = Arow[j], y = B[(int64_t)j * ncol + b];
        v = mk(x.x, x.y) * mk(y.x, y.y);
      }
      sC[c][q] = v;
    }
    __syncthreads();
#pragma unroll
    for (int ks = 0; ks < kTJ / 4; ++ks) {
      const int c = 4 * ks + lk;
      cpx<double> w[2], x[2];
#pragma unroll
      for (int u = 0; u < 2; ++u) w[u] = sW[c][qt + 16 * u + lr];
#pragma unroll
      for (int v = 0; v < 2; ++v) x[v] = sC[c][qb + 16 * v + lr];
#pragma unroll
      for (int u = 0; u < 2; ++u)
#pragma unroll
        for (int v = 0; v < 2; ++v) {
          // !TFAST: A operand W (rows t), B operand C (columns b); TFAST: A = C^T (rows b), B = W^T
          const cpx<double> p = TFAST ? x[v] : w[u], q = TFAST ? w[u] : x[v];
          re[u][v] = __builtin_amdgcn_mfma_f64_16x16x4f64(p.x, q.x, re[u][v], 0, 0, 0);
          re[u][v] = __builtin_amdgcn_mfma_f64_16x16x4f64(-p.y, q.y, re[u][v], 0, 0, 0);
          im[u][v] = __builtin_amdgcn_mfma_f64_16x16x4f64(p.x, q.y, im[u][v], 0, 0, 0);
          im[u][v] = __builtin_amdgcn_mfma_f64_16x16x4f64(p.y, q.x, im[u][v], 0, 0, 0);
        }
    }
    __syncthreads();
  }
#pragma unroll
  for (int u = 0; u < 2; ++u)
#pragma unroll
    for (int v = 0; v < 2; ++v)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = lk + 4 * r;  // MFMA output row of this lane's r-th value; column lr
        const int t = t0 + qt + 16 * u + (TFAST ? lr : row);
        const int b = b0 + qb + 16 * v + (TFAST ? row : lr);
        if (t < ns && b < ncol)
          out[(int64_t)t * so_t + (int64_t)a * so_a + (int64_t)b * so_b] = make_double2(re[u][v][r], im[u][v][r]);
      }
}

}  // namespace

extern "C" int64_t pyqed_spectral_work_bytes(int32_t k, int32_t ns) {
  if (k < 1 || ns < 1) return 0;
  return (int64_t)k * ns * (int64_t)sizeof(double2);
}

namespace {

// Both launches of one spectral sum, with the weights staged in W (>= k * ns complex doubles).
int launch_spectral(int32_t kind, int32_t nrow, int32_t k, int32_t ncol, int32_t ns, const void* A, const void* lam,
                    const void* B, const double* s, void* out, int64_t so_t, int64_t so_a, int64_t so_b, double2* W,
                    hipStream_t st) {
  const int64_t nw = (int64_t)ns * k;
  const bool wide = ncol >= 48;
  const int tbw = wide ? 4 * kTB : kTB;
  const dim3 grid((unsigned)((ns + kTT - 1) / kTT), (unsigned)((ncol + tbw - 1) / tbw), (unsigned)nrow);
  if (grid.y > 65535) return fail(PYQED_EINVAL, "spectral_sum: too many columns");
  hipLaunchKernelGGL(spectral_weights_kernel, dim3((unsigned)((nw + 255) / 256)), dim3(256), 0, st, (int)kind,
                     (int)ns, (int)k, (const double2*)lam, s, W);
  PQA_HIP(hipGetLastError());
  const bool tfast = so_t == 1;
  const double2 *Wc = (const double2*)W, *Ac = (const double2*)A, *Bc = (const double2*)B;
  double2* O = (double2*)out;
  if (wide && tfast)
    hipLaunchKernelGGL((spectral_mfma_kernel<true>), grid, dim3(kThreads), 0, st, (int)ns, (int)k, (int)ncol, Wc, Ac,
                       Bc, O, so_t, so_a, so_b);
  else if (wide)
    hipLaunchKernelGGL((spectral_mfma_kernel<false>), grid, dim3(kThreads), 0, st, (int)ns, (int)k, (int)ncol, Wc,
                       Ac, Bc, O, so_t, so_a, so_b);
  else if (tfast)
    hipLaunchKernelGGL((spectral_contract_kernel<1, true>), grid, dim3(kThreads), 0, st, (int)ns, (int)k,
                       (int)ncol, Wc, Ac, Bc, O, so_t, so_a, so_b);
  else
    hipLaunchKernelGGL((spectral_contract_kernel<1, false>), grid, dim3(kThreads), 0, st, (int)ns, (int)k,
                       (int)ncol, Wc, Ac, Bc, O, so_t, so_a, so_b);
  PQA_HIP(hipGetLastError());
  return PYQED_OK;
}

int check_sum_args(int32_t kind, int32_t nrow, int32_t k, int32_t ncol, int32_t ns, const void* A, const void* lam,
                   const void* B, const double* s, void* out) {
  if (kind != 0 && kind != 1) return fail(PYQED_EINVAL, "spectral_sum: kind must be 0 (exp) or 1 (resolvent)");
  if (nrow < 1 || k < 1 || ncol < 1 || ns < 1)
    return fail(PYQED_EINVAL, "spectral_sum: nrow, k, ncol and ns must be positive");
  if (nrow > 65535) return fail(PYQED_EINVAL, "spectral_sum: at most 65535 rows per call");
  if (!A || !lam || !B || !s || !out) return fail(PYQED_EINVAL, "spectral_sum: null pointer");
  return PYQED_OK;
}

// Weight staging of pyqed_spectral_sum_static: PYQED_SPECTRAL_STATIC_MAX complex doubles (16 MiB) in
// every GPU's copy of the library's code object, held until the process ends.  Host code never passes
// `scratch` itself to a kernel (on the host that name is only a stand-in, not a device address): it
// passes the address hipGetSymbolAddress gives for the current device, looked up once per device.
__device__ __align__(256) double2 scratch[PYQED_SPECTRAL_STATIC_MAX];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double2* g_scratch_dev[kMaxDevices] = {};

int scratch_of_current_device(double2** out) {
  int dev = -1;
  PQA_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices)
    return fail(PYQED_EINVAL, "spectral_sum_static: device index " + std::to_string(dev) + " out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_dev[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fail(PYQED_EHIP, std::string("spectral_sum_static: hipGetSymbolAddress(scratch): ") +
                                  hipGetErrorString(e));
    g_scratch_dev[dev] = (double2*)p;
  }
  *out = g_scratch_dev[dev];
  return PYQED_OK;
}

}  // namespace

extern "C" int pyqed_spectral_sum(int32_t kind, int32_t nrow, int32_t k, int32_t ncol, int32_t ns, const void* A,
                                  const void* lam, const void* B, const double* s, void* out, int64_t so_t,
                                  int64_t so_a, int64_t so_b, void* work, int64_t work_bytes, void* stream) {
  int rc = check_sum_args(kind, nrow, k, ncol, ns, A, lam, B, s, out);
  if (rc) return rc;
  if (!work) return fail(PYQED_EINVAL, "spectral_sum: null pointer");
  if (work_bytes < pyqed_spectral_work_bytes(k, ns))
    return fail(PYQED_EINVAL, "spectral_sum: workspace of " + std::to_string(work_bytes) + " bytes < " +
                                  std::to_string(pyqed_spectral_work_bytes(k, ns)) + " needed");
  return launch_spectral(kind, nrow, k, ncol, ns, A, lam, B, s, out, so_t, so_a, so_b, (double2*)work,
                         (hipStream_t)stream);
}

extern "C" int pyqed_spectral_sum_static(int32_t kind, int32_t nrow, int32_t k, int32_t ncol, int32_t ns,
                                         const void* A, const void* lam, const void* B, const double* s, void* out,
                                         int64_t so_t, int64_t so_a, int64_t so_b, void* stream) {
  int rc = check_sum_args(kind, nrow, k, ncol, ns, A, lam, B, s, out);
  if (rc) return rc;
  if ((int64_t)k * ns > (int64_t)PYQED_SPECTRAL_STATIC_MAX)
    return fail(PYQED_EINVAL, "spectral_sum_static: k * ns = " + std::to_string((int64_t)k * ns) +
                                  " weights exceed the static scratch of " +
                                  std::to_string((int64_t)PYQED_SPECTRAL_STATIC_MAX) +
                                  " (use pyqed_spectral_sum with a workspace)");
  double2* W = nullptr;
  if ((rc = scratch_of_current_device(&W))) return rc;
  // every weight the GEMM reads (t < ns, j < k) is written by the first launch: nothing to clear
  return launch_spectral(kind, nrow, k, ncol, ns, A, lam, B, s, out, so_t, so_a, so_b, W, (hipStream_t)stream);
}
