// Host-side driver for the AddressSanitizer / UndefinedBehaviorSanitizer build of libpyqed_amd's host
// code: the hierarchy table builder (pyqed_heom_tables_size / _build) on large shapes, checked against
// the defining relations of the tables, and pyqed_heom_plan_create on a 4096-instance batch of uneven
// hierarchies (64-bit element offsets of about 4.6e9), which builds every host structure of a plan and
// stops at its first device call when no GPU is present.  Exit code 0 when every check holds.
#include <cmath>
#include <complex>
#include <cstdio>
#include <random>
#include <vector>

#include "pyqed_amd.h"

static int failures = 0;
#define CHECK(c)                                                              \
  do {                                                                        \
    if (!(c)) {                                                               \
      std::fprintf(stderr, "check failed at line %d: %s\n", __LINE__, #c);   \
      ++failures;                                                             \
    }                                                                         \
  } while (0)

static void check_tables(int K, int L) {
  int64_t n = 0;
  CHECK(pyqed_heom_tables_size(K, L, &n) == PYQED_OK);
  std::vector<int64_t> keys((size_t)n * K);
  std::vector<int32_t> plus((size_t)n * K), minus((size_t)n * K);
  CHECK(pyqed_heom_tables_build(K, L, keys.data(), plus.data(), minus.data()) == PYQED_OK);
  for (int64_t a = 0; a < n; ++a) {
    int64_t tier = 0;
    for (int k = 0; k < K; ++k) tier += keys[(size_t)a * K + k];
    CHECK(tier <= L);
    for (int k = 0; k < K; ++k) {
      const int32_t p = plus[(size_t)a * K + k], m = minus[(size_t)a * K + k];
      CHECK((p >= 0) == (tier < L));
      CHECK((m >= 0) == (keys[(size_t)a * K + k] > 0));
      for (int j = 0; j < K; ++j) {
        if (p >= 0) CHECK(keys[(size_t)p * K + j] == keys[(size_t)a * K + j] + (j == k));
        if (m >= 0) CHECK(keys[(size_t)m * K + j] == keys[(size_t)a * K + j] - (j == k));
      }
    }
  }
}

int main() {
  const int shapes[][2] = {{1, 4}, {2, 5}, {3, 4}, {6, 3}, {14, 4}, {48, 3}, {32, 4}, {0, 3}, {5, 0}};
  for (auto& s : shapes) check_tables(s[0], s[1]);
  int64_t n = 0;
  CHECK(pyqed_heom_tables_size(40, 20, &n) != PYQED_OK);  // C(60, 20) > 2^31: rejected, no overflow

  // a batch like the 4096-instance workload: exciton chains, d in 2..16, depth 2..4, K = 2d
  std::mt19937_64 rng(3);
  const int ninst = 4096;
  std::vector<std::vector<double>> pool;
  std::vector<std::vector<int32_t>> modes;
  std::vector<pyqed_heom_instance> inst(ninst);
  auto keep = [&](std::vector<double> v) {
    pool.push_back(std::move(v));
    return pool.back().data();
  };
  for (int i = 0; i < ninst; ++i) {
    const int d = 2 + (int)(rng() % 15), L = 2 + (int)(rng() % 3), K = 2 * d;
    std::vector<double> H(2 * d * d, 0.0), q(2 * d * d, 0.0), ex(2 * K), el(2 * K), er(2 * K), ea(2 * K);
    for (int a = 0; a < d; ++a) H[2 * (a * d + a)] = std::sin(1.0 + a + i);
    for (int a = 0; a + 1 < d; ++a) H[2 * (a * d + a + 1)] = H[2 * ((a + 1) * d + a)] = -0.3;
    for (int m = 0; m < d; ++m) q[2 * (m * d + m)] = 1.0;
    for (int k = 0; k < K; ++k) {
      ex[2 * k] = (k % 2) ? 7.7 : 0.5;
      el[2 * k] = 0.1;
      el[2 * k + 1] = (k % 2) ? 0.0 : -0.02;
      er[2 * k] = 0.1;
      er[2 * k + 1] = -el[2 * k + 1];
      ea[2 * k] = std::abs(std::complex<double>(el[2 * k], el[2 * k + 1]));
    }
    std::vector<int32_t> md(K);
    for (int k = 0; k < K; ++k) md[k] = k / 2;
    modes.push_back(md);
    pyqed_heom_instance& s = inst[i];
    s = pyqed_heom_instance{};
    s.d = d;
    s.K = K;
    s.L = L;
    s.ncoup = d;
    s.mode = modes.back().data();
    s.H = keep(H);
    s.qdiag = keep(std::vector<double>(q.begin(), q.begin() + 2 * d * d));
    // diagonal couplings: one entry per site
    std::vector<double> qd(2 * d * d, 0.0);
    for (int m = 0; m < d; ++m) qd[2 * (m * d + m)] = 1.0;
    s.qdiag = keep(qd);
    s.expn = keep(ex);
    s.etal = keep(el);
    s.etar = keep(er);
    s.etaa = keep(ea);
  }
  for (int flags : {0, PYQED_PLAN_HERMITIAN | PYQED_PLAN_PACKED}) {
    pyqed_heom_plan* plan = nullptr;
    const int rc = pyqed_heom_plan_create(inst.data(), ninst, flags, &plan);
    CHECK(rc == PYQED_OK || rc == PYQED_EHIP);  // EHIP: no device in this process
    if (rc == PYQED_OK) {
      int64_t ne = 0, na = 0, w64 = 0, wmx = 0;
      CHECK(pyqed_heom_plan_query(plan, &ne, &na, &w64, &wmx) == PYQED_OK);
      CHECK(ne > (int64_t)1 << 32);  // the element offsets need 64 bits
      pyqed_heom_plan_destroy(plan);
    } else {
      CHECK(plan == nullptr);
    }
  }
  // rejected inputs: non-finite values, modes out of range, zero etaa
  {
    std::vector<double> H = {0.5, 0, 0.1, 0, 0.1, 0, -0.5, 0}, q = {1, 0, -1, 0}, ex = {1, 0, 6, 0},
                        el = {0.2, -0.05, 0.06, 0}, er = {0.2, 0.05, 0.06, 0}, ea = {0.2, 0, 0.06, 0};
    std::vector<int32_t> md = {0, 0};
    pyqed_heom_instance s{};
    s.d = 2;
    s.K = 2;
    s.L = 3;
    s.ncoup = 1;
    s.mode = md.data();
    s.H = H.data();
    s.qdiag = q.data();
    s.expn = ex.data();
    s.etal = el.data();
    s.etar = er.data();
    s.etaa = ea.data();
    pyqed_heom_plan* plan = nullptr;
    H[0] = NAN;
    CHECK(pyqed_heom_plan_create(&s, 1, 0, &plan) == PYQED_EINVAL && plan == nullptr);
    H[0] = 0.5;
    ea[0] = 0.0;
    CHECK(pyqed_heom_plan_create(&s, 1, 0, &plan) == PYQED_EINVAL);
    ea[0] = 0.2;
    md[1] = 3;
    CHECK(pyqed_heom_plan_create(&s, 1, 0, &plan) == PYQED_EINVAL);
  }
  std::printf("host driver: %d failed checks\n", failures);
  return failures == 0 ? 0 : 1;
}
