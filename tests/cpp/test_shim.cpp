// test_shim.cpp -- exercises the C++ drop-in `Simplex` the way the reference's driver does
// (GPU_Tester.cu:119-127, 244-281: one A, b; C = N copies of an objective; results and
// statuses read back), on the reference's hand LPs (parseBenchmark.cpp:128-319).
// Prints one line per check and "ALL OK" at the end; exits non-zero on any failure.
#include <cmath>
#include <cstdio>
#include <vector>

#include "blpg/simplex.hpp"

static int failures = 0;
#define CHECK(cond, ...)                          \
  do {                                            \
    if (!(cond)) {                                \
      ++failures;                                 \
      std::printf("FAIL %s: ", #cond);            \
      std::printf(__VA_ARGS__);                   \
      std::printf("\n");                          \
    }                                             \
  } while (0)

struct Hand {
  const char* name;
  int m, n;
  std::vector<double> A, b, c;
  double opt;
};

int main() {
  const std::vector<Hand> hand = {
      {"parseSmallLP", 2, 2, {2, -1, 1, -5}, {2, -4}, {2, -1}, 2.0},
      {"parseSmallLP2", 3, 2, {1, 1, -1, 1, 5, 4}, {10, -3, 35}, {5, 6}, 355.0 / 9.0},
      {"parseSmallLP3", 3, 2, {1, -1, -1, -1, -1, 4}, {8, -3, 2}, {1, 3}, 64.0 / 3.0},
      {"parseSmallLP4", 3, 3, {1, 1, -1, -3, 1, 0, 1, -2, -2}, {7, -8, 0}, {2, 0, -6}, 28.0 / 3.0},
      {"parseSmallLP5", 3, 2, {4, -1, 2, 1, -5, 2}, {8, 10, 2}, {1, 1}, 8.0},
  };
  const int N = 1000;
  for (const Hand& h : hand) {
    blpg::matrix<double> A(h.m, h.n), C(N, h.n);
    for (int i = 0; i < h.m; ++i)
      for (int j = 0; j < h.n; ++j) A(i, j) = h.A[i * h.n + j];
    for (int k = 0; k < N; ++k)
      for (int j = 0; j < h.n; ++j) C(k, j) = h.c[j];
    Simplex s(N, 0);
    s.setConstratint(A, h.b);
    s.ComputeLP(C, 10);
    const std::vector<double> r = s.getResultAll();
    const std::vector<int> st = s.getStatusAll();
    CHECK(s.getNo_OF_Simplx() == N, "%s count", h.name);
    int bad = 0;
    for (int k = 0; k < N; ++k) bad += (st[k] != 2 || std::fabs(r[k] - h.opt) > 1e-9);
    CHECK(bad == 0, "%s: %d of %d wrong (r[0]=%.12g status %d)", h.name, bad, N, r[0], st[0]);
    CHECK(std::fabs(s.getResult(N) - h.opt) < 1e-9, "%s getResult", h.name);
    CHECK(s.getStatus(1) == 2, "%s getStatus", h.name);
    std::printf("%s: %d LPs, objective %.12g, status %d, pivots %d\n", h.name, N, r[0], st[0],
                s.getPivots(1));
  }
  // unbounded and infeasible statuses (6 and 4)
  {
    blpg::matrix<double> A(1, 2), C(2, 2);
    A(0, 0) = 1; A(0, 1) = -1;
    C(0, 0) = 1; C(0, 1) = 1;   // max x1 + x2 s.t. x1 - x2 <= 1: unbounded
    C(1, 0) = -1; C(1, 1) = -1; // max -x1 - x2: optimal 0
    Simplex s(2);
    s.setConstratint(A, {1.0});
    s.ComputeLP(C);
    CHECK(s.getStatus(1) == 6, "unbounded status %d", s.getStatus(1));
    CHECK(s.getStatus(2) == 2 && std::fabs(s.getResult(2)) < 1e-12, "optimal 0");
  }
  {
    blpg::matrix<double> A(2, 1), C(1, 1);
    A(0, 0) = 1; A(1, 0) = -1;  // x <= -1 and -x <= -2 (x >= 2): infeasible
    C(0, 0) = 1;
    Simplex s(1);
    s.setConstratint(A, {-1.0, -2.0});
    s.ComputeLP(C);
    CHECK(s.getStatus(1) == 4, "infeasible status %d", s.getStatus(1));
  }
  // equality rows through the extension, multi-batch chunking on
  {
    blpg::matrix<double> A(2, 3), C(3, 3);
    const double a[6] = {1, 1, 1, 1, -1, 0};
    for (int e = 0; e < 6; ++e) A(e / 3, e % 3) = a[e];
    for (int k = 0; k < 3; ++k)
      for (int j = 0; j < 3; ++j) C(k, j) = (j == k) ? 1.0 : 0.0;
    Simplex s(3, 1);
    s.setConstraints(A, {4.0, 0.0}, {1, 1});  // x1+x2+x3 = 4, x1 = x2
    s.ComputeLP(C);
    CHECK(std::fabs(s.getResult(1) - 2.0) < 1e-12, "eq max x1 = %.12g", s.getResult(1));
    CHECK(std::fabs(s.getResult(3) - 4.0) < 1e-12, "eq max x3 = %.12g", s.getResult(3));
    const std::vector<double> x = s.getX(3);
    CHECK(std::fabs(x[2] - 4.0) < 1e-12, "eq x3");
  }
  if (failures) {
    std::printf("%d FAILURES\n", failures);
    return 1;
  }
  std::printf("ALL OK\n");
  return 0;
}
