// simplex.hpp -- C++ drop-in for the reference's batched solver class.
//
// Same class name, method names, argument meaning and conventions as the reference's
// `Simplex` (BLPG-Solver/src/simplex.cuh:37-83): construct with the number of LPs,
// setConstratint(A, b) once, ComputeLP(C, streams) with one objective row per LP
// (maximisation), then read getResultAll() / getStatusAll().  Differences a caller can
// rely on (all of them fixes of reference defects):
//   - statuses are decided by the solver: 2 optimal (the reference's "feasible and
//     optimal", simplex.cu:410), 6 unbounded (simplex.cu:408, never produced there),
//     4 infeasible (GLPK's GLP_NOFEAS; the reference prints a message, simplex.cu:821),
//     1 undefined (iteration limit / numerical failure);
//   - getResult(n) returns a double (the reference narrows to float, simplex.cu:442);
//   - no device reset between batches (simplex.cu:974); `number_of_streams` is accepted
//     and ignored (size bins and a work queue replace per-stream launches);
//   - b may be any sign and A may carry '=' / '>=' rows (setConstraints with senses).
// Matrices may be any type with size1(), size2() and operator()(i, j) -- including the
// reference's math::matrix<double> (boost::numeric::ublas) -- or blpg::matrix<double>.
#pragma once
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "blpg.h"

namespace blpg {

// Minimal row-major matrix with the accessors the reference's callers use
// (math::matrix<T>, BLPG-Solver/src/math/matrix.h:18-55).
template <typename T>
class matrix {
 public:
  matrix() = default;
  matrix(size_t r, size_t c) : r_(r), c_(c), d_(r * c) {}
  void resize(size_t r, size_t c) { r_ = r; c_ = c; d_.assign(r * c, T()); }
  size_t size1() const { return r_; }
  size_t size2() const { return c_; }
  T& operator()(size_t i, size_t j) { return d_[i * c_ + j]; }
  const T& operator()(size_t i, size_t j) const { return d_[i * c_ + j]; }
  T* data() { return d_.data(); }
  const T* data() const { return d_.data(); }

 private:
  size_t r_ = 0, c_ = 0;
  std::vector<T> d_;
};

}  // namespace blpg

class Simplex {
 public:
  // N_S: number of LPs per ComputeLP; multipleBatch >= 1: split batches that exceed
  // the device's free memory into chunks (the reference's multi-batch mode,
  // GPU_Tester.cu:284-387), without its per-chunk cudaDeviceReset.
  Simplex(unsigned int N_S, unsigned int multipleBatch = 0);
  ~Simplex();
  Simplex(const Simplex&) = delete;
  Simplex& operator=(const Simplex&) = delete;

  // constraints A x <= b shared by all LPs (simplex.cu:463-550)
  template <class Mat>
  void setConstratint(const Mat& A, const std::vector<double>& B) {
    std::vector<double> a(A.size1() * A.size2());
    for (size_t i = 0; i < A.size1(); ++i)
      for (size_t j = 0; j < A.size2(); ++j) a[i * A.size2() + j] = A(i, j);
    set_constraints_rowmajor((int)A.size1(), (int)A.size2(), a.data(), B, {});
  }
  // extension: per-row senses 0 '<=', 1 '=', 2 '>='
  template <class Mat>
  void setConstraints(const Mat& A, const std::vector<double>& B,
                      const std::vector<int8_t>& sense) {
    std::vector<double> a(A.size1() * A.size2());
    for (size_t i = 0; i < A.size1(); ++i)
      for (size_t j = 0; j < A.size2(); ++j) a[i * A.size2() + j] = A(i, j);
    set_constraints_rowmajor((int)A.size1(), (int)A.size2(), a.data(), B, sense);
  }

  // one objective per row of C (maximise), simplex.cu:552 + 880-890
  template <class Mat>
  void ComputeLP(const Mat& C1, unsigned int number_of_streams = 1) {
    (void)number_of_streams;
    std::vector<double> c(C1.size1() * C1.size2());
    for (size_t i = 0; i < C1.size1(); ++i)
      for (size_t j = 0; j < C1.size2(); ++j) c[i * C1.size2() + j] = C1(i, j);
    compute_rowmajor((int64_t)C1.size1(), (int)C1.size2(), c.data());
  }

  int getStatus(int n);                // 1-based, simplex.cu:403-416
  int getNo_OF_Simplx();               // simplex.cu:419-421
  std::vector<double> getResultAll();  // simplex.cu:424-436
  double getResult(int n);             // 1-based, simplex.cu:440-449
  std::vector<int> getStatusAll();     // simplex.cu:451-461

  // extensions: the optimal point, basis and pivot count of LP n (1-based)
  std::vector<double> getX(int n);
  std::vector<int32_t> getBasis(int n);
  int getPivots(int n);
  // solver options (precision, tolerances, kernel choice) for later ComputeLP calls
  void setOptions(const blpg_options& opt);

 private:
  void set_constraints_rowmajor(int m, int n, const double* A, const std::vector<double>& b,
                                const std::vector<int8_t>& sense);
  void compute_rowmajor(int64_t count, int n, const double* C);
  unsigned int number_of_LPs_;
  unsigned int multiple_batch_;
  int m_ = 0, n_ = 0;
  std::vector<double> A_, b_;
  std::vector<int8_t> sense_;
  std::vector<double> result_, x_;
  std::vector<int32_t> status_, basis_, pivots_;
  blpg_options opt_;
  blpg_ctx* ctx_ = nullptr;
};
