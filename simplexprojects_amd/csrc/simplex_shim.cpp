// simplex_shim.cpp -- the C++ drop-in `Simplex` (include/blpg/simplex.hpp) over the C ABI.
#include "../../include/blpg/simplex.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>

namespace {

int ref_status(int s) {
  switch (s) {
    case BLPG_OPTIMAL: return 2;    // simplex.cu:410 "feasible and optimal"
    case BLPG_UNBOUNDED: return 6;  // simplex.cu:408
    case BLPG_INFEASIBLE: return 4; // GLP_NOFEAS
    default: return 1;              // GLP_UNDEF
  }
}

void check(blpg_ctx* ctx, int rc, const char* what) {
  if (rc != BLPG_OK)
    throw std::runtime_error(std::string(what) + ": " + (ctx ? blpg_last_error(ctx) : "no context") +
                             " (code " + std::to_string(rc) + ")");
}

}  // namespace

Simplex::Simplex(unsigned int N_S, unsigned int multipleBatch)
    : number_of_LPs_(N_S), multiple_batch_(multipleBatch) {
  blpg_default_options(&opt_);
}

Simplex::~Simplex() {
  if (ctx_) blpg_destroy(ctx_);
}

void Simplex::setOptions(const blpg_options& opt) {
  opt_ = opt;
  if (ctx_) check(ctx_, blpg_set_options(ctx_, &opt_), "blpg_set_options");
}

void Simplex::set_constraints_rowmajor(int m, int n, const double* A, const std::vector<double>& b,
                                       const std::vector<int8_t>& sense) {
  if ((int)b.size() != m) throw std::invalid_argument("setConstratint: b must have A.size1() entries");
  if (!sense.empty() && (int)sense.size() != m)
    throw std::invalid_argument("setConstraints: sense must have A.size1() entries");
  m_ = m;
  n_ = n;
  A_.assign(A, A + (size_t)m * n);
  b_ = b;
  sense_ = sense;
}

void Simplex::compute_rowmajor(int64_t count, int n, const double* C) {
  if (m_ <= 0) throw std::logic_error("ComputeLP before setConstratint");
  if (n != n_) throw std::invalid_argument("ComputeLP: C.size2() must equal A.size2()");
  if (!ctx_) check(nullptr, blpg_create(&ctx_, &opt_), "blpg_create");
  result_.assign(count, 0.0);
  status_.assign(count, BLPG_UNDECIDED);
  pivots_.assign(count, 0);
  x_.assign((size_t)count * n_, 0.0);
  basis_.assign((size_t)count * m_, -1);
  // chunking (multi-batch mode): keep each chunk well inside the free device memory
  int64_t chunk = count;
  if (multiple_batch_ >= 1) {
    size_t free_b = 0, total_b = 0;
    if (hipMemGetInfo(&free_b, &total_b) == hipSuccess) {
      const double per_lp = 8.0 * ((double)m_ * n_ + m_ + 3.0 * n_ + 8) + 64.0;
      chunk = std::max<int64_t>(1, (int64_t)(0.25 * (double)free_b / per_lp));
    }
  }
  std::vector<blpg_result> out(count);
  for (int64_t k = 0; k < count; ++k) {
    std::memset(&out[k], 0, sizeof(blpg_result));
    out[k].x = x_.data() + k * n_;
    out[k].basis = basis_.data() + k * m_;
  }
  for (int64_t first = 0; first < count; first += chunk) {
    const int64_t cnt = std::min(chunk, count - first);
    check(ctx_,
          blpg_solve_shared_A(ctx_, m_, n_, A_.data(), b_.data(),
                              sense_.empty() ? nullptr : sense_.data(), C + first * n_, cnt,
                              out.data() + first),
          "blpg_solve_shared_A");
  }
  for (int64_t k = 0; k < count; ++k) {
    result_[k] = out[k].objective;
    status_[k] = out[k].status;
    pivots_[k] = out[k].pivots;
  }
}

int Simplex::getStatus(int n) {
  if (n < 1 || n > (int)status_.size()) throw std::out_of_range("getStatus: n is 1-based");
  return ref_status(status_[n - 1]);
}

int Simplex::getNo_OF_Simplx() {
  return status_.empty() ? (int)number_of_LPs_ : (int)status_.size();
}

std::vector<double> Simplex::getResultAll() { return result_; }

double Simplex::getResult(int n) {
  if (n < 1 || n > (int)result_.size()) throw std::out_of_range("getResult: n is 1-based");
  return result_[n - 1];
}

std::vector<int> Simplex::getStatusAll() {
  std::vector<int> s(status_.size());
  for (size_t k = 0; k < s.size(); ++k) s[k] = ref_status(status_[k]);
  return s;
}

std::vector<double> Simplex::getX(int n) {
  if (n < 1 || n > (int)status_.size()) throw std::out_of_range("getX: n is 1-based");
  return std::vector<double>(x_.begin() + (size_t)(n - 1) * n_, x_.begin() + (size_t)n * n_);
}

std::vector<int32_t> Simplex::getBasis(int n) {
  if (n < 1 || n > (int)status_.size()) throw std::out_of_range("getBasis: n is 1-based");
  return std::vector<int32_t>(basis_.begin() + (size_t)(n - 1) * m_, basis_.begin() + (size_t)n * m_);
}

int Simplex::getPivots(int n) {
  if (n < 1 || n > (int)pivots_.size()) throw std::out_of_range("getPivots: n is 1-based");
  return pivots_[n - 1];
}
