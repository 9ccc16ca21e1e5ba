// HipSparseMatrixSolver: the reference's SparseMatrixSolver contract
// (src/Math/SparseMatrixSolver.h:11-62) on top of include/phase_hip.h.
#include "HipSparseMatrixSolver.h"

#include <cstdio>

namespace phase_amd {

void SparseMatrixSolverBase::printStatus(const std::string &msg) const {
  std::printf("%s iterations = %d, error = %lf.\n", msg.c_str(), nIters(), error());
}

HipSparseMatrixSolver::HipSparseMatrixSolver(int device) {
  int rc = phase_hip_solver_create(&h_, device);
  if (rc != PHASE_HIP_OK)
    throw Exception("HipSparseMatrixSolver", "HipSparseMatrixSolver", phase_hip_last_error(nullptr));
}

HipSparseMatrixSolver::~HipSparseMatrixSolver() { phase_hip_solver_destroy(h_); }

void HipSparseMatrixSolver::check(int rc, const char *method) const {
  if (rc != PHASE_HIP_OK) throw Exception("HipSparseMatrixSolver", method, phase_hip_last_error(h_));
}

void HipSparseMatrixSolver::setup(const std::map<std::string, std::string> &parameters) {
  for (const auto &kv : parameters) check(phase_hip_solver_setup(h_, kv.first.c_str(), kv.second.c_str()), "setup");
}

void HipSparseMatrixSolver::setRank(int rank) {
  check(phase_hip_solver_set_rank(h_, rank), "setRank");
  rank_ = rank;
  x_.assign(rank, 0.);
}

void HipSparseMatrixSolver::set(const std::vector<Index> &rowPtr, const std::vector<Index> &colInds,
                                const std::vector<Scalar> &vals) {
  const int n = (int)rowPtr.size() - 1;
  if (n <= 0 || (int)colInds.size() < rowPtr.back() || (int)vals.size() < rowPtr.back())
    throw Exception("HipSparseMatrixSolver", "set", "inconsistent CSR arrays");
  check(phase_hip_solver_set(h_, n, rowPtr.data(), colInds.data(), vals.data()), "set");
  rank_ = n;
  x_.resize(n, 0.);
}

void HipSparseMatrixSolver::setGuess(const std::vector<Scalar> &x0) {
  check(phase_hip_solver_set_guess(h_, x0.data()), "setGuess");
}

void HipSparseMatrixSolver::setRhs(const std::vector<Scalar> &rhs) {
  check(phase_hip_solver_set_rhs(h_, rhs.data()), "setRhs");
}

Scalar HipSparseMatrixSolver::solve() {
  double relres = 0.;
  int rc = phase_hip_solver_solve(h_, &relres);
  // like Belos, a solve that stops at maxIters still returns its achieved
  // tolerance; every other failure is an Exception
  if (rc != PHASE_HIP_OK && rc != PHASE_HIP_ENOCONV) check(rc, "solve");
  x_.resize(rank_);
  check(phase_hip_solver_get_x(h_, x_.data()), "solve");
  return relres;
}

int HipSparseMatrixSolver::nIters() const { return phase_hip_solver_n_iters(h_); }

Scalar HipSparseMatrixSolver::error() const { return phase_hip_solver_error(h_); }

void HipSparseMatrixSolver::printStatus(const std::string &msg) const {
  std::printf("%s %s iterations = %d, error = %lf.\n", msg.c_str(), "Krylov (hip)", nIters(), error());
}

}  // namespace phase_amd
