// HipSparseMatrixSolver: the reference's SparseMatrixSolver contract
// (src/Math/SparseMatrixSolver.h:11-62) on top of include/phase_hip.h.
#include "HipSparseMatrixSolver.h"

#include <cstdio>
#include <memory>

namespace phase_amd {

// --- defaults of the base (src/Math/SparseMatrixSolver.cpp) ----------------

// Triplets grouped into rows, repeated (row, col) pairs summed in order of
// appearance (SparseMatrixSolver.cpp:5-31).
void SparseMatrixSolverBase::set(const std::vector<std::tuple<Index, Index, Scalar>> &entries) {
  CoefficientList coeffs;
  for (const auto &e : entries) {
    const Index row = std::get<0>(e), col = std::get<1>(e);
    const Scalar val = std::get<2>(e);
    if (row < 0) throw Exception("SparseMatrixSolver", "set", "negative row index");
    if ((size_t)row >= coeffs.size()) coeffs.resize(row + 1);
    bool found = false;
    for (Entry &q : coeffs[row])
      if (q.first == col) {
        q.second += val;
        found = true;
        break;
      }
    if (!found) coeffs[row].push_back(Entry(col, val));
  }
  set(coeffs);
}

Scalar SparseMatrixSolverBase::solve(const Vector &x0) {
  setGuess(x0);
  return solve();
}

Scalar SparseMatrixSolverBase::solveLeastSquares() {
  throw Exception("SparseMatrixSolver", "solveLeastSquares",
                  "least squares solver is not available for this sparse matrix solver type.");
}

void SparseMatrixSolverBase::printStatus(const std::string &msg) const {
  std::printf("%s iterations = %d, error = %lf.\n", msg.c_str(), nIters(), error());
}

// --- the HIP backend ---------------------------------------------------------

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

void HipSparseMatrixSolver::setRank(int rowRank, int colRank) {
  if (rowRank != colRank)
    throw Exception("HipSparseMatrixSolver", "setRank", "rectangular systems (" + std::to_string(rowRank) + " x " +
                                                            std::to_string(colRank) + ") are not supported");
  setRank(rowRank);
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

// Rows of a coefficient list to CSR (columns in order of first appearance,
// repeated columns summed); rows past the list up to the rank stay empty,
// which set() reports as a missing diagonal.
void HipSparseMatrixSolver::setRows(const CoefficientList &rows, const char *method) {
  const int n = rank_ > 0 ? rank_ : (int)rows.size();
  if ((int)rows.size() > n) throw Exception("HipSparseMatrixSolver", method, "more rows than the rank");
  std::vector<Index> rowPtr(1, 0), cols;
  std::vector<Scalar> vals;
  for (int r = 0; r < n; ++r) {
    const size_t start = cols.size();
    if (r < (int)rows.size())
      for (const Entry &e : rows[r]) {
        if (e.first < 0 || e.first >= n) throw Exception("HipSparseMatrixSolver", method, "column out of range");
        bool found = false;
        for (size_t q = start; q < cols.size() && !found; ++q)
          if (cols[q] == e.first) {
            vals[q] += e.second;
            found = true;
          }
        if (!found) {
          cols.push_back(e.first);
          vals.push_back(e.second);
        }
      }
    rowPtr.push_back((Index)cols.size());
  }
  set(rowPtr, cols, vals);
}

void HipSparseMatrixSolver::set(const CoefficientList &eqn) { setRows(eqn, "set"); }

void HipSparseMatrixSolver::set(const std::vector<SparseEntry> &entries) {
  CoefficientList rows(rank_ > 0 ? rank_ : 0);
  for (const SparseEntry &e : entries) {
    if (e.row < 0) throw Exception("HipSparseMatrixSolver", "set", "negative row index");
    if ((size_t)e.row >= rows.size()) rows.resize(e.row + 1);
    rows[e.row].push_back(Entry(e.col, e.val));
  }
  setRows(rows, "set");
}

void HipSparseMatrixSolver::setGuess(const Vector &x0) {
  if ((int)x0.size() != rank_) throw Exception("HipSparseMatrixSolver", "setGuess", "size differs from the rank");
  check(phase_hip_solver_set_guess(h_, x0.data().data()), "setGuess");
}

void HipSparseMatrixSolver::setRhs(const Vector &rhs) {
  if ((int)rhs.size() != rank_) throw Exception("HipSparseMatrixSolver", "setRhs", "size differs from the rank");
  check(phase_hip_solver_set_rhs(h_, rhs.data().data()), "setRhs");
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

std::shared_ptr<SparseMatrixSolverBase> createSparseMatrixSolver(const std::string &type, int device) {
  if (type == "hip") return std::make_shared<HipSparseMatrixSolver>(device);
  throw Exception("SparseMatrixSolverFactory", "create", "bad solver type \"" + type + "\".");
}

}  // namespace phase_amd
