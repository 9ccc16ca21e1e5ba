// C++ adapter: the reference's sparse-solver plugin interface implemented on
// the phase_hip C ABI.
//
// SparseMatrixSolverBase restates the virtual interface of the reference's
// SparseMatrixSolver (src/Math/SparseMatrixSolver.h:11-62) with the same
// method names, argument meaning and error behaviour, but without its Boost
// dependency: setup() takes key/value pairs where the reference takes the
// LinearAlgebra.<eqn> property tree.  A maintainer dropping the adapter into
// the reference derives HipSparseMatrixSolver from the reference's own class
// instead and forwards setup(ptree) keys to setup(key, value) -- the method
// bodies are unchanged.
#pragma once

#include <map>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/phase_hip.h"

namespace phase_amd {

using Index = int;      // src/Types/Types.h
using Scalar = double;

// Exception(class, method, description) -- src/System/Exception.h
class Exception : public std::runtime_error {
 public:
  Exception(const std::string &cls, const std::string &method, const std::string &desc)
      : std::runtime_error(cls + "::" + method + ": " + desc) {}
};

class SparseMatrixSolverBase {
 public:
  enum Type { EIGEN, TRILINOS_BELOS, TRILINOS_AMESOS2, TRILINOS_MUELU, HIP };
  virtual ~SparseMatrixSolverBase() = default;
  virtual Type type() const = 0;
  virtual void setRank(int rank) = 0;
  virtual void set(const std::vector<Index> &rowPtr, const std::vector<Index> &colInds,
                   const std::vector<Scalar> &vals) = 0;
  virtual void setGuess(const std::vector<Scalar> &x0) = 0;
  virtual void setRhs(const std::vector<Scalar> &rhs) = 0;
  virtual Scalar solve() = 0;
  virtual Scalar x(Index idx) const = 0;
  virtual void setup(const std::map<std::string, std::string> &parameters) {}
  virtual int nIters() const = 0;
  virtual Scalar error() const = 0;
  virtual bool supportsMPI() const = 0;
  virtual void printStatus(const std::string &msg) const;
};

class HipSparseMatrixSolver : public SparseMatrixSolverBase {
 public:
  explicit HipSparseMatrixSolver(int device = 0);
  ~HipSparseMatrixSolver() override;
  HipSparseMatrixSolver(const HipSparseMatrixSolver &) = delete;
  HipSparseMatrixSolver &operator=(const HipSparseMatrixSolver &) = delete;

  Type type() const override { return HIP; }
  void setRank(int rank) override;
  void set(const std::vector<Index> &rowPtr, const std::vector<Index> &colInds,
           const std::vector<Scalar> &vals) override;
  void setGuess(const std::vector<Scalar> &x0) override;
  void setRhs(const std::vector<Scalar> &rhs) override;
  Scalar solve() override;
  Scalar x(Index idx) const override { return x_.at(idx); }
  void setup(const std::map<std::string, std::string> &parameters) override;
  int nIters() const override;
  Scalar error() const override;
  bool supportsMPI() const override { return false; }
  void printStatus(const std::string &msg) const override;

  phase_hip_solver *handle() { return h_; }

 private:
  void check(int rc, const char *method) const;
  phase_hip_solver *h_ = nullptr;
  int rank_ = 0;
  std::vector<Scalar> x_;
};

}  // namespace phase_amd
