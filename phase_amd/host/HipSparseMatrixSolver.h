// C++ adapter: the reference's sparse-solver plugin interface implemented on
// the phase_hip C ABI.
//
// SparseMatrixSolverBase restates the virtual interface of the reference's
// SparseMatrixSolver (src/Math/SparseMatrixSolver.h:11-62) declaration by
// declaration -- every pure virtual with the reference's parameter types
// (Vector, CoefficientList, SparseEntry, the rectangular setRank) and the
// non-pure defaults of src/Math/SparseMatrixSolver.cpp (the triplet set(),
// solve(x0), solveLeastSquares(), printStatus()) -- with two differences
// only, both because Boost is not part of this build:
//   * setup() takes key/value pairs where the reference takes the
//     LinearAlgebra.<eqn> boost::property_tree::ptree;
//   * Type carries one more value, HIP.
// Vector and SparseEntry are minimal stand-ins with the members the interface
// uses (src/Math/Vector.h:10-56, src/Math/SparseEntry.h:6-17).  A maintainer
// dropping the adapter into the reference derives HipSparseMatrixSolver from
// the reference's own SparseMatrixSolver instead; the method bodies are
// unchanged.
#pragma once

#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <tuple>
#include <utility>
#include <vector>

#include "../../include/phase_hip.h"

namespace phase_amd {

using Index = int;      // src/Types/Types.h
using Size = std::size_t;
using Scalar = double;

// Exception(class, method, description) -- src/System/Exception.h
class Exception : public std::runtime_error {
 public:
  Exception(const std::string &cls, const std::string &method, const std::string &desc)
      : std::runtime_error(cls + "::" + method + ": " + desc) {}
};

// The parts of the reference's Vector (src/Math/Vector.h) the solver interface uses.
class Vector {
 public:
  explicit Vector(Size size = 0, Scalar val = 0.) : data_(size, val) {}
  Scalar &operator()(Index i) { return data_[i]; }
  Scalar operator()(Index i) const { return data_[i]; }
  size_t size() const { return data_.size(); }
  void resize(Size size) { data_.resize(size); }
  const std::vector<Scalar> &data() const { return data_; }

 private:
  std::vector<Scalar> data_;
};

// src/Math/SparseEntry.h
class SparseEntry {
 public:
  SparseEntry() {}
  SparseEntry(Index row, Index col, Scalar val) : row(row), col(col), val(val) {}
  Index row = 0, col = 0;
  Scalar val = 0.;
};

class SparseMatrixSolverBase {
 public:
  enum Type { EIGEN, TRILINOS_BELOS, TRILINOS_AMESOS2, TRILINOS_MUELU, HIP };

  typedef std::pair<Index, Scalar> Entry;
  typedef std::vector<Entry> Row;
  typedef std::vector<Row> CoefficientList;

  virtual ~SparseMatrixSolverBase() = default;
  virtual Type type() const = 0;
  virtual void setRank(int rank) = 0;
  virtual void setRank(int rowRank, int colRank) = 0;
  virtual void set(const std::vector<std::tuple<Index, Index, Scalar>> &entries);
  virtual void set(const CoefficientList &eqn) = 0;
  virtual void set(const std::vector<Index> &rowPtr, const std::vector<Index> &colInds,
                   const std::vector<Scalar> &vals) = 0;
  virtual void set(const std::vector<SparseEntry> &entries) = 0;
  virtual void setGuess(const Vector &x0) = 0;
  virtual void setRhs(const Vector &rhs) = 0;
  virtual Scalar solve() = 0;
  virtual Scalar solve(const Vector &x0);
  virtual Scalar solveLeastSquares();
  virtual Scalar x(Index idx) const = 0;
  virtual void setup(const std::map<std::string, std::string> &parameters) {}
  virtual int nIters() const = 0;
  virtual Scalar error() const = 0;
  virtual bool supportsMPI() const = 0;
  virtual void printStatus(const std::string &msg) const;

 protected:
  int nPreconUses_ = 1, maxPreconUses_ = 1;
};

class HipSparseMatrixSolver : public SparseMatrixSolverBase {
 public:
  explicit HipSparseMatrixSolver(int device = 0);
  ~HipSparseMatrixSolver() override;
  HipSparseMatrixSolver(const HipSparseMatrixSolver &) = delete;
  HipSparseMatrixSolver &operator=(const HipSparseMatrixSolver &) = delete;

  using SparseMatrixSolverBase::set;
  using SparseMatrixSolverBase::solve;

  Type type() const override { return HIP; }
  void setRank(int rank) override;
  // Square systems only: the reference reaches a rectangular rank only on
  // the way to solveLeastSquares(), which this backend does not offer.
  void setRank(int rowRank, int colRank) override;
  // Rows of (column, value) pairs; repeated columns of a row are summed, as
  // the reference's backends sum repeated insertions.
  void set(const CoefficientList &eqn) override;
  // Padded CSR, column -1 marks an empty slot (FiniteVolumeEquation::solve).
  void set(const std::vector<Index> &rowPtr, const std::vector<Index> &colInds,
           const std::vector<Scalar> &vals) override;
  void set(const std::vector<SparseEntry> &entries) override;
  void setGuess(const Vector &x0) override;
  void setRhs(const Vector &rhs) override;
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
  void setRows(const CoefficientList &rows, const char *method);
  phase_hip_solver *h_ = nullptr;
  int rank_ = 0;
  std::vector<Scalar> x_;
};

// SparseMatrixSolverFactory::create(const std::string&, ...) for this backend
// (src/Math/SparseMatrixSolverFactory.cpp:26-41): "hip" gives a
// HipSparseMatrixSolver on `device`; anything else is an Exception.
std::shared_ptr<SparseMatrixSolverBase> createSparseMatrixSolver(const std::string &type, int device = 0);

}  // namespace phase_amd
