// Drives phase_amd::HipSparseMatrixSolver through the reference's plugin
// interface the way the reference's callers do: created by name through a
// factory as a std::shared_ptr to the base class
// (SparseMatrixSolverFactory.cpp:26-41), then setRank, set (padded CSR with -1
// slots, as FiniteVolumeEquation<T>::solve passes it, FiniteVolumeEquation.tpp:64-86),
// setRhs(-rhs), solve, x(i) per row.  The same system is then handed over
// through every other set() overload of SparseMatrixSolver.h:21-36
// (CoefficientList, SparseEntry list, triplets) and solved with solve(x0).
// Usage: host_driver <in.bin> <out.bin> ; in.bin = int32 n, nslots, then
// rowPtr[n+1], colInd[nslots] (int32), vals[nslots], rhs[n] (float64).
// out.bin = for each of the four set() forms: error (float64), x[n].
#include <cstdio>
#include <fstream>
#include <memory>
#include <type_traits>
#include <vector>

#include "../../phase_amd/host/HipSparseMatrixSolver.h"

using phase_amd::Index;
using phase_amd::Scalar;
using phase_amd::SparseMatrixSolverBase;

static_assert(!std::is_abstract<phase_amd::HipSparseMatrixSolver>::value,
              "the adapter must override every pure virtual of the plugin interface");

template <class F>
static bool throws(F &&f, const char *what) {
  try {
    f();
  } catch (const phase_amd::Exception &e) {
    std::printf("exception ok (%s): %s\n", what, e.what());
    return true;
  }
  std::printf("no exception: %s\n", what);
  return false;
}

int main(int argc, char **argv) {
  if (argc < 3) return 2;
  std::ifstream in(argv[1], std::ios::binary);
  int n = 0, ns = 0;
  in.read((char *)&n, 4);
  in.read((char *)&ns, 4);
  std::vector<int> rp(n + 1), ci(ns);
  std::vector<double> v(ns), rhs(n);
  in.read((char *)rp.data(), 4 * (n + 1));
  in.read((char *)ci.data(), 4 * ns);
  in.read((char *)v.data(), 8 * ns);
  in.read((char *)rhs.data(), 8 * n);
  std::ofstream out(argv[2], std::ios::binary);
  try {
    std::shared_ptr<SparseMatrixSolverBase> s = phase_amd::createSparseMatrixSolver("hip", 0);
    s->setup({{"tolerance", "1e-10"}, {"maxIters", "5000"}, {"mode", "fp64"}});
    phase_amd::Vector b(n), zero(n, 0.);
    for (int i = 0; i < n; ++i) b(i) = -rhs[i];
    auto emit = [&](double err) {
      out.write((char *)&err, 8);
      for (int i = 0; i < n; ++i) {
        const double xi = s->x(i);
        out.write((char *)&xi, 8);
      }
    };
    // 1. padded CSR
    s->setRank(n);
    s->set(rp, ci, v);
    s->setRhs(b);
    emit(s->solve());
    s->printStatus("FiniteVolumeEquation pCorr:");
    // 2. CoefficientList
    SparseMatrixSolverBase::CoefficientList rows(n);
    std::vector<phase_amd::SparseEntry> entries;
    std::vector<std::tuple<Index, Index, Scalar>> triplets;
    for (int r = 0; r < n; ++r)
      for (int j = rp[r]; j < rp[r + 1]; ++j)
        if (ci[j] >= 0) {
          rows[r].push_back({ci[j], v[j]});
          entries.emplace_back(r, ci[j], v[j]);
          triplets.emplace_back(r, ci[j], v[j]);
        }
    s->setRank(n);
    s->set(rows);
    s->setRhs(b);
    emit(s->solve(zero));
    // 3. SparseEntry list
    s->set(entries);
    s->setRhs(b);
    emit(s->solve(zero));
    // 4. triplets (the base class's conversion to a CoefficientList)
    s->set(triplets);
    s->setRhs(b);
    emit(s->solve(zero));
    // error paths surface as the reference's Exception type
    bool ok = throws([&] { s->set(std::vector<Index>{0, 1}, std::vector<Index>{3}, std::vector<Scalar>{1.0}); },
                     "bad pattern");
    ok &= throws([&] { s->setRank(n, n + 1); }, "rectangular rank");
    ok &= throws([&] { s->solveLeastSquares(); }, "least squares");
    ok &= throws([&] { phase_amd::createSparseMatrixSolver("belos"); }, "unknown backend");
    ok &= throws([&] { s->setRhs(phase_amd::Vector(n + 1)); }, "rhs size");
    if (!ok) return 3;
    std::printf("type %d\n", (int)s->type());
  } catch (const phase_amd::Exception &e) {
    std::printf("%s\n", e.what());
    return 1;
  }
  return 0;
}
