// Drives phase_amd::HipSparseMatrixSolver the way FiniteVolumeEquation<T>::solve
// drives a SparseMatrixSolver (FiniteVolumeEquation.tpp:64-86): setRank, set
// (padded CSR with -1 slots), setRhs(-rhs), solve, x(i) per row.
// Usage: host_driver <in.bin> <out.bin> ; in.bin = int32 n, nslots, then
// rowPtr[n+1], colInd[nslots] (int32), vals[nslots], rhs[n] (float64).
#include <cstdio>
#include <fstream>
#include <vector>

#include "../../phase_amd/host/HipSparseMatrixSolver.h"

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
  try {
    phase_amd::HipSparseMatrixSolver s(0);
    s.setup({{"tolerance", "1e-10"}, {"maxIters", "5000"}, {"mode", "fp64"}});
    s.setRank(n);
    s.set(rp, ci, v);
    std::vector<double> b(n);
    for (int i = 0; i < n; ++i) b[i] = -rhs[i];
    s.setRhs(b);
    double err = s.solve();
    s.printStatus("FiniteVolumeEquation pCorr:");
    std::vector<double> x(n);
    for (int i = 0; i < n; ++i) x[i] = s.x(i);
    std::ofstream out(argv[2], std::ios::binary);
    out.write((char *)&err, 8);
    out.write((char *)x.data(), 8 * n);
    // error path: a bad pattern must surface as phase_amd::Exception
    try {
      s.set({0, 1}, {3}, {1.0});
      std::printf("no exception\n");
      return 3;
    } catch (const phase_amd::Exception &e) {
      std::printf("exception ok: %s\n", e.what());
    }
  } catch (const phase_amd::Exception &e) {
    std::printf("%s\n", e.what());
    return 1;
  }
  return 0;
}
