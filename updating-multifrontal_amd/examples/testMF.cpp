// testMF.cpp -- the reference's test flow (test_src/DE/testMF2D.cpp:124-278,
// test_src/DE/testMF3D.cpp:137-289) written against the drop-in headers, with assertions:
//   testMF <dim 2|3> <N> <width> <maxSideLength>
// builds the uniform-coefficient operator, factors, checks Apply(x) = Ax and Solve(Ax) = x,
// then doubles a corner block of the coefficient field, checks that the stale factorization's
// Apply is visibly wrong and that after Update it is exact again.
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <set>

#include "DE/mf.hpp"

using LinAlg::Sparse;
using LinAlg::Vec;

// -div(a grad u) with separator width 1, nodal padded coefficient a, Dirichlet via padding
static Sparse<double> laplacian(int dim, int N, const std::vector<double>& a) {
  int P = N + 2, n = 1;
  for (int d = 0; d < dim; d++) n *= N;
  std::vector<double> val;
  IdxVec ci, rp;
  auto A = [&](int i, int j, int k) { return a[((size_t)i * P + j) * (dim == 3 ? P : 1) + (dim == 3 ? k : 0)]; };
  int NZ = dim == 3 ? N : 1;
  for (int z = 0; z < NZ; z++)
    for (int j = 0; j < N; j++)
      for (int i = 0; i < N; i++) {
        rp.push_back((int)val.size());
        int ii = i + 1, jj = j + 1, zz = dim == 3 ? z + 1 : 0;
        int idx = i + j * N + z * N * N;
        double diag = 0;
        auto nb = [&](int di, int dj, int dk, int col) {
          double h = 0.5 * (A(ii, jj, zz) + A(ii + di, jj + dj, zz + dk));
          diag += h;
          if (col >= 0) { ci.push_back(col); val.push_back(-h); }
        };
        // ascending column order: -z, -y, -x, self, +x, +y, +z
        if (dim == 3) nb(0, 0, -1, z > 0 ? idx - N * N : -1);
        nb(0, -1, 0, j > 0 ? idx - N : -1);
        nb(-1, 0, 0, i > 0 ? idx - 1 : -1);
        size_t self = val.size();
        ci.push_back(idx); val.push_back(0);
        nb(1, 0, 0, i < N - 1 ? idx + 1 : -1);
        nb(0, 1, 0, j < N - 1 ? idx + N : -1);
        if (dim == 3) nb(0, 0, 1, z < N - 1 ? idx + N * N : -1);
        val[self] = diag;
      }
  rp.push_back((int)val.size());
  return Sparse<double>(val, ci, rp, n, n);
}

int main(int argc, char** argv) {
  int dim = argc > 1 ? atoi(argv[1]) : 2, N = argc > 2 ? atoi(argv[2]) : 127;
  int width = argc > 3 ? atoi(argv[3]) : 1, maxSide = argc > 4 ? atoi(argv[4]) : 8;
  if (width != 1) { printf("this driver builds width-1 operators only\n"); return 2; }
  int P = N + 2, n = 1;
  for (int d = 0; d < dim; d++) n *= N;
  size_t pad = (size_t)P * P * (dim == 3 ? P : 1);
  std::vector<double> a1(pad, 1.0), a2(pad, 1.0);
  int blk = dim == 2 ? 8 : 4;  // corner block doubled, as the reference's testProblemUpdating
  for (int i = 0; i < blk; i++)
    for (int j = 0; j < blk; j++)
      for (int k = 0; k < (dim == 3 ? blk : 1); k++) a2[((size_t)i * P + j) * (dim == 3 ? P : 1) + k] *= 2;
  Sparse<double> M1 = laplacian(dim, N, a1), M2 = laplacian(dim, N, a2);
  std::set<int> mod;
  for (int k = 0; k < M1.nnz(); k++)
    if (M1.val(k) != M2.val(k)) mod.insert(M1.col_ind(k));
  srand(1234);
  Vec<double> x(n);
  for (int i = 0; i < n; i++) x(i) = double(rand() % 1000000) / 1e6;

  auto run = [&](auto& F) {
    auto t0 = std::chrono::high_resolution_clock::now();
    F.Factor();
    auto t1 = std::chrono::high_resolution_clock::now();
    Vec<double> b = M1 * x;
    Vec<double> y = F.Apply(x);
    y -= b;
    double eapply = y.norm() / b.norm();
    Vec<double> s = F.Solve(b);
    s -= x;
    double esolve = s.norm() / x.norm();
    Vec<double> bt = M2 * x;
    Vec<double> y1 = F.Apply(x);
    y1 -= bt;
    double ebefore = y1.norm() / bt.norm();
    auto t2 = std::chrono::high_resolution_clock::now();
    F.Update(M2, mod);
    auto t3 = std::chrono::high_resolution_clock::now();
    Vec<double> y2 = F.Apply(x);
    y2 -= bt;
    double eafter = y2.norm() / bt.norm();
    printf("dim %d N %d n %d: factor %.3f s, update %.3f s (%zu modified DOFs)\n", dim, N, n,
           std::chrono::duration<double>(t1 - t0).count(), std::chrono::duration<double>(t3 - t2).count(), mod.size());
    printf("relative apply error %.3e, relative solve error %.3e\n", eapply, esolve);
    printf("relative apply error without updating %.3e, after updating %.3e\n", ebefore, eafter);
    bool ok = eapply < 1e-12 && esolve < 1e-10 && ebefore > 1e-6 && eafter < 1e-12;
    printf("%s\n", ok ? "PASS" : "FAIL");
    return ok ? 0 : 1;
  };
  try {
    if (dim == 2) {
      TreeFactor2D::MF2D<double> F(M1, N, N, width, maxSide);
      return run(F);
    }
    TreeFactor3D::MF3D<double> F(M1, N, N, N, width, maxSide);
    return run(F);
  } catch (std::exception& e) {
    printf("error: %s\n", e.what());
    return 3;
  }
}
