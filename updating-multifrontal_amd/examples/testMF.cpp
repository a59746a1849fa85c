// testMF.cpp -- the reference's test flow (test_src/DE/testMF2D.cpp:124-278,
// test_src/DE/testMF3D.cpp:137-289) written against the drop-in headers, with assertions:
//   testMF <dim 2|3> <N> <width> <maxSideLength>
// builds the operator (uniform coefficient for width 1, a random field for widths 2-3: then
// unsymmetric, factored with LU fronts), factors, checks Apply(x) = Ax and Solve(Ax) = x,
// then doubles a corner block of the coefficient field, checks that the stale factorization's
// Apply is visibly wrong and that after Update it is exact again.
#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <set>

#include "DE/mf.hpp"

using LinAlg::Sparse;
using LinAlg::Vec;

// -div(a grad u) with stencil width w (1-3), nodal padded coefficient a ((N + 2w)^dim, last index
// fastest), h = 1, Dirichlet via the padding: the entry to the node at offset +-k along an axis
// is cf_k (a_{k-1} + a_k) / 2 and the diagonal minus their sum -- the harness operator
// (test_include/DE/buildLaplacian.hpp).  With w > 1 and a variable field it is unsymmetric.
static Sparse<double> laplacian(int dim, int N, int w, const std::vector<double>& a) {
  static const double cf[4][4] = {{0, 0, 0, 0}, {0, -1, 0, 0}, {0, -4.0 / 3, 1.0 / 12, 0}, {0, -1.5, 3.0 / 20, -1.0 / 90}};
  const int P = N + 2 * w, NZ = dim == 3 ? N : 1;
  int n = 1;
  for (int d = 0; d < dim; d++) n *= N;
  auto A = [&](int i, int j, int k) { return a[((size_t)i * P + j) * (dim == 3 ? P : 1) + (dim == 3 ? k : 0)]; };
  std::vector<double> val;
  IdxVec ci, rp;
  const int str[3] = {1, N, N * N};
  for (int z = 0; z < NZ; z++)
    for (int j = 0; j < N; j++)
      for (int i = 0; i < N; i++) {
        rp.push_back((int)val.size());
        const int q[3] = {i, j, z}, pc[3] = {i + w, j + w, dim == 3 ? z + w : 0};
        const int idx = i + j * N + z * N * N;
        auto coef = [&](int ax, int s) {  // (a at offset s - sign(s) + a at offset s) / 2
          int p0[3] = {pc[0], pc[1], pc[2]}, p1[3] = {pc[0], pc[1], pc[2]};
          p0[ax] += s;
          p1[ax] += s > 0 ? s - 1 : s + 1;
          return 0.5 * (A(p0[0], p0[1], p0[2]) + A(p1[0], p1[1], p1[2]));
        };
        double diag = 0;
        for (int ax = 0; ax < dim; ax++)
          for (int k = 1; k <= w; k++) diag -= cf[w][k] * (coef(ax, -k) + coef(ax, k));
        // ascending columns: slowest axis first, negative offsets far to near, self, positive near to far
        for (int ax = dim - 1; ax >= 0; ax--)
          for (int k = w; k >= 1; k--)
            if (q[ax] - k >= 0) { ci.push_back(idx - k * str[ax]); val.push_back(cf[w][k] * coef(ax, -k)); }
        ci.push_back(idx);
        val.push_back(diag);
        for (int ax = 0; ax < dim; ax++)
          for (int k = 1; k <= w; k++)
            if (q[ax] + k < N) { ci.push_back(idx + k * str[ax]); val.push_back(cf[w][k] * coef(ax, k)); }
      }
  rp.push_back((int)val.size());
  // rows were emitted axis by axis; sort each row's columns ascending (the CSR convention)
  for (int r = 0; r < n; r++) {
    std::vector<std::pair<int, double>> row;
    for (int k = rp[r]; k < rp[r + 1]; k++) row.push_back({ci[k], val[k]});
    std::sort(row.begin(), row.end());
    for (int k = rp[r]; k < rp[r + 1]; k++) { ci[k] = row[k - rp[r]].first; val[k] = row[k - rp[r]].second; }
  }
  return Sparse<double>(val, ci, rp, n, n);
}

int main(int argc, char** argv) {
  int dim = argc > 1 ? atoi(argv[1]) : 2, N = argc > 2 ? atoi(argv[2]) : 127;
  int width = argc > 3 ? atoi(argv[3]) : 1, maxSide = argc > 4 ? atoi(argv[4]) : 8;
  if (width < 1 || width > 3) { printf("stencil width must be 1..3\n"); return 2; }
  int P = N + 2 * width, n = 1;
  for (int d = 0; d < dim; d++) n *= N;
  size_t pad = (size_t)P * P * (dim == 3 ? P : 1);
  // width 1: a = 1 (the reference's uniform test); wider stencils: a = 1 + U(0,1), which makes
  // the operator unsymmetric -- the reference's LU case, LU fronts here
  std::vector<double> a1(pad, 1.0);
  srand(4321);
  if (width > 1)
    for (double& v : a1) v = 1.0 + double(rand()) / RAND_MAX;
  std::vector<double> a2 = a1;
  int blk = dim == 2 ? 8 : 4;  // corner block doubled, as the reference's testProblemUpdating
  for (int i = 0; i < blk; i++)
    for (int j = 0; j < blk; j++)
      for (int k = 0; k < (dim == 3 ? blk : 1); k++) a2[((size_t)i * P + j) * (dim == 3 ? P : 1) + k] *= 2;
  Sparse<double> M1 = laplacian(dim, N, width, a1), M2 = laplacian(dim, N, width, a2);
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
