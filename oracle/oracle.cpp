// oracle.cpp -- CPU restatement of the reference multifrontal path.
//
// TEST INFRASTRUCTURE ONLY.  This file is the checker for the HIP library: it is loaded by
// tests/, by __graft_entry__.smoke() and by bench.py's cpu_baseline leg, never by the product.
//
// It restates, in plain sequential C++ with no external libraries, what the reference does:
//   * the nested-dissection planner           src/DE/mf2d.cpp:8-148, src/DE/mf3d.cpp:26-206
//   * grid geometry (edges/faces, children)   src/DE/tree2d_common.cpp:8-226, tree3d_common.cpp:9-328
//   * center points                           src/DE/mf2d.cpp:186-270, src/DE/mf3d.cpp:253-524
//   * separator / front assembly              src/DE/mf2d.cpp:274-446, src/DE/mf3d.cpp:528-702
//   * leaf / upper front factorization (LU)   src/DE/mf2d.cpp:504-753, src/DE/mf3d.cpp:770-1016
//   * extend-add scatters (find-based maps)   src/DE/mf2d.cpp:755-901, src/DE/mf3d.cpp:1018-1161
//   * Update's dirty-set marking              src/DE/mf2d.cpp:150-181, src/DE/mf3d.cpp:208-248
//   * Solve / Apply sweeps                    src/DE/mf2d.cpp:904-1054, src/DE/mf3d.cpp:1164-1313
//   * the test operator and modDOF rule       test_include/DE/buildLaplacian.hpp:10-132, :244-391;
//                                             test_src/DE/testMF2D.cpp:143-150
// Dense kernels restate LAPACK dgetrf/dgetrs semantics (partial pivoting, unit lower L),
// as called through include/LinAlg/dense.hpp:180-227.
//
// Pinning: every function here is checked by tests/test_oracle.py against the reference built
// from /root/reference (oracle/_ref, when present) and against committed fixtures in
// tests/golden/, plus closed forms (5-/7-point stencil, dense Schur complement, Solve(Ax)=x).
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <set>
#include <stdexcept>
#include <string>
#include <vector>

namespace orc {

typedef std::vector<int> Idx;

// ---------------------------------------------------------------- dense column-major matrix
struct Mat {
  int m = 0, n = 0;
  std::vector<double> a;
  std::vector<int> piv;  // LU pivots (0-based row swapped with at step k)
  Mat() {}
  Mat(int m_, int n_) : m(m_), n(n_), a((size_t)m_ * n_, 0.0) {}
  double& operator()(int i, int j) { return a[(size_t)j * m + i]; }
  double operator()(int i, int j) const { return a[(size_t)j * m + i]; }
};

// LU with partial pivoting, dgetrf semantics (include/LinAlg/dense.hpp:180-188).
// Returns 0, or k+1 when U(k,k) == 0 (LAPACK info convention).
static int lu(Mat& A) {
  int n = A.n;
  A.piv.assign(n, 0);
  int info = 0;
  for (int k = 0; k < n; k++) {
    int p = k;
    double best = std::fabs(A(k, k));
    for (int i = k + 1; i < n; i++)
      if (std::fabs(A(i, k)) > best) { best = std::fabs(A(i, k)); p = i; }
    A.piv[k] = p;
    if (A(p, k) == 0.0) { if (!info) info = k + 1; continue; }
    if (p != k)
      for (int j = 0; j < n; j++) std::swap(A(k, j), A(p, j));
    double inv = 1.0 / A(k, k);
    for (int i = k + 1; i < n; i++) A(i, k) *= inv;
    for (int j = k + 1; j < n; j++) {
      double akj = A(k, j);
      if (akj == 0.0) continue;
      double* cj = &A.a[(size_t)j * n];
      const double* ck = &A.a[(size_t)k * n];
      for (int i = k + 1; i < n; i++) cj[i] -= ck[i] * akj;
    }
  }
  return info;
}

// Solve A X = B in place (dgetrs 'N'), B is n x nrhs column-major (dense.hpp:190-212).
static void lu_solve(const Mat& A, double* B, int nrhs, int ldb) {
  int n = A.n;
  for (int r = 0; r < nrhs; r++) {
    double* b = B + (size_t)r * ldb;
    for (int k = 0; k < n; k++) std::swap(b[k], b[A.piv[k]]);
    for (int j = 0; j < n; j++) {
      double bj = b[j];
      if (bj == 0.0) continue;
      for (int i = j + 1; i < n; i++) b[i] -= A(i, j) * bj;
    }
    for (int j = n - 1; j >= 0; j--) {
      b[j] /= A(j, j);
      double bj = b[j];
      if (bj == 0.0) continue;
      for (int i = 0; i < j; i++) b[i] -= A(i, j) * bj;
    }
  }
}

// Solve A^T X = B in place (dgetrs 'T').
static void lu_solve_t(const Mat& A, double* B, int nrhs, int ldb) {
  int n = A.n;
  for (int r = 0; r < nrhs; r++) {
    double* b = B + (size_t)r * ldb;
    // U^T y = b
    for (int j = 0; j < n; j++) {
      double s = b[j];
      for (int i = 0; i < j; i++) s -= A(i, j) * b[i];
      b[j] = s / A(j, j);
    }
    // L^T z = y
    for (int j = n - 1; j >= 0; j--) {
      double s = b[j];
      for (int i = j + 1; i < n; i++) s -= A(i, j) * b[i];
      b[j] = s;
    }
    for (int k = n - 1; k >= 0; k--) std::swap(b[k], b[A.piv[k]]);
  }
}

// XL = B * A^{-1} via transpose + getrs('T') + transpose (dense.hpp:214-227).
static void lu_solve_right(const Mat& A, Mat& B) {
  Mat T(B.n, B.m);
  for (int j = 0; j < B.n; j++)
    for (int i = 0; i < B.m; i++) T(j, i) = B(i, j);
  lu_solve_t(A, T.a.data(), T.n, T.m);
  for (int j = 0; j < B.n; j++)
    for (int i = 0; i < B.m; i++) B(i, j) = T(j, i);
}

static Mat matmul(const Mat& A, const Mat& B) {  // dgemm, C = A*B (src/LinAlg/dense.cpp:152-163)
  Mat C(A.m, B.n);
  for (int j = 0; j < B.n; j++)
    for (int k = 0; k < A.n; k++) {
      double b = B(k, j);
      if (b == 0.0) continue;
      for (int i = 0; i < A.m; i++) C(i, j) += A(i, k) * b;
    }
  return C;
}

// ---------------------------------------------------------------- CSR operator
struct Csr {
  int n = 0;
  std::vector<int> rp, ci;
  std::vector<double> v;
  // A_(i,j): first match in row i, else 0 (src/LinAlg/sparse.cpp:80-92)
  double at(int i, int j) const {
    for (int k = rp[i]; k < rp[i + 1]; k++)
      if (ci[k] == j) return v[k];
    return 0.0;
  }
  // A_(rows, cols) dense extraction (src/LinAlg/sparse.cpp:95-116).  Same result as the
  // reference's per-entry row scan (the last matching CSR entry of a row wins), via a column table.
  Mat sub(const Idx& rows, const Idx& cols) const {
    Mat B((int)rows.size(), (int)cols.size());
    if (rows.empty() || cols.empty()) return B;
    int maxc = *std::max_element(cols.begin(), cols.end());
    std::vector<int> cpos(maxc + 1, -1);
    for (int jj = 0; jj < (int)cols.size(); jj++) cpos[cols[jj]] = jj;
    for (int ii = 0; ii < (int)rows.size(); ii++) {
      int i = rows[ii];
      for (int k = rp[i]; k < rp[i + 1]; k++) {
        int j = ci[k];
        if (j <= maxc && cpos[j] >= 0) B(ii, cpos[j]) = v[k];
      }
    }
    return B;
  }
};

// ---------------------------------------------------------------- geometry
// 2D: src/DE/tree2d_common.cpp.  Edge index layout: rows of 2b-1 entries, T2B first (b-1), then L2R (b).
struct Geo2 {
  enum { T2B = 0, L2R = 1 };
  static int type(int e, int ell) { int b = 1 << ell; return (e % (2 * b - 1) < b - 1) ? T2B : L2R; }
  static int idx(int row, int col, int ell, int t) {
    int b = 1 << ell;
    return t == L2R ? row * (2 * b - 1) + col + b - 1 : row * (2 * b - 1) + col;
  }
  static void coords(int e, int ell, int& x, int& y) {
    int b = 1 << ell;
    if (type(e, ell) == L2R) e -= b - 1;
    x = e % (2 * b - 1);
    y = e / (2 * b - 1);
  }
  static int nseps(int ell) { int b = 1 << ell; return 2 * b * b - 2 * b; }
  static int nboxes(int ell) { int b = 1 << ell; return b * b; }
  static Idx nborSeps(int box, int ell) {  // tree2d_common.cpp:135-184
    int b = 1 << ell, x = box % b, y = box / b;
    Idx r;
    r.push_back(x > 0 ? idx(y, x - 1, ell, T2B) : -1);
    r.push_back(x < b - 1 ? idx(y, x, ell, T2B) : -1);
    r.push_back(y > 0 ? idx(y - 1, x, ell, L2R) : -1);
    r.push_back(y < b - 1 ? idx(y, x, ell, L2R) : -1);
    return r;
  }
  static Idx nborBoxes(int e, int ell) {  // tree2d_common.cpp:186-211
    int b = 1 << ell, x, y;
    coords(e, ell, x, y);
    if (type(e, ell) == T2B) return Idx{x + y * b, (x + 1) + y * b};
    return Idx{x + y * b, x + (y + 1) * b};
  }
  static Idx sepChildren(int e, int ell) {  // tree2d_common.cpp:91-119
    int x, y;
    coords(e, ell, x, y);
    if (type(e, ell) == L2R)
      return Idx{idx(2 * y + 1, 2 * x, ell + 1, L2R), idx(2 * y + 1, 2 * x + 1, ell + 1, L2R)};
    return Idx{idx(2 * y, 2 * x + 1, ell + 1, T2B), idx(2 * y + 1, 2 * x + 1, ell + 1, T2B)};
  }
  static Idx boxChildren(int box, int ell) {  // tree2d_common.cpp:121-131
    int b = 1 << ell, x = box % b, y = box / b, W = 2 * b;
    return Idx{2 * x + 2 * y * W, 2 * x + (2 * y + 1) * W, 2 * x + 1 + 2 * y * W, 2 * x + 1 + (2 * y + 1) * W};
  }
  static int parent(int box, int ell) {
    int b = 1 << ell, x = box % b, y = box / b;
    return x / 2 + (y / 2) * (b / 2);
  }
  // self (interior, parity 0) / shared (parity 1) child separators, std::set-sorted
  static Idx childSeps(int box, int ell, int parity) {  // tree2d_common.cpp:37-79
    std::set<int> s;
    for (int c : boxChildren(box, ell))
      for (int e : nborSeps(c, ell + 1)) {
        if (e == -1) continue;
        int x, y;
        coords(e, ell + 1, x, y);
        int t = type(e, ell + 1);
        if ((t == T2B && x % 2 == parity) || (t == L2R && y % 2 == parity)) s.insert(e);
      }
    return Idx(s.begin(), s.end());
  }
};

// 3D: src/DE/tree3d_common.cpp.  Per z-slice: rows of (X-normal b-1, Y-normal b) then b*b Z-normal.
struct Geo3 {
  enum { XN = 0, YN = 1, ZN = 2 };
  static int fps(int b) { return 2 * (b - 1) * b + b * b; }
  static int type(int f, int ell) {
    int b = 1 << ell;
    f %= fps(b);
    if (f >= 2 * (b - 1) * b) return ZN;
    f %= (2 * b - 1);
    return f < b - 1 ? XN : YN;
  }
  static int idx(int row, int col, int sli, int ell, int t) {
    int b = 1 << ell, F = fps(b);
    if (t == XN) return sli * F + row * (2 * b - 1) + col;
    if (t == YN) return sli * F + row * (2 * b - 1) + col + b - 1;
    return sli * F + row * b + col + 2 * (b - 1) * b;
  }
  static void coords(int f, int ell, int& x, int& y, int& z) {
    int b = 1 << ell, F = fps(b), t = type(f, ell);
    if (t == ZN) {
      f -= 2 * (b - 1) * b;
      z = f / F; f -= z * F; y = f / b; x = f - y * b;
      return;
    }
    if (t == YN) f -= b - 1;
    z = f / F; f -= z * F; y = f / (2 * b - 1); x = f - y * (2 * b - 1);
  }
  static int nseps(int ell) { int b = 1 << ell; return 3 * b * b * (b - 1); }
  static int nboxes(int ell) { int b = 1 << ell; return b * b * b; }
  static Idx nborSeps(int box, int ell) {  // tree3d_common.cpp:185-275
    int b = 1 << ell, x = box % b, y = (box / b) % b, z = box / (b * b);
    Idx r;
    r.push_back(x > 0 ? idx(y, x - 1, z, ell, XN) : -1);
    r.push_back(x < b - 1 ? idx(y, x, z, ell, XN) : -1);
    r.push_back(y > 0 ? idx(y - 1, x, z, ell, YN) : -1);
    r.push_back(y < b - 1 ? idx(y, x, z, ell, YN) : -1);
    r.push_back(z > 0 ? idx(y, x, z - 1, ell, ZN) : -1);
    r.push_back(z < b - 1 ? idx(y, x, z, ell, ZN) : -1);
    return r;
  }
  static Idx nborBoxes(int f, int ell) {  // tree3d_common.cpp:277-312
    int b = 1 << ell, x, y, z;
    coords(f, ell, x, y, z);
    int t = type(f, ell), me = x + y * b + z * b * b;
    if (t == XN) return Idx{me, me + 1};
    if (t == YN) return Idx{me, me + b};
    return Idx{me, me + b * b};
  }
  static Idx sepChildren(int f, int ell) {  // tree3d_common.cpp:127-164
    int x, y, z;
    coords(f, ell, x, y, z);
    int t = type(f, ell), e = ell + 1;
    int r = y, c = x, s = z;
    if (t == XN)
      return Idx{idx(2 * r, 2 * c + 1, 2 * s, e, XN), idx(2 * r, 2 * c + 1, 2 * s + 1, e, XN),
                 idx(2 * r + 1, 2 * c + 1, 2 * s, e, XN), idx(2 * r + 1, 2 * c + 1, 2 * s + 1, e, XN)};
    if (t == YN)
      return Idx{idx(2 * r + 1, 2 * c, 2 * s, e, YN), idx(2 * r + 1, 2 * c, 2 * s + 1, e, YN),
                 idx(2 * r + 1, 2 * c + 1, 2 * s, e, YN), idx(2 * r + 1, 2 * c + 1, 2 * s + 1, e, YN)};
    return Idx{idx(2 * r, 2 * c, 2 * s + 1, e, ZN), idx(2 * r + 1, 2 * c, 2 * s + 1, e, ZN),
               idx(2 * r, 2 * c + 1, 2 * s + 1, e, ZN), idx(2 * r + 1, 2 * c + 1, 2 * s + 1, e, ZN)};
  }
  static Idx boxChildren(int box, int ell) {  // tree3d_common.cpp:166-181
    int b = 1 << ell, x = box % b, y = (box / b) % b, z = box / (b * b), W = 2 * b;
    auto L = [&](int i, int j, int k) { return i + j * W + k * W * W; };
    return Idx{L(2 * x, 2 * y, 2 * z),         L(2 * x + 1, 2 * y, 2 * z),     L(2 * x, 2 * y + 1, 2 * z),
               L(2 * x, 2 * y, 2 * z + 1),     L(2 * x + 1, 2 * y + 1, 2 * z), L(2 * x + 1, 2 * y, 2 * z + 1),
               L(2 * x, 2 * y + 1, 2 * z + 1), L(2 * x + 1, 2 * y + 1, 2 * z + 1)};
  }
  static int parent(int box, int ell) {
    int b = 1 << ell, x = box % b, y = (box / b) % b, z = box / (b * b), h = b / 2;
    return x / 2 + (y / 2) * h + (z / 2) * h * h;
  }
  static Idx childSeps(int box, int ell, int parity) {  // tree3d_common.cpp:73-115
    std::set<int> s;
    for (int c : boxChildren(box, ell))
      for (int f : nborSeps(c, ell + 1)) {
        if (f == -1) continue;
        int x, y, z;
        coords(f, ell + 1, x, y, z);
        int t = type(f, ell + 1);
        if ((t == XN && x % 2 == parity) || (t == YN && y % 2 == parity) || (t == ZN && z % 2 == parity))
          s.insert(f);
      }
    return Idx(s.begin(), s.end());
  }
};

// ---------------------------------------------------------------- the multifrontal factorization
struct Box {
  Idx intDOFs, bdDOFs;
  Mat A22, A22Inv, XL, XR, S, A11post;
};
struct Sep {
  Idx DOFs;
  Mat A11;
};

struct Front {
  virtual ~Front() {}
  int dim = 2, N[3] = {1, 1, 1}, w = 1, L = 0, bw[3] = {0, 0, 0};
  Csr A;
  std::vector<std::vector<Box>> boxes;  // boxes[ell][b]
  std::vector<std::vector<Sep>> seps;   // seps[ell-1][e]: separators between level-ell boxes
  std::set<int> markedBoxes, markedSeps;
  // dirty-set trace of the last Factor/Update: per level, boxes factored and separators rebuilt
  std::vector<std::vector<int>> traceBoxes, traceSeps;
  int info = 0;
  bool factored = false;  // the reference's Update on an unfactored object reads unset fronts

  virtual int nboxes(int ell) const = 0;
  virtual int nseps(int ell) const = 0;
  virtual Idx nborSeps(int box, int ell) const = 0;
  virtual Idx nborBoxes(int e, int ell) const = 0;
  virtual Idx sepChildren(int e, int ell) const = 0;
  virtual Idx boxChildren(int box, int ell) const = 0;
  virtual int parent(int box, int ell) const = 0;
  virtual Idx childSeps(int box, int ell, int parity) const = 0;
  virtual Idx boxCenter(int box, int ell) const = 0;
  virtual Idx sepCenter(int e, int ell) const = 0;
  virtual int leafOfDof(int dof) const = 0;

  // nominal box width at level ell along axis a (mf2d.cpp:196-200)
  int BW(int a, int ell) const {
    int d = L - 1 - ell;
    return bw[a] * (1 << d) + ((1 << d) - 1) * w;
  }

  // ScatterAdd_ / ScatterReplace_ (mf2d.cpp:755-834): first-match find maps
  static void maps(const Idx& big, const Idx& small, Idx& t1, Idx& t2) {
    t1.clear(); t2.clear();
    for (int i = 0; i < (int)small.size(); i++) {
      auto it = std::find(big.begin(), big.end(), small[i]);
      if (it != big.end()) { t1.push_back((int)(it - big.begin())); t2.push_back(i); }
    }
  }
  static void scatterAdd(Mat& A, const Idx& big, const Mat& B, const Idx& small) {
    Idx t1, t2; maps(big, small, t1, t2);
    for (size_t i = 0; i < t1.size(); i++)
      for (size_t j = 0; j < t1.size(); j++) A(t1[i], t1[j]) += B(t2[i], t2[j]);
  }
  static void scatterReplace(Mat& A, const Idx& big, const Mat& B, const Idx& small) {
    Idx t1, t2; maps(big, small, t1, t2);
    for (size_t i = 0; i < t1.size(); i++)
      for (size_t j = 0; j < t1.size(); j++) A(t1[i], t1[j]) = B(t2[i], t2[j]);
  }
  static void scatterReplace2(Mat& A, const Idx& bigR, const Idx& bigC, const Mat& B, const Idx& small) {
    Idx r1, r2, c1, c2; maps(bigR, small, r1, r2); maps(bigC, small, c1, c2);
    for (size_t i = 0; i < r1.size(); i++)
      for (size_t j = 0; j < c1.size(); j++) A(r1[i], c1[j]) = B(r2[i], c2[j]);
  }
  static int findPos(const Idx& v, int x) { return (int)(std::find(v.begin(), v.end(), x) - v.begin()); }
  // AddNewInteractions_ (mf2d.cpp:836-862)
  void addNew(Mat& M, const Idx& total, const Idx& nw) const {
    Idx table;
    std::vector<bool> flags(total.size());
    for (int x : nw) { int p = findPos(total, x); table.push_back(p); flags[p] = true; }
    for (int i = 0; i < (int)total.size(); i++)
      for (int j = 0; j < (int)nw.size(); j++) {
        M(i, table[j]) += A.at(total[i], nw[j]);
        if (table[j] != i && !flags[i]) M(table[j], i) += A.at(nw[j], total[i]);
      }
  }
  // AddNewRowInteractions_ / AddNewColInteractions_ (mf2d.cpp:864-901)
  void addNewRow(Mat& M, const Idx& rows, const Idx& cols, const Idx& nw) const {
    Idx table;
    for (int x : nw) table.push_back(findPos(rows, x));
    for (int i = 0; i < (int)nw.size(); i++)
      for (int j = 0; j < (int)cols.size(); j++) M(table[i], j) += A.at(nw[i], cols[j]);
  }
  void addNewCol(Mat& M, const Idx& rows, const Idx& cols, const Idx& nw) const {
    Idx table;
    for (int x : nw) table.push_back(findPos(cols, x));
    for (int i = 0; i < (int)rows.size(); i++)
      for (int j = 0; j < (int)nw.size(); j++) M(i, table[j]) += A.at(rows[i], nw[j]);
  }

  // Front elimination: LU, XL = A12 A22^-1, XR = A22^-1 A21, S = -A12 XR  (mf2d.cpp:559-567)
  void eliminate(Box& B, const Mat& A12, const Mat& A21) {
    B.A22Inv = B.A22;
    int r = lu(B.A22Inv);
    if (r && !info) info = r;
    B.XL = A12;
    lu_solve_right(B.A22Inv, B.XL);
    B.XR = A21;
    lu_solve(B.A22Inv, B.XR.a.data(), B.XR.n, B.XR.m);
    B.S = matmul(A12, B.XR);
    for (double& x : B.S.a) x *= -1.0;
  }
  static void addInPlace(Mat& A, const Mat& B) {
    for (size_t i = 0; i < A.a.size(); i++) A.a[i] += B.a[i];
  }

  // mf2d.cpp:504-626
  void factorLeaf() {
    int ell = L - 1;
    traceBoxes[ell] = Idx(markedBoxes.begin(), markedBoxes.end());
    for (int b : markedBoxes) {
      Box& B = boxes[ell][b];
      B.bdDOFs.clear();
      Idx ns = nborSeps(b, ell);
      for (int e : ns) {
        if (e == -1) continue;
        const Idx& d = seps[ell - 1][e].DOFs;
        B.bdDOFs.insert(B.bdDOFs.end(), d.begin(), d.end());
      }
      B.A22 = A.sub(B.intDOFs, B.intDOFs);
      Mat A21 = A.sub(B.intDOFs, B.bdDOFs);
      Mat A12 = A.sub(B.bdDOFs, B.intDOFs);
      eliminate(B, A12, A21);
      B.A11post = A.sub(B.bdDOFs, B.bdDOFs);
      addInPlace(B.A11post, B.S);
      markedSeps.insert(ns.begin(), ns.end());
    }
    markedBoxes.clear();
    std::set<int> next;
    Idx tr;
    for (int e : markedSeps) {
      if (e == -1) continue;
      tr.push_back(e);
      Sep& E = seps[ell - 1][e];
      E.A11 = A.sub(E.DOFs, E.DOFs);
      for (int bb : nborBoxes(e, ell)) {
        scatterAdd(E.A11, E.DOFs, boxes[ell][bb].S, boxes[ell][bb].bdDOFs);
        int p = parent(bb, ell);
        markedBoxes.insert(p);
        if (ell > 1) { Idx q = nborSeps(p, ell - 1); next.insert(q.begin(), q.end()); }
      }
    }
    traceSeps[ell] = tr;
    markedSeps = next;
  }

  // updateEdgeData_ / updateFaceData_ (mf2d.cpp:274-319)
  void updateSeps(int ell) {
    Idx tr;
    for (int e : markedSeps) {
      if (e == -1) continue;
      tr.push_back(e);
      Sep& E = seps[ell - 1][e];
      E.DOFs.clear();
      Idx ch = sepChildren(e, ell);
      for (int c : ch) { const Idx& d = seps[ell][c].DOFs; E.DOFs.insert(E.DOFs.end(), d.begin(), d.end()); }
      Idx cp = sepCenter(e, ell);
      E.DOFs.insert(E.DOFs.end(), cp.begin(), cp.end());
      E.A11 = Mat((int)E.DOFs.size(), (int)E.DOFs.size());
      addNew(E.A11, E.DOFs, cp);
      for (int c : ch) scatterReplace(E.A11, E.DOFs, seps[ell][c].A11, seps[ell][c].DOFs);
    }
    traceRebuilt[ell] = tr;
    markedSeps.clear();
  }
  std::vector<std::vector<int>> traceRebuilt;

  // updateBoxData_ (mf2d.cpp:322-446)
  void updateBoxes(int ell) {
    for (int b : markedBoxes) {
      Box& B = boxes[ell][b];
      Idx& self = B.intDOFs;
      Idx& nb = B.bdDOFs;
      self.clear(); nb.clear();
      Idx sc = childSeps(b, ell, 0);
      for (int e : sc) { const Idx& d = seps[ell][e].DOFs; self.insert(self.end(), d.begin(), d.end()); }
      Idx cp = boxCenter(b, ell);
      self.insert(self.end(), cp.begin(), cp.end());
      B.A22 = Mat((int)self.size(), (int)self.size());
      addNew(B.A22, self, cp);
      Idx ns = nborSeps(b, ell);
      for (int e : ns) {
        if (e == -1) continue;
        const Idx& d = seps[ell - 1][e].DOFs;
        nb.insert(nb.end(), d.begin(), d.end());
      }
      B.A11post = Mat((int)nb.size(), (int)nb.size());
      B.XL = Mat((int)nb.size(), (int)self.size());  // holds A12 until FactorUpperLevel_
      B.XR = Mat((int)self.size(), (int)nb.size());  // holds A21
      for (int e : ns) {
        if (e == -1) continue;
        Idx ecp = sepCenter(e, ell);
        addNewRow(B.XL, nb, self, ecp);
        addNewCol(B.XR, self, nb, ecp);
      }
      for (int c : boxChildren(b, ell)) {
        Box& C = boxes[ell + 1][c];
        scatterReplace2(B.XL, nb, self, C.A11post, C.bdDOFs);
        scatterReplace2(B.XR, self, nb, C.A11post, C.bdDOFs);
        scatterReplace(B.A22, self, C.A11post, C.bdDOFs);
        scatterReplace(B.A11post, nb, C.A11post, C.bdDOFs);
      }
      for (int e : childSeps(b, ell, 0)) scatterReplace(B.A22, self, seps[ell][e].A11, seps[ell][e].DOFs);
      for (int e : childSeps(b, ell, 1)) scatterReplace(B.A11post, nb, seps[ell][e].A11, seps[ell][e].DOFs);
    }
  }

  // FactorUpperLevel_ (mf2d.cpp:628-753)
  void factorUpper(int ell) {
    traceBoxes[ell] = Idx(markedBoxes.begin(), markedBoxes.end());
    if (ell == 0) {
      Box& R = boxes[0][0];
      R.A22Inv = R.A22;
      int r = lu(R.A22Inv);
      if (r && !info) info = r;
      return;
    }
    for (int b : markedBoxes) {
      Box& B = boxes[ell][b];
      Mat A12 = B.XL, A21 = B.XR;
      eliminate(B, A12, A21);
      addInPlace(B.A11post, B.S);
      Idx ns = nborSeps(b, ell);
      markedSeps.insert(ns.begin(), ns.end());
    }
    markedBoxes.clear();
    std::set<int> next;
    Idx tr;
    for (int e : markedSeps) {
      if (e == -1) continue;
      tr.push_back(e);
      Sep& E = seps[ell - 1][e];
      for (int bb : nborBoxes(e, ell)) {
        scatterAdd(E.A11, E.DOFs, boxes[ell][bb].S, boxes[ell][bb].bdDOFs);
        int p = parent(bb, ell);
        markedBoxes.insert(p);
        if (ell > 1) { Idx q = nborSeps(p, ell - 1); next.insert(q.begin(), q.end()); }
      }
    }
    traceSeps[ell] = tr;
    markedSeps = next;
  }

  // Factor (mf2d.cpp:449-500)
  void factor() {
    info = 0;
    traceBoxes.assign(L, Idx());
    traceSeps.assign(L, Idx());
    traceRebuilt.assign(L, Idx());
    factorLeaf();
    for (int ell = L - 2; ell >= 0; ell--) {
      if (ell != 0) updateSeps(ell);
      updateBoxes(ell);
      factorUpper(ell);
    }
  }

  // Update (mf2d.cpp:150-181); returns -1 if a modified DOF maps outside the leaf grid
  int update(const std::vector<double>& vals, const Idx& mod) {
    markedBoxes.clear();
    markedSeps.clear();
    A.v = vals;
    for (int d : mod) {
      int leaf = leafOfDof(d);
      if (leaf < 0) return -1;
      markedBoxes.insert(leaf);
    }
    factor();
    return 0;
  }

  // Solve (mf2d.cpp:976-1054): one right-hand side per sweep
  std::vector<double> solve(const std::vector<double>& x) const {
    std::vector<double> b(x);
    for (int ell = L - 1; ell > 0; ell--)
      for (int bi = 0; bi < nboxes(ell); bi++) {
        const Box& B = boxes[ell][bi];
        int n = (int)B.intDOFs.size(), m = (int)B.bdDOFs.size();
        std::vector<double> rd(n), sk(m);
        for (int i = 0; i < n; i++) rd[i] = b[B.intDOFs[i]];
        for (int i = 0; i < m; i++) sk[i] = b[B.bdDOFs[i]];
        for (int i = 0; i < m; i++) {  // b_sk -= XL * b_rd
          double s = 0;
          for (int j = 0; j < n; j++) s += B.XL(i, j) * rd[j];
          sk[i] -= s;
        }
        for (int i = 0; i < m; i++) b[B.bdDOFs[i]] = sk[i];
        lu_solve(B.A22Inv, rd.data(), 1, n);
        for (int i = 0; i < n; i++) b[B.intDOFs[i]] = rd[i];
      }
    {
      const Box& R = boxes[0][0];
      int n = (int)R.intDOFs.size();
      std::vector<double> rd(n);
      for (int i = 0; i < n; i++) rd[i] = b[R.intDOFs[i]];
      lu_solve(R.A22Inv, rd.data(), 1, n);
      for (int i = 0; i < n; i++) b[R.intDOFs[i]] = rd[i];
    }
    for (int ell = 1; ell < L; ell++)
      for (int bi = 0; bi < nboxes(ell); bi++) {
        const Box& B = boxes[ell][bi];
        int n = (int)B.intDOFs.size(), m = (int)B.bdDOFs.size();
        for (int i = 0; i < n; i++) {  // b_rd -= XR * b_sk
          double s = 0;
          for (int j = 0; j < m; j++) s += B.XR(i, j) * b[B.bdDOFs[j]];
          b[B.intDOFs[i]] -= s;
        }
      }
    return b;
  }

  // Apply (mf2d.cpp:904-974)
  std::vector<double> apply(const std::vector<double>& x) const {
    std::vector<double> b(x);
    for (int ell = L - 1; ell > 0; ell--)
      for (int bi = 0; bi < nboxes(ell); bi++) {
        const Box& B = boxes[ell][bi];
        int n = (int)B.intDOFs.size(), m = (int)B.bdDOFs.size();
        std::vector<double> rd(n), t(n);
        for (int i = 0; i < n; i++) {
          double s = 0;
          for (int j = 0; j < m; j++) s += B.XR(i, j) * b[B.bdDOFs[j]];
          rd[i] = b[B.intDOFs[i]] + s;
        }
        for (int i = 0; i < n; i++) {
          double s = 0;
          for (int j = 0; j < n; j++) s += B.A22(i, j) * rd[j];
          t[i] = s;
        }
        for (int i = 0; i < n; i++) b[B.intDOFs[i]] = t[i];
      }
    {
      const Box& R = boxes[0][0];
      int n = (int)R.intDOFs.size();
      std::vector<double> rd(n), t(n);
      for (int i = 0; i < n; i++) rd[i] = b[R.intDOFs[i]];
      for (int i = 0; i < n; i++) {
        double s = 0;
        for (int j = 0; j < n; j++) s += R.A22(i, j) * rd[j];
        t[i] = s;
      }
      for (int i = 0; i < n; i++) b[R.intDOFs[i]] = t[i];
    }
    for (int ell = 1; ell < L; ell++)
      for (int bi = 0; bi < nboxes(ell); bi++) {
        const Box& B = boxes[ell][bi];
        int n = (int)B.intDOFs.size(), m = (int)B.bdDOFs.size();
        for (int i = 0; i < m; i++) {
          double s = 0;
          for (int j = 0; j < n; j++) s += B.XL(i, j) * b[B.intDOFs[j]];
          b[B.bdDOFs[i]] += s;
        }
      }
    return b;
  }
};

static int levelsFor(int maxN, int w, int maxSide) {  // mf2d.cpp:23-31
  double num = maxN + w, den = maxSide + w;
  return (int)std::ceil(std::log2(num / den));
}

struct Front2 : Front {
  int nboxes(int ell) const override { return Geo2::nboxes(ell); }
  int nseps(int ell) const override { return Geo2::nseps(ell); }
  Idx nborSeps(int b, int ell) const override { return Geo2::nborSeps(b, ell); }
  Idx nborBoxes(int e, int ell) const override { return Geo2::nborBoxes(e, ell); }
  Idx sepChildren(int e, int ell) const override { return Geo2::sepChildren(e, ell); }
  Idx boxChildren(int b, int ell) const override { return Geo2::boxChildren(b, ell); }
  int parent(int b, int ell) const override { return Geo2::parent(b, ell); }
  Idx childSeps(int b, int ell, int p) const override { return Geo2::childSeps(b, ell, p); }
  int leafOfDof(int d) const override {  // mf2d.cpp:163-174
    int b = 1 << (L - 1);
    int i = d % N[0], j = d / N[0];
    int ib = i / (bw[0] + w), jb = j / (bw[1] + w);
    if (ib >= b || jb >= b) return -1;
    return ib + jb * b;
  }
  Idx boxCenter(int box, int ell) const override {  // mf2d.cpp:186-218
    int b = 1 << ell, col = box % b, row = box / b;
    int fx = col * (BW(0, ell) + w) + BW(0, ell + 1), fy = row * (BW(1, ell) + w) + BW(1, ell + 1);
    Idx r;
    for (int j = fy; j < fy + w; j++)
      for (int i = fx; i < fx + w; i++) r.push_back(i + j * N[0]);
    return r;
  }
  Idx sepCenter(int e, int ell) const override {  // mf2d.cpp:220-270
    int col, row;
    Geo2::coords(e, ell, col, row);
    int fx, fy;
    if (Geo2::type(e, ell) == Geo2::L2R) {
      fx = col * (BW(0, ell) + w) + BW(0, ell + 1);
      fy = row * (BW(1, ell) + w) + BW(1, ell);
    } else {
      fx = col * (BW(0, ell) + w) + BW(0, ell);
      fy = row * (BW(1, ell) + w) + BW(1, ell + 1);
    }
    Idx r;
    for (int j = fy; j < fy + w; j++)
      for (int i = fx; i < fx + w; i++) r.push_back(i + j * N[0]);
    return r;
  }
  // ctor (mf2d.cpp:8-148)
  void plan(int Nx, int Ny, int width, int maxSide) {
    dim = 2; N[0] = Nx; N[1] = Ny; N[2] = 1; w = width;
    L = levelsFor(std::max(Nx, Ny), width, maxSide);
    if (L <= 1)
      throw std::runtime_error("Based on the input parameters, the recursion depth is 1 level (i.e., do not recurse).  Try increasing the problem size or decreasing maxSideLength.");
    boxes.resize(L);
    seps.resize(L - 1);
    int b = 1 << (L - 1);
    boxes[L - 1].resize(b * b);
    seps[L - 2].resize(nseps(L - 1));
    bw[0] = (int)std::floor((double)(Nx - (b - 1) * w) / (double)b);
    bw[1] = (int)std::floor((double)(Ny - (b - 1) * w) / (double)b);
    if (Nx < bw[0] * (b - 1) + w * (b - 1) || Ny < bw[1] * (b - 1) + w * (b - 1))
      throw std::runtime_error("You found a bug!  The math for computing the number of points per box at a leaf level broke on this input.");
    for (int box = 0; box < b * b; box++) {
      markedBoxes.insert(box);
      int x = box % b, y = box / b;
      int fx = x * (bw[0] + w), fy = y * (bw[1] + w);
      int xUB = x == b - 1 ? Nx : fx + bw[0], yUB = y == b - 1 ? Ny : fy + bw[1];
      Idx& self = boxes[L - 1][box].intDOFs;
      for (int j = fy; j < yUB; j++)
        for (int i = fx; i < xUB; i++) self.push_back(i + j * Nx);
      Idx ns = nborSeps(box, L - 1);
      if (x > 0) {
        Idx& d = seps[L - 2][ns[0]].DOFs;
        for (int j = fy; j < yUB; j++)
          for (int i = fx - w; i < fx; i++) d.push_back(i + j * Nx);
      }
      if (y < b - 1) {
        Idx& d = seps[L - 2][ns[3]].DOFs;
        for (int j = yUB; j < yUB + w; j++)
          for (int i = fx; i < xUB; i++) d.push_back(i + j * Nx);
      }
    }
    for (int ell = L - 2; ell >= 0; ell--) {
      if (ell != 0) seps[ell - 1].resize(nseps(ell));
      boxes[ell].resize(nboxes(ell));
    }
  }
};

struct Front3 : Front {
  int nboxes(int ell) const override { return Geo3::nboxes(ell); }
  int nseps(int ell) const override { return Geo3::nseps(ell); }
  Idx nborSeps(int b, int ell) const override { return Geo3::nborSeps(b, ell); }
  Idx nborBoxes(int e, int ell) const override { return Geo3::nborBoxes(e, ell); }
  Idx sepChildren(int e, int ell) const override { return Geo3::sepChildren(e, ell); }
  Idx boxChildren(int b, int ell) const override { return Geo3::boxChildren(b, ell); }
  int parent(int b, int ell) const override { return Geo3::parent(b, ell); }
  Idx childSeps(int b, int ell, int p) const override { return Geo3::childSeps(b, ell, p); }
  int leafOfDof(int d) const override {  // mf3d.cpp:223-240
    int b = 1 << (L - 1);
    int i = d % N[0]; d /= N[0];
    int j = d % N[1]; d /= N[1];
    int k = d;
    int ib = i / (bw[0] + w), jb = j / (bw[1] + w), kb = k / (bw[2] + w);
    if (ib >= b || jb >= b || kb >= b) return -1;
    return ib + jb * b + kb * b * b;
  }
  void block(std::set<int>& s, int x0, int x1, int y0, int y1, int z0, int z1) const {
    for (int k = z0; k < z1; k++)
      for (int j = y0; j < y1; j++)
        for (int i = x0; i < x1; i++) s.insert(i + j * N[0] + k * N[0] * N[1]);
  }
  Idx boxCenter(int box, int ell) const override {  // mf3d.cpp:253-348
    int b = 1 << ell, col = box % b, row = (box / b) % b, sli = box / (b * b);
    int X = BW(0, ell), Y = BW(1, ell), Z = BW(2, ell);
    int Xb = BW(0, ell + 1), Yb = BW(1, ell + 1), Zb = BW(2, ell + 1);
    std::set<int> s;
    int fx = col * (X + w), fy = row * (Y + w) + Yb, fz = sli * (Z + w) + Zb;
    block(s, fx, col == b - 1 ? N[0] : fx + X, fy, fy + w, fz, fz + w);
    fx = col * (X + w) + Xb; fy = row * (Y + w); fz = sli * (Z + w) + Zb;
    block(s, fx, fx + w, fy, row == b - 1 ? N[1] : fy + Y, fz, fz + w);
    fx = col * (X + w) + Xb; fy = row * (Y + w) + Yb; fz = sli * (Z + w);
    block(s, fx, fx + w, fy, fy + w, fz, sli == b - 1 ? N[2] : fz + Z);
    return Idx(s.begin(), s.end());
  }
  Idx sepCenter(int f, int ell) const override {  // mf3d.cpp:350-524
    int b = 1 << ell, col, row, sli;
    Geo3::coords(f, ell, col, row, sli);
    int X = BW(0, ell), Y = BW(1, ell), Z = BW(2, ell);
    int Xb = BW(0, ell + 1), Yb = BW(1, ell + 1), Zb = BW(2, ell + 1);
    std::set<int> s;
    int t = Geo3::type(f, ell), fx, fy, fz;
    if (t == Geo3::YN) {
      fx = col * (X + w) + Xb; fy = row * (Y + w) + Y; fz = sli * (Z + w);
      block(s, fx, fx + w, fy, fy + w, fz, sli == b - 1 ? N[2] : fz + Z);
      fx = col * (X + w); fy = row * (Y + w) + Y; fz = sli * (Z + w) + Zb;
      block(s, fx, col == b - 1 ? N[0] : fx + X, fy, fy + w, fz, fz + w);
    } else if (t == Geo3::XN) {
      fx = col * (X + w) + X; fy = row * (Y + w) + Yb; fz = sli * (Z + w);
      block(s, fx, fx + w, fy, fy + w, fz, sli == b - 1 ? N[2] : fz + Z);
      fx = col * (X + w) + X; fy = row * (Y + w); fz = sli * (Z + w) + Zb;
      block(s, fx, fx + w, fy, row == b - 1 ? N[1] : fy + Y, fz, fz + w);
    } else {
      fx = col * (X + w); fy = row * (Y + w) + Yb; fz = sli * (Z + w) + Z;
      block(s, fx, col == b - 1 ? N[0] : fx + X, fy, fy + w, fz, fz + w);
      fx = col * (X + w) + Xb; fy = row * (Y + w); fz = sli * (Z + w) + Z;
      block(s, fx, fx + w, fy, row == b - 1 ? N[1] : fy + Y, fz, fz + w);
    }
    return Idx(s.begin(), s.end());
  }
  // ctor (mf3d.cpp:26-206)
  void plan(int Nx, int Ny, int Nz, int width, int maxSide) {
    dim = 3; N[0] = Nx; N[1] = Ny; N[2] = Nz; w = width;
    L = levelsFor(std::max(Nx, std::max(Ny, Nz)), width, maxSide);
    if (L <= 1)
      throw std::runtime_error("Based on the input parameters, the recursion depth is 1 level (i.e., do not recurse).  Try increasing the problem size or decreasing maxSideLength.");
    boxes.resize(L);
    seps.resize(L - 1);
    int b = 1 << (L - 1);
    boxes[L - 1].resize(b * b * b);
    seps[L - 2].resize(nseps(L - 1));
    for (int a = 0; a < 3; a++) bw[a] = (int)std::floor((double)(N[a] - (b - 1) * w) / (double)b);
    for (int a = 0; a < 3; a++)
      if (N[a] < bw[a] * (b - 1) + w * (b - 1))
        throw std::runtime_error("You found a bug!  The math for computing the number of points per box at a leaf level broke on this input.");
    for (int box = 0; box < b * b * b; box++) {
      markedBoxes.insert(box);
      int x = box % b, y = (box / b) % b, z = box / (b * b);
      int fx = x * (bw[0] + w), fy = y * (bw[1] + w), fz = z * (bw[2] + w);
      int xUB = x == b - 1 ? Nx : fx + bw[0], yUB = y == b - 1 ? Ny : fy + bw[1], zUB = z == b - 1 ? Nz : fz + bw[2];
      Idx& self = boxes[L - 1][box].intDOFs;
      for (int k = fz; k < zUB; k++)
        for (int j = fy; j < yUB; j++)
          for (int i = fx; i < xUB; i++) self.push_back(i + j * Nx + k * Nx * Ny);
      Idx ns = nborSeps(box, L - 1);
      auto fill = [&](Idx& d, int x0, int x1, int y0, int y1, int z0, int z1) {
        for (int k = z0; k < z1; k++)
          for (int j = y0; j < y1; j++)
            for (int i = x0; i < x1; i++) d.push_back(i + j * Nx + k * Nx * Ny);
      };
      if (x > 0) fill(seps[L - 2][ns[0]].DOFs, fx - w, fx, fy, yUB, fz, zUB);
      if (y < b - 1) fill(seps[L - 2][ns[3]].DOFs, fx, xUB, yUB, yUB + w, fz, zUB);
      if (z > 0) fill(seps[L - 2][ns[4]].DOFs, fx, xUB, fy, yUB, fz - w, fz);
    }
    for (int ell = L - 2; ell >= 0; ell--) {
      if (ell != 0) seps[ell - 1].resize(nseps(ell));
      boxes[ell].resize(nboxes(ell));
    }
  }
};

}  // namespace orc

// ======================================================================== C ABI for ctypes
using namespace orc;

static thread_local std::string g_err;

extern "C" {

const char* orc_last_error() { return g_err.c_str(); }

void* orc_create(int dim, int nx, int ny, int nz, int width, int maxSide, int n, const int* rp,
                 const int* ci, const double* val) {
  try {
    Front* F;
    if (dim == 2) {
      Front2* f = new Front2();
      f->plan(nx, ny, width, maxSide);
      F = f;
    } else {
      Front3* f = new Front3();
      f->plan(nx, ny, nz, width, maxSide);
      F = f;
    }
    F->A.n = n;
    F->A.rp.assign(rp, rp + n + 1);
    F->A.ci.assign(ci, ci + rp[n]);
    F->A.v.assign(val, val + rp[n]);
    return F;
  } catch (std::exception& e) {
    g_err = e.what();
    return nullptr;
  }
}
void orc_destroy(void* h) { delete (Front*)h; }
int orc_levels(void* h) { return ((Front*)h)->L; }
int orc_leaf_width(void* h, int a) { return ((Front*)h)->bw[a]; }
int orc_nboxes(void* h, int ell) { return ((Front*)h)->nboxes(ell); }
int orc_nseps(void* h, int ell) { return ((Front*)h)->nseps(ell); }
int orc_box_dofs(void* h, int ell, int box, int which, int* out) {
  const Box& B = ((Front*)h)->boxes[ell][box];
  const Idx& v = which == 0 ? B.intDOFs : B.bdDOFs;
  if (out) std::copy(v.begin(), v.end(), out);
  return (int)v.size();
}
int orc_sep_dofs(void* h, int ell, int e, int* out) {
  const Idx& v = ((Front*)h)->seps[ell - 1][e].DOFs;
  if (out) std::copy(v.begin(), v.end(), out);
  return (int)v.size();
}
int orc_factor(void* h) {
  // plan() marked every leaf, as the constructor does (mf2d.cpp:74); a later call refactors
  // from the marks the previous call left behind, as the reference's Factor() does.
  Front* F = (Front*)h;
  F->factor();
  F->factored = true;
  return F->info;
}
int orc_update(void* h, const double* val, int nmod, const int* mod) {
  Front* F = (Front*)h;
  if (!F->factored) { g_err = "update before factor"; return -2; }
  std::vector<double> v(val, val + F->A.rp[F->A.n]);
  Idx m(mod, mod + nmod);
  int r = F->update(v, m);
  if (r < 0) { g_err = "modified DOF maps outside the leaf box grid"; return r; }
  return F->info;
}
// which: 0 boxes factored at level ell, 1 separators that received Schur contributions at level ell,
//        2 separators rebuilt from children at level ell
int orc_trace(void* h, int ell, int which, int* out) {
  Front* F = (Front*)h;
  const Idx& v = which == 0 ? F->traceBoxes[ell] : which == 1 ? F->traceSeps[ell] : F->traceRebuilt[ell];
  if (out) std::copy(v.begin(), v.end(), out);
  return (int)v.size();
}
// which: 0 A22, 1 XL, 2 XR, 3 S (SchurComp), 4 A11post
int orc_box_block(void* h, int ell, int box, int which, double* out, int* m, int* n) {
  const Box& B = ((Front*)h)->boxes[ell][box];
  const Mat* M = which == 0 ? &B.A22 : which == 1 ? &B.XL : which == 2 ? &B.XR : which == 3 ? &B.S : &B.A11post;
  *m = M->m; *n = M->n;
  if (out) std::copy(M->a.begin(), M->a.end(), out);
  return 0;
}
int orc_sep_block(void* h, int ell, int e, double* out, int* k) {
  const Sep& E = ((Front*)h)->seps[ell - 1][e];
  *k = E.A11.m;
  if (out) std::copy(E.A11.a.begin(), E.A11.a.end(), out);
  return 0;
}
void orc_solve(void* h, const double* b, double* x) {
  Front* F = (Front*)h;
  std::vector<double> v(b, b + F->A.n);
  v = F->solve(v);
  std::copy(v.begin(), v.end(), x);
}
void orc_apply(void* h, const double* x, double* y) {
  Front* F = (Front*)h;
  std::vector<double> v(x, x + F->A.n);
  v = F->apply(v);
  std::copy(v.begin(), v.end(), y);
}

// buildLaplacian 2D (test_include/DE/buildLaplacian.hpp:10-132).  coeff is the padded nodal
// field, (nx+2w) x (ny+2w), stored as coeff[ii*(ny+2w)+jj] for A[ii][jj].  Pass null outputs
// to get nnz.  Returns nnz.
int orc_laplacian_2d(int nx, int ny, const double* coeff, int w, double hx, double hy, int* rp, int* ci,
                     double* val) {
  static const double C[3][4] = {{0, -1, 0, 0}, {0, -4.0 / 3.0, 1.0 / 12.0, 0}, {0, -1.5, 3.0 / 20.0, -1.0 / 90.0}};
  const double* c = C[w - 1];
  int P = ny + 2 * w;
  auto A = [&](int ii, int jj) { return coeff[(size_t)ii * P + jj]; };
  double hhx = hx * hx, hhy = hy * hy;
  int off = 0;
  auto push = [&](int col, double v) {
    if (ci) { ci[off] = col; val[off] = v; }
    off++;
  };
  for (int j = 0; j < ny; j++)
    for (int i = 0; i < nx; i++) {
      if (rp) rp[i + j * nx] = off;
      int ii = i + w, jj = j + w, idx = i + j * nx;
      for (int k = w; k > 0; k--)
        if (j - k >= 0) push(i + (j - k) * nx, c[k] * (0.5 * (A(ii, jj - k) + A(ii, jj - k + 1))) / hhy);
      for (int k = w; k > 0; k--)
        if (i - k >= 0) push((i - k) + j * nx, c[k] * (0.5 * (A(ii - k, jj) + A(ii - k + 1, jj))) / hhx);
      double v = 0;
      for (int k = w; k > 0; k--) {
        double a1 = 0.5 * (A(ii - k, jj) + A(ii - k + 1, jj));
        double a2 = 0.5 * (A(ii + k, jj) + A(ii + k - 1, jj));
        double a3 = 0.5 * (A(ii, jj - k) + A(ii, jj - k + 1));
        double a4 = 0.5 * (A(ii, jj + k) + A(ii, jj + k - 1));
        v -= c[k] * (a1 + a2) / hhx;
        v -= c[k] * (a3 + a4) / hhy;
      }
      push(idx, v);
      for (int k = 1; k <= w; k++)
        if (i + k <= nx - 1) push((i + k) + j * nx, c[k] * (0.5 * (A(ii + k, jj) + A(ii + k - 1, jj))) / hhx);
      for (int k = 1; k <= w; k++)
        if (j + k <= ny - 1) push(i + (j + k) * nx, c[k] * (0.5 * (A(ii, jj + k) + A(ii, jj + k - 1))) / hhy);
    }
  if (rp) rp[nx * ny] = off;
  return off;
}

// buildLaplacian 3D (test_include/DE/buildLaplacian.hpp:244-391); coeff[(ii*(ny+2w)+jj)*(nz+2w)+zz]
int orc_laplacian_3d(int nx, int ny, int nz, const double* coeff, int w, double hx, double hy, double hz, int* rp,
                     int* ci, double* val) {
  static const double C[3][4] = {{0, -1, 0, 0}, {0, -4.0 / 3.0, 1.0 / 12.0, 0}, {0, -1.5, 3.0 / 20.0, -1.0 / 90.0}};
  const double* c = C[w - 1];
  int PY = ny + 2 * w, PZ = nz + 2 * w;
  auto A = [&](int ii, int jj, int zz) { return coeff[((size_t)ii * PY + jj) * PZ + zz]; };
  double hhx = hx * hx, hhy = hy * hy, hhz = hz * hz;
  int off = 0;
  auto push = [&](int col, double v) {
    if (ci) { ci[off] = col; val[off] = v; }
    off++;
  };
  for (int z = 0; z < nz; z++)
    for (int j = 0; j < ny; j++)
      for (int i = 0; i < nx; i++) {
        int idx = i + j * nx + z * nx * ny;
        if (rp) rp[idx] = off;
        int ii = i + w, jj = j + w, zz = z + w;
        for (int k = w; k > 0; k--)
          if (z - k >= 0) push(i + j * nx + (z - k) * nx * ny, c[k] * (0.5 * (A(ii, jj, zz - k) + A(ii, jj, zz - k + 1))) / hhz);
        for (int k = w; k > 0; k--)
          if (j - k >= 0) push(i + (j - k) * nx + z * nx * ny, c[k] * (0.5 * (A(ii, jj - k, zz) + A(ii, jj - k + 1, zz))) / hhy);
        for (int k = w; k > 0; k--)
          if (i - k >= 0) push((i - k) + j * nx + z * nx * ny, c[k] * (0.5 * (A(ii - k, jj, zz) + A(ii - k + 1, jj, zz))) / hhx);
        double v = 0;
        for (int k = w; k > 0; k--) {
          double a1 = 0.5 * (A(ii - k, jj, zz) + A(ii - k + 1, jj, zz));
          double a2 = 0.5 * (A(ii + k, jj, zz) + A(ii + k - 1, jj, zz));
          double a3 = 0.5 * (A(ii, jj - k, zz) + A(ii, jj - k + 1, zz));
          double a4 = 0.5 * (A(ii, jj + k, zz) + A(ii, jj + k - 1, zz));
          double a5 = 0.5 * (A(ii, jj, zz - k) + A(ii, jj, zz - k + 1));
          double a6 = 0.5 * (A(ii, jj, zz + k) + A(ii, jj, zz + k - 1));
          v -= c[k] * (a1 + a2) / hhx;
          v -= c[k] * (a3 + a4) / hhy;
          v -= c[k] * (a5 + a6) / hhz;
        }
        push(idx, v);
        for (int k = 1; k <= w; k++)
          if (i + k <= nx - 1) push((i + k) + j * nx + z * nx * ny, c[k] * (0.5 * (A(ii + k, jj, zz) + A(ii + k - 1, jj, zz))) / hhx);
        for (int k = 1; k <= w; k++)
          if (j + k <= ny - 1) push(i + (j + k) * nx + z * nx * ny, c[k] * (0.5 * (A(ii, jj + k, zz) + A(ii, jj + k - 1, zz))) / hhy);
        for (int k = 1; k <= w; k++)
          if (z + k <= nz - 1) push(i + j * nx + (z + k) * nx * ny, c[k] * (0.5 * (A(ii, jj, zz + k) + A(ii, jj, zz + k - 1))) / hhz);
      }
  if (rp) rp[nx * ny * nz] = off;
  return off;
}

// modDOF rule: columns of CSR entries whose values differ (test_src/DE/testMF2D.cpp:143-150).
// Writes the sorted unique set; returns its size.
int orc_moddofs(int nnz, const int* ci, const double* v1, const double* v2, int* out) {
  std::set<int> s;
  for (int k = 0; k < nnz; k++)
    if (v1[k] != v2[k]) s.insert(ci[k]);
  if (out) std::copy(s.begin(), s.end(), out);
  return (int)s.size();
}

// CSR SpMV (src/LinAlg/sparse.cpp:63-77)
void orc_spmv(int n, const int* rp, const int* ci, const double* v, const double* x, double* y) {
  for (int i = 0; i < n; i++) {
    double s = 0;
    for (int k = rp[i]; k < rp[i + 1]; k++) s += v[k] * x[ci[k]];
    y[i] = s;
  }
}

}  // extern "C"
