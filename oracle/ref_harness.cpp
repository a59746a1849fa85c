// ref_harness.cpp -- C ABI around the reference implementation, for checking the oracle.
//
// TEST INFRASTRUCTURE ONLY.  Built by oracle/Makefile from the reference sources where they lie
// (-I /root/reference/include/..., the reference's headers #include its .cpp files), linked
// against the host MKL that the reference's own build uses (CMakeLists.txt:34-39), with output
// only into oracle/_ref/.  No reference source is copied into this repository.
//
// Private members are opened with `#define private public` after every standard header is
// already included, so the harness can read the elimination tree, the per-front blocks and the
// dirty sets.  The traced factorization calls the reference's own private level routines in the
// order of MF2D::Factor (src/DE/mf2d.cpp:449-500) / MF3D::Factor (src/DE/mf3d.cpp:705-766),
// snapshotting markedBoxes_ / markedEdges_ between them.  The traced update repeats the leaf
// marking of MF2D::Update (src/DE/mf2d.cpp:161-174) / MF3D::Update (src/DE/mf3d.cpp:219-241)
// before the traced factorization; the untraced path calls the reference's Update itself.
#include <algorithm>
#include <cassert>
#include <chrono>
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <iterator>
#include <limits.h>
#include <set>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#define private public
#include "linalg.hpp"
#include "mf.hpp"
#include "buildLaplacian.hpp"
#undef private

using namespace LinAlg;

namespace {
struct Handle {
  int dim;
  TreeFactor2D::MF2D<double>* f2 = nullptr;
  TreeFactor3D::MF3D<double>* f3 = nullptr;
  int n = 0;
  std::vector<std::vector<int>> traceBoxes, traceSeps, traceRebuilt;
  ~Handle() { delete f2; delete f3; }
};
std::string g_err;

Sparse<double> mkcsr(int n, const int* rp, const int* ci, const double* v) {
  int nnz = rp[n];
  std::vector<double> val(v, v + nnz);
  IdxVec col(ci, ci + nnz), row(rp, rp + n + 1);
  return Sparse<double>(val, col, row, n, n);
}

template <class F, class NborSeps>
void traced_factor(Handle* h, F& f, NborSeps nborSeps, std::set<int>& markedSeps) {
  int L = f.levels_;
  h->traceBoxes.assign(L, {});
  h->traceSeps.assign(L, {});
  h->traceRebuilt.assign(L, {});
  auto sepsOf = [&](const std::set<int>& boxes, int ell) {
    std::set<int> s;
    for (int b : boxes)
      for (int e : nborSeps(b, ell))
        if (e != -1) s.insert(e);
    return std::vector<int>(s.begin(), s.end());
  };
  h->traceBoxes[L - 1] = std::vector<int>(f.markedBoxes_.begin(), f.markedBoxes_.end());
  h->traceSeps[L - 1] = sepsOf(f.markedBoxes_, L - 1);
  f.FactorLeafLevel_();
  for (int ell = L - 2; ell >= 0; ell--) {
    if (ell != 0) {
      std::vector<int> r;
      for (int e : markedSeps)
        if (e != -1) r.push_back(e);
      h->traceRebuilt[ell] = r;
      if (h->dim == 2) h->f2->updateEdgeData_(ell); else h->f3->updateFaceData_(ell);
    }
    h->traceBoxes[ell] = std::vector<int>(f.markedBoxes_.begin(), f.markedBoxes_.end());
    if (ell != 0) h->traceSeps[ell] = sepsOf(f.markedBoxes_, ell);
    f.updateBoxData_(ell);
    f.FactorUpperLevel_(ell);
  }
}

void traced(Handle* h) {
  if (h->dim == 2)
    traced_factor(h, *h->f2, [](int b, int ell) { return TreeFactor2D::getNborEdgesOfBox_(b, ell); },
                  h->f2->markedEdges_);
  else
    traced_factor(h, *h->f3, [](int b, int ell) { return TreeFactor3D::getNborFacesOfBox_(b, ell); },
                  h->f3->markedFaces_);
}

template <class Box>
const Dense<double>& pick(const Box& B, int which) {
  switch (which) {
    case 0: return B.A22;
    case 1: return B.XL;
    case 2: return B.XR;
    case 3: return B.SchurComp;
    case 5: return B.A22Inv;
    default: return B.A11post;
  }
}
void copyDense(const Dense<double>& M, double* out, int* m, int* n) {
  *m = M.m_; *n = M.n_;
  if (out && M.m_ * M.n_ > 0) std::memcpy(out, M.data_.data(), sizeof(double) * (size_t)M.m_ * M.n_);
}
}  // namespace

extern "C" {

const char* ref_last_error() { return g_err.c_str(); }

void* ref_create(int dim, int nx, int ny, int nz, int width, int maxSide, int n, const int* rp, const int* ci,
                 const double* val) {
  try {
    Handle* h = new Handle();
    h->dim = dim;
    h->n = n;
    Sparse<double> A = mkcsr(n, rp, ci, val);
    if (dim == 2) h->f2 = new TreeFactor2D::MF2D<double>(A, nx, ny, width, maxSide);
    else h->f3 = new TreeFactor3D::MF3D<double>(A, nx, ny, nz, width, maxSide);
    return h;
  } catch (std::exception& e) {
    g_err = e.what();
    return nullptr;
  }
}
void ref_destroy(void* p) { delete (Handle*)p; }
int ref_levels(void* p) { Handle* h = (Handle*)p; return h->dim == 2 ? h->f2->levels_ : h->f3->levels_; }
int ref_leaf_width(void* p, int a) {
  Handle* h = (Handle*)p;
  if (h->dim == 2) return a == 0 ? h->f2->boxWidthXBottom_ : h->f2->boxWidthYBottom_;
  return a == 0 ? h->f3->boxWidthXBottom_ : a == 1 ? h->f3->boxWidthYBottom_ : h->f3->boxWidthZBottom_;
}
int ref_nboxes(void* p, int ell) {
  Handle* h = (Handle*)p;
  return h->dim == 2 ? (int)h->f2->boxData_[ell].size() : (int)h->f3->boxData_[ell].size();
}
int ref_nseps(void* p, int ell) {
  Handle* h = (Handle*)p;
  return h->dim == 2 ? (int)h->f2->edgeData_[ell - 1].size() : (int)h->f3->faceData_[ell - 1].size();
}
int ref_box_dofs(void* p, int ell, int box, int which, int* out) {
  Handle* h = (Handle*)p;
  const IdxVec& v = h->dim == 2 ? (which == 0 ? h->f2->boxData_[ell][box].intDOFs : h->f2->boxData_[ell][box].bdDOFs)
                                : (which == 0 ? h->f3->boxData_[ell][box].intDOFs : h->f3->boxData_[ell][box].bdDOFs);
  if (out) std::copy(v.begin(), v.end(), out);
  return (int)v.size();
}
int ref_sep_dofs(void* p, int ell, int e, int* out) {
  Handle* h = (Handle*)p;
  const IdxVec& v = h->dim == 2 ? h->f2->edgeData_[ell - 1][e].DOFs : h->f3->faceData_[ell - 1][e].DOFs;
  if (out) std::copy(v.begin(), v.end(), out);
  return (int)v.size();
}
// Reference Factor(), optionally traced.
int ref_factor(void* p, int trace) {
  Handle* h = (Handle*)p;
  if (trace) traced(h);
  else if (h->dim == 2) h->f2->Factor();
  else h->f3->Factor();
  return 0;
}
// Reference Update(A2, modDOFs), optionally traced.
int ref_update(void* p, const int* rp, const int* ci, const double* val, int nmod, const int* mod, int trace) {
  Handle* h = (Handle*)p;
  Sparse<double> A = mkcsr(h->n, rp, ci, val);
  std::set<int> m(mod, mod + nmod);
  if (!trace) {
    if (h->dim == 2) h->f2->Update(A, m); else h->f3->Update(A, m);
    return 0;
  }
  if (h->dim == 2) {
    auto& f = *h->f2;
    f.markedBoxes_.clear(); f.markedEdges_.clear(); f.A_ = A;
    int b = 1 << (f.levels_ - 1);
    for (int idx : m) {
      int i = idx % f.Nx_, j = (idx - i) / f.Nx_;
      int ib = i / (f.boxWidthXBottom_ + f.width_), jb = j / (f.boxWidthYBottom_ + f.width_);
      f.markedBoxes_.insert(TreeFactor2D::TensorIdx2D(ib, jb, b, b).LinearIndex());
    }
  } else {
    auto& f = *h->f3;
    f.markedBoxes_.clear(); f.markedFaces_.clear(); f.A_ = A;
    int b = 1 << (f.levels_ - 1);
    for (int idx : m) {
      int i = idx % f.Nx_; idx /= f.Nx_;
      int j = idx % f.Ny_; idx /= f.Ny_;
      int k = idx;
      int ib = i / (f.boxWidthXBottom_ + f.width_), jb = j / (f.boxWidthYBottom_ + f.width_),
          kb = k / (f.boxWidthZBottom_ + f.width_);
      f.markedBoxes_.insert(TreeFactor3D::TensorIdx3D(ib, jb, kb, b, b, b).LinearIndex());
    }
  }
  traced(h);
  return 0;
}
int ref_trace(void* p, int ell, int which, int* out) {
  Handle* h = (Handle*)p;
  const std::vector<int>& v = which == 0 ? h->traceBoxes[ell] : which == 1 ? h->traceSeps[ell] : h->traceRebuilt[ell];
  if (out) std::copy(v.begin(), v.end(), out);
  return (int)v.size();
}
int ref_box_block(void* p, int ell, int box, int which, double* out, int* m, int* n) {
  Handle* h = (Handle*)p;
  if (h->dim == 2) copyDense(pick(h->f2->boxData_[ell][box], which), out, m, n);
  else copyDense(pick(h->f3->boxData_[ell][box], which), out, m, n);
  return 0;
}
int ref_sep_block(void* p, int ell, int e, double* out, int* k) {
  Handle* h = (Handle*)p;
  int n;
  if (h->dim == 2) copyDense(h->f2->edgeData_[ell - 1][e].A11, out, k, &n);
  else copyDense(h->f3->faceData_[ell - 1][e].A11, out, k, &n);
  return 0;
}
void ref_solve(void* p, const double* b, double* x) {
  Handle* h = (Handle*)p;
  Vec<double> v(h->n);
  for (int i = 0; i < h->n; i++) v(i) = b[i];
  Vec<double> r = h->dim == 2 ? h->f2->Solve(v) : h->f3->Solve(v);
  for (int i = 0; i < h->n; i++) x[i] = r(i);
}
void ref_apply(void* p, const double* xin, double* y) {
  Handle* h = (Handle*)p;
  Vec<double> v(h->n);
  for (int i = 0; i < h->n; i++) v(i) = xin[i];
  Vec<double> r = h->dim == 2 ? h->f2->Apply(v) : h->f3->Apply(v);
  for (int i = 0; i < h->n; i++) y[i] = r(i);
}
// The reference harness operator, buildLaplacian (test_include/DE/buildLaplacian.hpp:10, :244).
// coeff layout as in the oracle: A[ii][jj] at coeff[ii*(ny+2w)+jj] (3D: [(ii*(ny+2w)+jj)*(nz+2w)+zz]).
int ref_laplacian(int dim, int nx, int ny, int nz, const double* coeff, int w, double h, int* rp, int* ci,
                  double* val) {
  Sparse<double> M;
  if (dim == 2) {
    std::vector<std::vector<double>> A(nx + 2 * w, std::vector<double>(ny + 2 * w));
    for (int i = 0; i < nx + 2 * w; i++)
      for (int j = 0; j < ny + 2 * w; j++) A[i][j] = coeff[(size_t)i * (ny + 2 * w) + j];
    M = buildLaplacian(nx, ny, A, w, h, h);
  } else {
    std::vector<std::vector<std::vector<double>>> A(
        nx + 2 * w, std::vector<std::vector<double>>(ny + 2 * w, std::vector<double>(nz + 2 * w)));
    for (int i = 0; i < nx + 2 * w; i++)
      for (int j = 0; j < ny + 2 * w; j++)
        for (int k = 0; k < nz + 2 * w; k++) A[i][j][k] = coeff[((size_t)i * (ny + 2 * w) + j) * (nz + 2 * w) + k];
    M = buildLaplacian(nx, ny, nz, A, w, h, h, h);
  }
  int n = M.m(), nnz = M.nnz();
  if (rp) {
    for (int i = 0; i <= n; i++) rp[i] = M.row_ptr(i);
    for (int k = 0; k < nnz; k++) { ci[k] = M.col_ind(k); val[k] = M.val(k); }
  }
  return nnz;
}

// ---- the reference's LinAlg free functions (include/LinAlg/dense.hpp:167-227, :281-310) on
// column-major arrays, for checking the drop-in header's versions.
static Dense<double> to_dense(int m, int n, const double* a) {
  Dense<double> A(m, n);
  for (int j = 0; j < n; j++)
    for (int i = 0; i < m; i++) A(i, j) = a[(size_t)j * m + i];
  return A;
}
static void from_dense(const Dense<double>& A, double* a) {
  for (int j = 0; j < A.n_; j++)
    for (int i = 0; i < A.m_; i++) a[(size_t)j * A.m_ + i] = A.data_[(size_t)j * A.m_ + i];
}
// LU(A) in place (n x n), ipiv (1-based) out.
void ref_dense_lu(int n, double* a, int* ipiv) {
  Dense<double> A = to_dense(n, n, a);
  LU(A);
  from_dense(A, a);
  for (int i = 0; i < n; i++) ipiv[i] = A.ipiv_[i];
}
// After ref_dense_lu: mode 0 LUSolve (b: n x nrhs), 1 LUSolveT (b: nrhs x n, solves X A = B),
// 2 LUSolveVec (b: n).
void ref_dense_lusolve(int n, const double* lu, const int* ipiv, int nrhs, double* b, int mode) {
  Dense<double> A = to_dense(n, n, lu);
  A.ipiv_.assign(ipiv, ipiv + n);
  if (mode == 2) {
    Vec<double> v(n);
    for (int i = 0; i < n; i++) v(i) = b[i];
    LUSolveVec(A, v);
    for (int i = 0; i < n; i++) b[i] = v(i);
    return;
  }
  Dense<double> B = mode == 0 ? to_dense(n, nrhs, b) : to_dense(nrhs, n, b);
  if (mode == 0) LUSolve(A, B);
  else LUSolveT(A, B);
  from_dense(B, b);
}
// C := alpha op(A) op(B) + beta C (GEMM); c is cm x cn, updated in place.
void ref_dense_gemm(double alpha, int am, int an, const double* a, char ta, int bm, int bn, const double* b, char tb,
                    double beta, int cm, int cn, double* c) {
  Dense<double> A = to_dense(am, an, a), B = to_dense(bm, bn, b), C = to_dense(cm, cn, c);
  GEMM(alpha, A, ta, B, tb, beta, C);
  from_dense(C, c);
}
// transpose(A, AT): at is n x m.
void ref_dense_transpose(int m, int n, const double* a, double* at) {
  Dense<double> A = to_dense(m, n, a), AT;
  transpose(A, AT);
  from_dense(AT, at);
}
// The reference's CSR text reader, Sparse(filename) (src/LinAlg/sparse.cpp:5-44): returns nnz
// and writes m, n; with rp non-null also the arrays (the caller sizes them from a first call).
// Returns -1 when the reader throws (its message in ref_last_error).
int ref_read_csr(const char* path, int* m, int* n, int* rp, int* ci, double* val) {
  try {
    Sparse<double> A{std::string(path)};
    *m = A.m();
    *n = A.n();
    const int nnz = A.nnz();
    if (rp) {
      for (int i = 0; i <= *m; i++) rp[i] = A.row_ptr(i);
      for (int k = 0; k < nnz; k++) { ci[k] = A.col_ind(k); val[k] = A.val(k); }
    }
    return nnz;
  } catch (const std::exception& e) {
    g_err = e.what();
    return -1;
  }
}

}  // extern "C"
