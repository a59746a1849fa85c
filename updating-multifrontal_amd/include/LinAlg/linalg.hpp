// linalg.hpp -- drop-in LinAlg types for code written against the reference's
// include/LinAlg/linalg.hpp: Dense (column-major), Vec and Sparse (CSR), header-only, host-side.
//
// Names, constructors, operators and storage order follow include/LinAlg/dense.hpp:46-133,
// include/LinAlg/vec.hpp:34-90 and include/LinAlg/sparse.hpp:33-83.  These are host containers
// for data going in and out of the factorization; the factorization itself runs on the GPU
// (include/DE/mf.hpp).  Scalar = double.
#pragma once
#ifndef UMF_LINALG_HPP
#define UMF_LINALG_HPP 1

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <fstream>
#include <iostream>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

typedef std::vector<int> IdxVec;

namespace LinAlg {

template <class Scalar>
class Vec {
 public:
  Vec() : n_(0) {}
  explicit Vec(int n) : n_(n), data_(n < 0 ? 0 : n, Scalar(0)) {
    if (n < 0) throw std::domain_error("negative vector length");
  }
  Vec(const Vec& v, const IdxVec& idxs) : n_((int)idxs.size()), data_(idxs.size()) {  // vec.cpp:41-52
    for (int i = 0; i < n_; i++) data_[i] = v.data_[idxs[i]];
  }
  int n() const { return n_; }
  void resize(int n) { n_ = n; data_.resize(n); }
  Scalar& operator()(int i) { return data_[i]; }
  const Scalar& operator()(int i) const { return data_[i]; }
  Vec operator()(const IdxVec& idxs) { return Vec(*this, idxs); }
  void set(const IdxVec& idxs, const Vec& vals) {  // vec.cpp:4-12
    for (size_t i = 0; i < idxs.size(); i++) data_[idxs[i]] = vals.data_[i];
  }
  Vec& operator+=(const Vec& a) { for (int i = 0; i < n_; i++) data_[i] += a.data_[i]; return *this; }
  Vec& operator-=(const Vec& a) { for (int i = 0; i < n_; i++) data_[i] -= a.data_[i]; return *this; }
  Vec& operator*=(const Scalar& s) { for (Scalar& x : data_) x *= s; return *this; }
  Vec& operator/=(const Scalar& s) { for (Scalar& x : data_) x /= s; return *this; }
  double norm() const {  // 2-norm (vec.cpp:134-138)
    double s = 0;
    for (const Scalar& x : data_) s += x * x;
    return std::sqrt(s);
  }
  Scalar* data() { return data_.data(); }
  const Scalar* data() const { return data_.data(); }

 private:
  int n_;
  std::vector<Scalar> data_;
};

template <class Scalar>
class Dense {
 public:
  Dense(int m = 0, int n = 0) : m_(m), n_(n), data_((size_t)(m < 0 ? 0 : m) * (n < 0 ? 0 : n), Scalar(0)) {
    if (m < 0 || n < 0) throw std::domain_error("negative matrix dimension");
  }
  Dense(const Dense& A, const IdxVec& rows, const IdxVec& cols)  // dense.cpp:68-86
      : m_((int)rows.size()), n_((int)cols.size()), data_(rows.size() * cols.size()) {
    for (int j = 0; j < n_; j++)
      for (int i = 0; i < m_; i++) (*this)(i, j) = A(rows[i], cols[j]);
  }
  void resize(int m, int n) { m_ = m; n_ = n; data_.assign((size_t)m * n, Scalar(0)); }
  int m() const { return m_; }
  int n() const { return n_; }
  Scalar& operator()(int i, int j) { return data_[(size_t)j * m_ + i]; }
  const Scalar& operator()(int i, int j) const { return data_[(size_t)j * m_ + i]; }
  Dense operator()(const IdxVec& rows, const IdxVec& cols) { return Dense(*this, rows, cols); }
  Dense& operator+=(const Dense& A) { for (size_t k = 0; k < data_.size(); k++) data_[k] += A.data_[k]; return *this; }
  Dense& operator-=(const Dense& A) { for (size_t k = 0; k < data_.size(); k++) data_[k] -= A.data_[k]; return *this; }
  Dense& operator*=(const Scalar& s) { for (Scalar& x : data_) x *= s; return *this; }
  Dense& operator/=(const Scalar& s) { for (Scalar& x : data_) x /= s; return *this; }
  Dense operator*(const Dense& B) const {  // C = A B
    if (n_ != B.m_) throw std::domain_error("dimension mismatch");
    Dense C(m_, B.n_);
    for (int j = 0; j < B.n_; j++)
      for (int k = 0; k < n_; k++) {
        Scalar b = B(k, j);
        for (int i = 0; i < m_; i++) C(i, j) += (*this)(i, k) * b;
      }
    return C;
  }
  Vec<Scalar> operator*(const Vec<Scalar>& v) const {
    Vec<Scalar> y(m_);
    for (int k = 0; k < n_; k++)
      for (int i = 0; i < m_; i++) y(i) += (*this)(i, k) * v(k);
    return y;
  }
  Dense T() const {
    Dense R(n_, m_);
    for (int j = 0; j < n_; j++)
      for (int i = 0; i < m_; i++) R(j, i) = (*this)(i, j);
    return R;
  }
  double norm() const {  // Frobenius
    double s = 0;
    for (const Scalar& x : data_) s += x * x;
    return std::sqrt(s);
  }
  Scalar* data() { return data_.data(); }
  const Scalar* data() const { return data_.data(); }
  // row interchanges of the last LU() of this matrix, 1-based (LAPACK's ipiv)
  const IdxVec& ipiv() const { return ipiv_; }
  IdxVec& ipiv() { return ipiv_; }

 private:
  int m_, n_;
  std::vector<Scalar> data_;  // column-major, ld = m_
  IdxVec ipiv_;
};

template <class Scalar>
class Sparse {
 public:
  Sparse() : m_(0), n_(0), nnz_(0) {}
  Sparse(const std::vector<Scalar>& val, const IdxVec& col_ind, const IdxVec& row_ptr, int m, int n)
      : m_(m), n_(n), nnz_(row_ptr.empty() ? 0 : row_ptr[m]), val_(val), col_ind_(col_ind), row_ptr_(row_ptr) {}
  // CSR text file (sparse.cpp:5-44): whitespace-separated "m n nnz", then nnz values, nnz column
  // indices and m+1 row pointers; nnz() is taken from the last row pointer.
  explicit Sparse(const std::string& filename) : m_(0), n_(0), nnz_(0) {
    std::ifstream file(filename);
    if (!file.is_open()) throw std::runtime_error("failed to open file");
    int nnz = 0;
    file >> m_ >> n_ >> nnz;
    if (!file || m_ < 0 || n_ < 0 || nnz < 0) throw std::runtime_error("malformed CSR file header");
    val_.resize(nnz);
    col_ind_.resize(nnz);
    row_ptr_.resize((size_t)m_ + 1);
    for (int i = 0; i < nnz; i++) {
      double a;
      file >> a;
      val_[i] = (Scalar)a;
    }
    for (int i = 0; i < nnz; i++) file >> col_ind_[i];
    for (int i = 0; i <= m_; i++) file >> row_ptr_[i];
    if (!file) throw std::runtime_error("truncated CSR file");
    nnz_ = row_ptr_[m_];
  }
  // The same format written back (full double precision), for on-disk operator exchange.
  void write(const std::string& filename) const {
    std::ofstream file(filename);
    if (!file.is_open()) throw std::runtime_error("failed to open file");
    file.precision(17);
    file << m_ << " " << n_ << " " << nnz_ << "\n";
    for (int i = 0; i < nnz_; i++) file << val_[i] << (i + 1 < nnz_ ? ' ' : '\n');
    for (int i = 0; i < nnz_; i++) file << col_ind_[i] << (i + 1 < nnz_ ? ' ' : '\n');
    for (int i = 0; i <= m_; i++) file << row_ptr_[i] << (i < m_ ? ' ' : '\n');
    if (!file) throw std::runtime_error("failed to write file");
  }
  int m() const { return m_; }
  int n() const { return n_; }
  int nnz() const { return nnz_; }
  Scalar val(int i) const { return val_[i]; }
  int col_ind(int i) const { return col_ind_[i]; }
  int row_ptr(int i) const { return row_ptr_[i]; }
  const Scalar operator()(int i, int j) const {  // first match, else 0 (sparse.cpp:80-92)
    for (int k = row_ptr_[i]; k < row_ptr_[i + 1]; k++)
      if (col_ind_[k] == j) return val_[k];
    return Scalar(0);
  }
  Dense<Scalar> operator()(const IdxVec& rows, const IdxVec& cols) const {
    Dense<Scalar> B((int)rows.size(), (int)cols.size());
    for (size_t ii = 0; ii < rows.size(); ii++)
      for (size_t jj = 0; jj < cols.size(); jj++) B((int)ii, (int)jj) = (*this)(rows[ii], cols[jj]);
    return B;
  }
  Vec<Scalar> operator*(const Vec<Scalar>& v) const {  // CSR SpMV (sparse.cpp:63-77)
    Vec<Scalar> y(m_);
    for (int i = 0; i < m_; i++) {
      Scalar s = 0;
      for (int k = row_ptr_[i]; k < row_ptr_[i + 1]; k++) s += val_[k] * v(col_ind_[k]);
      y(i) = s;
    }
    return y;
  }
  double norm() const {
    double s = 0;
    for (const Scalar& x : val_) s += x * x;
    return std::sqrt(s);
  }
  const int* row_ptr_data() const { return row_ptr_.data(); }
  const int* col_ind_data() const { return col_ind_.data(); }
  const Scalar* val_data() const { return val_.data(); }
 private:
  int m_, n_, nnz_;
  std::vector<Scalar> val_;
  IdxVec col_ind_, row_ptr_;
};

// ---- free functions of the reference's LinAlg (include/LinAlg/dense.hpp:167-227, :281-310),
// host-side, Scalar = double.  Inline (the reference defines them non-inline in a header, so two
// translation units that include it collide).  A numerical failure the reference meets with
// assert(info == 0) -- an exactly singular pivot -- throws std::runtime_error here.

// AT = A^T, AT resized to n x m (dense.hpp:167-176).
inline void transpose(Dense<double>& A, Dense<double>& AT) {
  AT.resize(A.n(), A.m());
  for (int j = 0; j < A.n(); j++)
    for (int i = 0; i < A.m(); i++) AT(j, i) = A(i, j);
}

// In-place LU with partial pivoting, P A = L U (L unit lower, U upper, both stored in A; row
// interchanges in A.ipiv(), 1-based) -- the reference's LU, LAPACK dgetrf (dense.hpp:180-188).
// The pivot of column k is the first row of largest |a_ik| at or below the diagonal.
inline void LU(Dense<double>& A) {
  const int m = A.m(), n = A.n();
  if (m != n) throw std::domain_error("LU: matrix is not square");
  IdxVec& ip = A.ipiv();
  ip.assign(n, 0);
  int singular = 0;
  for (int k = 0; k < n; k++) {
    int p = k;
    double best = std::fabs(A(k, k));
    for (int i = k + 1; i < n; i++)
      if (std::fabs(A(i, k)) > best) {
        best = std::fabs(A(i, k));
        p = i;
      }
    ip[k] = p + 1;
    if (p != k)
      for (int j = 0; j < n; j++) std::swap(A(k, j), A(p, j));
    const double d = A(k, k);
    if (d == 0.0) {
      if (!singular) singular = k + 1;
      continue;  // LAPACK goes on with the column left as it is
    }
    for (int i = k + 1; i < n; i++) A(i, k) /= d;
    for (int j = k + 1; j < n; j++) {
      const double u = A(k, j);
      if (u != 0.0)
        for (int i = k + 1; i < n; i++) A(i, j) -= A(i, k) * u;
    }
  }
  if (singular) throw std::runtime_error("LU: U(" + std::to_string(singular) + "," + std::to_string(singular) +
                                         ") is exactly zero (singular matrix)");
}

namespace detail {
// B (n x nrhs, column-major, ld) <- A^{-1} B with A = P^T L U from LU().
inline void lu_solve_cols(const Dense<double>& A, double* b, int ld, int nrhs) {
  const int n = A.m();
  const IdxVec& ip = A.ipiv();
  if ((int)ip.size() != n) throw std::runtime_error("LUSolve: the matrix was not factored by LU()");
  for (int c = 0; c < nrhs; c++) {
    double* x = b + (size_t)c * ld;
    for (int k = 0; k < n; k++)
      if (ip[k] - 1 != k) std::swap(x[k], x[ip[k] - 1]);
    for (int j = 0; j < n; j++) {
      const double v = x[j];
      if (v != 0.0)
        for (int i = j + 1; i < n; i++) x[i] -= A(i, j) * v;
    }
    for (int j = n - 1; j >= 0; j--) {
      x[j] /= A(j, j);
      const double v = x[j];
      if (v != 0.0)
        for (int i = 0; i < j; i++) x[i] -= A(i, j) * v;
    }
  }
}
// B (n x nrhs) <- A^{-T} B.
inline void lu_solve_cols_t(const Dense<double>& A, double* b, int ld, int nrhs) {
  const int n = A.m();
  const IdxVec& ip = A.ipiv();
  if ((int)ip.size() != n) throw std::runtime_error("LUSolveT: the matrix was not factored by LU()");
  for (int c = 0; c < nrhs; c++) {
    double* x = b + (size_t)c * ld;
    for (int j = 0; j < n; j++) {  // U^T y = b
      double s = x[j];
      for (int i = 0; i < j; i++) s -= A(i, j) * x[i];
      x[j] = s / A(j, j);
    }
    for (int j = n - 1; j >= 0; j--) {  // L^T z = y
      double s = x[j];
      for (int i = j + 1; i < n; i++) s -= A(i, j) * x[i];
      x[j] = s;
    }
    for (int k = n - 1; k >= 0; k--)
      if (ip[k] - 1 != k) std::swap(x[k], x[ip[k] - 1]);
  }
}
}  // namespace detail

// B <- A^{-1} B after LU(A) (dgetrs 'N', dense.hpp:190-200).
inline void LUSolve(Dense<double>& A, Dense<double>& B) {
  if (A.m() != A.n() || B.m() != A.m()) throw std::domain_error("LUSolve: dimension mismatch");
  detail::lu_solve_cols(A, B.data(), B.m(), B.n());
}
// b <- A^{-1} b after LU(A) (dense.hpp:202-212).
inline void LUSolveVec(Dense<double>& A, Vec<double>& b) {
  if (A.m() != A.n() || b.n() != A.m()) throw std::domain_error("LUSolveVec: dimension mismatch");
  detail::lu_solve_cols(A, b.data(), b.n(), 1);
}
// Solves X A = B for X after LU(A), B (m x n) overwritten by X (dgetrs 'T' on B^T,
// dense.hpp:214-227).
inline void LUSolveT(Dense<double>& A, Dense<double>& B) {
  if (A.m() != A.n() || B.n() != A.m()) throw std::domain_error("LUSolveT: dimension mismatch");
  Dense<double> X;
  transpose(B, X);
  detail::lu_solve_cols_t(A, X.data(), X.m(), X.n());
  transpose(X, B);
}

// C := alpha op(A) op(B) + beta C, op(X) = X or X^T for trans 'N'/'n' or otherwise; C keeps its
// size (it must be m x n); nothing happens when m, n or k is 0 (dense.hpp:281-310).
inline void GEMM(double alpha, const Dense<double>& A, char transa, const Dense<double>& B, char transb, double beta,
                 Dense<double>& C) {
  const bool ta = !(transa == 'N' || transa == 'n'), tb = !(transb == 'N' || transb == 'n');
  const int m = ta ? A.n() : A.m(), k = ta ? A.m() : A.n(), n = tb ? B.m() : B.n();
  if (m == 0 || n == 0 || k == 0) return;
  if ((tb ? B.n() : B.m()) != k || C.m() != m || C.n() != n) throw std::domain_error("GEMM: dimension mismatch");
  for (int j = 0; j < n; j++)
    for (int i = 0; i < m; i++) {
      double s = 0.0;
      for (int l = 0; l < k; l++) s += (ta ? A(l, i) : A(i, l)) * (tb ? B(j, l) : B(l, j));
      C(i, j) = (beta == 0.0 ? 0.0 : beta * C(i, j)) + alpha * s;
    }
}

}  // namespace LinAlg

#endif
