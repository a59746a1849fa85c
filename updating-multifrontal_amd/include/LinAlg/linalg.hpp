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
  void print() const {
    for (const Scalar& x : data_) std::cout << x << ' ';
    std::cout << std::endl;
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
  void print() const {
    for (int i = 0; i < m_; i++) {
      for (int j = 0; j < n_; j++) std::cout << (*this)(i, j) << ' ';
      std::cout << std::endl;
    }
  }
  Scalar* data() { return data_.data(); }
  const Scalar* data() const { return data_.data(); }

 private:
  int m_, n_;
  std::vector<Scalar> data_;  // column-major, ld = m_
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
  // Printing (sparse.cpp:180-215): triples, the dense matrix, and 1-based "i,j,a," triples to
  // data.csv for MATLAB.
  void print() const {
    std::cout << m_ << " " << n_ << std::endl;
    for (int i = 0; i < m_; i++)
      for (int k = row_ptr_[i]; k < row_ptr_[i + 1]; k++)
        std::cout << "(i,j,a) = (" << i << "," << col_ind_[k] << "," << val_[k] << ")" << std::endl;
    std::cout << std::flush;
  }
  void printDense() const {
    std::cout << m_ << " " << n_ << std::endl;
    for (int i = 0; i < m_; i++) {
      for (int j = 0; j < n_; j++) std::cout << (*this)(i, j) << ' ';
      std::cout << std::endl;
    }
  }
  void printMatlab() const {
    std::ofstream f("data.csv");
    for (int i = 0; i < m_; i++)
      for (int k = row_ptr_[i]; k < row_ptr_[i + 1]; k++) f << i + 1 << "," << col_ind_[k] + 1 << "," << val_[k] << ",";
  }

 private:
  int m_, n_, nnz_;
  std::vector<Scalar> val_;
  IdxVec col_ind_, row_ptr_;
};

}  // namespace LinAlg

#endif
