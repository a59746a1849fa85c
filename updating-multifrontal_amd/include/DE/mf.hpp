// mf.hpp -- drop-in TreeFactor2D::MF2D / TreeFactor3D::MF3D over libumf_hip.so.
//
// Same class names, constructor arguments and public methods as the reference's
// include/DE/mf.hpp:43-80 and :149-188:
//   MF2D(const Sparse& A, int Nx, int Ny, int width = 1, int maxSideLength = 4)
//   MF3D(const Sparse& A, int Nx, int Ny, int Nz, int width = 1, int maxSideLength = 4)
//   void Factor();  void Update(const Sparse& A, std::set<int> modifiedDOFs);
//   Vec Solve(const Vec& b);  Vec Apply(const Vec& x);
// plus Solve/Apply on a Dense n x nrhs block.  The factorization lives in GPU memory; every
// call is synchronous.  Errors: the constructor throws std::runtime_error with the reference's
// geometry messages (src/DE/mf2d.cpp:42, :64); a non-positive pivot, which the reference meets
// with assert(info == 0), throws std::runtime_error naming the front and pivot.
// Link with -L<pkg> -lumf_hip; include this directory and ../LinAlg.
#pragma once
#ifndef UMF_MF_HPP
#define UMF_MF_HPP 1

#include <set>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

#include "../../../include/umf.h"
#include "../LinAlg/linalg.hpp"

namespace umf_detail {

inline void check(int rc) {
  if (rc == UMF_OK) return;
  std::string msg = umf_last_error();
  if (rc == UMF_EINVAL || rc == UMF_EPATTERN || rc == UMF_EUPDATEMAP) throw std::invalid_argument(msg);
  throw std::runtime_error(msg);
}

template <class Scalar>
class Front {
  static_assert(std::is_same<Scalar, double>::value, "the multifrontal factorization is FP64 (Scalar = double)");

 public:
  Front(const LinAlg::Sparse<Scalar>& A, int dim, int Nx, int Ny, int Nz, int width, int maxSideLength)
      : n_(A.m()) {
    check(umf_plan_create(dim, Nx, Ny, Nz, width, maxSideLength, A.m(), A.row_ptr_data(), A.col_ind_data(),
                          A.val_data(), &plan_));
  }
  ~Front() { umf_plan_destroy(plan_); }
  Front(const Front&) = delete;
  Front& operator=(const Front&) = delete;

  void Factor() { check(umf_factor(plan_)); }

  // Precision form (no reference counterpart): Precision::FP32 keeps and factors the fronts in
  // FP32 on FP32 MFMA, and Solve refines in FP64 to backward error <= tol.  BF16X3 is FP32
  // with the Schur and trailing-update GEMMs on BF16 MFMA over two bfloat16 pieces of each
  // operand (3 products, FP32 accumulation).  FP32T64 keeps the top three separator levels in
  // FP64 and the levels below in FP32, refined as FP32.  Factor() again after switching.
  enum class Precision {
    FP64 = UMF_PREC_FP64,
    FP32 = UMF_PREC_FP32,
    BF16X3 = UMF_PREC_BF16X3,
    FP32T64 = UMF_PREC_FP32T64
  };
  void SetPrecision(Precision p, double tol = 1e-10, int maxIter = 50) {
    check(umf_set_precision(plan_, (int)p, tol, maxIter));
  }

  void Update(const LinAlg::Sparse<Scalar>& A, std::set<int> modifiedDOFs) {
    std::vector<int> mod(modifiedDOFs.begin(), modifiedDOFs.end());
    check(umf_update(plan_, A.row_ptr_data(), A.col_ind_data(), A.val_data(), (int)mod.size(), mod.data()));
  }

  LinAlg::Vec<Scalar> Solve(const LinAlg::Vec<Scalar>& b) {
    LinAlg::Vec<Scalar> x(n_);
    check(umf_solve(plan_, b.data(), x.data(), 1, n_));
    return x;
  }
  LinAlg::Vec<Scalar> Apply(const LinAlg::Vec<Scalar>& x) {
    LinAlg::Vec<Scalar> y(n_);
    check(umf_apply(plan_, x.data(), y.data(), 1, n_));
    return y;
  }
  // many right-hand sides at once (n x nrhs, column-major)
  LinAlg::Dense<Scalar> Solve(const LinAlg::Dense<Scalar>& B) {
    LinAlg::Dense<Scalar> X(B.m(), B.n());
    check(umf_solve(plan_, B.data(), X.data(), B.n(), n_));
    return X;
  }
  LinAlg::Dense<Scalar> Apply(const LinAlg::Dense<Scalar>& B) {
    LinAlg::Dense<Scalar> Y(B.m(), B.n());
    check(umf_apply(plan_, B.data(), Y.data(), B.n(), n_));
    return Y;
  }

  // Front factorization: LU (partial pivoting within each front's interior rows, as the
  // reference's every front) or Cholesky.  Operators with asymmetric values get LU fronts at
  // construction; SetLU(true) makes a symmetric (e.g. indefinite) operator use them too.
  void SetLU(bool on) { check(umf_set_lu(plan_, on ? 1 : 0)); }
  bool IsLU() const { return umf_get_lu(plan_) == 1; }

  // Plan files (no reference counterpart): geometry, current operator values and, when
  // factored, the factors; the file constructors of MF2D / MF3D restore them without
  // refactoring (umf_plan_save / umf_plan_load).
  void Save(const std::string& path) { check(umf_plan_save(plan_, path.c_str())); }

  umf_plan* plan() { return plan_; }
  umf_stats stats() const {
    umf_stats s;
    check(umf_get_stats(plan_, &s));
    return s;
  }

 protected:
  Front(const std::string& planFile, int dim) : n_(0) {
    check(umf_plan_load(planFile.c_str(), &plan_));
    const bool is3d = umf_leaf_width(plan_, 2) >= 0;
    if (is3d != (dim == 3)) {
      umf_plan_destroy(plan_);
      plan_ = nullptr;
      throw std::invalid_argument(planFile + ": the plan is " + (is3d ? "3D" : "2D"));
    }
    const int rc = umf_batch_info(plan_, 0, nullptr, &n_, nullptr, nullptr);
    check(rc < 0 ? rc : UMF_OK);
  }

 private:
  int n_;
  umf_plan* plan_ = nullptr;
};

}  // namespace umf_detail

namespace TreeFactor2D {
template <class Scalar>
class MF2D : public umf_detail::Front<Scalar> {
 public:
  MF2D(const LinAlg::Sparse<Scalar>& A, int Nx, int Ny, int width = 1, int maxSideLength = 4)
      : umf_detail::Front<Scalar>(A, 2, Nx, Ny, 1, width, maxSideLength) {}
  explicit MF2D(const std::string& planFile) : umf_detail::Front<Scalar>(planFile, 2) {}
};
}  // namespace TreeFactor2D

namespace TreeFactor3D {
template <class Scalar>
class MF3D : public umf_detail::Front<Scalar> {
 public:
  MF3D(const LinAlg::Sparse<Scalar>& A, int Nx, int Ny, int Nz, int width = 1, int maxSideLength = 4)
      : umf_detail::Front<Scalar>(A, 3, Nx, Ny, Nz, width, maxSideLength) {}
  explicit MF3D(const std::string& planFile) : umf_detail::Front<Scalar>(planFile, 3) {}
};
}  // namespace TreeFactor3D

#endif
