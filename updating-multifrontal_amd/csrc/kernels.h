// kernels.h -- device-side descriptors shared by the host runtime and the gfx950 kernels.
#pragma once
#include <hip/hip_runtime.h>

#include "plan.h"

namespace umf {

// One front as the kernels see it.  The front F is (n+m) x (n+m), column-major, ld = N = n+m,
// at arena + off:  rows/cols [0,n) = intDOFs (pivot block A22), [n,N) = bdDOFs.
// After factorization: L11 in the lower part of F[0:n,0:n], L21 = A12 L11^{-T} in F[n:N,0:n],
// F[n:N,n:N] (lower) = A11 + S off the diagonal separator blocks and S = -L21 L21^T on them.
struct FrontDev {
  i64 off;      // arena offset of F
  i64 int_off;  // DOF list offsets
  i64 bd_off;
  i64 x_off;    // solve scratch offset: N x 16 right-hand-side block, row-major
  i64 dinv_off; // inverted 64x64 diagonal blocks of L11, column-major, zero above the diagonal
  int n, m, N, nsep;
  int sep_pos[6], sep_len[6];
  int flag_off;  // first of the ceil(n/64) block flags of the blocked solve sweeps
  int fly_off;   // its entries by panel tile in the fly table (leaf fronts), -1 = none
  // LU fronts (unsymmetric operators): F[0:n,0:n] = L\U of P A22 (L unit lower), F[0:n,n:N] =
  // U12 = L^{-1} P A21, F[n:N,0:n] = L21 = A12 U^{-1}, F[n:N,n:N] = A11 - L21 U12; dinv holds the
  // inverted 64x64 diagonal blocks of L, uinv_off those of U (in the LU buffer), ut_off the
  // front's 64-row slice of transposed U rows used by the trailing update (N x 64)
  i64 uinv_off;
  i64 ut_off;
};
struct SepDev {
  i64 off;  // arena offset of the k x k A11
  i64 dof_off;
  int k;
  int box[2];  // front ids
  int pos[2];  // offsets inside each neighbour's bdDOFs
};
struct MatDev {
  i64 off;
  int ld;
  int cols;
  int lev;  // tree level of the matrix (merged level in a batch)
};

}  // namespace umf
