// kernels.hip -- hand-written gfx950 kernels of the multifrontal factor / update / solve.
//
// Every kernel works on a batch: a list of front (or separator) ids of one tree level, one
// front per blockIdx.y (or blockIdx.x), so a whole level runs in one launch per stage.
//   assembly   k_zero_lower, k_assemble, k_sep_add          (HBM-bound gathers / scatters)
//   factor     k_chol_diag, k_chol_panel, k_gemm<0>, k_gemm<1>   (blocked Cholesky of the
//              pivot columns, left- or right-looking, then the Schur update; FP64 MFMA)
//   solve      k_solve_fwd, k_solve_bwd, k_sep_acc           (multi-RHS level sweeps)
//   apply      k_apply_up, k_apply_down
//   residual   k_spmv
#include <hip/hip_runtime.h>

#include "kernels.h"

namespace umf {

typedef double d4 __attribute__((ext_vector_type(4)));

#define NB 64  // panel width of the blocked factorization and of the solve blocks

// ----------------------------------------------------------------------------- assembly
// Zero the lower triangle (row >= col) of each listed matrix.
__global__ __launch_bounds__(256) void k_zero_lower(const int* __restrict__ ids, const MatDev* __restrict__ mats,
                                                    double* __restrict__ A) {
  const MatDev M = mats[ids[blockIdx.y]];
  double* a = A + M.off;
  for (int c = blockIdx.x; c < M.cols; c += gridDim.x) {
    double* col = a + (i64)c * M.ld;
    for (int r = c + threadIdx.x; r < M.ld; r += 256) col[r] = 0.0;
  }
}

// CSR scatter triples and 64x64 tiles of block copies into each listed matrix.  All writes of
// one matrix are to distinct entries (the planner emits disjoint programs), so no ordering or
// atomics are needed; the matrix must have been zeroed by k_zero_lower.
__global__ __launch_bounds__(256) void k_assemble(const int* __restrict__ ids, const MatDev* __restrict__ mats,
                                                  double* __restrict__ A, const i64* __restrict__ tri_off,
                                                  const Triple* __restrict__ tri, const i64* __restrict__ tile_off,
                                                  const Tile* __restrict__ tiles, const BlockOp* __restrict__ ops,
                                                  const double* __restrict__ csr_val) {
  const int mid = ids[blockIdx.y];
  const MatDev M = mats[mid];
  double* a = A + M.off;
  const i64 t0 = tri_off[mid], t1 = tri_off[mid + 1];
  for (i64 i = t0 + (i64)blockIdx.x * 256 + threadIdx.x; i < t1; i += (i64)gridDim.x * 256) {
    const Triple t = tri[i];
    a[(i64)t.col * M.ld + t.row] = csr_val[t.k];
  }
  const i64 b0 = tile_off[mid], b1 = tile_off[mid + 1];
  for (i64 ti = b0 + blockIdx.x; ti < b1; ti += gridDim.x) {
    const Tile T = tiles[ti];
    const BlockOp op = ops[T.op];
    const MatDev S = mats[op.src];
    const double* s = A + S.off;
    const bool trans = op.flags & 1, lower = op.flags & 2;
    for (int e = threadIdx.x; e < 64 * 64; e += 256) {
      int r = T.r0 + (e & 63), c = T.c0 + (e >> 6);
      if (r >= op.nr || c >= op.nc) continue;
      if (lower && r < c) continue;
      double v = trans ? s[(i64)(op.sc0 + r) * S.ld + (op.sr0 + c)] : s[(i64)(op.sc0 + c) * S.ld + (op.sr0 + r)];
      a[(i64)(op.dc0 + c) * M.ld + (op.dr0 + r)] = v;
    }
  }
}

// Separator Schur accumulation: A11 = (A11 + S_box0) + S_box1 on the lower triangle, the two
// neighbour boxes in ascending index order as the reference's ScatterAdd_ calls
// (src/DE/mf2d.cpp:599-606).  S is read from each box's trailing diagonal separator block.
__global__ __launch_bounds__(256) void k_sep_add(const int* __restrict__ ids, const SepDev* __restrict__ seps,
                                                 const FrontDev* __restrict__ fr, double* __restrict__ A) {
  const SepDev S = seps[ids[blockIdx.y]];
  const FrontDev B0 = fr[S.box[0]], B1 = fr[S.box[1]];
  double* a = A + S.off;
  const double* s0 = A + B0.off + (i64)(B0.n + S.pos[0]) * B0.N + B0.n + S.pos[0];
  const double* s1 = A + B1.off + (i64)(B1.n + S.pos[1]) * B1.N + B1.n + S.pos[1];
  for (int c = blockIdx.x; c < S.k; c += gridDim.x)
    for (int r = c + threadIdx.x; r < S.k; r += 256) {
      double v = a[(i64)c * S.k + r] + s0[(i64)c * B0.N + r];
      a[(i64)c * S.k + r] = v + s1[(i64)c * B1.N + r];
    }
}

// ----------------------------------------------------------------------------- MFMA tile GEMM
// acc += P Q^T for one 64 x 64 tile with v_mfma_f64_16x16x4 (4 waves as 2 x 2, each 32 x 32 =
// 2 x 2 blocks of 16 x 16).  P(i,k) = P[k*ldp + i], Q(j,k) = Q[k*ldq + j]; rows i >= pn and
// j >= qn read as zero; k in [0, K).  The MFMA A operand is Q and B is P, so accumulator r of a
// lane holds C(i, j) with i = 16-block row + lane%16 and j = 16-block col + lane/16 + 4r
// (v_mfma_f64_16x16x4 leaves D[(lane/16) + 4r][lane%16] in accumulator r), and the 16 lanes of a
// group cover 16 consecutive rows of C.
struct TileSmem {
  double As[16][64];
  double Bs[16][64];
};

__device__ __forceinline__ void tile_gemm_nt(d4 (&acc)[2][2], const double* __restrict__ P, int ldp, int pn,
                                             const double* __restrict__ Q, int ldq, int qn, int K, TileSmem& sm) {
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  const int wr = w >> 1, wc = w & 1;
  const int li = t & 63, lk = t >> 6;
  const bool prow = li < pn, qrow = li < qn;
  double pa[4], qb[4];
  auto fetch = [&](int kc) {
#pragma unroll
    for (int q = 0; q < 4; q++) {
      const int k = kc + lk + 4 * q;
      const bool kin = k < K;
      pa[q] = (kin && prow) ? P[(i64)k * ldp + li] : 0.0;
      qb[q] = (kin && qrow) ? Q[(i64)k * ldq + li] : 0.0;
    }
  };
  if (K <= 0) return;
  fetch(0);
  for (int kc = 0; kc < K; kc += 16) {
    __syncthreads();
#pragma unroll
    for (int q = 0; q < 4; q++) {
      sm.As[lk + 4 * q][li] = pa[q];
      sm.Bs[lk + 4 * q][li] = qb[q];
    }
    __syncthreads();
    if (kc + 16 < K) fetch(kc + 16);
#pragma unroll
    for (int k4 = 0; k4 < 4; k4++) {
      const int kr = k4 * 4 + (lane >> 4);
#pragma unroll
      for (int bi = 0; bi < 2; bi++) {
        const double bv = sm.As[kr][wr * 32 + bi * 16 + (lane & 15)];
#pragma unroll
        for (int bj = 0; bj < 2; bj++) {
          const double av = sm.Bs[kr][wc * 32 + bj * 16 + (lane & 15)];
          acc[bi][bj] = __builtin_amdgcn_mfma_f64_16x16x4f64(av, bv, acc[bi][bj], 0, 0, 0);
        }
      }
    }
  }
  __syncthreads();
}

// tile-local coordinates of accumulator (bi, bj, r) of this thread
__device__ __forceinline__ int acc_row(int bi) { return ((threadIdx.x >> 6) >> 1) * 32 + bi * 16 + (threadIdx.x & 15); }
__device__ __forceinline__ int acc_col(int bj, int r) {
  return ((threadIdx.x >> 6) & 1) * 32 + bj * 16 + ((threadIdx.x & 63) >> 4) + 4 * r;
}

// ----------------------------------------------------------------------------- factorization
// Blocked partial Cholesky of the pivot columns [0, n) of each front, 64 columns per block.
//   k_chol_diag:  the diagonal block (already carrying every earlier block's update) is
//                 factored in LDS, T = L_jj L_jj^T, together with Linv_jj = L_jj^{-1}.
//   k_chol_panel: rows below the block: X = (A(rows, jb block) - L(rows, 0:K) L(jb.., 0:K)^T)
//                 Linv_jj^T, i.e. the panel TRSM as a product with the inverted diagonal block.
//                 Left-looking (K = jb): the off-diagonal tiles are formed once, from the
//                 untouched input; the workgroup then applies its rank-64 update to the one
//                 future diagonal tile it owns (rows = cols = its row range).  Right-looking
//                 (K = 0): the trailing region was updated by k_gemm<0>.
// A non-positive (or NaN) pivot records the first failing (front, column) in err[0..2].
// acc += P Q^T for 16 x 16 blocks held in LDS, one wave, v_mfma_f64_16x16x4 (K = 16):
// P(i,k) = P[i*psi + k*psk], Q(j,k) = Q[j*qsj + k*qsk]; accumulator r of a lane is
// C(lane%16, lane/16 + 4r).  PM / QM mark a triangular operand whose other triangle is not
// stored as zeros: 1 = entries with k > row index read as zero, 2 = entries with k < row index.
template <int PM = 0, int QM = 0>
__device__ __forceinline__ void mma16(d4& acc, const double* P, int psi, int psk, const double* Q, int qsj, int qsk) {
  const int lane = threadIdx.x & 63, i = lane & 15, kq = lane >> 4;
#pragma unroll
  for (int k4 = 0; k4 < 4; k4++) {
    const int k = 4 * k4 + kq;
    double pv = P[i * psi + k * psk], qv = Q[i * qsj + k * qsk];
    if ((PM == 1 && k > i) || (PM == 2 && k < i)) pv = 0.0;
    if ((QM == 1 && k > i) || (QM == 2 && k < i)) qv = 0.0;
    acc = __builtin_amdgcn_mfma_f64_16x16x4f64(qv, pv, acc, 0, 0, 0);
  }
}

// 1/sqrt(d) from the hardware estimate and two Newton steps, y <- y (3/2 - d y^2 / 2), which
// carry the estimate to full double precision.
__device__ __forceinline__ double rsqrt_nr(double d) {
  double y = __builtin_amdgcn_rsq(d);
#pragma unroll
  for (int it = 0; it < 2; it++) {
    const double e = fma(-0.5 * d * y, y, 0.5);
    y = fma(y, e, y);
  }
  return y;
}

// Cholesky of the 64 x 64 diagonal block at (jb, jb) of each front with its inverse, one
// workgroup per front, as 4 x 4 blocks of 16 x 16 in LDS:
//   for k: eliminate D_kk on [D_kk | I] without scaling the pivot rows (one thread per element,
//   one barrier per column: step j reads column j and row j of the inverse part and writes only
//   later columns / rows), then scale: L_kk = D' diag(r), Linv_kk = diag(r) V with r_j =
//   1/sqrt(pivot j); panel L_ik = A_ik Linv_kk^T and trailing A_ij -= L_ik L_jk^T on MFMA, one
//   16-block per wave; then the off-diagonal blocks of the inverse,
//   Linv_ij = -Linv_ii sum_k L_ik Linv_kj, by distance i - j.
// T[c*LD + r] holds L(r, c) for r >= c; the inverse is stored row-major in the strict upper
// triangle: Linv(r, c) at U[r*LD + c], U = T + LD, for r >= c.
#define LD (NB + 1)
__global__ __launch_bounds__(256) void k_chol_diag(const int* __restrict__ ids, const FrontDev* __restrict__ fr,
                                                   double* __restrict__ A, int jb, int* __restrict__ err) {
  const int fid = ids[blockIdx.x];
  const FrontDev F = fr[fid];
  if (jb >= F.n) return;
  const int kb = min(NB, F.n - jb);
  __shared__ double T[LD * LD];
  __shared__ double rs[NB];
  double* U = T + LD;
  const int t = threadIdx.x, w = t >> 6, lane = t & 63;
  double* blk = A + F.off + (i64)jb * F.N + jb;
  for (int e = t; e < NB * NB; e += 256) {
    const int i = e & 63, j = e >> 6;
    double v = 0.0;  // strict upper part: the inverse, initially zero
    if (i >= j) v = (i < kb && j < kb) ? blk[(i64)j * F.N + i] : (i == j ? 1.0 : 0.0);  // identity padding
    T[j * LD + i] = v;
  }
  if (t < NB) T[NB * LD + t] = 0.0;
  __syncthreads();
  if (t < NB) U[t * LD + t] = 1.0;
  const int er = t & 15, ec = t >> 4;  // element of a 16 x 16 block
  for (int k = 0; k < 4; k++) {
    const int K0 = 16 * k;
    double* Dk = T + K0 * LD + K0;  // D(r, c) at Dk[c*LD + r]
    double* Vk = U + K0 * LD + K0;  // V(r, c) at Vk[r*LD + c]
    for (int j = 0; j < 16; j++) {
      __syncthreads();
      const double d = Dk[j * LD + j];
      const double ar = Dk[j * LD + er], ac = Dk[j * LD + ec], vj = Vk[j * LD + ec];
      const double r = rsqrt_nr(d);
      if (t == 0) {
        rs[K0 + j] = r;
        if (K0 + j < kb && !(d > 0.0) && atomicCAS(err, 0, 1) == 0) {
          err[1] = fid;
          err[2] = jb + K0 + j;
        }
      }
      const double lr = ar * r;
      if (er > j && ec > j && ec <= er) Dk[ec * LD + er] -= lr * (ac * r);
      if (er > j && ec <= j) Vk[er * LD + ec] -= (lr * r) * vj;
    }
    __syncthreads();
    if (er >= ec) {
      Dk[ec * LD + er] *= rs[K0 + ec];
      Vk[er * LD + ec] *= rs[K0 + er];
    }
    __syncthreads();
    // (b) panel blocks below: L_ik = A_ik Linv_kk^T, wave w -> block i = k+1+w
    {
      const int i = k + 1 + w;
      d4 acc = (d4){0.0, 0.0, 0.0, 0.0};
      if (i < 4) mma16<0, 1>(acc, T + K0 * LD + 16 * i, 1, LD, Vk, LD, 1);
      __syncthreads();
      if (i < 4)
#pragma unroll
        for (int q = 0; q < 4; q++) T[(K0 + (lane >> 4) + 4 * q) * LD + 16 * i + (lane & 15)] = acc[q];
    }
    __syncthreads();
    // (c) trailing blocks A_ij -= L_ik L_jk^T for k < j <= i
    for (int p = w; p < 6; p += 4) {
      int cnt = 0, bi = -1, bj = -1;
      for (int i2 = k + 1; i2 < 4; i2++)
        for (int j2 = k + 1; j2 <= i2; j2++) {
          if (cnt == p) { bi = i2; bj = j2; }
          cnt++;
        }
      if (bi < 0) continue;
      d4 acc = (d4){0.0, 0.0, 0.0, 0.0};
      mma16(acc, T + K0 * LD + 16 * bi, 1, LD, T + K0 * LD + 16 * bj, 1, LD);
      // a diagonal block keeps its strict upper triangle, which holds the inverse
#pragma unroll
      for (int q = 0; q < 4; q++)
        if (bi != bj || (lane & 15) >= (lane >> 4) + 4 * q)
          T[(16 * bj + (lane >> 4) + 4 * q) * LD + 16 * bi + (lane & 15)] -= acc[q];
    }
  }
  __syncthreads();
  // inverse off-diagonal blocks, by distance; S = sum_kk L_i,kk Linv_kk,j is staged in the
  // (not yet computed) destination block Linv_ij.
  for (int dist = 1; dist < 4; dist++) {
    const int i = w + dist, j = w;
    double* Sij = U + 16 * i * LD + 16 * j;  // S(i', j') at Sij[i'*LD + j']
    d4 acc = (d4){0.0, 0.0, 0.0, 0.0};
    if (i < 4) {  // P(i', k) = L(16i+i', 16kk+k), Q(j', k) = Linv(16kk+k, 16j+j'), lower for kk = j
      mma16<0, 2>(acc, T + 16 * j * LD + 16 * i, 1, LD, U + 16 * j * LD + 16 * j, 1, LD);
      for (int kk = j + 1; kk < i; kk++) mma16(acc, T + 16 * kk * LD + 16 * i, 1, LD, U + 16 * kk * LD + 16 * j, 1, LD);
    }
    if (i < 4)
#pragma unroll
      for (int q = 0; q < 4; q++) Sij[(lane & 15) * LD + (lane >> 4) + 4 * q] = acc[q];
    __syncthreads();
    d4 a2 = (d4){0.0, 0.0, 0.0, 0.0};
    if (i < 4)  // P(i', k) = Linv_ii(i', k), Q(j', k) = S(k, j')
      mma16<1, 0>(a2, U + 16 * i * LD + 16 * i, LD, 1, Sij, 1, LD);
    __syncthreads();
    if (i < 4)
#pragma unroll
      for (int q = 0; q < 4; q++) Sij[(lane & 15) * LD + (lane >> 4) + 4 * q] = -a2[q];
    __syncthreads();
  }
  for (int e = t; e < NB * NB; e += 256) {
    const int i = e & 63, j = e >> 6;
    if (i < kb && j < kb && i >= j) blk[(i64)j * F.N + i] = T[j * LD + i];
  }
  double* dinv = A + F.dinv_off + (i64)(jb / NB) * NB * NB;
  for (int e = t; e < NB * NB; e += 256) {
    const int r = e & 63, c = e >> 6;
    dinv[e] = r >= c ? U[r * LD + c] : 0.0;  // column-major, zero above the diagonal
  }
}

__global__ __launch_bounds__(256) void k_chol_panel(const int* __restrict__ ids, const FrontDev* __restrict__ fr,
                                                    double* __restrict__ A, int jb, int left) {
  const FrontDev F = fr[ids[blockIdx.y]];
  if (jb >= F.n) return;
  const int kb = min(NB, F.n - jb);
  const int row0 = jb + kb + NB * blockIdx.x;
  if (row0 >= F.N) return;
  const int nrow = min(NB, F.N - row0);
  __shared__ union {
    TileSmem g;
    struct {
      double X[NB * (NB + 1)];
    } c;
  } sm;
  const int t = threadIdx.x, lane = t & 63;
  const int wr = (t >> 6) >> 1, wc = (t >> 6) & 1;
  d4 acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; i++)
#pragma unroll
    for (int j = 0; j < 2; j++) acc[i][j] = (d4){0.0, 0.0, 0.0, 0.0};
  const int K = left ? jb : 0;
  tile_gemm_nt(acc, A + F.off + row0, F.N, nrow, A + F.off + jb, F.N, kb, K, sm.g);
  double* pan = A + F.off + (i64)jb * F.N + row0;
#pragma unroll
  for (int bi = 0; bi < 2; bi++)
#pragma unroll
    for (int bj = 0; bj < 2; bj++)
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int i = acc_row(bi), j = acc_col(bj, r);
        sm.c.X[j * (NB + 1) + i] = (i < nrow && j < kb) ? pan[(i64)j * F.N + i] - acc[bi][bj][r] : 0.0;
      }
  const double* dinv = A + F.dinv_off + (i64)(jb / NB) * NB * NB;  // Linv(j, k) at dinv[k*NB + j]
  __syncthreads();
  // Y = X Linv^T: Y(i,j) = sum_k X(i,k) Linv(j,k)
#pragma unroll
  for (int i = 0; i < 2; i++)
#pragma unroll
    for (int j = 0; j < 2; j++) acc[i][j] = (d4){0.0, 0.0, 0.0, 0.0};
#pragma unroll
  for (int k4 = 0; k4 < NB / 4; k4++) {
    const int k = k4 * 4 + (lane >> 4);
#pragma unroll
    for (int bi = 0; bi < 2; bi++) {
      const double bv = sm.c.X[k * (NB + 1) + wr * 32 + bi * 16 + (lane & 15)];
#pragma unroll
      for (int bj = 0; bj < 2; bj++) {
        const double av = dinv[k * NB + wc * 32 + bj * 16 + (lane & 15)];
        acc[bi][bj] = __builtin_amdgcn_mfma_f64_16x16x4f64(av, bv, acc[bi][bj], 0, 0, 0);
      }
    }
  }
#pragma unroll
  for (int bi = 0; bi < 2; bi++)
#pragma unroll
    for (int bj = 0; bj < 2; bj++)
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int i = acc_row(bi), j = acc_col(bj, r);
        if (i < nrow && j < kb) pan[(i64)j * F.N + i] = acc[bi][bj][r];
      }
  if (!left || row0 >= F.n) return;
  // rank-64 update of the future diagonal tile D = A(row0.., row0..): D -= Y Y^T (lower)
  __syncthreads();
#pragma unroll
  for (int bi = 0; bi < 2; bi++)
#pragma unroll
    for (int bj = 0; bj < 2; bj++)
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int i = acc_row(bi), j = acc_col(bj, r);
        sm.c.X[j * (NB + 1) + i] = (i < nrow && j < kb) ? acc[bi][bj][r] : 0.0;
      }
  __syncthreads();
#pragma unroll
  for (int i = 0; i < 2; i++)
#pragma unroll
    for (int j = 0; j < 2; j++) acc[i][j] = (d4){0.0, 0.0, 0.0, 0.0};
#pragma unroll
  for (int k4 = 0; k4 < NB / 4; k4++) {
    const int k = k4 * 4 + (lane >> 4);
#pragma unroll
    for (int bi = 0; bi < 2; bi++) {
      const double bv = sm.c.X[k * (NB + 1) + wr * 32 + bi * 16 + (lane & 15)];
#pragma unroll
      for (int bj = 0; bj < 2; bj++) {
        const double av = sm.c.X[k * (NB + 1) + wc * 32 + bj * 16 + (lane & 15)];
        acc[bi][bj] = __builtin_amdgcn_mfma_f64_16x16x4f64(av, bv, acc[bi][bj], 0, 0, 0);
      }
    }
  }
  const int dn = min(NB, F.n - row0);
  double* D = A + F.off + (i64)row0 * F.N + row0;
#pragma unroll
  for (int bi = 0; bi < 2; bi++)
#pragma unroll
    for (int bj = 0; bj < 2; bj++)
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int i = acc_row(bi), j = acc_col(bj, r);
        if (i < dn && j < dn && i >= j) D[(i64)j * F.N + i] -= acc[bi][bj][r];
      }
}

// Lower-triangular tile enumeration of the region rows [s, rend) x cols [s, cend), 64x64
// tiles: first the T(T+1)/2 tiles of the square part, then the full rows below it.
__device__ __forceinline__ bool tile_of(int x, int s, int rend, int cend, int& row0, int& col0) {
  const int T = (cend - s + 63) >> 6, RT = (rend - s + 63) >> 6;
  const int tri = T * (T + 1) / 2;
  if (x >= tri + (RT - T) * T) return false;
  int rt, ct;
  if (x < tri) {
    rt = (int)((sqrt(8.0 * x + 1.0) - 1.0) * 0.5);
    while (rt * (rt + 1) / 2 > x) rt--;
    while ((rt + 1) * (rt + 2) / 2 <= x) rt++;
    ct = x - rt * (rt + 1) / 2;
  } else {
    x -= tri;
    rt = T + x / T;
    ct = x % T;
  }
  row0 = s + 64 * rt;
  col0 = s + 64 * ct;
  return true;
}

// C -= P Q^T on the lower part of a region, one 64x64 tile per workgroup.
// MODE 0: right-looking trailing update after panel k0: P = F[rows, k0:k0+kb],
//         Q = F[cols, k0:k0+kb], C = F[s:N, s:n] lower, s = k0+kb.
// MODE 1: Schur complement: P = Q = L21 = F[n:N, 0:n], K = n; C = F[n:N, n:N] lower gets
//         A11 + S, except on the diagonal separator blocks, which get S alone.
template <int MODE>
__global__ __launch_bounds__(256) void k_gemm(const int* __restrict__ ids, const FrontDev* __restrict__ fr,
                                              double* __restrict__ A, int k0) {
  const FrontDev F = fr[ids[blockIdx.y]];
  int s, K, kbase, rend, cend;
  if (MODE == 0) {
    if (k0 >= F.n) return;
    const int kb = min(NB, F.n - k0);
    s = k0 + kb;
    if (s >= F.n) return;
    K = kb; kbase = k0; rend = F.N; cend = F.n;
  } else {
    if (F.m == 0) return;
    s = F.n; K = F.n; kbase = 0; rend = F.N; cend = F.N;
  }
  int row0, col0;
  if (!tile_of(blockIdx.x, s, rend, cend, row0, col0)) return;
  __shared__ TileSmem sm;
  d4 acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; i++)
#pragma unroll
    for (int j = 0; j < 2; j++) acc[i][j] = (d4){0.0, 0.0, 0.0, 0.0};
  const double* base = A + F.off + (i64)kbase * F.N;
  tile_gemm_nt(acc, base + row0, F.N, rend - row0, base + col0, F.N, cend - col0, K, sm);
  double* C = A + F.off;
#pragma unroll
  for (int bi = 0; bi < 2; bi++)
#pragma unroll
    for (int bj = 0; bj < 2; bj++)
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int i = row0 + acc_row(bi), j = col0 + acc_col(bj, r);
        if (i >= rend || j >= cend || i < j) continue;
        double* c = C + (i64)j * F.N + i;
        if (MODE == 0) {
          *c -= acc[bi][bj][r];
        } else {
          const int ti = i - F.n, tj = j - F.n;
          int ri = -1, rj = -2;
#pragma unroll
          for (int q = 0; q < 6; q++) {
            if (q < F.nsep) {
              if (ti >= F.sep_pos[q] && ti < F.sep_pos[q] + F.sep_len[q]) ri = q;
              if (tj >= F.sep_pos[q] && tj < F.sep_pos[q] + F.sep_len[q]) rj = q;
            }
          }
          const double sv = -acc[bi][bj][r];
          *c = (ri == rj) ? sv : (*c + sv);
        }
      }
}

template __global__ void k_gemm<0>(const int*, const FrontDev*, double*, int);
template __global__ void k_gemm<1>(const int*, const FrontDev*, double*, int);

// ----------------------------------------------------------------------------- solve
// The sweeps work on W, an n_dof x 16 row-major right-hand-side block (W[dof*16 + r]), and on a
// per-front scratch X (N x 16, row-major).  All block products run on v_mfma_f64_16x16x4 with a
// 16-row block of L (or Linv) as the A operand and a 4 x 16 slice of X as B, so accumulator r
// of a lane is row (lane/16) + 4r, right-hand side lane%16.
#define RB 16
#define GU 4  // 16-row groups in flight per wave

__device__ __forceinline__ void gather_rhs(const FrontDev& F, double* __restrict__ X, const double* __restrict__ W,
                                           const int* __restrict__ di, const int* __restrict__ db) {
  for (int e = threadIdx.x; e < F.N * RB; e += 256) {
    const int i = e >> 4, r = e & 15;
    const int d = i < F.n ? di[i] : (db ? db[i - F.n] : -1);
    X[e] = d >= 0 ? W[(i64)d * RB + r] : 0.0;
  }
}

// Forward sweep of one front: X[0:n] <- L11^{-1} W[int] block by block (x_j = Linv_j b_j, then
// b_below -= L(below, j) x_j), X[n:N] <- -L21 X[0:n]; X[0:n] goes back to W[int].
__global__ __launch_bounds__(256) void k_solve_fwd(const int* __restrict__ ids, const FrontDev* __restrict__ fr,
                                                   const double* __restrict__ A, const int* __restrict__ dofs,
                                                   double* __restrict__ W, double* __restrict__ xs) {
  const FrontDev F = fr[ids[blockIdx.x]];
  const double* L = A + F.off;
  double* X = xs + F.x_off;
  const int* di = dofs + F.int_off;
  const int t = threadIdx.x, lane = t & 63, w = t >> 6, c16 = lane & 15, q = lane >> 4;
  gather_rhs(F, X, W, di, nullptr);
  __shared__ double xb[NB][RB];
  for (int jb = 0; jb < F.n; jb += NB) {
    const int kb = min(NB, F.n - jb);
    __syncthreads();
    {  // x_j = Linv_j b_j: wave w -> rows 16w..16w+15
      const double* Li = A + F.dinv_off + (i64)(jb / NB) * NB * NB;
      d4 acc = (d4){0.0, 0.0, 0.0, 0.0};
#pragma unroll
      for (int k4 = 0; k4 < NB / 4; k4++) {
        const int k = 4 * k4 + q;
        const double av = Li[k * NB + 16 * w + c16];
        const double bv = k < kb ? X[(i64)(jb + k) * RB + c16] : 0.0;
        acc = __builtin_amdgcn_mfma_f64_16x16x4f64(av, bv, acc, 0, 0, 0);
      }
      __syncthreads();
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int i = 16 * w + q + 4 * r;
        xb[i][c16] = acc[r];
        if (i < kb) X[(i64)(jb + i) * RB + c16] = acc[r];
      }
    }
    __syncthreads();
    // rows below, GU 16-row groups per wave in flight; out-of-range rows read row N-1 and are
    // never written
    for (int g0 = jb + kb + 16 * w; g0 < F.N; g0 += 64 * GU) {
      double av[GU][NB / 4];
#pragma unroll
      for (int u = 0; u < GU; u++) {
        const int row = min(g0 + 64 * u + c16, F.N - 1);
#pragma unroll
        for (int k4 = 0; k4 < NB / 4; k4++) {
          const int k = 4 * k4 + q;
          av[u][k4] = k < kb ? L[(i64)(jb + k) * F.N + row] : 0.0;
        }
      }
#pragma unroll
      for (int u = 0; u < GU; u++) {
        d4 acc = (d4){0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int k4 = 0; k4 < NB / 4; k4++)
          acc = __builtin_amdgcn_mfma_f64_16x16x4f64(av[u][k4], xb[4 * k4 + q][c16], acc, 0, 0, 0);
#pragma unroll
        for (int r = 0; r < 4; r++) {
          const int i = g0 + 64 * u + q + 4 * r;
          if (i < F.N) X[(i64)i * RB + c16] -= acc[r];
        }
      }
    }
  }
  __syncthreads();
  for (int e = t; e < F.n * RB; e += 256) W[(i64)di[e >> 4] * RB + (e & 15)] = X[e];
}

// Separator accumulation of per-box contributions X[n+pos : n+pos+k] of both neighbour boxes
// into the work block, box 0 first: W = (W + X_b0) + X_b1.
__global__ __launch_bounds__(256) void k_sep_acc(const int* __restrict__ ids, const SepDev* __restrict__ seps,
                                                 const FrontDev* __restrict__ fr, const int* __restrict__ dofs,
                                                 double* __restrict__ W, const double* __restrict__ xs) {
  const SepDev S = seps[ids[blockIdx.x]];
  const FrontDev B0 = fr[S.box[0]], B1 = fr[S.box[1]];
  const int* d = dofs + S.dof_off;
  const double* x0 = xs + B0.x_off + (i64)(B0.n + S.pos[0]) * RB;
  const double* x1 = xs + B1.x_off + (i64)(B1.n + S.pos[1]) * RB;
  for (int e = threadIdx.x; e < S.k * RB; e += 256) {
    double* wp = W + (i64)d[e >> 4] * RB + (e & 15);
    const double v = *wp + x0[e];
    *wp = v + x1[e];
  }
}

// Backward sweep of one front: X[0:n] = W[int] (forward result y), X[n:N] = W[bd];
// block by block from the last: z_j = y_j - L(below, j)^T X(below), x_j = Linv_j^T z_j
// (= A22^{-1} b_rd - XR b_sk overall); X[0:n] goes back to W[int].
__global__ __launch_bounds__(256) void k_solve_bwd(const int* __restrict__ ids, const FrontDev* __restrict__ fr,
                                                   const double* __restrict__ A, const int* __restrict__ dofs,
                                                   double* __restrict__ W, double* __restrict__ xs) {
  const FrontDev F = fr[ids[blockIdx.x]];
  const double* L = A + F.off;
  double* X = xs + F.x_off;
  const int* di = dofs + F.int_off;
  const int* db = dofs + F.bd_off;
  const int t = threadIdx.x, lane = t & 63, w = t >> 6, c16 = lane & 15, q = lane >> 4;
  gather_rhs(F, X, W, di, db);
  __shared__ double zb[NB][RB];
  const int nblk = (F.n + NB - 1) / NB;
  for (int bk = nblk - 1; bk >= 0; bk--) {
    const int jb = bk * NB, kb = min(NB, F.n - jb);
    __syncthreads();
    // T(c, r) = sum_{i >= jb+kb} L(i, jb+c) X(i, r); wave w owns columns c in [16w, 16w+16),
    // streaming 4-row slices straight from the column-major L (lanes: column c16, row q)
    const int col = 16 * w + c16;
    const double* Lc = L + (i64)(jb + min(col, kb - 1)) * F.N;
    const bool cin = col < kb;
    d4 acc = (d4){0.0, 0.0, 0.0, 0.0};
    for (int i0 = jb + kb; i0 < F.N; i0 += 16 * GU) {
      double av[GU][4], bv[GU][4];
#pragma unroll
      for (int u = 0; u < GU; u++)
#pragma unroll
        for (int k4 = 0; k4 < 4; k4++) {
          const int i = i0 + 16 * u + 4 * k4 + q;
          const bool ok = cin && i < F.N;
          const int ic = min(i, F.N - 1);
          av[u][k4] = ok ? Lc[ic] : 0.0;
          bv[u][k4] = X[(i64)ic * RB + c16];
        }
#pragma unroll
      for (int u = 0; u < GU; u++)
#pragma unroll
        for (int k4 = 0; k4 < 4; k4++) acc = __builtin_amdgcn_mfma_f64_16x16x4f64(av[u][k4], bv[u][k4], acc, 0, 0, 0);
    }
#pragma unroll
    for (int r = 0; r < 4; r++) {
      const int c = 16 * w + q + 4 * r;
      zb[c][c16] = c < kb ? X[(i64)(jb + c) * RB + c16] - acc[r] : 0.0;
    }
    __syncthreads();
    {  // x_j = Linv_j^T z: x(i, r) = sum_c Linv(c, i) z(c, r); wave w -> rows 16w..16w+15
      const double* Li = A + F.dinv_off + (i64)bk * NB * NB;
      d4 a2 = (d4){0.0, 0.0, 0.0, 0.0};
#pragma unroll
      for (int k4 = 0; k4 < NB / 4; k4++) {
        const int c = 4 * k4 + q;
        a2 = __builtin_amdgcn_mfma_f64_16x16x4f64(Li[(16 * w + c16) * NB + c], zb[c][c16], a2, 0, 0, 0);
      }
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int i = 16 * w + q + 4 * r;
        if (i < kb) X[(i64)(jb + i) * RB + c16] = a2[r];
      }
    }
  }
  __syncthreads();
  for (int e = t; e < F.n * RB; e += 256) W[(i64)di[e >> 4] * RB + (e & 15)] = X[e];
}

// ---- blocked sweeps for large fronts: one workgroup per 64-row block, blocks of a front
// hand their results to one another through per-block flags (set to the launch's epoch).
// A workgroup only waits for blocks owned by workgroups with smaller ids, which the hardware
// dispatches first, and every wait is bounded: past the limit err[3] is set and the
// workgroup goes on, so the launch always drains.
#define SPIN_LIMIT (1 << 22)
// The flag is polled with relaxed agent-scope loads (no cache maintenance per poll); one
// acquire fence follows the successful poll.  Publication: every wave's stores are released at
// agent scope before the barrier, then one relaxed store sets the flag.
__device__ __forceinline__ void wait_block(const int* flag, int epoch, int* err) {
  if (threadIdx.x == 0) {
    int spins = 0;
    while (__hip_atomic_load(flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != epoch) {
      if (++spins > SPIN_LIMIT) {
        atomicCAS(err + 3, 0, 1);
        break;
      }
      __builtin_amdgcn_s_sleep(1);
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
  }
  __syncthreads();
}
__device__ __forceinline__ void post_block(int* flag, int epoch) {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
  __syncthreads();
  if (threadIdx.x == 0) __hip_atomic_store(flag, epoch, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Forward sweep of one front by blocks (blockIdx.x = block, blockIdx.y = front): interior block
// b waits for the blocks j < b, forms v = b_b - sum_j L_bj x_j and publishes x_b = Linv_b v
// (to X and to W[int]); boundary block b accumulates over every interior block and leaves
// X[rows] = -sum_j L_bj x_j, the contribution gathered by k_sep_acc.
__global__ __launch_bounds__(256) void k_fwd_blk(const int* __restrict__ ids, const FrontDev* __restrict__ fr,
                                                 const double* __restrict__ A, const int* __restrict__ dofs,
                                                 double* __restrict__ W, double* __restrict__ xs, int* flags, int epoch,
                                                 int* err) {
  const FrontDev F = fr[ids[blockIdx.y]];
  const int nbi = (F.n + NB - 1) / NB, nbt = nbi + (F.m + NB - 1) / NB;
  const int b = blockIdx.x;
  if (b >= nbt) return;
  const bool interior = b < nbi;
  const int r0 = interior ? NB * b : F.n + NB * (b - nbi);
  const int rend = min(interior ? F.n : F.N, r0 + NB);
  const double* L = A + F.off;
  double* X = xs + F.x_off;
  const int* di = dofs + F.int_off;
  int* fl = flags + F.flag_off;
  const int t = threadIdx.x, lane = t & 63, w = t >> 6, c16 = lane & 15, q = lane >> 4;
  __shared__ double vb[NB][RB];
  const int row = r0 + 16 * w + c16;  // L row supplied by this lane
  const bool rok = row < rend;
  const int rowc = rok ? row : r0;
  d4 acc = (d4){0.0, 0.0, 0.0, 0.0};
  const int jmax = interior ? b : nbi;
  for (int j = 0; j < jmax; j++) {
    wait_block(fl + j, epoch, err);
    const int jb = NB * j, kb = min(NB, F.n - jb);
    double av[NB / 4], bv[NB / 4];
#pragma unroll
    for (int k4 = 0; k4 < NB / 4; k4++) {
      const int k = 4 * k4 + q;
      const bool kin = k < kb;
      av[k4] = (rok && kin) ? L[(i64)(jb + k) * F.N + rowc] : 0.0;
      bv[k4] = kin ? X[(i64)(jb + k) * RB + c16] : 0.0;
    }
#pragma unroll
    for (int k4 = 0; k4 < NB / 4; k4++) acc = __builtin_amdgcn_mfma_f64_16x16x4f64(av[k4], bv[k4], acc, 0, 0, 0);
  }
  if (!interior) {
#pragma unroll
    for (int r = 0; r < 4; r++) {
      const int i = r0 + 16 * w + q + 4 * r;
      if (i < rend) X[(i64)i * RB + c16] = -acc[r];
    }
    return;
  }
#pragma unroll
  for (int r = 0; r < 4; r++) {
    const int i = 16 * w + q + 4 * r;
    vb[i][c16] = r0 + i < rend ? W[(i64)di[r0 + i] * RB + c16] - acc[r] : 0.0;
  }
  __syncthreads();
  const double* Li = A + F.dinv_off + (i64)b * NB * NB;  // Linv(i, c) at Li[c*NB + i]
  d4 a2 = (d4){0.0, 0.0, 0.0, 0.0};
#pragma unroll
  for (int k4 = 0; k4 < NB / 4; k4++) {
    const int c = 4 * k4 + q;
    a2 = __builtin_amdgcn_mfma_f64_16x16x4f64(Li[c * NB + 16 * w + c16], vb[c][c16], a2, 0, 0, 0);
  }
#pragma unroll
  for (int r = 0; r < 4; r++) {
    const int i = r0 + 16 * w + q + 4 * r;
    if (i < rend) {
      X[(i64)i * RB + c16] = a2[r];
      W[(i64)di[i] * RB + c16] = a2[r];
    }
  }
  post_block(fl + b, epoch);
}

// Backward sweep of one front by blocks, last block first (workgroup b owns block
// nbi-1-b): z = y_i - L(rows below, i)^T x(rows below), boundary rows (read from W[bd]) first,
// then the interior blocks below as they are published; x_i = Linv_i^T z goes to X and W[int].
__global__ __launch_bounds__(256) void k_bwd_blk(const int* __restrict__ ids, const FrontDev* __restrict__ fr,
                                                 const double* __restrict__ A, const int* __restrict__ dofs,
                                                 double* __restrict__ W, double* __restrict__ xs, int* flags, int epoch,
                                                 int* err) {
  const FrontDev F = fr[ids[blockIdx.y]];
  const int nbi = (F.n + NB - 1) / NB;
  if ((int)blockIdx.x >= nbi) return;
  const int ib = nbi - 1 - blockIdx.x;
  const int jb = NB * ib, kb = min(NB, F.n - jb);
  const double* L = A + F.off;
  double* X = xs + F.x_off;
  const int* di = dofs + F.int_off;
  const int* db = dofs + F.bd_off;
  int* fl = flags + F.flag_off;
  const int t = threadIdx.x, lane = t & 63, w = t >> 6, c16 = lane & 15, q = lane >> 4;
  __shared__ double zb[NB][RB];
  const int col = 16 * w + c16;  // column of the block supplied by this lane
  const bool cin = col < kb;
  const double* Lc = L + (i64)(jb + (cin ? col : 0)) * F.N;
  d4 acc = (d4){0.0, 0.0, 0.0, 0.0};
  for (int i0 = F.n; i0 < F.N; i0 += 16 * GU) {  // boundary rows: x = W[bd]
    double av[GU][4], bv[GU][4];
#pragma unroll
    for (int u = 0; u < GU; u++)
#pragma unroll
      for (int k4 = 0; k4 < 4; k4++) {
        const int i = i0 + 16 * u + 4 * k4 + q;
        const bool ok = i < F.N;
        av[u][k4] = (ok && cin) ? Lc[i] : 0.0;
        bv[u][k4] = ok ? W[(i64)db[i - F.n] * RB + c16] : 0.0;
      }
#pragma unroll
    for (int u = 0; u < GU; u++)
#pragma unroll
      for (int k4 = 0; k4 < 4; k4++) acc = __builtin_amdgcn_mfma_f64_16x16x4f64(av[u][k4], bv[u][k4], acc, 0, 0, 0);
  }
  for (int kbk = nbi - 1; kbk > ib; kbk--) {  // interior blocks below, in publication order
    wait_block(fl + kbk, epoch, err);
    const int k0 = NB * kbk, kn = min(NB, F.n - k0);
    double av[NB / 4], bv[NB / 4];
#pragma unroll
    for (int k4 = 0; k4 < NB / 4; k4++) {
      const int k = 4 * k4 + q;
      const bool kin = k < kn;
      av[k4] = (kin && cin) ? Lc[k0 + k] : 0.0;
      bv[k4] = kin ? X[(i64)(k0 + k) * RB + c16] : 0.0;
    }
#pragma unroll
    for (int k4 = 0; k4 < NB / 4; k4++) acc = __builtin_amdgcn_mfma_f64_16x16x4f64(av[k4], bv[k4], acc, 0, 0, 0);
  }
#pragma unroll
  for (int r = 0; r < 4; r++) {
    const int c = 16 * w + q + 4 * r;
    zb[c][c16] = c < kb ? W[(i64)di[jb + c] * RB + c16] - acc[r] : 0.0;
  }
  __syncthreads();
  const double* Li = A + F.dinv_off + (i64)ib * NB * NB;
  d4 a2 = (d4){0.0, 0.0, 0.0, 0.0};
#pragma unroll
  for (int k4 = 0; k4 < NB / 4; k4++) {
    const int c = 4 * k4 + q;
    a2 = __builtin_amdgcn_mfma_f64_16x16x4f64(Li[(16 * w + c16) * NB + c], zb[c][c16], a2, 0, 0, 0);
  }
#pragma unroll
  for (int r = 0; r < 4; r++) {
    const int i = 16 * w + q + 4 * r;
    if (i < kb) {
      X[(i64)(jb + i) * RB + c16] = a2[r];
      W[(i64)di[jb + i] * RB + c16] = a2[r];
    }
  }
  post_block(fl + ib, epoch);
}

// Column-major user block (n x nr, ld) <-> row-major work block (n x 16); unused columns zero.
__global__ __launch_bounds__(256) void k_to_rows(int n, const double* __restrict__ B, int ld, int nr,
                                                 double* __restrict__ W) {
  const i64 e = (i64)blockIdx.x * 256 + threadIdx.x;
  if (e >= (i64)n * RB) return;
  const int i = (int)(e >> 4), r = (int)(e & 15);
  W[e] = r < nr ? B[(i64)r * ld + i] : 0.0;
}
__global__ __launch_bounds__(256) void k_from_rows(int n, const double* __restrict__ W, double* __restrict__ B,
                                                   int ld, int nr) {
  const i64 e = (i64)blockIdx.x * 256 + threadIdx.x;
  if (e >= (i64)n * nr) return;
  const int r = (int)(e / n), i = (int)(e % n);
  B[(i64)r * ld + i] = W[(i64)i * RB + r];
}

// ----------------------------------------------------------------------------- apply
// Upward step of Apply for one front: u = [L11; L21]^T [x_rd; x_sk] (= L11^T b_rd + W b_sk),
// b_rd <- L11 u.  u is kept in X[0:n] for the downward step.
__global__ __launch_bounds__(256) void k_apply_up(const int* __restrict__ ids, const FrontDev* __restrict__ fr,
                                                  const double* __restrict__ A, const int* __restrict__ dofs,
                                                  double* __restrict__ W, double* __restrict__ xs, int nr) {
  const FrontDev F = fr[ids[blockIdx.x]];
  const double* L = A + F.off;
  double* X = xs + F.x_off;
  const int* di = dofs + F.int_off;
  const int* db = dofs + F.bd_off;
  const int t = threadIdx.x;
  gather_rhs(F, X, W, di, db);
  __syncthreads();
  // u_c = sum_{i >= c} L[i][c] X[i]: one column at a time, block reduction
  __shared__ double red[256];
  for (int c = 0; c < F.n; c++) {
    for (int r = 0; r < nr; r++) {
      double s = 0.0;
      for (int i = c + t; i < F.N; i += 256) s += L[(i64)c * F.N + i] * X[(i64)i * RB + r];
      red[t] = s;
      __syncthreads();
      for (int h = 128; h > 0; h >>= 1) {
        if (t < h) red[t] += red[t + h];
        __syncthreads();
      }
      if (t == 0) X[(i64)c * RB + r] = red[0];  // row c is no longer read by later columns
      __syncthreads();
    }
  }
  // b_rd = L11 u
  for (int i = t; i < F.n; i += 256)
    for (int r = 0; r < nr; r++) {
      double s = 0.0;
      for (int c = 0; c <= i; c++) s += L[(i64)c * F.N + i] * X[(i64)c * RB + r];
      W[(i64)di[i] * RB + r] = s;
    }
}

// Downward step of Apply: X[n:N] = L21 u (the reference's XL b_rd), accumulated by k_sep_acc.
__global__ __launch_bounds__(256) void k_apply_down(const int* __restrict__ ids, const FrontDev* __restrict__ fr,
                                                    const double* __restrict__ A, double* __restrict__ xs, int nr) {
  const FrontDev F = fr[ids[blockIdx.x]];
  const double* L = A + F.off;
  double* X = xs + F.x_off;
  for (int i = F.n + threadIdx.x; i < F.N; i += 256)
    for (int r = 0; r < RB; r++) {
      double s = 0.0;
      if (r < nr)
        for (int c = 0; c < F.n; c++) s += L[(i64)c * F.N + i] * X[(i64)c * RB + r];
      X[(i64)i * RB + r] = s;
    }
}

// ----------------------------------------------------------------------------- SpMV
__global__ __launch_bounds__(256) void k_spmv(int n, const int* __restrict__ rp, const int* __restrict__ ci,
                                              const double* __restrict__ v, const double* __restrict__ x, int ldx,
                                              double* __restrict__ y, int ldy, int nr) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const int k0 = rp[i], k1 = rp[i + 1];
  for (int r = 0; r < nr; r++) {
    double s = 0.0;
    for (int k = k0; k < k1; k++) s += v[k] * x[(i64)r * ldx + ci[k]];
    y[(i64)r * ldy + i] = s;
  }
}

}  // namespace umf
