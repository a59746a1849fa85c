// kernels.hip -- hand-written gfx950 kernels of the multifrontal factor / update / solve.
//
// Every kernel works on a batch: a list of front (or separator) ids of one tree level, one
// front per blockIdx.y (or blockIdx.x), so a whole level runs in one launch per stage.
//   assembly   k_zero_lower, k_assemble, k_sep_add          (HBM-bound gathers / scatters)
//   factor     k_potrf_diag, k_trsm_panel, k_gemm<0>, k_gemm<1>   (right-looking blocked
//              Cholesky of the pivot columns, then the Schur update on FP64 MFMA)
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

// ----------------------------------------------------------------------------- factorization
// Unblocked Cholesky of the kb x kb diagonal block at (k0,k0) of each front, in LDS.
// A non-positive (or NaN) pivot records the first failing (front, column) in err[0..2].
__global__ __launch_bounds__(256) void k_potrf_diag(const int* __restrict__ ids, const FrontDev* __restrict__ fr,
                                                    double* __restrict__ A, int k0, int* __restrict__ err) {
  const int fid = ids[blockIdx.x];
  const FrontDev F = fr[fid];
  if (k0 >= F.n) return;
  const int kb = min(NB, F.n - k0);
  __shared__ double T[NB * (NB + 1)];
  double* a = A + F.off + (i64)k0 * F.N + k0;
  const int t = threadIdx.x;
  for (int e = t; e < NB * NB; e += 256) {
    int r = e & 63, c = e >> 6;
    T[c * (NB + 1) + r] = (r < kb && c < kb && r >= c) ? a[(i64)c * F.N + r] : 0.0;
  }
  for (int j = 0; j < kb; j++) {
    __syncthreads();
    const double d = T[j * (NB + 1) + j];
    if (t == 0 && !(d > 0.0)) {
      if (atomicCAS(err, 0, 1) == 0) { err[1] = fid; err[2] = k0 + j; }
    }
    const double sd = sqrt(d);
    __syncthreads();
    for (int i = j + 1 + t; i < kb; i += 256) T[j * (NB + 1) + i] /= sd;
    if (t == 0) T[j * (NB + 1) + j] = sd;
    __syncthreads();
    const int cnt = kb - j - 1;
    for (int e = t; e < cnt * cnt; e += 256) {
      int rr = e % cnt, cc = e / cnt;
      if (rr >= cc) T[(j + 1 + cc) * (NB + 1) + j + 1 + rr] -= T[j * (NB + 1) + j + 1 + rr] * T[j * (NB + 1) + j + 1 + cc];
    }
  }
  __syncthreads();
  for (int e = t; e < NB * NB; e += 256) {
    int r = e & 63, c = e >> 6;
    if (r < kb && c < kb && r >= c) a[(i64)c * F.N + r] = T[c * (NB + 1) + r];
  }
}

// Panel TRSM: rows [k0+kb, N) of columns [k0, k0+kb) become X = A L_kk^{-T} (forward
// substitution per row, one row per thread, the factored diagonal block in LDS).  Covers the
// rows below the block inside A22 and all rows of A12, so L21 is produced along the way.
// FULL = every listed front has a full 64-wide panel at k0 (unconditional, unrolled stores);
// otherwise only fronts whose last panel is narrower are handled (stores by a short loop:
// unrolled conditional stores make the compiler spill the whole row).
template <bool FULL>
__global__ __launch_bounds__(256) void k_trsm_panel(const int* __restrict__ ids, const FrontDev* __restrict__ fr,
                                                    double* __restrict__ A, int k0) {
  const FrontDev F = fr[ids[blockIdx.y]];
  if (k0 >= F.n) return;
  const int kb = min(NB, F.n - k0);
  if (FULL != (kb == NB)) return;
  __shared__ double Ld[NB * (NB + 1)];
  __shared__ double Dinv[NB];
  const double* dblk = A + F.off + (i64)k0 * F.N + k0;
  const int t = threadIdx.x;
  for (int e = t; e < NB * NB; e += 256) {
    int r = e & 63, c = e >> 6;
    double v = (r < kb && c < kb && r >= c) ? dblk[(i64)c * F.N + r] : 0.0;
    if (r == c && r >= kb) v = 1.0;  // identity padding keeps the unrolled solve branch-free
    Ld[c * (NB + 1) + r] = v;
  }
  __syncthreads();
  if (t < NB) Dinv[t] = 1.0 / Ld[t * (NB + 1) + t];
  __syncthreads();
  const int row = k0 + kb + blockIdx.x * 256 + t;
  if (row >= F.N) return;
  // uniform column base + 32-bit per-thread offset
  double* p = A + F.off + (i64)k0 * F.N;
  const unsigned ld = F.N;
  // two 32-column halves: solve the first, eliminate it from the second, solve the second
  double x[32], y[32];
#pragma unroll
  for (int c = 0; c < 32; c++) x[c] = (FULL || c < kb) ? p[c * ld + row] : 0.0;
#pragma unroll
  for (int j = 0; j < 32; j++) {
    double v = x[j];
#pragma unroll
    for (int i = 0; i < j; i++) v -= x[i] * Ld[i * (NB + 1) + j];
    x[j] = v * Dinv[j];
  }
#pragma unroll
  for (int c = 0; c < 32; c++) y[c] = (FULL || c + 32 < kb) ? p[(c + 32) * ld + row] : 0.0;
#pragma unroll
  for (int i = 0; i < 32; i++) {
    const double xi = x[i];
#pragma unroll
    for (int j = 0; j < 32; j++) y[j] -= xi * Ld[i * (NB + 1) + 32 + j];
  }
#pragma unroll
  for (int j = 0; j < 32; j++) {
    double v = y[j];
#pragma unroll
    for (int i = 0; i < j; i++) v -= y[i] * Ld[(32 + i) * (NB + 1) + 32 + j];
    y[j] = v * Dinv[32 + j];
  }
  if (FULL) {
#pragma unroll
    for (int c = 0; c < 32; c++) p[c * ld + row] = x[c];
#pragma unroll
    for (int c = 0; c < 32; c++) p[(c + 32) * ld + row] = y[c];
  } else {
    for (int c = 0; c < kb && c < 32; c++) p[c * ld + row] = x[c];
    for (int c = 32; c < kb; c++) p[c * ld + row] = y[c - 32];
  }
}
template __global__ void k_trsm_panel<true>(const int*, const FrontDev*, double*, int);
template __global__ void k_trsm_panel<false>(const int*, const FrontDev*, double*, int);

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

// C -= P Q^T on 64x64 tiles with v_mfma_f64_16x16x4 (4 waves, 2x2 of 16x16 blocks each).
// MODE 0: trailing update of the pivot columns after panel k0: P = F[rows, k0:k0+kb],
//         Q = F[cols, k0:k0+kb], C = F[s:N, s:n] lower, s = k0+kb.
// MODE 1: Schur complement: P = Q = L21 = F[n:N, 0:n], K = n; C = F[n:N, n:N] lower gets
//         A11 + S, except on the diagonal separator blocks, which get S alone.
// The MFMA A operand is Q and B is P.  v_mfma_f64_16x16x4 leaves D[(lane/16) + 4r][lane%16]
// in accumulator r of a lane, so the 16 lanes of a group cover 16 consecutive rows of C
// (coalesced stores) and accumulator r is column (lane/16) + 4r of the 16-wide block.
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

  __shared__ double As[16][64];
  __shared__ double Bs[16][64];
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  const int wr = w >> 1, wc = w & 1;
  const double* base = A + F.off;
  d4 acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; i++)
#pragma unroll
    for (int j = 0; j < 2; j++) acc[i][j] = (d4){0.0, 0.0, 0.0, 0.0};

  // each thread stages 4 elements of each operand per 16-wide k chunk
  const int li = t & 63, lk = t >> 6;  // row within tile, k offset (0..3), +4q
  const bool prow = row0 + li < rend, qrow = col0 + li < cend;
  double pa[4], qb[4];
  auto fetch = [&](int kc) {
#pragma unroll
    for (int q = 0; q < 4; q++) {
      int k = kc + lk + 4 * q;
      bool kin = k < K;
      i64 colk = (i64)(kbase + k) * F.N;
      pa[q] = (kin && prow) ? base[colk + row0 + li] : 0.0;
      qb[q] = (kin && qrow) ? base[colk + col0 + li] : 0.0;
    }
  };
  fetch(0);
  for (int kc = 0; kc < K; kc += 16) {
    __syncthreads();
#pragma unroll
    for (int q = 0; q < 4; q++) {
      As[lk + 4 * q][li] = pa[q];
      Bs[lk + 4 * q][li] = qb[q];
    }
    __syncthreads();
    if (kc + 16 < K) fetch(kc + 16);
#pragma unroll
    for (int k4 = 0; k4 < 4; k4++) {
      const int kr = k4 * 4 + (lane >> 4);
#pragma unroll
      for (int bi = 0; bi < 2; bi++) {
        const double b = As[kr][wr * 32 + bi * 16 + (lane & 15)];
#pragma unroll
        for (int bj = 0; bj < 2; bj++) {
          const double a = Bs[kr][wc * 32 + bj * 16 + (lane & 15)];
          acc[bi][bj] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc[bi][bj], 0, 0, 0);
        }
      }
    }
  }
  double* C = A + F.off;
#pragma unroll
  for (int bi = 0; bi < 2; bi++)
#pragma unroll
    for (int bj = 0; bj < 2; bj++)
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int i = row0 + wr * 32 + bi * 16 + (lane & 15);
        const int j = col0 + wc * 32 + bj * 16 + (lane >> 4) + 4 * r;
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
// Work vector W is n_dof x R column-major (ld = ldw); per-front scratch X is N x R (ld = N).
#define RB 16  // right-hand sides per sweep pass

// Forward sweep of one front: X[0:n] <- L11^{-1} W[int], X[n:N] <- -L21 X[0:n];
// writes X[0:n] back to W[int] (the reference's A22^{-1} b_rd is finished by the backward sweep).
__global__ __launch_bounds__(256) void k_solve_fwd(const int* __restrict__ ids, const FrontDev* __restrict__ fr,
                                                   const double* __restrict__ A, const int* __restrict__ dofs,
                                                   double* __restrict__ W, int ldw, double* __restrict__ xs, int nr) {
  const FrontDev F = fr[ids[blockIdx.x]];
  const double* L = A + F.off;
  double* X = xs + F.x_off;
  const int* di = dofs + F.int_off;
  const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
  for (int e = t; e < F.N * nr; e += 256) {
    int i = e % F.N, r = e / F.N;
    X[(i64)r * F.N + i] = i < F.n ? W[(i64)r * ldw + di[i]] : 0.0;
  }
  __shared__ double Ld[NB * (NB + 1)];
  __shared__ double Xd[NB][RB];
  for (int jb = 0; jb < F.n; jb += NB) {
    const int kb = min(NB, F.n - jb);
    __syncthreads();
    for (int e = t; e < NB * NB; e += 256) {
      int r = e & 63, c = e >> 6;
      double v = (r < kb && c < kb && r >= c) ? L[(i64)(jb + c) * F.N + jb + r] : 0.0;
      if (r == c && r >= kb) v = 1.0;
      Ld[c * (NB + 1) + r] = v;
    }
    for (int e = t; e < NB * RB; e += 256) {
      int i = e & 63, r = e >> 6;
      Xd[i][r] = (i < kb && r < nr) ? X[(i64)r * F.N + jb + i] : 0.0;
    }
    __syncthreads();
    if (wv == 0) {
      // lane = (rhs r, quarter q): rows q, q+4, ..., q+60
      const int r = lane & 15, q = lane >> 4;
      double x[16];
#pragma unroll
      for (int s = 0; s < 16; s++) x[s] = Xd[4 * s + q][r];
#pragma unroll
      for (int j = 0; j < NB; j++) {
        const int sj = j >> 2, qj = j & 3;
        if (q == qj) x[sj] /= Ld[j * (NB + 1) + j];
        const double xj = __shfl(x[sj], qj * 16 + r, 64);
#pragma unroll
        for (int s = 0; s < 16; s++)
          if (4 * s + q > j) x[s] -= Ld[j * (NB + 1) + 4 * s + q] * xj;
      }
#pragma unroll
      for (int s = 0; s < 16; s++) Xd[4 * s + q][r] = x[s];
    }
    __syncthreads();
    for (int e = t; e < kb * nr; e += 256) {
      int i = e % kb, r = e / kb;
      X[(i64)r * F.N + jb + i] = Xd[i][r];
    }
    // rows below: X[i] -= L[i, jb:jb+kb] Xd
    for (int i = jb + kb + t; i < F.N; i += 256) {
      double acc[RB];
#pragma unroll
      for (int r = 0; r < RB; r++) acc[r] = 0.0;
      for (int c = 0; c < kb; c++) {
        const double l = L[(i64)(jb + c) * F.N + i];
#pragma unroll
        for (int r = 0; r < RB; r++) acc[r] += l * Xd[c][r];
      }
      for (int r = 0; r < nr; r++) X[(i64)r * F.N + i] -= acc[r];
    }
  }
  __syncthreads();
  for (int e = t; e < F.n * nr; e += 256) {
    int i = e % F.n, r = e / F.n;
    W[(i64)r * ldw + di[i]] = X[(i64)r * F.N + i];
  }
}

// Separator accumulation of per-box contributions X[n+pos : n+pos+k] of both neighbour boxes
// into the work vector, box 0 first: W = (W + X_b0) + X_b1.
__global__ __launch_bounds__(256) void k_sep_acc(const int* __restrict__ ids, const SepDev* __restrict__ seps,
                                                 const FrontDev* __restrict__ fr, const int* __restrict__ dofs,
                                                 double* __restrict__ W, int ldw, const double* __restrict__ xs,
                                                 int nr) {
  const SepDev S = seps[ids[blockIdx.x]];
  const FrontDev B0 = fr[S.box[0]], B1 = fr[S.box[1]];
  const int* d = dofs + S.dof_off;
  for (int e = threadIdx.x; e < S.k * nr; e += 256) {
    int i = e % S.k, r = e / S.k;
    double* w = W + (i64)r * ldw + d[i];
    double v = *w + xs[B0.x_off + (i64)r * B0.N + B0.n + S.pos[0] + i];
    *w = v + xs[B1.x_off + (i64)r * B1.N + B1.n + S.pos[1] + i];
  }
}

// Backward sweep of one front: X[0:n] = W[int] (forward result y), X[n:N] = W[bd];
// X[0:n] <- L11^{-T} (y - L21^T X[n:N]) = A22^{-1} b_rd - XR b_sk; writes X[0:n] to W[int].
__global__ __launch_bounds__(256) void k_solve_bwd(const int* __restrict__ ids, const FrontDev* __restrict__ fr,
                                                   const double* __restrict__ A, const int* __restrict__ dofs,
                                                   double* __restrict__ W, int ldw, double* __restrict__ xs, int nr) {
  const FrontDev F = fr[ids[blockIdx.x]];
  const double* L = A + F.off;
  double* X = xs + F.x_off;
  const int* di = dofs + F.int_off;
  const int* db = dofs + F.bd_off;
  const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
  for (int e = t; e < F.N * nr; e += 256) {
    int i = e % F.N, r = e / F.N;
    X[(i64)r * F.N + i] = W[(i64)r * ldw + (i < F.n ? di[i] : db[i - F.n])];
  }
  __shared__ double Lt[NB * (NB + 1)];
  __shared__ double Xt[NB][RB];
  __shared__ double Xd[NB][RB];
  const int nblk = (F.n + NB - 1) / NB;
  for (int bk = nblk - 1; bk >= 0; bk--) {
    const int jb = bk * NB, kb = min(NB, F.n - jb);
    // T = sum over rows below the block of L[i, jb+c] X[i, r]; thread -> (c, 4 rhs)
    const int c = t & 63, rg = (t >> 6) * 4;
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    for (int i0 = jb + kb; i0 < F.N; i0 += NB) {
      const int h = min(NB, F.N - i0);
      __syncthreads();
      for (int e = t; e < NB * NB; e += 256) {
        int ii = e & 63, cc = e >> 6;
        Lt[cc * (NB + 1) + ii] = (ii < h && cc < kb) ? L[(i64)(jb + cc) * F.N + i0 + ii] : 0.0;
      }
      for (int e = t; e < NB * RB; e += 256) {
        int ii = e & 63, r = e >> 6;
        Xt[ii][r] = (ii < h && r < nr) ? X[(i64)r * F.N + i0 + ii] : 0.0;
      }
      __syncthreads();
      for (int ii = 0; ii < NB; ii++) {
        const double l = Lt[c * (NB + 1) + ii];
#pragma unroll
        for (int q = 0; q < 4; q++) acc[q] += l * Xt[ii][rg + q];
      }
    }
    __syncthreads();
    for (int e = t; e < NB * NB; e += 256) {
      int r = e & 63, cc = e >> 6;
      double v = (r < kb && cc < kb && r >= cc) ? L[(i64)(jb + cc) * F.N + jb + r] : 0.0;
      if (r == cc && r >= kb) v = 1.0;
      Lt[cc * (NB + 1) + r] = v;
    }
#pragma unroll
    for (int q = 0; q < 4; q++) {
      const int r = rg + q;
      Xd[c][r] = (c < kb && r < nr) ? X[(i64)r * F.N + jb + c] - acc[q] : 0.0;
    }
    __syncthreads();
    if (wv == 0) {
      const int r = lane & 15, q = lane >> 4;
      double x[16];
#pragma unroll
      for (int s = 0; s < 16; s++) x[s] = Xd[4 * s + q][r];
#pragma unroll
      for (int j = NB - 1; j >= 0; j--) {
        const int sj = j >> 2, qj = j & 3;
        if (q == qj) x[sj] /= Lt[j * (NB + 1) + j];
        const double xj = __shfl(x[sj], qj * 16 + r, 64);
#pragma unroll
        for (int s = 0; s < 16; s++)
          if (4 * s + q < j) x[s] -= Lt[(4 * s + q) * (NB + 1) + j] * xj;
      }
#pragma unroll
      for (int s = 0; s < 16; s++) Xd[4 * s + q][r] = x[s];
    }
    __syncthreads();
    for (int e = t; e < kb * nr; e += 256) {
      int i = e % kb, r = e / kb;
      X[(i64)r * F.N + jb + i] = Xd[i][r];
    }
  }
  __syncthreads();
  for (int e = t; e < F.n * nr; e += 256) {
    int i = e % F.n, r = e / F.n;
    W[(i64)r * ldw + di[i]] = X[(i64)r * F.N + i];
  }
}

// ----------------------------------------------------------------------------- apply
// Upward step of Apply for one front: u = [L11; L21]^T [x_rd; x_sk] (= L11^T b_rd + W b_sk),
// b_rd <- L11 u.  u is kept in X[0:n] for the downward step.
__global__ __launch_bounds__(256) void k_apply_up(const int* __restrict__ ids, const FrontDev* __restrict__ fr,
                                                  const double* __restrict__ A, const int* __restrict__ dofs,
                                                  double* __restrict__ W, int ldw, double* __restrict__ xs, int nr) {
  const FrontDev F = fr[ids[blockIdx.x]];
  const double* L = A + F.off;
  double* X = xs + F.x_off;
  const int* di = dofs + F.int_off;
  const int* db = dofs + F.bd_off;
  const int t = threadIdx.x;
  for (int e = t; e < F.N * nr; e += 256) {
    int i = e % F.N, r = e / F.N;
    X[(i64)r * F.N + i] = W[(i64)r * ldw + (i < F.n ? di[i] : db[i - F.n])];
  }
  __syncthreads();
  // u_c = sum_{i >= c} L[i][c] X[i]: one column per wave-lane group via a strided reduction
  __shared__ double red[256];
  for (int c = 0; c < F.n; c++) {
    for (int r = 0; r < nr; r++) {
      double s = 0.0;
      for (int i = c + t; i < F.N; i += 256) s += L[(i64)c * F.N + i] * X[(i64)r * F.N + i];
      red[t] = s;
      __syncthreads();
      for (int h = 128; h > 0; h >>= 1) {
        if (t < h) red[t] += red[t + h];
        __syncthreads();
      }
      if (t == 0) X[(i64)r * F.N + c] = red[0];  // row c is no longer read by later columns
      __syncthreads();
    }
  }
  // b_rd = L11 u
  for (int i = t; i < F.n; i += 256)
    for (int r = 0; r < nr; r++) {
      double s = 0.0;
      for (int c = 0; c <= i; c++) s += L[(i64)c * F.N + i] * X[(i64)r * F.N + c];
      W[(i64)r * ldw + di[i]] = s;
    }
}

// Downward step of Apply: X[n:N] = L21 u (the reference's XL b_rd), accumulated by k_sep_acc.
__global__ __launch_bounds__(256) void k_apply_down(const int* __restrict__ ids, const FrontDev* __restrict__ fr,
                                                    const double* __restrict__ A, double* __restrict__ xs, int nr) {
  const FrontDev F = fr[ids[blockIdx.x]];
  const double* L = A + F.off;
  double* X = xs + F.x_off;
  for (int i = F.n + threadIdx.x; i < F.N; i += 256)
    for (int r = 0; r < nr; r++) {
      double s = 0.0;
      for (int c = 0; c < F.n; c++) s += L[(i64)c * F.N + i] * X[(i64)r * F.N + c];
      X[(i64)r * F.N + i] = s;
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
