// kernels.hip -- hand-written gfx950 kernels of the multifrontal factor / update / solve.
//
// Every kernel works on a batch: a list of front (or separator) ids of one tree level, one
// front per blockIdx.y (or blockIdx.x), so a whole level runs in one launch per stage.
//   assembly   k_zero_lower, k_assemble, k_sep_add          (HBM-bound gathers / scatters)
//   factor     k_chol_diag, k_chol_dpanel, k_chol_panel, k_gemm_group, k_gemm_trail,
//              k_gemm_schur (blocked Cholesky of the pivot columns, left- or right-looking,
//              then the Schur update; MFMA)
//   LU fronts  k_zero_full, k_lu_panel, k_lu_swap, k_lu_trsm, k_lu_gemm, k_lu_perm,
//              k_sep_add_full, k_lu_bwd, k_lu_apply_up (unsymmetric operators)
//   solve      k_solve_fwd(_r), k_solve_bwd, k_fwd_blk, k_bwd_blk, k_sep_acc (multi-RHS sweeps)
//   apply      k_apply_up, k_apply_down
//   residual   k_spmv, k_spmv_dot, k_pc_in/out, k_cg_*, k_red_final, k_rowabs (FP64 refinement)
//   operator   k_box_scale, k_lap_rows, k_lap_collect (coefficient-change requests)
#include <hip/hip_runtime.h>

#include "kernels.h"

// accumulator-row layout of v_mfma_f32_16x16x4 on gfx950: accumulator r of lane l holds row
// 4 (l/16) + r (v_mfma_f64_16x16x4 holds row l/16 + 4r)
#define MROW_F32(q, r) (4 * (q) + (r))

namespace umf {

// Every factor / solve kernel is a template over the storage type T of the fronts: double (the
// FP64 form) or float (the FP32-MFMA form, a preconditioner for FP64 iterative refinement).
// The MFMA is v_mfma_f64_16x16x4 or v_mfma_f32_16x16x4 with the same operand layout: lane l
// supplies A[l%16][l/16] and B[l/16][l%16]; accumulator r of lane l holds D[mrow(l/16, r)][l%16],
// where the two instructions differ in mrow.
template <class T> struct Vec4;
template <> struct Vec4<double> { typedef double type __attribute__((ext_vector_type(4))); };
template <> struct Vec4<float> { typedef float type __attribute__((ext_vector_type(4))); };
template <class T> using V4 = typename Vec4<T>::type;
template <class T> __device__ __forceinline__ V4<T> vzero() {
  V4<T> v;
  v[0] = v[1] = v[2] = v[3] = T(0);
  return v;
}
__device__ __forceinline__ V4<double> mfma4(double a, double b, V4<double> c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ V4<float> mfma4(float a, float b, V4<float> c) {
  return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}
// row within the 16 x 16 result of accumulator r of a lane in group q = lane / 16
template <class T> __device__ __forceinline__ int mrow(int q, int r);
template <> __device__ __forceinline__ int mrow<double>(int q, int r) { return q + 4 * r; }
template <> __device__ __forceinline__ int mrow<float>(int q, int r) { return MROW_F32(q, r); }

#define NB 64  // panel width of the blocked factorization and of the solve blocks
#define SPIN_LIMIT (1 << 22)  // polls of a bounded wait between workgroups of one launch

// ----------------------------------------------------------------------------- assembly
// Zero n ints (the factorization's error flags and arrival counters at its start: a kernel
// node, so that a captured factorization holds no memset node).
__global__ void k_zero_ints(int* __restrict__ p, int n) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = 0;
}

// Zero the lower triangle (row >= col) of each listed matrix: one wave per column (4 columns per
// workgroup step), so short columns do not leave most of a 256-thread workgroup idle.
template <class T>
__global__ __launch_bounds__(256) void k_zero_lower(const int* __restrict__ ids, const MatDev* __restrict__ mats,
                                                    T* __restrict__ A) {
  const MatDev M = mats[ids[blockIdx.y]];
  T* a = A + M.off;
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  for (int c = 4 * blockIdx.x + w; c < M.cols; c += 4 * gridDim.x) {
    T* col = a + (i64)c * M.ld;
    for (int r = c + lane; r < M.ld; r += 64) col[r] = T(0);
  }
}

// Leaf fronts whose below-diagonal panel tiles take their entries in the panel step (FrontDev::
// fly_off): zero only what is read before it is written -- the lower triangles of the 64-column
// blocks' diagonal blocks and of the trailing block.
// Work unit u < nb = ceil(n/64): diagonal block u, 64 x 64 per workgroup; u >= nb: trailing
// columns n + 4(u - nb) .. n + 4(u - nb) + 3, one wave each.
template <class T>
__global__ __launch_bounds__(256) void k_zero_fly(const int* __restrict__ ids, const FrontDev* __restrict__ fr,
                                                  T* __restrict__ A) {
  const FrontDev F = fr[ids[blockIdx.y]];
  T* a = A + F.off;
  const int nb = (F.n + NB - 1) / NB;
  for (int u = blockIdx.x; u < nb + (F.m + 3) / 4; u += gridDim.x) {
    if (u < nb) {
      const int c0 = NB * u, kb = min(NB, F.n - c0);
      T* blk = a + (i64)c0 * F.N + c0;
#pragma unroll 4
      for (int e = threadIdx.x; e < NB * NB; e += 256) {
        const int i = e & (NB - 1), j = e / NB;
        if (i >= j && i < kb && j < kb) blk[(i64)j * F.N + i] = T(0);
      }
    } else {  // trailing columns, one wave each
      for (int c = F.n + 4 * (u - nb) + (threadIdx.x >> 6); c < min(F.N, F.n + 4 * (u - nb + 1)); c += 4) {
        T* col = a + (i64)c * F.N;
        for (int r = c + (threadIdx.x & 63); r < F.N; r += 64) col[r] = T(0);
      }
    }
  }
}

// CSR scatter triples and 64x64 tiles of block copies into each listed matrix.  All writes of
// one matrix are to distinct entries (the planner emits disjoint programs), so no ordering or
// atomics are needed; the matrix must have been zeroed by k_zero_lower.
// Block sources of level lev32 or deeper are read from the FP32 pool A32 (the FP64 top levels of
// the mixed form assembling from the FP32 levels below them); lev32 = INT_MAX otherwise.
template <class T>
__global__ __launch_bounds__(256) void k_assemble(const int* __restrict__ ids, const MatDev* __restrict__ mats,
                                                  T* __restrict__ A, const i64* __restrict__ tri_off,
                                                  const Triple* __restrict__ tri, const i64* __restrict__ tile_off,
                                                  const Tile* __restrict__ tiles, const BlockOp* __restrict__ ops,
                                                  const double* __restrict__ csr_val, const float* __restrict__ A32,
                                                  int lev32) {
  const int mid = ids[blockIdx.y];
  const MatDev M = mats[mid];
  T* a = A + M.off;
  const i64 t0 = tri_off[mid], t1 = tri_off[mid + 1];
  for (i64 i = t0 + (i64)blockIdx.x * 256 + threadIdx.x; i < t1; i += (i64)gridDim.x * 256) {
    const Triple t = tri[i];
    a[(i64)t.col * M.ld + t.row] = csr_val[t.k];
  }
  const i64 b0 = tile_off[mid], b1 = tile_off[mid + 1];
  for (i64 ti = b0 + blockIdx.x; ti < b1; ti += gridDim.x) {
    const Tile tl = tiles[ti];
    const BlockOp op = ops[tl.op];
    const MatDev S = mats[op.src];
    const bool src32 = S.lev >= lev32;
    const T* s = A + S.off;
    const float* s32 = src32 ? A32 + S.off : nullptr;
    const bool trans = op.flags & 1, lower = op.flags & 2;
    for (int e = threadIdx.x; e < 64 * 64; e += 256) {
      int r = tl.r0 + (e & 63), c = tl.c0 + (e >> 6);
      if (r >= op.nr || c >= op.nc) continue;
      if (lower && r < c) continue;
      const i64 si = trans ? (i64)(op.sc0 + r) * S.ld + (op.sr0 + c) : (i64)(op.sc0 + c) * S.ld + (op.sr0 + r);
      const T v = src32 ? (T)s32[si] : s[si];
      a[(i64)(op.dc0 + c) * M.ld + (op.dr0 + r)] = v;
    }
  }
}

// Separator Schur accumulation: A11 = (A11 + S_box0) + S_box1 on the lower triangle, the two
// neighbour boxes in ascending index order as the reference's ScatterAdd_ calls
// (src/DE/mf2d.cpp:599-606).  S is read from each box's trailing diagonal separator block.
template <class T>
__global__ __launch_bounds__(256) void k_sep_add(const int* __restrict__ ids, const SepDev* __restrict__ seps,
                                                 const FrontDev* __restrict__ fr, T* __restrict__ A) {
  const SepDev S = seps[ids[blockIdx.y]];
  const FrontDev B0 = fr[S.box[0]], B1 = fr[S.box[1]];
  T* a = A + S.off;
  const T* s0 = A + B0.off + (i64)(B0.n + S.pos[0]) * B0.N + B0.n + S.pos[0];
  const T* s1 = A + B1.off + (i64)(B1.n + S.pos[1]) * B1.N + B1.n + S.pos[1];
  // one wave per column, 4 columns per workgroup step
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  for (int c = 4 * blockIdx.x + w; c < S.k; c += 4 * gridDim.x)
    for (int r = c + lane; r < S.k; r += 64) {
      T v = a[(i64)c * S.k + r] + s0[(i64)c * B0.N + r];
      a[(i64)c * S.k + r] = v + s1[(i64)c * B1.N + r];
    }
}

// ----------------------------------------------------------------------------- MFMA tile GEMMs
// Orders a wave's LDS accesses before its later ones, as seen by the other lanes of the same
// wave (a wave's LDS operations complete in order; the fences keep the compiler from moving
// accesses across this point).
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Cross-lane moves of a float or double (as 32-bit halves): lane_read takes lane l's value (l
// wave-uniform, readlane), lane_pull takes lane l's value for a per-lane l (ds_bpermute),
// quad_bcast takes the value of lane q of the lane's own group of 4 (DPP quad permute).
template <class T>
__device__ __forceinline__ T lane_read(T v, int l) {
  if constexpr (sizeof(T) == 4) {
    return __builtin_bit_cast(T, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), l));
  } else {
    const unsigned long long b = __builtin_bit_cast(unsigned long long, v);
    const unsigned lo = __builtin_amdgcn_readlane((int)(unsigned)b, l);
    const unsigned hi = __builtin_amdgcn_readlane((int)(unsigned)(b >> 32), l);
    return __builtin_bit_cast(T, ((unsigned long long)hi << 32) | lo);
  }
}
template <class T>
__device__ __forceinline__ T lane_pull(T v, int l) {
  if constexpr (sizeof(T) == 4) {
    return __builtin_bit_cast(T, __builtin_amdgcn_ds_bpermute(4 * l, __builtin_bit_cast(int, v)));
  } else {
    const unsigned long long b = __builtin_bit_cast(unsigned long long, v);
    const unsigned lo = __builtin_amdgcn_ds_bpermute(4 * l, (int)(unsigned)b);
    const unsigned hi = __builtin_amdgcn_ds_bpermute(4 * l, (int)(unsigned)(b >> 32));
    return __builtin_bit_cast(T, ((unsigned long long)hi << 32) | lo);
  }
}
template <int Q>
__device__ __forceinline__ int quad_bcast_i(int v) {
  return __builtin_amdgcn_mov_dpp(v, Q * 0x55, 0xF, 0xF, false);
}
template <class T>
__device__ __forceinline__ T quad_bcast(T v, int q) {  // q: a constant after unrolling
  auto one = [&](int x) {
    switch (q) {
      case 0: return quad_bcast_i<0>(x);
      case 1: return quad_bcast_i<1>(x);
      case 2: return quad_bcast_i<2>(x);
      default: return quad_bcast_i<3>(x);
    }
  };
  if constexpr (sizeof(T) == 4) {
    return __builtin_bit_cast(T, one(__builtin_bit_cast(int, v)));
  } else {
    const unsigned long long b = __builtin_bit_cast(unsigned long long, v);
    const unsigned lo = one((int)(unsigned)b), hi = one((int)(unsigned)(b >> 32));
    return __builtin_bit_cast(T, ((unsigned long long)hi << 32) | lo);
  }
}

// the wave index as a scalar, so that branches on it are uniform
__device__ __forceinline__ int wave_id() { return __builtin_amdgcn_readfirstlane(threadIdx.x >> 6); }

// Operand loads of the tile GEMMs are buffer loads: one resource per K chunk covers the chunk's
// valid columns (base = the chunk's first column, size = its valid columns x ld), so columns at
// or beyond K read as 0 without a branch, and each lane's byte offset within the chunk is fixed
// for the whole loop.  Rows past the operand's end are clamped to its last row: their products
// land only in accumulators outside the tile, which the epilogues skip.
__device__ __forceinline__ const void* wave_uniform(const void* p) {
  const unsigned long long v = (unsigned long long)p;
  const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)v);
  const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(v >> 32));
  return (const void*)(((unsigned long long)hi << 32) | lo);
}
// buffer resource over `bytes` bytes at p (gfx9 raw buffer: stride 0, dword 3 = 0x00020000)
__device__ __forceinline__ __amdgpu_buffer_rsrc_t chunk_rsrc(const void* p, int bytes) {
  return __builtin_amdgcn_make_buffer_rsrc((void*)wave_uniform(p), 0, __builtin_amdgcn_readfirstlane(bytes),
                                           0x00020000);
}
template <class T>
__device__ __forceinline__ T buf_load(__amdgpu_buffer_rsrc_t r, int off);
template <>
__device__ __forceinline__ float buf_load<float>(__amdgpu_buffer_rsrc_t r, int off) {
  return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, off, 0, 0));
}
template <>
__device__ __forceinline__ double buf_load<double>(__amdgpu_buffer_rsrc_t r, int off) {
  return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, 0));
}

// acc += P Q^T for one TS x TS tile (TS = 64 or 128), P(i,k) = P[k*ldp + i], Q(j,k) = Q[k*ldq + j],
// k in [0, K), on 16x16x4 MFMAs: 4 waves as 2 x 2, each (TS/2)^2 = (TS/32)^2 blocks of 16 x 16.
// The MFMA A operand is Q and B is P, so accumulator r of a lane holds C(i, j) with i = 16-block
// row + lane%16 and j = 16-block col + mrow(lane/16, r), and the 16 lanes of a group cover 16
// consecutive rows of C.  Per 4-deep k step a wave reads TS/32 + TS/32 operand fragments from
// LDS for (TS/32)^2 MFMAs; the next step's fragments are read before this step's MFMAs issue, so
// their LDS latency hides behind them.  Two LDS stages: chunk c is computed from stage c&1 while chunk
// c+1 (loaded into registers one chunk earlier) is written to the other stage and chunk c+2 is
// fetched, one barrier per 16-deep chunk.  P and Q start at the tile's first row; pn, qn >= 1.
// FP32 stage rows are padded by 16 floats: the fragment reads of one half-wave take rows kr and
// kr + 1 (lane groups q = 0, 1) at the same columns, which then fall on the other 16 of the 32
// LDS banks instead of the same ones (FP64 reads go 16 lanes at a time, one row: no conflict).
template <class T>
constexpr int kStagePad = sizeof(T) == 4 ? 16 : 0;
// column stride of the 64-column X blocks of the panel kernels, read as X[k * stride + row] with
// k = 4 k4 + lane / 16: FP32 strides of 16 mod 32 banks keep a half-wave's two k rows apart
template <class T>
constexpr int kXs = sizeof(T) == 4 ? NB + 16 : NB + 1;
template <class T, int TS>
struct TileSmemT {
  T As[2][16][TS + kStagePad<T>];
  T Bs[2][16][TS + kStagePad<T>];
};

// LOWER (P = Q, only the lower triangle of the product wanted): the wave above the diagonal
// (wr < wc) only stages operands.
template <class T, int TS, bool LOWER = false, int DEPTH = 1, bool PRIO = false>
__device__ __forceinline__ void tile_gemm_nt_t(V4<T> (&acc)[TS / 32][TS / 32], const T* __restrict__ P, int ldp,
                                               int pn, const T* __restrict__ Q, int ldq, int qn, int K,
                                               TileSmemT<T, TS>& sm) {
  static_assert(DEPTH == 1 || DEPTH == 2, "one or two chunks fetched ahead");
  constexpr int NL = TS / 64, NBK = TS / 32;
  const int t = threadIdx.x, lane = t & 63;
  const int w = __builtin_amdgcn_readfirstlane(t >> 6);
  const int wr = w >> 1, wc = w & 1;
  const int li = lane, lk = w;
  T pa[DEPTH][4][NL], qb[DEPTH][4][NL];  // DEPTH register sets of fetched operands
  int po[4][NL], qo[4][NL];  // byte offsets of this lane's elements within a chunk
#pragma unroll
  for (int q = 0; q < 4; q++)
#pragma unroll
    for (int u = 0; u < NL; u++) {
      const int row = li + 64 * u;
      po[q][u] = ((lk + 4 * q) * ldp + min(row, pn - 1)) * (int)sizeof(T);
      qo[q][u] = ((lk + 4 * q) * ldq + min(row, qn - 1)) * (int)sizeof(T);
    }
  auto fetch = [&](int kc, auto set) {
    constexpr int S = decltype(set)::value;
    const int nk = max(min(16, K - kc), 0);
    const __amdgpu_buffer_rsrc_t rp = chunk_rsrc(P + (i64)kc * ldp, nk * ldp * (int)sizeof(T));
    const __amdgpu_buffer_rsrc_t rq = chunk_rsrc(Q + (i64)kc * ldq, nk * ldq * (int)sizeof(T));
#pragma unroll
    for (int q = 0; q < 4; q++)
#pragma unroll
      for (int u = 0; u < NL; u++) {
        pa[S][q][u] = buf_load<T>(rp, po[q][u]);
        qb[S][q][u] = buf_load<T>(rq, qo[q][u]);
      }
  };
  auto stage = [&](int b, auto set) {
    constexpr int S = decltype(set)::value;
#pragma unroll
    for (int q = 0; q < 4; q++)
#pragma unroll
      for (int u = 0; u < NL; u++) {
        sm.As[b][lk + 4 * q][li + 64 * u] = pa[S][q][u];
        sm.Bs[b][lk + 4 * q][li + 64 * u] = qb[S][q][u];
      }
  };
  if (K <= 0) return;
  using S0 = std::integral_constant<int, 0>;
  using S1 = std::integral_constant<int, DEPTH - 1>;
  const bool mma = !LOWER || wr >= wc;
  // chunk kc from LDS stage b; then the chunk after it (register set `next`) is staged and the
  // chunk DEPTH + 1 ahead is fetched into that set
  auto step = [&](int kc, int b, auto next) {
    T av[2][NBK], bv[2][NBK];
    auto rd = [&](int k4, int z) {
      const int kr = k4 * 4 + (lane >> 4);
#pragma unroll
      for (int x = 0; x < NBK; x++) {
        bv[z][x] = sm.As[b][kr][wr * (TS / 2) + x * 16 + (lane & 15)];
        av[z][x] = sm.Bs[b][kr][wc * (TS / 2) + x * 16 + (lane & 15)];
      }
    };
    if (mma) rd(0, 0);
    // PRIO (the panels' 64-tile GEMMs, four tiles per CU): issue priority to the wave in its MFMA
    // phase (FP64 3D 127^3 panels 306 -> 296 ms per step; the stage GEMMs' tiles and the diagonal
    // blocks' SYRK measured slower with it)
    if constexpr (PRIO) __builtin_amdgcn_s_setprio(1);
#pragma unroll
    for (int k4 = 0; k4 < 4 && mma; k4++) {
      if (k4 < 3) rd(k4 + 1, (k4 + 1) & 1);
      __builtin_amdgcn_sched_barrier(0);  // the next fragments' reads stay ahead of these MFMAs
#pragma unroll
      for (int bi = 0; bi < NBK; bi++)
#pragma unroll
        for (int bj = 0; bj < NBK; bj++) acc[bi][bj] = mfma4(av[k4 & 1][bj], bv[k4 & 1][bi], acc[bi][bj]);
    }
    if constexpr (PRIO) __builtin_amdgcn_s_setprio(0);
    stage(b ^ 1, next);  // past the last chunk: a harmless write of zeros to the unused stage
    fetch(kc + 16 * (DEPTH + 1), next);
    __syncthreads();
  };
  __syncthreads();  // the stages may still be read by an earlier use of the shared memory
  fetch(0, S0{});
  if constexpr (DEPTH == 2) fetch(16, S1{});
  stage(0, S0{});
  fetch(16 * DEPTH, S0{});
  __syncthreads();
  if constexpr (DEPTH == 1) {
    int b = 0;
    for (int kc = 0; kc < K; kc += 16, b ^= 1) step(kc, b, S0{});
  } else {
    for (int kc = 0; kc < K; kc += 32) {
      step(kc, 0, S1{});  // the next chunk sits in set 1
      if (kc + 16 < K) step(kc + 16, 1, S0{});
    }
  }
}

// tile_gemm_nt_t for a 128 x 128 tile on 8 waves (512 threads): waves as 2 (rows) x 4 (columns),
// wave w holding rows 64 (w / 4) .. + 64 and columns 32 (w % 4) .. + 32 (acc[bi][bj]: 16-row
// block bi, 16-column block bj).  The same stages, chunks, operand loads and fragment prefetch;
// per 4-deep step a wave reads 4 P and 2 Q fragments for 8 MFMAs.  Twice the waves per tile at
// half the accumulators each, so a SIMD holds four waves of two tiles instead of two.
template <class T, int DEPTH = 1>
__device__ __forceinline__ void tile_gemm_nt_8(V4<T> (&acc)[4][2], const T* __restrict__ P, int ldp, int pn,
                                               const T* __restrict__ Q, int ldq, int qn, int K,
                                               TileSmemT<T, 128>& sm) {
  static_assert(DEPTH == 1 || DEPTH == 2, "one or two chunks fetched ahead");
  const int lane = threadIdx.x & 63;
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int wr = w >> 2, wc = w & 3;
  T pa[DEPTH][2][2], qb[DEPTH][2][2];
  int po[2][2], qo[2][2];  // byte offsets of this lane's elements within a chunk: k = w + 8q, row = lane + 64u
#pragma unroll
  for (int q = 0; q < 2; q++)
#pragma unroll
    for (int u = 0; u < 2; u++) {
      po[q][u] = ((w + 8 * q) * ldp + min(lane + 64 * u, pn - 1)) * (int)sizeof(T);
      qo[q][u] = ((w + 8 * q) * ldq + min(lane + 64 * u, qn - 1)) * (int)sizeof(T);
    }
  auto fetch = [&](int kc, auto set) {
    constexpr int S = decltype(set)::value;
    const int nk = max(min(16, K - kc), 0);
    const __amdgpu_buffer_rsrc_t rp = chunk_rsrc(P + (i64)kc * ldp, nk * ldp * (int)sizeof(T));
    const __amdgpu_buffer_rsrc_t rq = chunk_rsrc(Q + (i64)kc * ldq, nk * ldq * (int)sizeof(T));
#pragma unroll
    for (int q = 0; q < 2; q++)
#pragma unroll
      for (int u = 0; u < 2; u++) {
        pa[S][q][u] = buf_load<T>(rp, po[q][u]);
        qb[S][q][u] = buf_load<T>(rq, qo[q][u]);
      }
  };
  auto stage = [&](int b, auto set) {
    constexpr int S = decltype(set)::value;
#pragma unroll
    for (int q = 0; q < 2; q++)
#pragma unroll
      for (int u = 0; u < 2; u++) {
        sm.As[b][w + 8 * q][lane + 64 * u] = pa[S][q][u];
        sm.Bs[b][w + 8 * q][lane + 64 * u] = qb[S][q][u];
      }
  };
  if (K <= 0) return;
  using S0 = std::integral_constant<int, 0>;
  using S1 = std::integral_constant<int, DEPTH - 1>;
  auto step = [&](int kc, int b, auto next) {
    T av[2][2], bv[2][4];
    auto rd = [&](int k4, int z) {
      const int kr = k4 * 4 + (lane >> 4);
#pragma unroll
      for (int x = 0; x < 4; x++) bv[z][x] = sm.As[b][kr][wr * 64 + x * 16 + (lane & 15)];
#pragma unroll
      for (int y = 0; y < 2; y++) av[z][y] = sm.Bs[b][kr][wc * 32 + y * 16 + (lane & 15)];
    };
    rd(0, 0);
#pragma unroll
    for (int k4 = 0; k4 < 4; k4++) {
      if (k4 < 3) rd(k4 + 1, (k4 + 1) & 1);
      __builtin_amdgcn_sched_barrier(0);  // the next fragments' reads stay ahead of these MFMAs
#pragma unroll
      for (int bi = 0; bi < 4; bi++)
#pragma unroll
        for (int bj = 0; bj < 2; bj++) acc[bi][bj] = mfma4(av[k4 & 1][bj], bv[k4 & 1][bi], acc[bi][bj]);
    }
    // issue priority to the wave staging and fetching the next chunks (four waves per SIMD: FP64
    // 3D 127^3 trailing update 1504 -> 1482 ms per step; the 4-wave FP32 128-tile and the 64 x 128
    // paired-step tiles measured slower with it, in either phase)
    __builtin_amdgcn_s_setprio(1);
    stage(b ^ 1, next);  // past the last chunk: a harmless write of zeros to the unused stage
    fetch(kc + 16 * (DEPTH + 1), next);
    __builtin_amdgcn_s_setprio(0);
    __syncthreads();
  };
  __syncthreads();  // the stages may still be read by an earlier use of the shared memory
  fetch(0, S0{});
  if constexpr (DEPTH == 2) fetch(16, S1{});
  stage(0, S0{});
  fetch(16 * DEPTH, S0{});
  __syncthreads();
  if constexpr (DEPTH == 1) {
    int b = 0;
    for (int kc = 0; kc < K; kc += 16, b ^= 1) step(kc, b, S0{});
  } else {
    for (int kc = 0; kc < K; kc += 32) {
      step(kc, 0, S1{});  // the next chunk sits in set 1
      if (kc + 16 < K) step(kc + 16, 1, S0{});
    }
  }
}

// BF16-split variant of tile_gemm_nt_t for FP32 operands (the split precision form).  Each FP32
// value x is written to LDS as two bfloat16 pieces, x ~ x_0 + x_1 (x_0 = bf16(x),
// x_1 = bf16(x - x_0); the difference is exact in FP32), and P Q^T is formed on
// v_mfma_f32_16x16x32_bf16 with FP32 accumulation from the three piece products
// x_1 y_0 + x_0 y_1 + x_0 y_0 (the small ones first; x_1 y_1 is dropped): about 16 significant
// bits per product.  bfloat16 has FP32's exponent range, so the pieces need no scaling.  (A
// third piece and six products reach FP32's 24 bits but measured slower than FP32 MFMA.)
// The instruction's operand layout: lane l supplies row l%16 of A (column l%16 of B) with 8
// consecutive k of group l/16 of the 32-deep step, one 16-byte LDS read; accumulator r of lane
// l holds D[4 (l/16) + r][l%16], the layout of v_mfma_f32_16x16x4.  LDS rows hold 32 k as four
// 16-byte groups plus one group of padding (80 B), so the 16-byte reads of 16 consecutive rows
// fall on distinct banks.  Two LDS stages, one barrier per 32-deep chunk, as in tile_gemm_nt_t;
// the loop body is one basic block, and scheduling groups interleave the MFMAs with the next
// chunk's split, LDS writes and buffer loads.
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

__device__ __forceinline__ V4<float> mfma_bf(bf16x8 a, bf16x8 b, V4<float> c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

template <int TS>
struct TileSmemBF {
  bf16x8 P[2][2][TS][5];  // [stage][piece][row][group of 8 k]
  bf16x8 Q[2][2][TS][5];
};

__device__ __forceinline__ void split_bf(const float (&x)[8], bf16x8 (&pc)[2]) {
#pragma unroll
  for (int j = 0; j < 8; j++) {
    const __bf16 h = (__bf16)x[j];
    pc[0][j] = h;
    pc[1][j] = (__bf16)(x[j] - (float)h);
  }
}

template <int TS>
__device__ __forceinline__ void tile_gemm_nt_bf(V4<float> (&acc)[TS / 32][TS / 32], const float* __restrict__ P,
                                                int ldp, int pn, const float* __restrict__ Q, int ldq, int qn, int K,
                                                TileSmemBF<TS>& sm) {
  constexpr int NL = TS / 64, NBK = TS / 32, NPC = 2;
  const int t = threadIdx.x, lane = t & 63;
  const int w = __builtin_amdgcn_readfirstlane(t >> 6);
  const int wr = w >> 1, wc = w & 1;
  const int li = lane, kg = w;  // staging: row li (+64u), k = 8 kg .. 8 kg + 7 of the chunk
  // two register sets of fetched operands, so a chunk's loads have two chunks of work to land
  float pa[2][NL][8], qb[2][NL][8];
  int po[NL][8], qo[NL][8];
#pragma unroll
  for (int u = 0; u < NL; u++)
#pragma unroll
    for (int j = 0; j < 8; j++) {
      const int row = li + 64 * u;
      po[u][j] = ((8 * kg + j) * ldp + min(row, pn - 1)) * 4;
      qo[u][j] = ((8 * kg + j) * ldq + min(row, qn - 1)) * 4;
    }
  auto fetch = [&](int kc, auto set) {
    constexpr int S = decltype(set)::value;
    const int nk = max(min(32, K - kc), 0);
    const __amdgpu_buffer_rsrc_t rp = chunk_rsrc(P + (i64)kc * ldp, nk * ldp * 4);
    const __amdgpu_buffer_rsrc_t rq = chunk_rsrc(Q + (i64)kc * ldq, nk * ldq * 4);
#pragma unroll
    for (int j = 0; j < 8; j++)
#pragma unroll
      for (int u = 0; u < NL; u++) {
        pa[S][u][j] = buf_load<float>(rp, po[u][j]);
        qb[S][u][j] = buf_load<float>(rq, qo[u][j]);
      }
  };
  auto stage = [&](int b, auto set) {
    constexpr int S = decltype(set)::value;
#pragma unroll
    for (int u = 0; u < NL; u++) {
      bf16x8 pp[NPC], qp[NPC];
      split_bf(pa[S][u], pp);
      split_bf(qb[S][u], qp);
#pragma unroll
      for (int p = 0; p < NPC; p++) {
        sm.P[b][p][li + 64 * u][kg] = pp[p];
        sm.Q[b][p][li + 64 * u][kg] = qp[p];
      }
    }
  };
  if (K <= 0) return;
  using S0 = std::integral_constant<int, 0>;
  using S1 = std::integral_constant<int, 1>;
  __syncthreads();  // the stages may still be read by an earlier use of the shared memory
  fetch(0, S0{});
  fetch(32, S1{});
  stage(0, S0{});
  fetch(64, S0{});
  __syncthreads();
  const int q = lane >> 4, c16 = lane & 15;
  // chunk kc (LDS stage b): its products, then chunk kc+32 (register set `next`) is staged and
  // chunk kc+96 is fetched into that set
  auto step = [&](int kc, int b, auto next) {
    bf16x8 av[NPC][NBK], bv[NPC][NBK];
#pragma unroll
    for (int p = 0; p < NPC; p++)
#pragma unroll
      for (int x = 0; x < NBK; x++) {
        bv[p][x] = sm.P[b][p][wr * (TS / 2) + x * 16 + c16][q];
        av[p][x] = sm.Q[b][p][wc * (TS / 2) + x * 16 + c16][q];
      }
#pragma unroll
    for (int s = NPC - 1; s >= 0; s--)
#pragma unroll
      for (int i = 0; i <= s; i++)
#pragma unroll
        for (int bi = 0; bi < NBK; bi++)
#pragma unroll
          for (int bj = 0; bj < NBK; bj++) acc[bi][bj] = mfma_bf(av[i][bj], bv[s - i][bi], acc[bi][bj]);
    stage(b ^ 1, next);  // past the last chunk: a harmless write of zeros to the unused stage
    fetch(kc + 96, next);
    constexpr int NM = NBK * NBK * 3;  // MFMAs per chunk
#pragma unroll
    for (int g = 0; g < 8; g++) {
      __builtin_amdgcn_sched_group_barrier(0x008, NM / 8, 0);  // MFMA
      __builtin_amdgcn_sched_group_barrier(0x002, 12, 0);      // VALU
      __builtin_amdgcn_sched_group_barrier(0x200, 1, 0);       // LDS write
      __builtin_amdgcn_sched_group_barrier(0x020, 4, 0);       // buffer load
    }
    __syncthreads();
  };
  for (int kc = 0; kc < K; kc += 64) {
    step(kc, 0, S1{});  // chunk kc+32 sits in set 1
    if (kc + 32 < K) step(kc + 32, 1, S0{});
  }
}

// tile-local coordinates of accumulator (bi, bj, r) of this thread
__device__ __forceinline__ int acc_row(int bi) { return ((threadIdx.x >> 6) >> 1) * 32 + bi * 16 + (threadIdx.x & 15); }
template <class T>
__device__ __forceinline__ int acc_col(int bj, int r) {
  return ((threadIdx.x >> 6) & 1) * 32 + bj * 16 + mrow<T>((threadIdx.x & 63) >> 4, r);
}

// The 64 x 64 tile at p (column stride ld, nr x nc valid, nr, nc >= 1) in the accumulator
// layout, entries outside it zero.  Every load is unconditional at a clamped address, so the
// loads are in flight together instead of one guarded load and wait per entry.  `on` false
// gives zeros without reading.
template <class T>
__device__ __forceinline__ void tile_read(T (&v)[2][2][4], const T* __restrict__ p, i64 ld, int nr, int nc, bool on) {
#pragma unroll
  for (int bi = 0; bi < 2; bi++)
#pragma unroll
    for (int bj = 0; bj < 2; bj++)
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int i = acc_row(bi), j = acc_col<T>(bj, r);
        const T x = on ? p[(i64)min(j, nc - 1) * ld + min(i, nr - 1)] : T(0);
        v[bi][bj][r] = (i < nr && j < nc) ? x : T(0);
      }
}

// Workgroup -> (tile, front id index).  xt <= 0: 2-D grid (tile = blockIdx.x, front =
// blockIdx.y; xt < 0 is the GEMMs' band order, see gemm_body).  xt > 0: 1-D grid of
// 8 * ceil(nf / 8) * xt workgroups laid out so that the xt tiles of a front get workgroup ids
// congruent mod 8; ids are dealt to the 8 XCDs round-robin, so a front's tiles run on one XCD
// and the operands they share are fetched into one L2.
__device__ __forceinline__ bool wg_tile(int xt, int nf, int& tile, int& f) {
  if (xt <= 0) {
    tile = blockIdx.x;
    f = blockIdx.y;
    return true;
  }
  const int id = blockIdx.x, x = id & 7, slot = id >> 3, fl = slot / xt;
  tile = slot - fl * xt;
  f = fl * 8 + x;
  return f < nf;
}

// ----------------------------------------------------------------------------- factorization
// Blocked partial Cholesky of the pivot columns [0, n) of each front, 64 columns per block.
//   k_chol_diag:  the diagonal block is factored in LDS, T = L_jj L_jj^T, together with
//                 Linv_jj = L_jj^{-1}.  Left-looking it first takes the update by every
//                 finished column, D = A(jb.., jb..) - L(jb.., 0:jb) L(jb.., 0:jb)^T;
//                 right-looking the block already carries it.
//   k_chol_panel: rows below the block: X = (A(rows, jb block) - L(rows, 0:K) L(jb.., 0:K)^T)
//                 Linv_jj^T, i.e. the panel TRSM as a product with the inverted diagonal block.
//                 Left-looking (K = jb): every tile is formed once, from the untouched input.
//                 Right-looking (K = 0): the trailing region was updated by k_gemm_trail.
// A non-positive (or NaN) pivot records the first failing (front, column) in err[0..2].
// acc += P Q^T for 16 x 16 blocks held in LDS, one wave, one MFMA per 4 k (K = 16):
// P(i,k) = P[i*psi + k*psk], Q(j,k) = Q[j*qsj + k*qsk]; accumulator r of a lane is
// C(lane%16, mrow(lane/16, r)).  PM / QM mark a triangular operand whose other triangle is not
// stored as zeros: 1 = entries with k > row index read as zero, 2 = entries with k < row index.
template <int PM = 0, int QM = 0, class T>
__device__ __forceinline__ void mma16(V4<T>& acc, const T* P, int psi, int psk, const T* Q, int qsj, int qsk) {
  const int lane = threadIdx.x & 63, i = lane & 15, kq = lane >> 4;
#pragma unroll
  for (int k4 = 0; k4 < 4; k4++) {
    const int k = 4 * k4 + kq;
    T pv = P[i * psi + k * psk], qv = Q[i * qsj + k * qsk];
    if ((PM == 1 && k > i) || (PM == 2 && k < i)) pv = T(0);
    if ((QM == 1 && k > i) || (QM == 2 && k < i)) qv = T(0);
    acc = mfma4(qv, pv, acc);
  }
}

// 1/sqrt(d) from the hardware estimate and Newton steps, y <- y (3/2 - d y^2 / 2): two in
// double, one in float, which carry the estimate to full precision of the type.
__device__ __forceinline__ float rsqrt_nr(float d) {
  const float y = __builtin_amdgcn_rsqf(d);
  return fmaf(y, fmaf(-0.5f * d * y, y, 0.5f), y);
}
// 1/d from the hardware estimate and Newton steps y <- y (2 - d y): two in double, one in float.
__device__ __forceinline__ float rcp_nr(float d) {
  const float y = __builtin_amdgcn_rcpf(d);
  return fmaf(y, fmaf(-d, y, 1.0f), y);
}
__device__ __forceinline__ double rcp_nr(double d) {
  double y = __builtin_amdgcn_rcp(d);
#pragma unroll
  for (int it = 0; it < 2; it++) y = fma(y, fma(-d, y, 1.0), y);
  return y;
}
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
//   for k: one wave eliminates D_kk on [D_kk | I] without scaling the pivot rows (elements in
//   registers, 4 per lane; each step's column of D and row of the inverse part V move between
//   lanes by readlane, DPP and ds_bpermute, with no LDS round trip or barrier per pivot), then
//   scales: L_kk = D' diag(r), Linv_kk = diag(r) V with r_j = 1/sqrt(pivot j); panel
//   L_ik = A_ik Linv_kk^T and trailing A_ij -= L_ik L_jk^T on MFMA, one 16-block per wave, and
//   row k of the inverse's off-diagonal blocks, Linv_kj = -Linv_kk sum_i L_ki Linv_ij.  Wave 0
//   updates only the next diagonal block and starts its elimination at once; the other waves do
//   the rest of step k (other trailing blocks, L stores, the inverse row) meanwhile.
// Ts[c*LD + r] holds L(r, c) for r >= c; the inverse is stored row-major in the strict upper
// triangle: Linv(r, c) at U[r*LD + c], U = Ts + LD, for r >= c.
#define LD (NB + 1)
template <class T>
struct DiagSmem {
  T Ts[LD * LD];
};

// Factor the block held in Ts (lower triangle D(i, j) at Ts[j*LD + i], kb x kb, identity
// padding beyond kb, strict upper triangle zero): L to blk (column stride ld), Linv to dinv,
// unless `store` is false (L and Linv then stay in LDS only).  fid / jb name the front and the
// block's first column in an error report.
template <class T>
__device__ __forceinline__ void diag_core(DiagSmem<T>& sm, T* __restrict__ blk, i64 ld, int kb, T* __restrict__ dinv,
                          int* __restrict__ err, int fid, int jb, bool store = true) {
  T* Ts = sm.Ts;
  T* U = Ts + LD;
  const int t = threadIdx.x, w = t >> 6, lane = t & 63;
  if (t < NB) Ts[NB * LD + t] = T(0);
  __syncthreads();
  for (int k = 0; k < 4; k++) {
    const int K0 = 16 * k;
    T* Dk = Ts + K0 * LD + K0;  // D(r, c) at Dk[c*LD + r]
    T* Vk = U + K0 * LD + K0;  // V(r, c) at Vk[r*LD + c]
    // (a) wave 0, all in registers: lane (er, cg) = (lane/4, lane%4) holds D(er, c) and V(er, c),
    //     c = cg + 4u (entries above the diagonal ride along unused).  Step j: the pivot D(j, j)
    //     comes by readlane, D(er, j) by a broadcast within the lane's quad, D(c, j) and V(j, c)
    //     by ds_bpermute; rows er > j take m = D(er, j) / D(j, j) (1/d by reciprocal + Newton),
    //     columns c > j of D take -m D(c, j), columns c <= j of V take -m V(j, c).  Whether
    //     c > j is known per u at compile time except for one u per step.
    if (w == 0) {
      const int er = lane >> 2, cg = lane & 3;
      T dv[4], vv[4];
#pragma unroll
      for (int u = 0; u < 4; u++) {
        dv[u] = Dk[(cg + 4 * u) * LD + er];
        vv[u] = cg + 4 * u == er ? T(1) : T(0);
      }
#pragma unroll
      for (int j = 0; j < 16; j++) {
        const int uj = j >> 2, qj = j & 3;
        const T d = lane_read(dv[uj], 4 * j + qj);
        const T ar = quad_bcast(dv[uj], qj);
        T x[4], y[4];  // D(c, j) for c > j, V(j, c) for c <= j
#pragma unroll
        for (int u = 0; u < 4; u++) {
          if (4 * u + 3 > j) x[u] = lane_pull(dv[uj], 4 * (cg + 4 * u) + qj);
          if (4 * u <= j) y[u] = lane_pull(vv[u], 4 * j + cg);
        }
        const T m = er > j ? ar * rcp_nr(d) : T(0);
#pragma unroll
        for (int u = 0; u < 4; u++) {
          if (4 * u > j) {
            dv[u] -= m * x[u];
          } else if (4 * u + 3 <= j) {
            vv[u] -= m * y[u];
          } else {
            const bool cd = cg + 4 * u > j;
            dv[u] -= (cd ? m : T(0)) * x[u];
            vv[u] -= (cd ? T(0) : m) * y[u];
          }
        }
      }
      // pivots: D'(c, c) is dv[u] of lane (c, cg), D'(er, er) is dv[er/4] of lane (er, er%4)
      T own = dv[0];
#pragma unroll
      for (int u = 1; u < 4; u++)
        if ((er >> 2) == u) own = dv[u];
      const T pr = lane_pull(own, 4 * er + (er & 3));
      const unsigned long long bad = __ballot(cg == 0 && K0 + er < kb && !(pr > T(0)));
      if (bad && lane == 0 && atomicCAS(err, 0, 1) == 0) {  // the first failing pivot, in order
        err[1] = fid;
        err[2] = jb + K0 + (int)(__builtin_ctzll(bad) >> 2);
      }
      // L = D' diag(r), Linv = diag(r) V, r_j = 1/sqrt(pivot j)
      const T rr = rsqrt_nr(pr);
#pragma unroll
      for (int u = 0; u < 4; u++) {
        const int c = cg + 4 * u;
        const T rc = rsqrt_nr(lane_pull(dv[u], 4 * c + cg));
        if (c <= er) {
          Dk[c * LD + er] = dv[u] * rc;
          Vk[er * LD + c] = vv[u] * rr;
        }
      }
    }
    __syncthreads();
    // (b) panel blocks below: L_ik = A_ik Linv_kk^T, wave w -> block i = k+1+w (each wave reads
    //     and rewrites its own block)
    {
      const int i = k + 1 + w;
      if (i < 4) {
        V4<T> acc = vzero<T>();
        mma16<0, 1>(acc, Ts + K0 * LD + 16 * i, 1, LD, Vk, LD, 1);
#pragma unroll
        for (int q = 0; q < 4; q++) Ts[(K0 + mrow<T>(lane >> 4, q)) * LD + 16 * i + (lane & 15)] = acc[q];
      }
    }
    __syncthreads();
    // (c) trailing blocks A_ij -= L_ik L_jk^T for k < j <= i.  Wave 0 takes the next diagonal
    //     block (k+1, k+1), the only one the next elimination reads, and goes on to that
    //     elimination without a barrier; waves 1-3 take the other blocks and store columns
    //     K0..K0+15 of L meanwhile.
    auto trail = [&](int bi, int bj) {
      V4<T> acc = vzero<T>();
      mma16(acc, Ts + K0 * LD + 16 * bi, 1, LD, Ts + K0 * LD + 16 * bj, 1, LD);
      // a diagonal block keeps its strict upper triangle, which holds the inverse
#pragma unroll
      for (int q = 0; q < 4; q++)
        if (bi != bj || (lane & 15) >= mrow<T>(lane >> 4, q))
          Ts[(16 * bj + mrow<T>(lane >> 4, q)) * LD + 16 * bi + (lane & 15)] -= acc[q];
    };
    if (w == 0) {
      if (k < 3) trail(k + 1, k + 1);
      wave_lds_sync();
    } else {
      if (store)
        for (int e = t - 64; e < 16 * NB; e += 192) {
          const int i = e & 63, j = K0 + (e >> 6);
          if (i < kb && j < kb && i >= j) blk[(i64)j * ld + i] = Ts[j * LD + i];
        }
      for (int p = w - 1;; p += 3) {
        int cnt = 0, bi = -1, bj = -1;
        for (int i2 = k + 1; i2 < 4; i2++)
          for (int j2 = k + 1; j2 <= i2; j2++) {
            if (i2 == k + 1 && j2 == k + 1) continue;
            if (cnt == p) {
              bi = i2;
              bj = j2;
            }
            cnt++;
          }
        if (bi < 0) break;
        trail(bi, bj);
      }
      // row k of the inverse's off-diagonal blocks, one per wave: Linv_kj = -Linv_kk S with
      // S = sum_{kk=j}^{k-1} L_k,kk Linv_kk,j (earlier rows are complete, Linv_kk came from (a));
      // S is staged in the destination block
      const int j = w - 1;
      if (j < k) {
        T* Skj = U + 16 * k * LD + 16 * j;  // S(i', j') at Skj[i'*LD + j']
        V4<T> acc = vzero<T>();  // P(i', c) = L(16k+i', 16kk+c), Q(j', c) = Linv(16kk+c, 16j+j')
        mma16<0, 2>(acc, Ts + 16 * j * LD + 16 * k, 1, LD, U + 16 * j * LD + 16 * j, 1, LD);
        for (int kk = j + 1; kk < k; kk++)
          mma16(acc, Ts + 16 * kk * LD + 16 * k, 1, LD, U + 16 * kk * LD + 16 * j, 1, LD);
#pragma unroll
        for (int q = 0; q < 4; q++) Skj[(lane & 15) * LD + mrow<T>(lane >> 4, q)] = acc[q];
        wave_lds_sync();
        V4<T> a2 = vzero<T>();  // P(i', c) = Linv_kk(i', c), Q(j', c) = S(c, j')
        mma16<1, 0>(a2, U + 16 * k * LD + 16 * k, LD, 1, Skj, 1, LD);
        wave_lds_sync();
#pragma unroll
        for (int q = 0; q < 4; q++) Skj[(lane & 15) * LD + mrow<T>(lane >> 4, q)] = -a2[q];
      }
    }
  }
  __syncthreads();
  if (store)
#pragma unroll
    for (int e = t; e < NB * NB; e += 256) {
      const int r = e & 63, c = e >> 6;
      dinv[e] = r >= c ? U[r * LD + c] : T(0);  // column-major, zero above the diagonal
    }
}

template <class T>
union DiagBlockSmem {
  TileSmemT<T, 64> g;  // the update's stages, before the block is loaded
  DiagSmem<T> d;
};

// The 64 x 64 diagonal block of block column jb of front F (fid in error reports): left = 1
// (left-looking levels) first takes the update by every finished column,
// D = A(jb.., jb..) - L(jb.., 0:jb) L(jb.., 0:jb)^T, on MFMA (the panel steps leave the later
// diagonal blocks untouched); then L_jj and Linv_jj (diag_core).
template <class T>
__device__ __forceinline__ void diag_block(const FrontDev& F, T* __restrict__ A, int jb, int left,
                                           int* __restrict__ err, int fid, DiagBlockSmem<T>& sm) {
  const int kb = min(NB, F.n - jb);
  T* blk = A + F.off + (i64)jb * F.N + jb;
  V4<T> acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; i++)
#pragma unroll
    for (int j = 0; j < 2; j++) acc[i][j] = vzero<T>();
  T bv[2][2][4];  // the block, read before the update's GEMM
  tile_read(bv, blk, F.N, kb, kb, true);
  if (left && jb > 0) {
    const T* Lr = A + F.off + jb;  // rows jb.., columns 0:jb
    tile_gemm_nt_t<T, 64, true>(acc, Lr, F.N, kb, Lr, F.N, kb, jb, sm.g);
  }
  // strict upper part: the inverse, initially zero; identity padding beyond kb
#pragma unroll
  for (int bi = 0; bi < 2; bi++)
#pragma unroll
    for (int bj = 0; bj < 2; bj++)
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int i = acc_row(bi), j = acc_col<T>(bj, r);
        T v = T(0);
        if (i >= j) v = (i < kb && j < kb) ? bv[bi][bj][r] - acc[bi][bj][r] : (i == j ? T(1) : T(0));
        sm.d.Ts[j * LD + i] = v;
      }
  diag_core(sm.d, blk, F.N, kb, A + F.dinv_off + (i64)(jb / NB) * NB * NB, err, fid, jb);
}

template <class T>
__global__ __launch_bounds__(256, 4) void k_chol_diag(const int* __restrict__ ids, const FrontDev* __restrict__ fr,
                                                   T* __restrict__ A, int jb, int left, int* __restrict__ err) {
  const int fid = ids[blockIdx.x];
  const FrontDev F = fr[fid];
  if (jb >= F.n) return;
  __shared__ DiagBlockSmem<T> sm;
  diag_block<T>(F, A, jb, left, err, fid, sm);
}

// Persistent launches over a queue of items (fronts of one level, largest first): a grid of as
// many workgroups as the chip holds at once, each taking the next item until the queue is empty.
// Next item of the queue counter q (items 0..cnt-1), or -1 when it is empty.  Every workgroup
// fetches until its first empty answer, so a launch makes exactly cnt + gridDim.x fetches; the
// last one resets the counter for the next launch on the stream.
__device__ __forceinline__ int queue_next(int* __restrict__ q, int cnt) {
  __shared__ int item;
  __syncthreads();  // the previous item's LDS work is done
  if (threadIdx.x == 0) {
    const int v = atomicAdd(q, 1);
    if (v == cnt + (int)gridDim.x - 1) atomicExch(q, 0);
    item = v < cnt ? v : -1;
  }
  __syncthreads();
  return item;
}

// Small fronts assembled from the operator alone (no child blocks: leaf fronts) with one pivot
// block (1 <= n <= 64) and m <= LEAF_M boundary rows, one workgroup per front (a persistent
// launch over the level's queue of them, largest first), the whole factorization in LDS: the pivot block and the panel rows are scattered from the front's CSR
// triples, the block is factored with its inverse (diag_core: L11 to the front, Linv to its
// dinv block), the panel becomes L21 = A12 Linv^T, and the boundary block, scattered into the
// freed diagonal buffer, takes its Schur complement S = A11 - L21 L21^T.  The front is written
// once (lower triangle: L11, L21, S) and never read: the zero / assemble / diagonal / panel /
// Schur launches of the general path, with their HBM round trips of the front, collapse into
// one.  Which fronts take this kernel is a property of the plan, so their arithmetic does not
// depend on the pass that factors them.
#define LEAF_M 48
#define LDP (LEAF_M + 1)
template <class T>
__global__ __launch_bounds__(256) void k_leaf(const int* __restrict__ ids, int cnt, int* __restrict__ queue,
                                              const FrontDev* __restrict__ fr, T* __restrict__ A,
                                              const i64* __restrict__ tri_off, const Triple* __restrict__ tri,
                                              const double* __restrict__ csr_val, int* __restrict__ err) {
  __shared__ DiagSmem<T> sd;
  __shared__ T Pn[NB * LDP];  // Pn[c*LDP + r]: A12(r, c), then L21(r, c); r < m, c < 64
  for (int item; (item = queue_next(queue, cnt)) >= 0;) {
    const int fid = ids[item];
    const FrontDev F = fr[fid];
    T* Ts = sd.Ts;
    T* U = Ts + LD;
    const int t = threadIdx.x, w = t >> 6, lane = t & 63, c16 = lane & 15, q4 = lane >> 4;
    const int n = F.n, m = F.m;
    T* a = A + F.off;
    // pivot block: strict upper triangle zero, identity beyond n; panel rows zero
    for (int e = t; e < LD * LD; e += 256) {
      const int c = e / LD, r = e - c * LD;
      Ts[e] = (r == c && r >= n && r < NB) ? T(1) : T(0);
    }
    for (int e = t; e < NB * LDP; e += 256) Pn[e] = T(0);
    __syncthreads();
    const i64 t0 = tri_off[fid], t1 = tri_off[fid + 1];
    for (i64 i = t0 + t; i < t1; i += 256) {
      const Triple x = tri[i];
      if (x.col >= n) continue;  // the boundary block comes later, the A21 block is not used
      if (x.row < n) {
        if (x.row >= x.col) Ts[x.col * LD + x.row] = (T)csr_val[x.k];
      } else {
        Pn[x.col * LDP + (x.row - n)] = (T)csr_val[x.k];
      }
    }
    __syncthreads();
    diag_core(sd, a, F.N, n, A + F.dinv_off, err, fid, 0);
    __syncthreads();
    if (m == 0) continue;
    const int nk = (n + 15) >> 4;  // 16-column blocks of the pivot columns
    // panel: wave w forms rows [16w, 16w + 16) of L21 = A12 Linv^T (Linv(r, c) at U[r*LD + c])
    if (16 * w < m) {
      V4<T> acc[4];
  #pragma unroll
      for (int cj = 0; cj < 4; cj++) {
        acc[cj] = vzero<T>();
        if (cj >= nk) continue;
  #pragma unroll
        for (int kb = 0; kb <= cj; kb++) {
          if (kb == cj)
            mma16<0, 1>(acc[cj], Pn + 16 * kb * LDP + 16 * w, 1, LDP, U + 16 * cj * LD + 16 * kb, LD, 1);
          else
            mma16<0, 0>(acc[cj], Pn + 16 * kb * LDP + 16 * w, 1, LDP, U + 16 * cj * LD + 16 * kb, LD, 1);
        }
      }
      const int r = 16 * w + c16;
  #pragma unroll
      for (int cj = 0; cj < 4; cj++)
  #pragma unroll
        for (int qq = 0; qq < 4; qq++) {
          const int c = 16 * cj + mrow<T>(q4, qq);
          const T v = c < n ? acc[cj][qq] : T(0);
          Pn[c * LDP + r] = v;  // this wave's rows only: no other wave reads them before the barrier
          if (c < n && r < m) a[(i64)c * F.N + n + r] = v;
        }
    }
    __syncthreads();
    // boundary block A11 (lower triangle) into the diagonal buffer
    for (int e = t; e < LD * LD; e += 256) Ts[e] = T(0);
    __syncthreads();
    for (i64 i = t0 + t; i < t1; i += 256) {
      const Triple x = tri[i];
      if (x.col >= n && x.row >= x.col) Ts[(x.col - n) * LD + (x.row - n)] = (T)csr_val[x.k];
    }
    __syncthreads();
    // S = A11 - L21 L21^T, lower 16 x 16 tiles dealt round-robin to the waves
    const int mt = (m + 15) >> 4;
    for (int tl = w; tl < mt * (mt + 1) / 2; tl += 4) {
      int bi = 0;
      while ((bi + 1) * (bi + 2) / 2 <= tl) bi++;
      const int bj = tl - bi * (bi + 1) / 2;
      V4<T> acc = vzero<T>();
      for (int kb = 0; kb < nk; kb++)
        mma16(acc, Pn + 16 * kb * LDP + 16 * bi, 1, LDP, Pn + 16 * kb * LDP + 16 * bj, 1, LDP);
      const int r = 16 * bi + c16;
  #pragma unroll
      for (int qq = 0; qq < 4; qq++) {
        const int c = 16 * bj + mrow<T>(q4, qq);
        if (r < m && c <= r) a[(i64)(n + c) * F.N + n + r] = Ts[c * LD + r] - acc[qq];
      }
    }
  }
}

// Right-looking levels: the diagonal block's factorization and the panel below it in one
// launch, one workgroup per 64-row tile of rows [jb + kb, N) (at least one per front).  Every
// workgroup factors the diagonal block in its own LDS -- the same input and the same arithmetic,
// so the copies agree bit for bit -- and applies the inverse to its row tile.  The front's last
// workgroup also stores L_jj and Linv_jj, after every workgroup of the front has loaded the
// block: each one counts itself in arrive[front] once its loads are done, and the last one
// (dispatched after the others, so the wait is short and always ends) waits for the full count,
// then resets it for the next launch.
// Lazy column groups (kstart < jb or dcap > 0): the group's earlier block columns [kstart, jb)
// have not been applied to this block column, so each row tile first takes the GEMM over them
// (left-looking inside the group), and a row tile that is the diagonal block of a later block
// column below dcap takes this block's rank-64 update, so that the later step finds its
// diagonal block complete.  The group's trailing update at its end is unchanged.
template <class T>
__global__ __launch_bounds__(256) void k_chol_dpanel(const int* __restrict__ ids, const FrontDev* __restrict__ fr,
                                                     T* __restrict__ A, int jb, int kstart, int dcap, int xt, int nf,
                                                     int* __restrict__ arrive, int* __restrict__ err) {
  int tile, fi;
  if (!wg_tile(xt, nf, tile, fi)) return;
  const int fid = ids[fi];
  const FrontDev F = fr[fid];
  if (jb >= F.n) return;
  const int kb = min(NB, F.n - jb), s0 = jb + kb;
  const int nt = (F.N - s0 + NB - 1) / NB, nw = max(nt, 1);
  if (tile >= nw) return;
  __shared__ union {
    TileSmemT<T, 64> g;  // the in-group GEMM's stages, before the block is loaded
    struct {
      DiagSmem<T> sd;
      T X[NB * kXs<T>];
    } f;
  } sm;
  DiagSmem<T>& sd = sm.f.sd;
  T* X = sm.f.X;
  const int t = threadIdx.x, lane = t & 63, c16 = lane & 15, q = lane >> 4;
  const int w = wave_id(), wr = w >> 1, wc = w & 1;
  T* blk = A + F.off + (i64)jb * F.N + jb;
  const int row0 = __builtin_amdgcn_readfirstlane(s0 + NB * tile), nrow = tile < nt ? min(NB, F.N - row0) : 0;
  T* pan = A + F.off + (i64)jb * F.N + row0;
  V4<T> acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; i++)
#pragma unroll
    for (int j = 0; j < 2; j++) acc[i][j] = vzero<T>();
  // the row tile and the diagonal block, read before the in-group GEMM
  T pv[2][2][4], bv[2][2][4];
  tile_read(pv, pan, F.N, max(nrow, 1), kb, nrow > 0);
  tile_read(bv, blk, F.N, kb, kb, true);
  const int K = nrow > 0 ? jb - kstart : 0;
  if (K > 0) {
    const T* G = A + F.off + (i64)kstart * F.N;
    tile_gemm_nt_t<T, 64, false, 1, true>(acc, G + row0, F.N, nrow, G + jb, F.N, kb, K, sm.g);
  }
  // X = the tile minus the GEMM (its stages are done: the GEMM ends on a barrier)
#pragma unroll
  for (int bi = 0; bi < 2; bi++)
#pragma unroll
    for (int bj = 0; bj < 2; bj++)
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int i = acc_row(bi), j = acc_col<T>(bj, r);
        X[j * kXs<T> + i] = (i < nrow && j < kb) ? pv[bi][bj][r] - acc[bi][bj][r] : T(0);
        T v = T(0);  // strict upper part: the inverse, initially zero; identity padding
        if (i >= j) v = (i < kb && j < kb) ? bv[bi][bj][r] : (i == j ? T(1) : T(0));
        sd.Ts[j * LD + i] = v;
      }
  __syncthreads();  // every load of this workgroup has returned
  const bool writer = tile == nw - 1;
  if (t == 0) {
    __hip_atomic_fetch_add(arrive + fid, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (writer) {
      int spins = 0;
      while (__hip_atomic_load(arrive + fid, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != nw) {
        if (++spins > SPIN_LIMIT) {
          atomicCAS(err + 3, 0, 1);
          break;
        }
        __builtin_amdgcn_s_sleep(1);
      }
      __hip_atomic_store(arrive + fid, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
  }
  diag_core(sd, blk, F.N, kb, A + F.dinv_off + (i64)(jb / NB) * NB * NB, err, fid, jb, writer);
  __syncthreads();
  if (nrow == 0) return;
  // Y = X Linv^T with Linv(j, k) = U[j*LD + k] for k <= j (the entries above are L's)
  const T* U = sd.Ts + LD;
#pragma unroll
  for (int i = 0; i < 2; i++)
#pragma unroll
    for (int j = 0; j < 2; j++) acc[i][j] = vzero<T>();
#pragma unroll
  for (int k4 = 0; k4 < NB / 4; k4++) {
    const int k = k4 * 4 + q;
#pragma unroll
    for (int bi = 0; bi < 2; bi++) {
      const T bv = X[k * kXs<T> + wr * 32 + bi * 16 + c16];
#pragma unroll
      for (int bj = 0; bj < 2; bj++) {
        const int j = wc * 32 + bj * 16 + c16;
        const T av = k <= j ? U[j * LD + k] : T(0);
        acc[bi][bj] = mfma4(av, bv, acc[bi][bj]);
      }
    }
  }
#pragma unroll
  for (int bi = 0; bi < 2; bi++)
#pragma unroll
    for (int bj = 0; bj < 2; bj++)
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int i = acc_row(bi), j = acc_col<T>(bj, r);
        if (i < nrow && j < kb) pan[(i64)j * F.N + i] = acc[bi][bj][r];
      }
  if (row0 >= min(F.n, dcap)) return;
  // rank-64 update of the later diagonal block D = A(row0.., row0..) of the group: D -= Y Y^T,
  // lower triangle only (the wave above the diagonal and the upper block of the diagonal waves
  // skip)
  __syncthreads();  // every wave is done reading X
#pragma unroll
  for (int bi = 0; bi < 2; bi++)
#pragma unroll
    for (int bj = 0; bj < 2; bj++)
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int i = acc_row(bi), j = acc_col<T>(bj, r);
        X[j * kXs<T> + i] = (i < nrow && j < kb) ? acc[bi][bj][r] : T(0);
      }
  __syncthreads();
  if (wr < wc) return;
#pragma unroll
  for (int i = 0; i < 2; i++)
#pragma unroll
    for (int j = 0; j < 2; j++) acc[i][j] = vzero<T>();
  auto syrk = [&](auto diag_wave) {
#pragma unroll
    for (int k4 = 0; k4 < NB / 4; k4++) {
      const int k = k4 * 4 + q;
#pragma unroll
      for (int bi = 0; bi < 2; bi++) {
        const T bv = X[k * kXs<T> + wr * 32 + bi * 16 + c16];
#pragma unroll
        for (int bj = 0; bj < 2; bj++) {
          if (decltype(diag_wave)::value && bj > bi) continue;
          const T av = X[k * kXs<T> + wc * 32 + bj * 16 + c16];
          acc[bi][bj] = mfma4(av, bv, acc[bi][bj]);
        }
      }
    }
  };
  const int dn = min(NB, F.n - row0);
  T* D = A + F.off + (i64)row0 * F.N + row0;
  T dv[2][2][4];  // read while the product runs
  tile_read(dv, D, F.N, dn, dn, true);
  if (wr == wc)
    syrk(std::true_type{});
  else
    syrk(std::false_type{});
#pragma unroll
  for (int bi = 0; bi < 2; bi++)
#pragma unroll
    for (int bj = 0; bj < 2; bj++)
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int i = acc_row(bi), j = acc_col<T>(bj, r);
        if (i < dn && j < dn && i >= j) D[(i64)j * F.N + i] = dv[bi][bj][r] - acc[bi][bj][r];
      }
}

// Y = X Linv^T for one 64 x 64 tile on 4 waves as 2 x 2 (acc_row / acc_col layout): X in LDS
// (X(i, k) at X[k * kXs + i], zero outside the tile), Linv(j, k) at dinv[k * NB + j] (zero above
// the diagonal).  Linv(j, k) = 0 for k > j and X(i, k) = 0 for k >= kb: block column bj of a
// wave needs k < min(32 wc + 16 bj + 16, kb).  Partial tiles (kb or nrow < 64): a product outside
// the block is multiplied by zero instead of skipped by a branch, so the inverse's loads stay
// unconditional and are issued ahead together (X is zero there and the inverse finite, and adding
// +-0 to an accumulator that started at +0 changes no bit).
template <class T>
__device__ __forceinline__ void trsm64(V4<T> (&acc)[2][2], const T* __restrict__ X, const T* __restrict__ dinv,
                                       int kb, int nrow) {
  const int lane = threadIdx.x & 63, c16 = lane & 15, q = lane >> 4;
  const int w = wave_id(), wr = w >> 1, wc = w & 1;
#pragma unroll
  for (int i = 0; i < 2; i++)
#pragma unroll
    for (int j = 0; j < 2; j++) acc[i][j] = vzero<T>();
  auto trsm = [&](auto half, auto guard) {
    constexpr int WC = decltype(half)::value;
    const int k4end = (kb + 3) / 4;
#pragma unroll
    for (int k4 = 0; k4 < 8 * WC + 8; k4++) {
      const int k = k4 * 4 + q;
#pragma unroll
      for (int bi = 0; bi < 2; bi++) {
        const T bv = X[k * kXs<T> + wr * 32 + bi * 16 + c16];
#pragma unroll
        for (int bj = 0; bj < 2; bj++) {
          if (k4 >= 8 * WC + 4 * bj + 4) continue;
          const T av = dinv[k * NB + WC * 32 + bj * 16 + c16];
          const bool on = !decltype(guard)::value || (k4 < k4end && wr * 32 + bi * 16 < nrow);
          acc[bi][bj] = mfma4(av, bv * (on ? T(1) : T(0)), acc[bi][bj]);
        }
      }
    }
  };
  const bool full = kb == NB && nrow == NB;
  if (wc == 0) {
    if (full)
      trsm(std::integral_constant<int, 0>{}, std::false_type{});
    else
      trsm(std::integral_constant<int, 0>{}, std::true_type{});
  } else {
    if (full)
      trsm(std::integral_constant<int, 1>{}, std::false_type{});
    else
      trsm(std::integral_constant<int, 1>{}, std::true_type{});
  }
}

template <class T>
union PanelSmem64 {
  TileSmemT<T, 64> g;
  struct {
    T X[NB * kXs<T>];
  } c;
};

// One panel row tile of block column jb (rows row0 = jb + kb + 64 tile .. + 64 of front F):
// X = A(rows, jb block) - L(rows, 0:K) L(jb block, 0:K)^T (K = jb left-looking, 0
// right-looking), CSR entries added for leaf fronts taking them in the panel step (fly), then
// L(rows, jb block) = X Linv_jj^T.  Returns without work for a tile past the front.
template <class T>
__device__ __forceinline__ void panel_tile(const FrontDev& F, T* __restrict__ A, int jb, int left, int tile, int fly,
                                           const int* __restrict__ fly_tab, const Triple* __restrict__ fly_tri,
                                           const double* __restrict__ csr_val, PanelSmem64<T>& sm) {
  const int kb = min(NB, F.n - jb);
  const int s0 = jb + kb;  // first row below the block
  if (s0 + NB * tile >= F.N) return;
  const int t = threadIdx.x;
  V4<T> acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; i++)
#pragma unroll
    for (int j = 0; j < 2; j++) acc[i][j] = vzero<T>();
  const int row0 = __builtin_amdgcn_readfirstlane(s0 + NB * tile);
  const int nrow = min(NB, F.N - row0);
  const int K = left ? jb : 0;
  T* pan = A + F.off + (i64)jb * F.N + row0;
  // the tile's input: the assembled front, read before the GEMM (all loads in flight together,
  // at clamped addresses), or (fly) its CSR entries added below
  const bool from_csr = fly && F.fly_off >= 0;
  T pv[2][2][4];
  tile_read(pv, pan, F.N, nrow, kb, !from_csr);
  tile_gemm_nt_t<T, 64, false, 1, true>(acc, A + F.off + row0, F.N, nrow, A + F.off + jb, F.N, kb, K, sm.g);
#pragma unroll
  for (int bi = 0; bi < 2; bi++)
#pragma unroll
    for (int bj = 0; bj < 2; bj++)
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int i = acc_row(bi), j = acc_col<T>(bj, r);
        sm.c.X[j * kXs<T> + i] = (i < nrow && j < kb) ? pv[bi][bj][r] - acc[bi][bj][r] : T(0);
      }
  const T* dinv = A + F.dinv_off + (i64)(jb / NB) * NB * NB;  // Linv(j, k) at dinv[k*NB + j]
  __syncthreads();
  if (from_csr) {  // entries (row, col) of this tile: col in [jb, jb + kb), row in [row0, row0 + 64)
    const int* tb = fly_tab + F.fly_off + (jb / NB) * ((F.N + NB - 1) / NB + 1) + tile;
    for (int e = tb[0] + t; e < tb[1]; e += 256) {
      const Triple tr = fly_tri[e];
      sm.c.X[(tr.col - jb) * kXs<T> + (tr.row - row0)] += T(csr_val[tr.k]);
    }
    __syncthreads();
  }
  trsm64<T>(acc, sm.c.X, dinv, kb, nrow);
#pragma unroll
  for (int bi = 0; bi < 2; bi++)
#pragma unroll
    for (int bj = 0; bj < 2; bj++)
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int i = acc_row(bi), j = acc_col<T>(bj, r);
        if (i < nrow && j < kb) pan[(i64)j * F.N + i] = acc[bi][bj][r];
      }
}

// The left-looking GEMM runs on the native MFMA of T in every precision form: BF16 pieces
// (tile_gemm_nt_bf) measured slower here, where the GEMMs are short and the 40 KB of split
// stages halve the workgroups per CU.
template <class T>
__global__ __launch_bounds__(256, 4) void k_chol_panel(const int* __restrict__ ids, const FrontDev* __restrict__ fr,
                                                    T* __restrict__ A, int jb, int left, int xt, int nf, int fly,
                                                    const int* __restrict__ fly_tab,
                                                    const Triple* __restrict__ fly_tri,
                                                    const double* __restrict__ csr_val) {
  int tile, fi;
  if (!wg_tile(xt, nf, tile, fi)) return;
  const FrontDev F = fr[ids[fi]];
  if (jb >= F.n) return;
  __shared__ PanelSmem64<T> sm;
  panel_tile<T>(F, A, jb, left, tile, fly, fly_tab, fly_tri, csr_val, sm);
}

// ---- paired left-looking steps: block columns jb and jb + 64 (both full) in three launches,
// k_chol_diag (jb), k_chol_diag2 (jb), k_chol_panel2 (jb), instead of four.  The panel below both
// blocks reads its rows' finished columns L(rows, 0:jb) once for 128 columns instead of once per
// 64 (the left-looking panel's dominant traffic).  Every entry gets the same operations in the
// same order as in two single steps, so the factors are bit-identical.

// Stages of the 64 x 128 panel GEMM (P: 64 rows, Q: 128 rows), padded as TileSmemT.
template <class T>
struct TileSmemR {
  T As[2][16][64 + kStagePad<T>];
  T Bs[2][16][128 + kStagePad<T>];
};

// acc += P Q^T for a 64-row x 128-column tile, k in [0, K): the 4 waves as 2 x 2, wave (wr, wc)
// holds rows 32 wr .. + 32 and columns 64 wc .. + 64 (acc[bi][bj]: 16-row block bi, 16-column
// block bj).  Stages, operand loads and fragment prefetch as tile_gemm_nt_t; per 4-deep step a
// wave reads 2 P and 4 Q fragments for 8 MFMAs.
template <class T>
__device__ __forceinline__ void tile_gemm_nt_r(V4<T> (&acc)[2][4], const T* __restrict__ P, int ldp, int pn,
                                               const T* __restrict__ Q, int ldq, int qn, int K, TileSmemR<T>& sm) {
  if (K <= 0) return;
  const int lane = threadIdx.x & 63, w = wave_id(), wr = w >> 1, wc = w & 1;
  T pa[4], qb[4][2];
  int po[4], qo[4][2];  // byte offsets of this lane's elements within a chunk: k = w + 4q
#pragma unroll
  for (int q = 0; q < 4; q++) {
    po[q] = ((w + 4 * q) * ldp + min(lane, pn - 1)) * (int)sizeof(T);
#pragma unroll
    for (int u = 0; u < 2; u++) qo[q][u] = ((w + 4 * q) * ldq + min(lane + 64 * u, qn - 1)) * (int)sizeof(T);
  }
  auto fetch = [&](int kc) {
    const int nk = max(min(16, K - kc), 0);
    const __amdgpu_buffer_rsrc_t rp = chunk_rsrc(P + (i64)kc * ldp, nk * ldp * (int)sizeof(T));
    const __amdgpu_buffer_rsrc_t rq = chunk_rsrc(Q + (i64)kc * ldq, nk * ldq * (int)sizeof(T));
#pragma unroll
    for (int q = 0; q < 4; q++) {
      pa[q] = buf_load<T>(rp, po[q]);
#pragma unroll
      for (int u = 0; u < 2; u++) qb[q][u] = buf_load<T>(rq, qo[q][u]);
    }
  };
  auto stage = [&](int b) {
#pragma unroll
    for (int q = 0; q < 4; q++) {
      sm.As[b][w + 4 * q][lane] = pa[q];
#pragma unroll
      for (int u = 0; u < 2; u++) sm.Bs[b][w + 4 * q][lane + 64 * u] = qb[q][u];
    }
  };
  auto step = [&](int b) {
    T av[2][4], bv[2][2];
    auto rd = [&](int k4, int z) {
      const int kr = k4 * 4 + (lane >> 4);
#pragma unroll
      for (int x = 0; x < 2; x++) bv[z][x] = sm.As[b][kr][wr * 32 + x * 16 + (lane & 15)];
#pragma unroll
      for (int y = 0; y < 4; y++) av[z][y] = sm.Bs[b][kr][wc * 64 + y * 16 + (lane & 15)];
    };
    rd(0, 0);
    // FP64: issue priority to the wave in its MFMA phase (paired FP64 panels 3D 127^3 296 ms
    // single-step -> 283 ms; FP32 measured slower with it)
    if constexpr (sizeof(T) == 8) __builtin_amdgcn_s_setprio(1);
#pragma unroll
    for (int k4 = 0; k4 < 4; k4++) {
      if (k4 < 3) rd(k4 + 1, (k4 + 1) & 1);
      __builtin_amdgcn_sched_barrier(0);  // the next fragments' reads stay ahead of these MFMAs
#pragma unroll
      for (int bi = 0; bi < 2; bi++)
#pragma unroll
        for (int bj = 0; bj < 4; bj++) acc[bi][bj] = mfma4(av[k4 & 1][bj], bv[k4 & 1][bi], acc[bi][bj]);
    }
    if constexpr (sizeof(T) == 8) __builtin_amdgcn_s_setprio(0);
  };
  __syncthreads();  // the stages may still be read by an earlier use of the shared memory
  fetch(0);
  stage(0);
  fetch(16);
  __syncthreads();
  int b = 0;
  for (int kc = 0; kc < K; kc += 16, b ^= 1) {
    step(b);
    stage(b ^ 1);  // past the last chunk: a harmless write of zeros to the unused stage
    fetch(kc + 32);
    __syncthreads();
  }
}

// k_chol_diag2: one workgroup per front: the first panel tile of block jb (rows jb2 = jb + 64 ..
// jb + 128: T0 = L(jb2 block, jb block)) and the diagonal block of jb2, from one 64 x 128 GEMM
// over the finished columns [0, jb) (P = L(jb2 block, 0:jb), Q = L(jb:jb + 128, 0:jb): the T0
// tile's update and the jb2 block's update, both rows read once); T0 = X Linv_1^T; the jb2
// block's update continues over k in [jb, jb2) with T0 as both operands; then its L and Linv
// (diag_core).  The same operations in the same order as k_chol_panel's tile 0 and k_chol_diag.
template <class T>
__global__ __launch_bounds__(256, sizeof(T) == 8 ? 3 : 4) void k_chol_diag2(
    const int* __restrict__ ids, const FrontDev* __restrict__ fr, T* __restrict__ A, int jb, int fly,
    const int* __restrict__ fly_tab, const Triple* __restrict__ fly_tri, const double* __restrict__ csr_val,
    int* __restrict__ err) {
  const int fid = ids[blockIdx.x];
  const FrontDev F = fr[fid];
  if (jb + 2 * NB > F.n) return;
  const int jb2 = jb + NB;
  __shared__ union {
    TileSmemR<T> g;
    struct {
      T X[NB * kXs<T>];
    } c;
    DiagSmem<T> d;
  } sm;
  const int t = threadIdx.x, lane = t & 63, c16 = lane & 15, q = lane >> 4;
  const int w = wave_id(), wr = w >> 1, wc = w & 1;
  V4<T> acc[2][4];
#pragma unroll
  for (int i = 0; i < 2; i++)
#pragma unroll
    for (int j = 0; j < 4; j++) acc[i][j] = vzero<T>();
  const bool from_csr = fly && F.fly_off >= 0;
  T* pan = A + F.off + (i64)jb * F.N + jb2;  // the T0 tile; the jb2 block at pan + 64 F.N
  tile_gemm_nt_r<T>(acc, A + F.off + jb2, F.N, NB, A + F.off + jb, F.N, 2 * NB, jb, sm.g);
  // T0: X = A(jb2 block, jb block) - acc (waves wc = 0; the stages are free after the GEMM)
  if (wc == 0) {
    T pv[2][4][4];
#pragma unroll
    for (int bi = 0; bi < 2; bi++)
#pragma unroll
      for (int bj = 0; bj < 4; bj++)
#pragma unroll
        for (int r = 0; r < 4; r++) {
          const int i = wr * 32 + bi * 16 + c16, j = bj * 16 + mrow<T>(q, r);
          pv[bi][bj][r] = from_csr ? T(0) : pan[(i64)j * F.N + i];
        }
#pragma unroll
    for (int bi = 0; bi < 2; bi++)
#pragma unroll
      for (int bj = 0; bj < 4; bj++)
#pragma unroll
        for (int r = 0; r < 4; r++) {
          const int i = wr * 32 + bi * 16 + c16, j = bj * 16 + mrow<T>(q, r);
          sm.c.X[j * kXs<T> + i] = pv[bi][bj][r] - acc[bi][bj][r];
        }
  }
  __syncthreads();
  if (from_csr) {  // CSR entries of the 64-row tile 0 below jb's block
    const int* tb = fly_tab + F.fly_off + (jb / NB) * ((F.N + NB - 1) / NB + 1);
    for (int e = tb[0] + t; e < tb[1]; e += 256) {
      const Triple tr = fly_tri[e];
      sm.c.X[(tr.col - jb) * kXs<T> + (tr.row - jb2)] += T(csr_val[tr.k]);
    }
    __syncthreads();
  }
  V4<T> y[2][2];
  trsm64<T>(y, sm.c.X, A + F.dinv_off + (i64)(jb / NB) * NB * NB, NB, NB);
#pragma unroll
  for (int bi = 0; bi < 2; bi++)
#pragma unroll
    for (int bj = 0; bj < 2; bj++)
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int i = acc_row(bi), j = acc_col<T>(bj, r);
        pan[(i64)j * F.N + i] = y[bi][bj][r];
      }
  __syncthreads();  // every wave is done reading X
#pragma unroll
  for (int bi = 0; bi < 2; bi++)
#pragma unroll
    for (int bj = 0; bj < 2; bj++)
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int i = acc_row(bi), j = acc_col<T>(bj, r);
        sm.c.X[j * kXs<T> + i] = y[bi][bj][r];  // T0 in the X layout
      }
  __syncthreads();
  // the jb2 block (waves wc = 1, rows 32 wr .. + 32, all 64 columns): its update continues over k
  // in [jb, jb2) with P(i, k) = T0(i, k), Q(j, k) = T0(j, k); the block's current values read
  T dv[2][4][4];
  if (wc == 1) {
#pragma unroll
    for (int k4 = 0; k4 < NB / 4; k4++) {
      const int k = k4 * 4 + q;
      T av[4], bv[2];
#pragma unroll
      for (int bj = 0; bj < 4; bj++) av[bj] = sm.c.X[k * kXs<T> + bj * 16 + c16];
#pragma unroll
      for (int bi = 0; bi < 2; bi++) bv[bi] = sm.c.X[k * kXs<T> + wr * 32 + bi * 16 + c16];
#pragma unroll
      for (int bi = 0; bi < 2; bi++)
#pragma unroll
        for (int bj = 0; bj < 4; bj++) acc[bi][bj] = mfma4(av[bj], bv[bi], acc[bi][bj]);
    }
#pragma unroll
    for (int bi = 0; bi < 2; bi++)
#pragma unroll
      for (int bj = 0; bj < 4; bj++)
#pragma unroll
        for (int r = 0; r < 4; r++) {
          const int i = wr * 32 + bi * 16 + c16, j = bj * 16 + mrow<T>(q, r);
          dv[bi][bj][r] = pan[(i64)(NB + j) * F.N + i];  // A(jb2 + i, jb2 + j)
        }
  }
  __syncthreads();  // every wave is done reading T0
  if (wc == 1) {  // the lower triangle of D into the diagonal factorization's LDS block
#pragma unroll
    for (int bi = 0; bi < 2; bi++)
#pragma unroll
      for (int bj = 0; bj < 4; bj++)
#pragma unroll
        for (int r = 0; r < 4; r++) {
          const int i = wr * 32 + bi * 16 + c16, j = bj * 16 + mrow<T>(q, r);
          sm.d.Ts[j * LD + i] = i >= j ? dv[bi][bj][r] - acc[bi][bj][r] : T(0);
        }
  }
  diag_core(sm.d, pan + (i64)NB * F.N, F.N, NB, A + F.dinv_off + (i64)(jb2 / NB) * NB * NB, err, fid, jb2);
}

// k_chol_panel2: the panel rows below both blocks (row0 = jb + 128 + 64 tile), one 64-row tile per
// workgroup.  X = A(rows, jb:jb + 128) - L(rows, 0:jb) L(jb:jb + 128, 0:jb)^T (one GEMM for both
// blocks); block jb: Y1 = X1 Linv_1^T; block jb2: the accumulation continues over k in [jb, jb2)
// with Y1 and T0 = L(jb2 block, jb block) as operands, X2 = A2 - acc, Y2 = X2 Linv_2^T.
template <class T>
__global__ __launch_bounds__(256, sizeof(T) == 8 ? 3 : 4) void k_chol_panel2(const int* __restrict__ ids,
                                                                            const FrontDev* __restrict__ fr,
                                                     T* __restrict__ A, int jb, int xt, int nf, int fly,
                                                     const int* __restrict__ fly_tab,
                                                     const Triple* __restrict__ fly_tri,
                                                     const double* __restrict__ csr_val) {
  int tile, fi;
  if (!wg_tile(xt, nf, tile, fi)) return;
  const FrontDev F = fr[ids[fi]];
  if (jb + 2 * NB > F.n) return;
  const int jb2 = jb + NB, s0 = jb + 2 * NB;
  if (s0 + NB * tile >= F.N) return;
  __shared__ union {
    TileSmemR<T> g;
    struct {
      T X[NB * kXs<T>];
    } c;
  } sm;
  const int t = threadIdx.x, lane = t & 63, c16 = lane & 15, q = lane >> 4;
  const int w = wave_id(), wr = w >> 1, wc = w & 1;
  V4<T> acc[2][4];
#pragma unroll
  for (int i = 0; i < 2; i++)
#pragma unroll
    for (int j = 0; j < 4; j++) acc[i][j] = vzero<T>();
  const int row0 = __builtin_amdgcn_readfirstlane(s0 + NB * tile);
  const int nrow = min(NB, F.N - row0);
  const bool from_csr = fly && F.fly_off >= 0;
  T* pan = A + F.off + (i64)jb * F.N + row0;  // X1 / Y1; X2 / Y2 at pan + 64 F.N
  tile_gemm_nt_r<T>(acc, A + F.off + row0, F.N, nrow, A + F.off + jb, F.N, 2 * NB, jb, sm.g);
  // X = (the wave's 32 rows x 64 columns of the tile's current values, block wc, read at clamped
  // addresses) - acc, into the X block; read only when needed (registers)
  auto x_out = [&]() {
    T pv[2][4][4];
#pragma unroll
    for (int bi = 0; bi < 2; bi++)
#pragma unroll
      for (int bj = 0; bj < 4; bj++)
#pragma unroll
        for (int r = 0; r < 4; r++) {
          const int i = wr * 32 + bi * 16 + c16, j = bj * 16 + mrow<T>(q, r);
          const T x = from_csr ? T(0) : pan[(i64)(wc * NB + j) * F.N + min(i, nrow - 1)];
          pv[bi][bj][r] = i < nrow ? x : T(0);
        }
#pragma unroll
    for (int bi = 0; bi < 2; bi++)
#pragma unroll
      for (int bj = 0; bj < 4; bj++)
#pragma unroll
        for (int r = 0; r < 4; r++) {
          const int i = wr * 32 + bi * 16 + c16, j = bj * 16 + mrow<T>(q, r);
          sm.c.X[j * kXs<T> + i] = i < nrow ? pv[bi][bj][r] - acc[bi][bj][r] : T(0);
        }
  };
  const int NT = (F.N + NB - 1) / NB;
  auto add_csr = [&](int cb, int t64) {  // CSR entries of 64-row tile t64 of block column cb
    const int* tb = fly_tab + F.fly_off + cb * (NT + 1);
    for (int e = tb[t64] + t; e < tb[t64 + 1]; e += 256) {
      const Triple tr = fly_tri[e];
      sm.c.X[(tr.col - cb * NB) * kXs<T> + (tr.row - row0)] += T(csr_val[tr.k]);
    }
  };
  // block jb: X1 (waves wc = 0), its CSR entries (64-row tile `tile + 1` below jb's block), Y1
  // (the GEMM ends on a barrier: the stages are free)
  if (wc == 0) x_out();
  __syncthreads();
  if (from_csr) {
    add_csr(jb / NB, tile + 1);
    __syncthreads();
  }
  V4<T> y[2][2];
  trsm64<T>(y, sm.c.X, A + F.dinv_off + (i64)(jb / NB) * NB * NB, NB, nrow);
#pragma unroll
  for (int bi = 0; bi < 2; bi++)
#pragma unroll
    for (int bj = 0; bj < 2; bj++)
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int i = acc_row(bi), j = acc_col<T>(bj, r);
        if (i < nrow) pan[(i64)j * F.N + i] = y[bi][bj][r];
      }
  __syncthreads();  // every wave is done reading X1
#pragma unroll
  for (int bi = 0; bi < 2; bi++)
#pragma unroll
    for (int bj = 0; bj < 2; bj++)
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int i = acc_row(bi), j = acc_col<T>(bj, r);
        sm.c.X[j * kXs<T> + i] = i < nrow ? y[bi][bj][r] : T(0);  // Y1 in the X layout
      }
  __syncthreads();
  // block jb2 (waves wc = 1): the accumulation continues over k in [jb, jb2): P(i, k) = Y1(i, k),
  // Q(j, k) = T0(j, k) = L(jb2 + j, jb + k)
  if (wc == 1) {
    const T* T0 = A + F.off + (i64)jb * F.N + jb2;
#pragma unroll
    for (int k4 = 0; k4 < NB / 4; k4++) {
      const int k = k4 * 4 + q;
      T av[4], bv[2];
#pragma unroll
      for (int bj = 0; bj < 4; bj++) av[bj] = T0[(i64)k * F.N + bj * 16 + c16];
#pragma unroll
      for (int bi = 0; bi < 2; bi++) bv[bi] = sm.c.X[k * kXs<T> + wr * 32 + bi * 16 + c16];
#pragma unroll
      for (int bi = 0; bi < 2; bi++)
#pragma unroll
        for (int bj = 0; bj < 4; bj++) acc[bi][bj] = mfma4(av[bj], bv[bi], acc[bi][bj]);
    }
  }
  __syncthreads();  // every wave is done reading Y1
  if (wc == 1) x_out();
  __syncthreads();
  if (from_csr) {
    add_csr(jb2 / NB, tile);
    __syncthreads();
  }
  trsm64<T>(y, sm.c.X, A + F.dinv_off + (i64)(jb2 / NB) * NB * NB, NB, nrow);
#pragma unroll
  for (int bi = 0; bi < 2; bi++)
#pragma unroll
    for (int bj = 0; bj < 2; bj++)
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int i = acc_row(bi), j = acc_col<T>(bj, r);
        if (i < nrow) pan[(i64)(NB + j) * F.N + i] = y[bi][bj][r];
      }
}

// Lower-triangular tile enumeration of the region rows [s, rend) x cols [s, cend), TS x TS
// tiles: first the T(T+1)/2 tiles of the square part, then the full rows below it.
template <int TS = 64>
__device__ __forceinline__ bool tile_of(int x, int s, int rend, int cend, int& row0, int& col0) {
  const int T = (cend - s + TS - 1) / TS, RT = (rend - s + TS - 1) / TS;
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
  row0 = s + TS * rt;
  col0 = s + TS * ct;
  return true;
}

// The same tile set as tile_of in band order: bands of 8 tile rows, each band column by column
// (the band's tiles of one tile column are consecutive).  A run of 64 consecutive indices then
// covers about 8 x 8 tiles, which share 8 row and 8 column operand panels.
template <int TS = 64>
__device__ __forceinline__ bool tile_of_band(int x, int s, int rend, int cend, int& row0, int& col0) {
  const int T = (cend - s + TS - 1) / TS, RT = (rend - s + TS - 1) / TS;
  const int tri = T * (T + 1) / 2;
  if (x >= tri + (RT - T) * T) return false;
  int R;  // tile rows wholly before x
  if (x < tri) {
    R = (int)((sqrt(8.0 * x + 1.0) - 1.0) * 0.5);
    while (R * (R + 1) / 2 > x) R--;
    while ((R + 1) * (R + 2) / 2 <= x) R++;
  } else {
    R = T + (x - tri) / T;
  }
  const int r0 = R & ~7, r1 = min(r0 + 8, RT) - 1, h = r1 - r0 + 1;
  int y = x - (r0 <= T ? r0 * (r0 + 1) / 2 : tri + (r0 - T) * T);
  const int F = min(r0 + 1, T);  // columns that every row of the band reaches
  int rt, ct;
  if (y < F * h) {
    ct = y / h;
    rt = r0 + y - ct * h;
  } else {  // the band's triangle: column c holds rows c..r1
    y -= F * h;
    int c = r0 + 1, cnt = r1 - c + 1;
    while (y >= cnt) {
      y -= cnt;
      c++;
      cnt--;
    }
    ct = c;
    rt = c + y;
  }
  row0 = s + TS * rt;
  col0 = s + TS * ct;
  return true;
}

// C -= P Q^T on the lower part of a region, one 64x64 tile per workgroup.
// MODE 0: right-looking update by the finished columns [k0, k0+K), K = min(kw, n - k0):
//         P = F[rows, k0:k0+K], Q = F[cols, k0:k0+K], C = F[s:N, s:min(n, ccap)] lower,
//         s = k0+K (kw = 64, ccap = the group end: the update inside a column group;
//         kw = group width, ccap = n: the group's trailing update).
// MODE 1: Schur complement: P = Q = L21 = F[n:N, 0:n], K = n; C = F[n:N, n:N] lower gets
//         A11 + S, except on the diagonal separator blocks, which get S alone.
// SPLIT = false: the native MFMA of T; true (T = float): the product on BF16 pieces
// (tile_gemm_nt_bf).
// NW = 8 (TS = 128, native MFMA): the tile on 8 waves (tile_gemm_nt_8).
template <class T, int MODE, int TS, bool SPLIT, int NW = 4>
__device__ __forceinline__ void gemm_body(const int* __restrict__ ids, const FrontDev* __restrict__ fr,
                                          T* __restrict__ A, int k0, int kw, int ccap, int xt, int nf) {
  static_assert(!SPLIT || std::is_same<T, float>::value, "the BF16 split takes FP32 fronts");
  static_assert(NW == 4 || (NW == 8 && TS == 128 && !SPLIT), "8 waves: 128-tiles on the native MFMA");
  int tile, fi;
  if (!wg_tile(xt, nf, tile, fi)) return;
  const FrontDev F = fr[ids[fi]];
  int s, K, kbase, rend, cend;
  if (MODE == 0) {  // columns [k0, k0 + kw) applied to rows [s, N) x columns [s, min(n, ccap))
    if (k0 >= F.n) return;
    K = min(kw, F.n - k0);
    s = k0 + K;
    cend = min(F.n, ccap);
    if (s >= cend) return;
    kbase = k0; rend = F.N;
  } else {
    if (F.m == 0) return;
    s = F.n; K = F.n; kbase = 0; rend = F.N; cend = F.N;
  }
  int row0, col0;
  if (xt < 0) {
    // band order, the front's tiles dealt to the 8 XCDs in contiguous eighths: the grid's x
    // dimension is a multiple of 8, so workgroup (x, y) runs on XCD x mod 8, and the tiles an
    // XCD holds at one time lie in one or two bands and share their operand panels in its L2
    const int TC = (cend - s + TS - 1) / TS, TR = (rend - s + TS - 1) / TS;
    const int nt = TC * (TC + 1) / 2 + (TR - TC) * TC, chunk = (nt + 7) >> 3, j = tile >> 3;
    if (j >= chunk) return;
    if (!tile_of_band<TS>((tile & 7) * chunk + j, s, rend, cend, row0, col0)) return;
  } else if (!tile_of<TS>(tile, s, rend, cend, row0, col0)) {
    return;
  }
  row0 = __builtin_amdgcn_readfirstlane(row0);  // wave-uniform: scalar operand addresses
  col0 = __builtin_amdgcn_readfirstlane(col0);
  // a wave's sub-tile: NBR x NBC blocks of 16 x 16 at (WR, WC) in the tile
  constexpr int NBR = TS / 32, NBC = NW == 8 ? 2 : TS / 32;
  V4<T> acc[NBR][NBC];
#pragma unroll
  for (int i = 0; i < NBR; i++)
#pragma unroll
    for (int j = 0; j < NBC; j++) acc[i][j] = vzero<T>();
  // The tile's current values are read with unconditional loads at clamped (always valid)
  // addresses -- entries outside the region are read and discarded -- so that the loads are in
  // flight together instead of one guarded load and wait per entry.  64 x 64 tiles read theirs
  // before the GEMM (its operands are other columns), 128 x 128 tiles 16 per lane after it.
  T* C = A + F.off;
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int WR = NW == 8 ? (w >> 2) * (TS / 2) : (w >> 1) * (TS / 2);
  const int WC = NW == 8 ? (w & 3) * (TS / 4) : (w & 1) * (TS / 2);
  constexpr int BR = TS == 64 ? NBR : 1;  // 16-row blocks read per batch
  T cv[BR][NBC][4];
  auto read_c = [&](int bi0) {
#pragma unroll
    for (int b = 0; b < BR; b++)
#pragma unroll
      for (int bj = 0; bj < NBC; bj++)
#pragma unroll
        for (int r = 0; r < 4; r++) {
          const int i = min(row0 + WR + (bi0 + b) * 16 + (lane & 15), rend - 1);
          const int j = min(col0 + WC + bj * 16 + mrow<T>(lane >> 4, r), cend - 1);
          cv[b][bj][r] = C[(i64)j * F.N + i];
        }
  };
  if constexpr (TS == 64) read_c(0);
  const T* base = A + F.off + (i64)kbase * F.N;
  if constexpr (NW == 8) {
    constexpr int DEPTH = sizeof(T) == 4 ? 2 : 1;
    __shared__ TileSmemT<T, 128> sm;
    tile_gemm_nt_8<T, DEPTH>(acc, base + row0, F.N, rend - row0, base + col0, F.N, cend - col0, K, sm);
  } else if constexpr (!SPLIT) {
    // FP32 fronts fetch two chunks ahead (FP64 tiles would lose occupancy to the registers)
    constexpr int DEPTH = sizeof(T) == 4 ? 2 : 1;
    __shared__ TileSmemT<T, TS> sm;
    tile_gemm_nt_t<T, TS, false, DEPTH>(acc, base + row0, F.N, rend - row0, base + col0, F.N, cend - col0, K, sm);
  } else {
    __shared__ TileSmemBF<TS> sm;
    tile_gemm_nt_bf<TS>(acc, base + row0, F.N, rend - row0, base + col0, F.N, cend - col0, K, sm);
  }
#pragma unroll
  for (int bi0 = 0; bi0 < NBR; bi0 += BR) {
    if constexpr (TS != 64) read_c(bi0);
#pragma unroll
    for (int b = 0; b < BR; b++)
#pragma unroll
      for (int bj = 0; bj < NBC; bj++)
#pragma unroll
        for (int r = 0; r < 4; r++) {
          const int bi = bi0 + b;
          const int i = row0 + WR + bi * 16 + (lane & 15);
          const int j = col0 + WC + bj * 16 + mrow<T>(lane >> 4, r);
          if (i >= rend || j >= cend || i < j) continue;
          T* c = C + (i64)j * F.N + i;
          const T sv = -acc[bi][bj][r];
          if (MODE == 0) {
            *c = cv[b][bj][r] + sv;
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
            *c = (ri == rj) ? sv : (cv[b][bj][r] + sv);
          }
        }
  }
}

// The three GEMM stages as kernels of their own names (the profiles and the runtime's timing
// name the same launches): the group update (K = 64 or 256 columns of a 256-column group), the
// trailing update (K = 256 / the outer group width), and the Schur complement (K = n).
template <class T, int TS = 64, bool SPLIT = false>
__global__ __launch_bounds__(256, TS == 128 ? 2 : 1) void k_gemm_group(const int* __restrict__ ids,
                                                                   const FrontDev* __restrict__ fr, T* __restrict__ A,
                                                                   int k0, int kw, int ccap, int xt, int nf) {
  gemm_body<T, 0, TS, SPLIT>(ids, fr, A, k0, kw, ccap, xt, nf);
}
// NW = 8: the 128-tile on 8 waves, four waves per SIMD (two tiles per CU)
template <class T, int TS = 64, bool SPLIT = false, int NW = 4>
__global__ __launch_bounds__(64 * NW, NW == 8 ? 4 : TS == 128 ? 2 : 1) void k_gemm_trail(
    const int* __restrict__ ids, const FrontDev* __restrict__ fr, T* __restrict__ A, int k0, int kw, int ccap, int xt,
    int nf) {
  gemm_body<T, 0, TS, SPLIT, NW>(ids, fr, A, k0, kw, ccap, xt, nf);
}
template <class T, int TS = 64, bool SPLIT = false, int NW = 4>
__global__ __launch_bounds__(64 * NW, NW == 8 ? 4 : TS == 128 ? 2 : 1) void k_gemm_schur(
    const int* __restrict__ ids, const FrontDev* __restrict__ fr, T* __restrict__ A, int k0, int kw, int ccap, int xt,
    int nf) {
  gemm_body<T, 1, TS, SPLIT, NW>(ids, fr, A, k0, kw, ccap, xt, nf);
}


// ----------------------------------------------------------------------------- solve
// The sweeps work on W, an n_dof x 16 row-major right-hand-side block (W[dof*16 + r]), and on a
// per-front scratch X (N x 16, row-major).  All block products run on 16x16x4 MFMAs with a
// 16-row block of L (or Linv) as the A operand and a 4 x 16 slice of X as B, so accumulator r
// of a lane is row mrow(lane/16, r), right-hand side lane%16.
#define RB 16
#define GU 4  // 16-row groups in flight per wave
#define FG 2  // the same in k_solve_fwd (4 took 189 VGPRs in FP32: 2 waves per SIMD)

// *p when ok, else 0, with the load unconditional: p is valid either way (the caller clamps the
// index), so the loads of an unrolled loop are issued together instead of one branch and wait
// per load.
template <class T>
__device__ __forceinline__ T ld_or0(const T* p, bool ok) {
  const T v = *p;
  return ok ? v : T(0);
}

template <class T>
__device__ __forceinline__ void gather_rhs(const FrontDev& F, T* __restrict__ X, const T* __restrict__ W,
                                           const int* __restrict__ di, const int* __restrict__ db) {
  // interior rows, then boundary rows (zero without db); unrolled so that several rows' loads
  // are in flight together
  const int ni = F.n * RB, na = F.N * RB;
#pragma unroll 4
  for (int e = threadIdx.x; e < ni; e += 256) X[e] = W[(i64)di[e >> 4] * RB + (e & 15)];
  if (db) {
#pragma unroll 4
    for (int e = ni + threadIdx.x; e < na; e += 256) X[e] = W[(i64)db[(e >> 4) - F.n] * RB + (e & 15)];
  } else {
    for (int e = ni + threadIdx.x; e < na; e += 256) X[e] = T(0);
  }
}

// Forward sweep of one front: X[0:n] <- L11^{-1} W[int] block by block (x_j = Linv_j b_j, then
// b_below -= L(below, j) x_j), X[n:N] <- -L21 X[0:n]; X[0:n] goes back to W[int].
// perm (LU fronts): the interior rows are gathered in pivot order, b_rd -> P b_rd, from the
// front's permuted DOF list perm + int_off (L then is unit lower; dinv holds its inverted blocks).
template <class T>
__global__ __launch_bounds__(256, sizeof(T) == 4 ? 4 : 1) void k_solve_fwd(const int* __restrict__ ids, const FrontDev* __restrict__ fr,
                                                   const T* __restrict__ A, const int* __restrict__ dofs,
                                                   T* __restrict__ W, T* __restrict__ xs,
                                                   const int* __restrict__ perm) {
  const FrontDev F = fr[ids[blockIdx.x]];
  const T* L = A + F.off;
  T* X = xs + F.x_off;
  const int* di = dofs + F.int_off;
  const int t = threadIdx.x, lane = t & 63, w = t >> 6, c16 = lane & 15, q = lane >> 4;
  gather_rhs(F, X, W, perm ? perm + F.int_off : di, nullptr);
  __shared__ T xb[NB][RB];
  for (int jb = 0; jb < F.n; jb += NB) {
    const int kb = min(NB, F.n - jb);
    __syncthreads();
    {  // x_j = Linv_j b_j: wave w -> rows 16w..16w+15; Linv is lower triangular, so the 4-column
       // slices right of the wave's rows, and those past kb, hold only zeros and are skipped
      const T* Li = A + F.dinv_off + (i64)(jb / NB) * NB * NB;
      V4<T> acc = vzero<T>();
      const int k4end = min(4 * (w + 1), (kb + 3) >> 2);
#pragma unroll
      for (int k4 = 0; k4 < NB / 4; k4++) {
        if (k4 >= k4end) break;
        const int k = 4 * k4 + q;
        const T av = Li[k * NB + 16 * w + c16];
        const T bv = ld_or0(X + (i64)(jb + min(k, kb - 1)) * RB + c16, k < kb);
        acc = mfma4(av, bv, acc);
      }
      __syncthreads();
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int i = 16 * w + mrow<T>(q, r);
        xb[i][c16] = acc[r];
        if (i < kb) X[(i64)(jb + i) * RB + c16] = acc[r];
      }
    }
    __syncthreads();
    // rows below, FG 16-row groups per wave in flight; out-of-range rows read row N-1 and are
    // never written
    for (int g0 = jb + kb + 16 * w; g0 < F.N; g0 += 64 * FG) {
      T av[FG][NB / 4], xv[FG][4];
#pragma unroll
      for (int u = 0; u < FG; u++) {
        const int row = min(g0 + 64 * u + c16, F.N - 1);
#pragma unroll
        for (int k4 = 0; k4 < NB / 4; k4++) {
          const int k = 4 * k4 + q;
          av[u][k4] = ld_or0(L + (i64)(jb + min(k, kb - 1)) * F.N + row, k < kb);
        }
#pragma unroll
        for (int r = 0; r < 4; r++) xv[u][r] = X[(i64)min(g0 + 64 * u + mrow<T>(q, r), F.N - 1) * RB + c16];
      }
#pragma unroll
      for (int u = 0; u < FG; u++) {
        V4<T> acc = vzero<T>();
#pragma unroll
        for (int k4 = 0; k4 < NB / 4; k4++)
          acc = mfma4(av[u][k4], xb[4 * k4 + q][c16], acc);
#pragma unroll
        for (int r = 0; r < 4; r++) {
          const int i = g0 + 64 * u + mrow<T>(q, r);
          if (i < F.N) X[(i64)i * RB + c16] = xv[u][r] - acc[r];
        }
      }
    }
  }
  __syncthreads();
#pragma unroll 4
  for (int e = t; e < F.n * RB; e += 256) W[(i64)di[e >> 4] * RB + (e & 15)] = X[e];
}

// Separator accumulation of per-box contributions X[n+pos : n+pos+k] of both neighbour boxes
// into the work block, box 0 first: W = (W + X_b0) + X_b1.
template <class T>
__global__ __launch_bounds__(256) void k_sep_acc(const int* __restrict__ ids, const SepDev* __restrict__ seps,
                                                 const FrontDev* __restrict__ fr, const int* __restrict__ dofs,
                                                 T* __restrict__ W, const T* __restrict__ xs) {
  const SepDev S = seps[ids[blockIdx.x]];
  const FrontDev B0 = fr[S.box[0]], B1 = fr[S.box[1]];
  const int* d = dofs + S.dof_off;
  const T* x0 = xs + B0.x_off + (i64)(B0.n + S.pos[0]) * RB;
  const T* x1 = xs + B1.x_off + (i64)(B1.n + S.pos[1]) * RB;
  for (int e = threadIdx.x; e < S.k * RB; e += 256) {
    T* wp = W + (i64)d[e >> 4] * RB + (e & 15);
    const T v = *wp + x0[e];
    *wp = v + x1[e];
  }
}

// Backward sweep of one front: X[0:n] = W[int] (forward result y), X[n:N] = W[bd];
// block by block from the last: z_j = y_j - L(below, j)^T X(below), x_j = Linv_j^T z_j
// (= A22^{-1} b_rd - XR b_sk overall); X[0:n] goes back to W[int].
template <class T>
__global__ __launch_bounds__(256) void k_solve_bwd(const int* __restrict__ ids, const FrontDev* __restrict__ fr,
                                                   const T* __restrict__ A, const int* __restrict__ dofs,
                                                   T* __restrict__ W, T* __restrict__ xs) {
  const FrontDev F = fr[ids[blockIdx.x]];
  const T* L = A + F.off;
  T* X = xs + F.x_off;
  const int* di = dofs + F.int_off;
  const int* db = dofs + F.bd_off;
  const int t = threadIdx.x, lane = t & 63, w = t >> 6, c16 = lane & 15, q = lane >> 4;
  gather_rhs(F, X, W, di, db);
  __shared__ T zb[NB][RB];
  const int nblk = (F.n + NB - 1) / NB;
  for (int bk = nblk - 1; bk >= 0; bk--) {
    const int jb = bk * NB, kb = min(NB, F.n - jb);
    __syncthreads();
    // T(c, r) = sum_{i >= jb+kb} L(i, jb+c) X(i, r); wave w owns columns c in [16w, 16w+16),
    // streaming 4-row slices straight from the column-major L (lanes: column c16, row q)
    const int col = 16 * w + c16;
    const T* Lc = L + (i64)(jb + min(col, kb - 1)) * F.N;
    const bool cin = col < kb;
    V4<T> acc = vzero<T>();
    for (int i0 = jb + kb; i0 < F.N; i0 += 16 * GU) {
      T av[GU][4], bv[GU][4];
#pragma unroll
      for (int u = 0; u < GU; u++)
#pragma unroll
        for (int k4 = 0; k4 < 4; k4++) {
          const int i = i0 + 16 * u + 4 * k4 + q;
          const bool ok = cin && i < F.N;
          const int ic = min(i, F.N - 1);
          av[u][k4] = ld_or0(Lc + ic, ok);
          bv[u][k4] = X[(i64)ic * RB + c16];
        }
#pragma unroll
      for (int u = 0; u < GU; u++)
#pragma unroll
        for (int k4 = 0; k4 < 4; k4++) acc = mfma4(av[u][k4], bv[u][k4], acc);
    }
#pragma unroll
    for (int r = 0; r < 4; r++) {
      const int c = 16 * w + mrow<T>(q, r);
      zb[c][c16] = ld_or0(X + (i64)(jb + min(c, kb - 1)) * RB + c16, c < kb) - (c < kb ? acc[r] : T(0));
    }
    __syncthreads();
    {  // x_j = Linv_j^T z: x(i, r) = sum_c Linv(c, i) z(c, r); wave w -> rows 16w..16w+15; the
       // 4-row slices above the wave's rows (Linv lower triangular) and past kb are zero: skipped
      const T* Li = A + F.dinv_off + (i64)bk * NB * NB;
      V4<T> a2 = vzero<T>();
      const int k4end = (kb + 3) >> 2;
#pragma unroll
      for (int k4 = 0; k4 < NB / 4; k4++) {
        if (k4 < 4 * w || k4 >= k4end) continue;
        const int c = 4 * k4 + q;
        a2 = mfma4(Li[(16 * w + c16) * NB + c], zb[c][c16], a2);
      }
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int i = 16 * w + mrow<T>(q, r);
        if (i < kb) X[(i64)(jb + i) * RB + c16] = a2[r];
      }
    }
  }
  __syncthreads();
#pragma unroll 4
  for (int e = t; e < F.n * RB; e += 256) W[(i64)di[e >> 4] * RB + (e & 15)] = X[e];
}

// ---- sweeps of small fronts (at most XLR rows), persistent launches over the level's queue of
// them, with the front's right-hand-side block in LDS instead of the global scratch: the only global traffic is the factor, the work block W and
// (forward) the boundary rows k_sep_acc reads; the steps synchronize with LDS-only barriers,
// so the next diagonal block's inverse, loaded during a step, stays in flight across them.
#define XLR 320   // rows of X held in LDS (2D leaf fronts of maxSideLength 8 have N = 285)
#define XLR2 1152  // FP32 fronts up to this many rows: the larger instance (74 KB of LDS)

// Workgroup barrier for LDS data: LDS writes before it are seen by LDS reads after it; global
// loads in flight stay in flight (__syncthreads would wait for them).
__device__ __forceinline__ void lds_barrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// Forward sweep (as k_solve_fwd) of fronts with N <= XR.
template <class T, int XR = XLR>
__global__ __launch_bounds__(256) void k_solve_fwd_l(const int* __restrict__ ids, int cnt, int* __restrict__ queue,
                                                     const FrontDev* __restrict__ fr, const T* __restrict__ A,
                                                     const int* __restrict__ dofs, T* __restrict__ W,
                                                     T* __restrict__ xs) {
  __shared__ T X[XR * RB];
  __shared__ T xb[NB][RB];
  for (int item; (item = queue_next(queue, cnt)) >= 0;) {
    const FrontDev F = fr[ids[cnt - 1 - item]];  // the list ascends in size: largest first
    const T* L = A + F.off;
    const int* di = dofs + F.int_off;
    const int t = threadIdx.x, lane = t & 63, w = wave_id(), c16 = lane & 15, q = lane >> 4;
  #pragma unroll 4
    for (int e = t; e < F.n * RB; e += 256) X[e] = W[(i64)di[e >> 4] * RB + (e & 15)];
    for (int e = F.n * RB + t; e < F.N * RB; e += 256) X[e] = T(0);
    // the wave's slices of a diagonal block's inverse (lower triangular: slices right of the
    // wave's rows and past kb are zero and not loaded)
    T li[NB / 4];
    auto load_li = [&](int jb) {
      const T* Li = A + F.dinv_off + (i64)(jb / NB) * NB * NB;
      const int k4end = min(4 * (w + 1), (min(NB, F.n - jb) + 3) >> 2);
  #pragma unroll
      for (int k4 = 0; k4 < NB / 4; k4++)
        if (k4 < k4end) li[k4] = Li[(4 * k4 + q) * NB + 16 * w + c16];
    };
    load_li(0);
    for (int jb = 0; jb < F.n; jb += NB) {
      const int kb = min(NB, F.n - jb);
      const int k4end = min(4 * (w + 1), (kb + 3) >> 2);
      lds_barrier();
      V4<T> acc = vzero<T>();  // x_j = Linv_j b_j: wave w -> rows 16w..16w+15
  #pragma unroll
      for (int k4 = 0; k4 < NB / 4; k4++) {
        if (k4 >= k4end) break;
        const int k = 4 * k4 + q;
        acc = mfma4(li[k4], k < kb ? X[(jb + k) * RB + c16] : T(0), acc);
      }
      if (jb + NB < F.n) load_li(jb + NB);
      lds_barrier();
  #pragma unroll
      for (int r = 0; r < 4; r++) {
        const int i = 16 * w + mrow<T>(q, r);
        xb[i][c16] = acc[r];
        if (i < kb) X[(jb + i) * RB + c16] = acc[r];
      }
      lds_barrier();
      // rows below: X(i) -= L(i, block) x_j, one 16-row group per wave at a time (the inverse's
      // slices in flight take the registers of a second)
      constexpr int G = 1;
      for (int g0 = jb + kb + 16 * w; g0 < F.N; g0 += 64 * G) {
        T av[G][NB / 4];
  #pragma unroll
        for (int u = 0; u < G; u++) {
          const int row = min(g0 + 64 * u + c16, F.N - 1);
  #pragma unroll
          for (int k4 = 0; k4 < NB / 4; k4++) {
            const int k = 4 * k4 + q;
            av[u][k4] = ld_or0(L + (i64)(jb + min(k, kb - 1)) * F.N + row, k < kb);
          }
        }
  #pragma unroll
        for (int u = 0; u < G; u++) {
          V4<T> a2 = vzero<T>();
  #pragma unroll
          for (int k4 = 0; k4 < NB / 4; k4++) a2 = mfma4(av[u][k4], xb[4 * k4 + q][c16], a2);
  #pragma unroll
          for (int r = 0; r < 4; r++) {
            const int i = g0 + 64 * u + mrow<T>(q, r);
            if (i < F.N) X[i * RB + c16] -= a2[r];
          }
        }
      }
    }
    lds_barrier();
    T* Xg = xs + F.x_off;  // boundary rows for k_sep_acc
  #pragma unroll 4
    for (int e = t; e < F.n * RB; e += 256) W[(i64)di[e >> 4] * RB + (e & 15)] = X[e];
    for (int e = F.n * RB + t; e < F.N * RB; e += 256) Xg[e] = X[e];
  }
}

// Backward sweep (as k_solve_bwd) of fronts with N <= XR.
template <class T, int XR = XLR>
__global__ __launch_bounds__(256) void k_solve_bwd_l(const int* __restrict__ ids, int cnt, int* __restrict__ queue,
                                                     const FrontDev* __restrict__ fr, const T* __restrict__ A,
                                                     const int* __restrict__ dofs, T* __restrict__ W) {
  __shared__ T X[XR * RB];
  __shared__ T zb[NB][RB];
  for (int item; (item = queue_next(queue, cnt)) >= 0;) {
    const FrontDev F = fr[ids[cnt - 1 - item]];  // the list ascends in size: largest first
    const T* L = A + F.off;
    const int* di = dofs + F.int_off;
    const int* db = dofs + F.bd_off;
    const int t = threadIdx.x, lane = t & 63, w = wave_id(), c16 = lane & 15, q = lane >> 4;
  #pragma unroll 4
    for (int e = t; e < F.n * RB; e += 256) X[e] = W[(i64)di[e >> 4] * RB + (e & 15)];
  #pragma unroll 4
    for (int e = F.n * RB + t; e < F.N * RB; e += 256) X[e] = W[(i64)db[(e >> 4) - F.n] * RB + (e & 15)];
    const int nblk = (F.n + NB - 1) / NB;
    for (int bk = nblk - 1; bk >= 0; bk--) {
      const int jb = bk * NB, kb = min(NB, F.n - jb);
      const int k4lo = 4 * w, k4end = (kb + 3) >> 2;
      // the wave's slices of Linv_j^T (rows 16w.., columns c >= 16w), loaded first: in flight
      // during the product below
      const T* Li = A + F.dinv_off + (i64)bk * NB * NB;
      T li[NB / 4];
  #pragma unroll
      for (int k4 = 0; k4 < NB / 4; k4++)
        if (k4 >= k4lo && k4 < k4end) li[k4] = Li[(16 * w + c16) * NB + 4 * k4 + q];
      lds_barrier();
      // T(c, r) = sum_{i >= jb+kb} L(i, jb+c) X(i, r); wave w owns columns [16w, 16w+16)
      const int col = 16 * w + c16;
      const T* Lc = L + (i64)(jb + min(col, kb - 1)) * F.N;
      const bool cin = col < kb;
      V4<T> acc = vzero<T>();
      for (int i0 = jb + kb; i0 < F.N; i0 += 16 * GU) {
        T av[GU][4], bv[GU][4];
  #pragma unroll
        for (int u = 0; u < GU; u++)
  #pragma unroll
          for (int k4 = 0; k4 < 4; k4++) {
            const int i = i0 + 16 * u + 4 * k4 + q;
            const int ic = min(i, F.N - 1);
            av[u][k4] = ld_or0(Lc + ic, cin && i < F.N);
            bv[u][k4] = X[ic * RB + c16];
          }
  #pragma unroll
        for (int u = 0; u < GU; u++)
  #pragma unroll
          for (int k4 = 0; k4 < 4; k4++) acc = mfma4(av[u][k4], bv[u][k4], acc);
      }
  #pragma unroll
      for (int r = 0; r < 4; r++) {
        const int c = 16 * w + mrow<T>(q, r);
        zb[c][c16] = c < kb ? X[(jb + c) * RB + c16] - acc[r] : T(0);
      }
      lds_barrier();
      {  // x_j = Linv_j^T z: wave w -> rows 16w..16w+15
        V4<T> a2 = vzero<T>();
  #pragma unroll
        for (int k4 = 0; k4 < NB / 4; k4++) {
          if (k4 < k4lo || k4 >= k4end) continue;
          a2 = mfma4(li[k4], zb[4 * k4 + q][c16], a2);
        }
  #pragma unroll
        for (int r = 0; r < 4; r++) {
          const int i = 16 * w + mrow<T>(q, r);
          if (i < kb) X[(jb + i) * RB + c16] = a2[r];
        }
      }
    }
    lds_barrier();
  #pragma unroll 4
    for (int e = t; e < F.n * RB; e += 256) W[(i64)di[e >> 4] * RB + (e & 15)] = X[e];
  }
}

// ---- forward sweep with the right-hand-side rows in registers (fronts of at most 64 G rows):
// wave w owns rows 16w..16w+15 of every 64-row chunk of the front; chunk c of those rows lives
// in a lane's accumulator c (row 64c + 16w + mrow(lane/16, r), right-hand side lane%16), so no
// row is written back before it is final.
// S(rows) = sum_j L(rows, block j) x_j accumulates over the blocks; block j's own rows
// go through LDS once (v = b_j - S_j, x_j = Linv_j v); at the end X[n:N] = -S (the boundary
// contributions k_sep_acc gathers) and W[int] holds y.
template <class T, int G>
__global__ __launch_bounds__(256, 2) void k_solve_fwd_r(const int* __restrict__ ids, const FrontDev* __restrict__ fr,
                                                     const T* __restrict__ A, const int* __restrict__ dofs,
                                                     T* __restrict__ W, T* __restrict__ xs) {
  const FrontDev F = fr[ids[blockIdx.x]];
  const T* L = A + F.off;
  T* X = xs + F.x_off;
  const int* di = dofs + F.int_off;
  const int t = threadIdx.x, lane = t & 63, w = t >> 6, c16 = lane & 15, q = lane >> 4;
  __shared__ T sb[NB][RB];  // v = b_j - S_j
  __shared__ T xb[NB][RB];  // x_j
  V4<T> S[G];
#pragma unroll
  for (int c = 0; c < G; c++) S[c] = vzero<T>();
  const int nch = (F.N + NB - 1) / NB;
  for (int j = 0, jb = 0; jb < F.n; j++, jb += NB) {
    const int kb = min(NB, F.n - jb);
#pragma unroll
    for (int c = 0; c < G; c++)
      if (c == j)
#pragma unroll
        for (int r = 0; r < 4; r++) {
          const int i = 16 * w + mrow<T>(q, r);
          sb[i][c16] = ld_or0(W + (i64)di[jb + min(i, kb - 1)] * RB + c16, i < kb) - (i < kb ? S[c][r] : T(0));
        }
    __syncthreads();
    {  // x_j = Linv_j v: wave w -> rows 16w..16w+15
      const T* Li = A + F.dinv_off + (i64)j * NB * NB;  // Linv(i, c) at Li[c*NB + i]
      V4<T> a = vzero<T>();
#pragma unroll
      for (int k4 = 0; k4 < NB / 4; k4++) {
        const int c = 4 * k4 + q;
        a = mfma4(Li[c * NB + 16 * w + c16], sb[c][c16], a);
      }
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int i = 16 * w + mrow<T>(q, r);
        xb[i][c16] = a[r];
        if (i < kb) W[(i64)di[jb + i] * RB + c16] = a[r];
      }
    }
    __syncthreads();
    // S(rows below the block) += L(rows, block j) x_j
#pragma unroll
    for (int c = 0; c < G; c++) {
      if (c < j || c >= nch) continue;
      const int row = NB * c + 16 * w + c16;
      const bool rok = row < F.N && row >= jb + kb;
      const int rowc = min(row, F.N - 1);
      T av[NB / 4];
#pragma unroll
      for (int k4 = 0; k4 < NB / 4; k4++) {
        const int k = 4 * k4 + q;
        av[k4] = ld_or0(L + (i64)(jb + min(k, kb - 1)) * F.N + rowc, rok && k < kb);
      }
#pragma unroll
      for (int k4 = 0; k4 < NB / 4; k4++) S[c] = mfma4(av[k4], xb[4 * k4 + q][c16], S[c]);
    }
  }
#pragma unroll
  for (int c = 0; c < G; c++)
#pragma unroll
    for (int r = 0; r < 4; r++) {
      const int i = NB * c + 16 * w + mrow<T>(q, r);
      if (i >= F.n && i < F.N) X[(i64)i * RB + c16] = -S[c][r];
    }
}

// ---- blocked sweeps for large fronts: one workgroup per 64-row block, blocks of a front
// hand their results to one another inside the launch.
// The grid is one-dimensional and interleaved over the launch's fronts: workgroup id
// s * nf + y owns slot s of front y (forward: interior slots first, then the boundary slots),
// so every front's chain starts in the first wave of workgroups.  A workgroup only waits for
// workgroups with smaller ids, which the hardware dispatches first, and every wait is bounded:
// past the limit err[3] is set and the workgroup goes on, so the launch always drains.
//
// Publication, SYNC = 0 (the default): flag-in-data.  A block's result x_b (64 rows x 16
// columns) is written as 64-bit words, each holding one 32-bit piece of a value in its low
// half and the launch's epoch in its high half (FP64 values take two words), with relaxed
// agent-scope atomic stores; a consumer polls the words themselves with relaxed agent-scope
// atomic loads until every word it needs carries the epoch.  A 64-bit atomic access is
// single-copy atomic, so a word that carries the epoch carries this launch's value: the
// handoff is ordered by the memory model alone, needs no flag, no store completion wait and
// no workgroup barrier on either side, and costs one store-to-load latency per block.
// SYNC = 1: the result goes to X with relaxed atomic stores, then a per-block flag is set
// after an agent-scope release fence and read before an agent-scope acquire fence (the L2
// write-back / invalidate they imply makes it slower); the parity tests compare the two.
template <class T>
constexpr int kLLW = (int)sizeof(T) / 4;  // 64-bit words per published value
constexpr int kLLBlock = NB * RB * 2;      // words per published block (room for FP64)

__device__ __forceinline__ void ll_put32(unsigned long long* p, unsigned v, int epoch) {
  __hip_atomic_store(p, ((unsigned long long)(unsigned)epoch << 32) | v, __ATOMIC_RELAXED,
                     __HIP_MEMORY_SCOPE_AGENT);
}
template <class T>
__device__ __forceinline__ void ll_put(unsigned long long* p, T v, int epoch) {
  if constexpr (sizeof(T) == 4) {
    ll_put32(p, __builtin_bit_cast(unsigned, v), epoch);
  } else {
    const unsigned long long b = __builtin_bit_cast(unsigned long long, v);
    ll_put32(p, (unsigned)b, epoch);
    ll_put32(p + 1, (unsigned)(b >> 32), epoch);
  }
}
// this lane's 4 values of wave w's quarter of a published 64 x 16 block (row 16 w + 4 s + q,
// column c16; rows >= kn read as 0), polled until every word carries `epoch`.  Each poll
// issues all of the lane's loads together (rows past kn load row kn - 1 and are ignored).
template <class T>
__device__ __forceinline__ void ll_quarter(const unsigned long long* blk, int w, int kn, int q, int c16, int epoch,
                                           T (&bv)[4], int* err) {
  unsigned long long* base = const_cast<unsigned long long*>(blk) + c16 * kLLW<T>;
  for (int spins = 0;;) {
    unsigned long long w0[4], w1[4];
#pragma unroll
    for (int s4 = 0; s4 < 4; s4++) {
      unsigned long long* a = base + min(16 * w + 4 * s4 + q, kn - 1) * (RB * kLLW<T>);
      w0[s4] = __hip_atomic_load(a, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      if constexpr (sizeof(T) == 8) w1[s4] = __hip_atomic_load(a + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    bool ok = true;
#pragma unroll
    for (int s4 = 0; s4 < 4; s4++) {
      const bool live = 16 * w + 4 * s4 + q < kn;
      if constexpr (sizeof(T) == 4) {
        ok = ok && (!live || (int)(w0[s4] >> 32) == epoch);
        bv[s4] = live ? __builtin_bit_cast(T, (unsigned)w0[s4]) : T(0);
      } else {
        ok = ok && (!live || ((int)(w0[s4] >> 32) == epoch && (int)(w1[s4] >> 32) == epoch));
        bv[s4] = live ? __builtin_bit_cast(T, (w0[s4] & 0xffffffffull) | (w1[s4] << 32)) : T(0);
      }
    }
    if (ok) break;
    if (++spins > SPIN_LIMIT) {
      atomicCAS(err + 3, 0, 1);
      break;
    }
    __builtin_amdgcn_s_sleep(1);
  }
}
__device__ __forceinline__ void wait_block_fenced(const int* flag, int epoch, int* err) {
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
__device__ __forceinline__ void post_block_fenced(int* flag, int epoch) {
  __syncthreads();
  if (threadIdx.x == 0) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    __hip_atomic_store(flag, epoch, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
}
template <class T>
__device__ __forceinline__ T ld_pub(const T* p) {
  return __hip_atomic_load(const_cast<T*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <class T>
__device__ __forceinline__ void st_pub(T* p, T v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// acc[rb] += (the factor tile av: 64 output rows x this wave's 16 of block j's rows) x (those
// published rows: 16 w + 4 s + q from k0, column c16).  Each wave takes a quarter of the block's
// rows for all 64 output rows, so a workgroup reads each published block once.
template <class T, int SYNC>
__device__ __forceinline__ void acc_block(const unsigned long long* ll, const T* X, const int* fl, int j, int k0,
                                          int kn, int w, int q, int c16, int epoch, const T (&av)[4][4],
                                          V4<T> (&acc)[4], int* err) {
  T bv[4];
  if constexpr (SYNC == 0) {
    ll_quarter<T>(ll + (i64)j * kLLBlock, w, kn, q, c16, epoch, bv, err);
  } else {
    wait_block_fenced(fl + j, epoch, err);
#pragma unroll
    for (int s4 = 0; s4 < 4; s4++) {
      const int r = 16 * w + 4 * s4 + q;
      const T v = ld_pub(X + (i64)(k0 + min(r, kn - 1)) * RB + c16);
      bv[s4] = r < kn ? v : T(0);
    }
  }
#pragma unroll
  for (int rb = 0; rb < 4; rb++)
#pragma unroll
    for (int s4 = 0; s4 < 4; s4++) acc[rb] = mfma4(av[rb][s4], bv[s4], acc[rb]);
}
// sum of the four waves' partial 64 x 16 blocks (red: [wave][row][column]); wave w gets rows
// 16 w + mrow(q, r) in acc[w]'s layout.  Every wave of the workgroup calls it.
template <class T>
__device__ __forceinline__ V4<T> reduce_waves(T (*red)[NB][RB], const V4<T> (&acc)[4], int w, int q, int c16) {
#pragma unroll
  for (int rb = 0; rb < 4; rb++)
#pragma unroll
    for (int r = 0; r < 4; r++) red[w][16 * rb + mrow<T>(q, r)][c16] = acc[rb][r];
  __syncthreads();
  V4<T> s = vzero<T>();
#pragma unroll
  for (int r = 0; r < 4; r++) {
    const int i = 16 * w + mrow<T>(q, r);
    s[r] = ((red[0][i][c16] + red[1][i][c16]) + red[2][i][c16]) + red[3][i][c16];
  }
  return s;
}

// Forward sweep of one front by blocks: interior block b waits for the blocks j < b, forms
// v = b_b - sum_j L_bj x_j and publishes x_b = Linv_b v (and writes it to W[int]); boundary
// block b accumulates over every interior block and leaves X[rows] = -sum_j L_bj x_j, the
// contribution gathered by k_sep_acc.  nbi_g: the most interior blocks of a front of the
// launch (the grid holds nf x (nbi_g + the most boundary blocks) workgroups); ll: the launch's
// published blocks, kLLBlock words per (front, block).
template <class T, int SYNC>
__global__ __launch_bounds__(256, sizeof(T) == 4 ? 4 : 3) void k_fwd_blk(const int* __restrict__ ids, int nf, int nbi_g,
                                                 const FrontDev* __restrict__ fr, const T* __restrict__ A,
                                                 const int* __restrict__ dofs, T* __restrict__ W, T* __restrict__ xs,
                                                 int* flags, unsigned long long* ll, int epoch, int* err) {
  const int id = blockIdx.x;
  const bool interior = id < nf * nbi_g;
  const int rid = interior ? id : id - nf * nbi_g;
  const int slot = rid / nf, y = rid - slot * nf;
  const FrontDev F = fr[ids[y]];
  const int nbi = (F.n + NB - 1) / NB, nbb = (F.m + NB - 1) / NB;
  if (interior ? slot >= nbi : slot >= nbb) return;
  const int b = interior ? slot : nbi + slot;
  const int r0 = interior ? NB * b : F.n + NB * slot;
  const int rend = min(interior ? F.n : F.N, r0 + NB);
  const T* L = A + F.off;
  T* X = xs + F.x_off;
  const int* di = dofs + F.int_off;
  int* fl = flags + F.flag_off;
  unsigned long long* llf = ll + (i64)y * nbi_g * kLLBlock;
  const int t = threadIdx.x, lane = t & 63, w = t >> 6, c16 = lane & 15, q = lane >> 4;
  __shared__ T red[4][NB][RB];
  __shared__ T vb[NB][RB];
  int rowc[4];  // L rows supplied by this lane, one per 16-row block (clamped; `rok` masks)
  bool rok[4];
#pragma unroll
  for (int rb = 0; rb < 4; rb++) {
    const int row = r0 + 16 * rb + c16;
    rok[rb] = row < rend;
    rowc[rb] = rok[rb] ? row : r0;
  }
  V4<T> acc[4];
#pragma unroll
  for (int rb = 0; rb < 4; rb++) acc[rb] = vzero<T>();
  const int jmax = interior ? b : nbi;
  // the factor data (this block's inverse) and this block's right-hand-side rows (written by no
  // other block of the launch; kept in vb, each lane its own entries) are read before any wait
  T li[NB / 4];
  if (interior) {
    const T* Li = A + F.dinv_off + (i64)b * NB * NB;  // Linv(i, c) at Li[c*NB + i]
#pragma unroll
    for (int k4 = 0; k4 < NB / 4; k4++) li[k4] = Li[(4 * k4 + q) * NB + 16 * w + c16];
#pragma unroll
    for (int r = 0; r < 4; r++)
      vb[16 * w + mrow<T>(q, r)][c16] = W[(i64)di[min(r0 + 16 * w + mrow<T>(q, r), rend - 1)] * RB + c16];
  }
  for (int j = 0; j < jmax; j++) {
    const int jb = NB * j, kb = min(NB, F.n - jb);
    T av[4][4];
#pragma unroll
    for (int s4 = 0; s4 < 4; s4++) {  // the factor tile: issued before block j's wait
      const int k = 16 * w + 4 * s4 + q;
      const T* Lk = L + (i64)(jb + min(k, kb - 1)) * F.N;
#pragma unroll
      for (int rb = 0; rb < 4; rb++) av[rb][s4] = ld_or0(Lk + rowc[rb], rok[rb] && k < kb);
    }
    acc_block<T, SYNC>(llf, X, fl, j, jb, kb, w, q, c16, epoch, av, acc, err);
  }
  const V4<T> tot = reduce_waves<T>(red, acc, w, q, c16);
  if (!interior) {
#pragma unroll
    for (int r = 0; r < 4; r++) {
      const int i = r0 + 16 * w + mrow<T>(q, r);
      if (i < rend) X[(i64)i * RB + c16] = -tot[r];
    }
    return;
  }
#pragma unroll
  for (int r = 0; r < 4; r++) {
    const int i = 16 * w + mrow<T>(q, r);
    vb[i][c16] = r0 + i < rend ? vb[i][c16] - tot[r] : T(0);
  }
  __syncthreads();
  V4<T> a2 = vzero<T>();
#pragma unroll
  for (int k4 = 0; k4 < NB / 4; k4++) a2 = mfma4(li[k4], vb[4 * k4 + q][c16], a2);
  unsigned long long* dst = llf + (i64)b * kLLBlock;
#pragma unroll
  for (int r = 0; r < 4; r++) {
    const int i = 16 * w + mrow<T>(q, r);
    if (r0 + i < rend) {
      if constexpr (SYNC == 0)
        ll_put<T>(dst + (i * RB + c16) * kLLW<T>, a2[r], epoch);
      else
        st_pub(X + (i64)(r0 + i) * RB + c16, a2[r]);
    }
  }
  if constexpr (SYNC == 1) post_block_fenced(fl + b, epoch);
#pragma unroll
  for (int r = 0; r < 4; r++) {  // after the publication: W is read by later launches only
    const int i = 16 * w + mrow<T>(q, r);
    if (r0 + i < rend) W[(i64)di[r0 + i] * RB + c16] = a2[r];
  }
}

// Backward sweep of one front by blocks, last block first (slot s owns block nbi-1-s):
// z = y_i - L(rows below, i)^T x(rows below), boundary rows (read from W[bd]) first, then the
// interior blocks below as they are published; x_i = Linv_i^T z is published and written to
// W[int].  Over the boundary rows each wave takes 16 of the block's columns; over the interior
// blocks below, as in the forward sweep, each wave takes a quarter of every published block for
// all 64 columns, and the waves' partial sums are added at the end.
template <class T, int SYNC>
__global__ __launch_bounds__(256, sizeof(T) == 4 ? 4 : 3) void k_bwd_blk(const int* __restrict__ ids, int nf, int nbi_g,
                                                 const FrontDev* __restrict__ fr, const T* __restrict__ A,
                                                 const int* __restrict__ dofs, T* __restrict__ W, T* __restrict__ xs,
                                                 int* flags, unsigned long long* ll, int epoch, int* err) {
  const int slot = blockIdx.x / nf, y = blockIdx.x - slot * nf;
  const FrontDev F = fr[ids[y]];
  const int nbi = (F.n + NB - 1) / NB;
  if (slot >= nbi) return;
  const int ib = nbi - 1 - slot;
  const int jb = NB * ib, kb = min(NB, F.n - jb);
  const T* L = A + F.off;
  T* X = xs + F.x_off;
  const int* di = dofs + F.int_off;
  const int* db = dofs + F.bd_off;
  int* fl = flags + F.flag_off;
  unsigned long long* llf = ll + (i64)y * nbi_g * kLLBlock;
  const int t = threadIdx.x, lane = t & 63, w = t >> 6, c16 = lane & 15, q = lane >> 4;
  __shared__ T red[4][NB][RB];
  __shared__ T zb[NB][RB];
  const T* Lc[4];  // the block's columns supplied by this lane, one per 16-column block
  bool cin[4];
#pragma unroll
  for (int rb = 0; rb < 4; rb++) {
    const int col = 16 * rb + c16;
    cin[rb] = col < kb;
    Lc[rb] = L + (i64)(jb + (cin[rb] ? col : 0)) * F.N;
  }
  // this block's right-hand-side rows (written by no other block of the launch), kept in zb
  // (each lane its own entries) until the end
  T wv[4];
#pragma unroll
  for (int r = 0; r < 4; r++) wv[r] = W[(i64)di[jb + min(16 * w + mrow<T>(q, r), kb - 1)] * RB + c16];
  // boundary rows (x = W[bd]): wave w takes the block's columns 16 w .. 16 w + 15 over all rows
  V4<T> accb = vzero<T>();
  {
    const T* Lw = Lc[0] + (i64)16 * w * F.N;  // column jb + 16 w + c16 (clamped to jb when past kb)
    const bool cw = 16 * w + c16 < kb;
    if (!cw) Lw = L + (i64)jb * F.N;
    for (int i0 = F.n; i0 < F.N; i0 += 16 * GU) {
      T av[GU][4], bv[GU][4];
#pragma unroll
      for (int u = 0; u < GU; u++)
#pragma unroll
        for (int k4 = 0; k4 < 4; k4++) {
          const int i = i0 + 16 * u + 4 * k4 + q;
          const bool ok = i < F.N;
          const int ic = min(i, F.N - 1);
          av[u][k4] = ld_or0(Lw + ic, ok && cw);
          bv[u][k4] = ld_or0(W + (i64)db[ic - F.n] * RB + c16, ok);
        }
#pragma unroll
      for (int u = 0; u < GU; u++)
#pragma unroll
        for (int k4 = 0; k4 < 4; k4++) accb = mfma4(av[u][k4], bv[u][k4], accb);
    }
  }
#pragma unroll
  for (int r = 0; r < 4; r++) zb[16 * w + mrow<T>(q, r)][c16] = wv[r] - accb[r];
  V4<T> acc[4];
#pragma unroll
  for (int rb = 0; rb < 4; rb++) acc[rb] = vzero<T>();
  // this block's inverse (factor data, final before the launch) is read before any wait
  T li[NB / 4];
  {
    const T* Li = A + F.dinv_off + (i64)ib * NB * NB;
#pragma unroll
    for (int k4 = 0; k4 < NB / 4; k4++) li[k4] = Li[(16 * w + c16) * NB + 4 * k4 + q];
  }
  for (int kbk = nbi - 1; kbk > ib; kbk--) {  // interior blocks below, in publication order
    const int k0 = NB * kbk, kn = min(NB, F.n - k0);
    T av[4][4];
#pragma unroll
    for (int s4 = 0; s4 < 4; s4++) {  // the factor tile: issued before block kbk's wait
      const int k = 16 * w + 4 * s4 + q;
#pragma unroll
      for (int rb = 0; rb < 4; rb++) av[rb][s4] = ld_or0(Lc[rb] + k0 + min(k, kn - 1), k < kn && cin[rb]);
    }
    acc_block<T, SYNC>(llf, X, fl, kbk, k0, kn, w, q, c16, epoch, av, acc, err);
  }
  const V4<T> tot = reduce_waves<T>(red, acc, w, q, c16);
#pragma unroll
  for (int r = 0; r < 4; r++) {
    const int c = 16 * w + mrow<T>(q, r);
    zb[c][c16] = c < kb ? zb[c][c16] - tot[r] : T(0);
  }
  __syncthreads();
  V4<T> a2 = vzero<T>();
#pragma unroll
  for (int k4 = 0; k4 < NB / 4; k4++) a2 = mfma4(li[k4], zb[4 * k4 + q][c16], a2);
  unsigned long long* dst = llf + (i64)ib * kLLBlock;
#pragma unroll
  for (int r = 0; r < 4; r++) {
    const int i = 16 * w + mrow<T>(q, r);
    if (i < kb) {
      if constexpr (SYNC == 0)
        ll_put<T>(dst + (i * RB + c16) * kLLW<T>, a2[r], epoch);
      else
        st_pub(X + (i64)(jb + i) * RB + c16, a2[r]);
    }
  }
  if constexpr (SYNC == 1) post_block_fenced(fl + ib, epoch);
#pragma unroll
  for (int r = 0; r < 4; r++) {  // after the publication: W is read by later launches only
    const int i = 16 * w + mrow<T>(q, r);
    if (i < kb) W[(i64)di[jb + i] * RB + c16] = a2[r];
  }
}

// Column-major user block (n x nr, ld) <-> row-major work block (n x 16); unused columns zero.
// One workgroup per 64 rows, transposed through LDS so both sides are read and written in
// contiguous runs (64 rows of a column; 64 x 16 consecutive work-block entries).
template <class T>
__global__ __launch_bounds__(256) void k_to_rows(int n, const double* __restrict__ B, int ld, int nr,
                                                 T* __restrict__ W) {
  __shared__ double s[64][RB + 1];
  const int i0 = blockIdx.x * 64;
  for (int e = threadIdx.x; e < 64 * RB; e += 256) {
    const int r = e >> 6, li = e & 63;
    s[li][r] = (r < nr && i0 + li < n) ? B[(i64)r * ld + i0 + li] : 0.0;
  }
  __syncthreads();
  for (int e = threadIdx.x; e < 64 * RB; e += 256) {
    const int li = e / RB, r = e % RB;
    if (i0 + li < n) W[(i64)(i0 + li) * RB + r] = (T)s[li][r];
  }
}
template <class T>
__global__ __launch_bounds__(256) void k_from_rows(int n, const T* __restrict__ W, double* __restrict__ B,
                                                   int ld, int nr) {
  __shared__ double s[64][RB + 1];
  const int i0 = blockIdx.x * 64;
  for (int e = threadIdx.x; e < 64 * RB; e += 256) {
    const int li = e / RB, r = e % RB;
    s[li][r] = i0 + li < n ? (double)W[(i64)(i0 + li) * RB + r] : 0.0;
  }
  __syncthreads();
  for (int e = threadIdx.x; e < 64 * nr; e += 256) {
    const int r = e >> 6, li = e & 63;
    if (i0 + li < n) B[(i64)r * ld + i0 + li] = s[li][r];
  }
}
// Element-wise conversion between the FP64 and the storage-type work blocks.
template <class To, class From>
__global__ __launch_bounds__(256) void k_cvt(i64 n, const From* __restrict__ a, To* __restrict__ b) {
  const i64 e = (i64)blockIdx.x * 256 + threadIdx.x;
  if (e < n) b[e] = (To)a[e];
}

// ----------------------------------------------------------------------------- apply
// Upward step of Apply for one front: u = [L11; L21]^T [x_rd; x_sk] (= L11^T b_rd + W b_sk),
// b_rd <- L11 u.  u is kept in X[0:n] for the downward step.
template <class T>
__global__ __launch_bounds__(256) void k_apply_up(const int* __restrict__ ids, const FrontDev* __restrict__ fr,
                                                  const T* __restrict__ A, const int* __restrict__ dofs,
                                                  T* __restrict__ W, T* __restrict__ xs, int nr) {
  const FrontDev F = fr[ids[blockIdx.x]];
  const T* L = A + F.off;
  T* X = xs + F.x_off;
  const int* di = dofs + F.int_off;
  const int* db = dofs + F.bd_off;
  const int t = threadIdx.x;
  gather_rhs(F, X, W, di, db);
  __syncthreads();
  // u_c = sum_{i >= c} L[i][c] X[i]: one column at a time, block reduction
  __shared__ T red[256];
  for (int c = 0; c < F.n; c++) {
    for (int r = 0; r < nr; r++) {
      T s = T(0);
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
      T s = T(0);
      for (int c = 0; c <= i; c++) s += L[(i64)c * F.N + i] * X[(i64)c * RB + r];
      W[(i64)di[i] * RB + r] = s;
    }
}

// Downward step of Apply: X[n:N] = L21 u (the reference's XL b_rd), accumulated by k_sep_acc.
template <class T>
__global__ __launch_bounds__(256) void k_apply_down(const int* __restrict__ ids, const FrontDev* __restrict__ fr,
                                                    const T* __restrict__ A, T* __restrict__ xs, int nr) {
  const FrontDev F = fr[ids[blockIdx.x]];
  const T* L = A + F.off;
  T* X = xs + F.x_off;
  for (int i = F.n + threadIdx.x; i < F.N; i += 256)
    for (int r = 0; r < RB; r++) {
      T s = T(0);
      if (r < nr)
        for (int c = 0; c < F.n; c++) s += L[(i64)c * F.N + i] * X[(i64)c * RB + r];
      X[(i64)i * RB + r] = s;
    }
}

// ----------------------------------------------------------------------------- LU fronts
// Unsymmetric (pattern-symmetric) operators: each front is factored by LU with partial pivoting
// restricted to its interior rows, as the reference factors A22 with dgetrf and forms
// XL = A12 A22^{-1}, XR = A22^{-1} A21, S = -A12 XR (include/LinAlg/dense.hpp:180-227,
// src/DE/mf2d.cpp:541-576).  Right-looking over 64-column panels of the whole N x N front:
//   k_lu_panel: pivot search in column c among rows [c, n), interchange within the panel,
//               scale rows (c, N) (the boundary rows give L21 = A12 U^{-1}), rank-1 update of
//               the panel; then the panel's diagonal block's unit-lower and upper inverses;
//   k_lu_swap:  the panel's interchanges applied to every other column of the front;
//   k_lu_trsm:  U row block = Linv_kk A(panel rows, later columns), also stored transposed;
//   k_lu_gemm:  trailing update A(s0.., s0..) -= L(s0.., panel) U(panel, s0..) on MFMA, the
//               boundary block accumulating S.
// FP64 only.  A zero pivot records the first failing (front, column) in err[0..2].

// Zero every entry of each listed matrix.
template <class T>
__global__ __launch_bounds__(256) void k_zero_full(const int* __restrict__ ids, const MatDev* __restrict__ mats,
                                                   T* __restrict__ A) {
  const MatDev M = mats[ids[blockIdx.y]];
  T* a = A + M.off;
  for (int c = blockIdx.x; c < M.cols; c += gridDim.x) {
    T* col = a + (i64)c * M.ld;
    for (int r = threadIdx.x; r < M.ld; r += 256) col[r] = T(0);
  }
}

// Separator Schur accumulation on the full k x k matrix, box 0 first (ScatterAdd_,
// src/DE/mf2d.cpp:599-606).
template <class T>
__global__ __launch_bounds__(256) void k_sep_add_full(const int* __restrict__ ids, const SepDev* __restrict__ seps,
                                                      const FrontDev* __restrict__ fr, T* __restrict__ A) {
  const SepDev S = seps[ids[blockIdx.y]];
  const FrontDev B0 = fr[S.box[0]], B1 = fr[S.box[1]];
  T* a = A + S.off;
  const T* s0 = A + B0.off + (i64)(B0.n + S.pos[0]) * B0.N + B0.n + S.pos[0];
  const T* s1 = A + B1.off + (i64)(B1.n + S.pos[1]) * B1.N + B1.n + S.pos[1];
  for (int c = blockIdx.x; c < S.k; c += gridDim.x)
    for (int r = threadIdx.x; r < S.k; r += 256) {
      T v = a[(i64)c * S.k + r] + s0[(i64)c * B0.N + r];
      a[(i64)c * S.k + r] = v + s1[(i64)c * B1.N + r];
    }
}

__global__ __launch_bounds__(256) void k_lu_panel(const int* __restrict__ ids, const FrontDev* __restrict__ fr,
                                                  double* __restrict__ A, int k0, int* __restrict__ ipiv,
                                                  double* __restrict__ uinv, int* __restrict__ err) {
  const int fid = ids[blockIdx.x];
  const FrontDev F = fr[fid];
  if (k0 >= F.n) return;
  const int kb = min(NB, F.n - k0), N = F.N, t = threadIdx.x;
  double* a = A + F.off;
  __shared__ double rv[256];
  __shared__ int ri[256];
  __shared__ double D[NB][NB + 1];
  for (int j = 0; j < kb; j++) {
    const int c = k0 + j;
    double* col = a + (i64)c * N;
    // pivot: the first row of largest |a_ic| among the interior rows [c, n)
    double best = -1.0;
    int bi = F.n;
    for (int i = c + t; i < F.n; i += 256) {
      const double v = fabs(col[i]);
      if (v > best) {
        best = v;
        bi = i;
      }
    }
    rv[t] = best;
    ri[t] = bi;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
      if (t < h && (rv[t + h] > rv[t] || (rv[t + h] == rv[t] && ri[t + h] < ri[t]))) {
        rv[t] = rv[t + h];
        ri[t] = ri[t + h];
      }
      __syncthreads();
    }
    const int p = ri[0] < F.n ? ri[0] : c;  // (a column of NaNs keeps its row)
    __syncthreads();
    if (p != c && t < kb) {  // interchange within the panel
      double* pc = a + (i64)(k0 + t) * N;
      const double x = pc[c];
      pc[c] = pc[p];
      pc[p] = x;
    }
    if (t == 0) ipiv[F.int_off + c] = p;
    __syncthreads();
    const double d = col[c];
    if (d == 0.0) {
      if (t == 0 && atomicCAS(err, 0, 1) == 0) {
        err[1] = fid;
        err[2] = c;
      }
      continue;  // the column stays as it is (LAPACK goes on; the result is reported invalid)
    }
    for (int i = c + 1 + t; i < N; i += 256) col[i] /= d;
    __syncthreads();
    for (int cc = c + 1; cc < k0 + kb; cc++) {
      double* pc = a + (i64)cc * N;
      const double u = pc[c];
      if (u != 0.0)
        for (int i = c + 1 + t; i < N; i += 256) pc[i] -= col[i] * u;
    }
    __syncthreads();
  }
  // inverses of the panel's diagonal block: unit lower (dinv) and upper (uinv), column-major,
  // zero outside their triangle and beyond kb
  for (int e = t; e < NB * NB; e += 256) {
    const int r = e & (NB - 1), cc = e / NB;
    D[r][cc] = (r < kb && cc < kb) ? a[(i64)(k0 + cc) * N + k0 + r] : 0.0;
  }
  __syncthreads();
  double* Li = A + F.dinv_off + (i64)(k0 / NB) * NB * NB;
  double* Ui = uinv + F.uinv_off + (i64)(k0 / NB) * NB * NB;
  if (t < NB) {  // column t of L^{-1}: forward substitution
    const int cc = t;
    for (int i = 0; i < NB; i++) {
      double x = 0.0;
      if (cc < kb && i >= cc && i < kb) {
        x = i == cc ? 1.0 : 0.0;
        for (int k = cc; k < i; k++) x -= D[i][k] * Li[(i64)cc * NB + k];
      }
      Li[(i64)cc * NB + i] = x;
    }
  } else if (t < 2 * NB) {  // column t - 64 of U^{-1}: back substitution
    const int cc = t - NB;
    for (int i = NB - 1; i >= 0; i--) {
      double x = 0.0;
      if (cc < kb && i <= cc) {
        x = i == cc ? 1.0 : 0.0;
        for (int k = i + 1; k <= cc; k++) x -= D[i][k] * Ui[(i64)cc * NB + k];
        x /= D[i][i];
      }
      Ui[(i64)cc * NB + i] = x;
    }
  }
}

// The panel's interchanges on the front's other columns [0, k0) and [k0 + kb, N), one column per
// thread.
__global__ __launch_bounds__(256) void k_lu_swap(const int* __restrict__ ids, const FrontDev* __restrict__ fr,
                                                 double* __restrict__ A, int k0, const int* __restrict__ ipiv) {
  const FrontDev F = fr[ids[blockIdx.y]];
  if (k0 >= F.n) return;
  const int kb = min(NB, F.n - k0);
  const int cc = blockIdx.x * 256 + threadIdx.x;
  if (cc >= F.N || (cc >= k0 && cc < k0 + kb)) return;
  double* col = A + F.off + (i64)cc * F.N;
  for (int j = 0; j < kb; j++) {
    const int c = k0 + j, p = ipiv[F.int_off + c];
    if (p != c) {
      const double x = col[c];
      col[c] = col[p];
      col[p] = x;
    }
  }
}

// U row block: A(k0.., cols) <- Linv_kk A(k0.., cols) for 64 columns per workgroup
// (blockIdx.x) of [k0 + kb, N); the result also goes transposed to ut (ut[c + k N] = U(k0+k, c)).
__global__ __launch_bounds__(256) void k_lu_trsm(const int* __restrict__ ids, const FrontDev* __restrict__ fr,
                                                 double* __restrict__ A, int k0, double* __restrict__ ut) {
  const FrontDev F = fr[ids[blockIdx.y]];
  if (k0 >= F.n) return;
  const int kb = min(NB, F.n - k0), s0 = k0 + kb, N = F.N;
  const int c0 = s0 + NB * blockIdx.x;
  if (c0 >= N) return;
  const int nc = min(NB, N - c0), t = threadIdx.x;
  __shared__ double Li[NB][NB + 1], B[NB][NB + 1];
  const double* Lg = A + F.dinv_off + (i64)(k0 / NB) * NB * NB;
  double* a = A + F.off;
  for (int e = t; e < NB * NB; e += 256) {
    const int r = e & (NB - 1), cc = e / NB;
    Li[r][cc] = Lg[(i64)cc * NB + r];
    B[r][cc] = (r < kb && cc < nc) ? a[(i64)(c0 + cc) * N + k0 + r] : 0.0;
  }
  __syncthreads();
  const int cc = t & (NB - 1), r0 = (t >> 6) * 16;
  if (cc >= nc) return;
  double* ucol = ut + F.ut_off + c0 + cc;
  for (int r = r0; r < r0 + 16 && r < kb; r++) {
    double s = 0.0;
    for (int k = 0; k <= r; k++) s += Li[r][k] * B[k][cc];
    a[(i64)(c0 + cc) * N + k0 + r] = s;
    ucol[(i64)r * N] = s;
  }
}

// Trailing update of the full region [s0, N) x [s0, N), s0 = k0 + kb: C -= L(rows, panel)
// U(panel, cols), one 64 x 64 tile per workgroup (blockIdx.x = tile, row-major over the region's
// tile grid), the product on FP64 MFMA (tile_gemm_nt_t with P = the panel's L columns and Q = the
// transposed U rows).
__global__ __launch_bounds__(256) void k_lu_gemm(const int* __restrict__ ids, const FrontDev* __restrict__ fr,
                                                 double* __restrict__ A, int k0, const double* __restrict__ ut) {
  const FrontDev F = fr[ids[blockIdx.y]];
  if (k0 >= F.n) return;
  const int kb = min(NB, F.n - k0), s0 = k0 + kb, N = F.N;
  const int T = (N - s0 + NB - 1) / NB;
  if ((int)blockIdx.x >= T * T) return;
  const int row0 = __builtin_amdgcn_readfirstlane(s0 + NB * ((int)blockIdx.x / T));
  const int col0 = __builtin_amdgcn_readfirstlane(s0 + NB * ((int)blockIdx.x % T));
  const int nr = min(NB, N - row0), nc = min(NB, N - col0);
  __shared__ TileSmemT<double, 64> sm;
  double* a = A + F.off;
  V4<double> acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; i++)
#pragma unroll
    for (int j = 0; j < 2; j++) acc[i][j] = vzero<double>();
  double cv[2][2][4];
  double* C = a + (i64)col0 * N + row0;
  tile_read(cv, C, N, nr, nc, true);
  tile_gemm_nt_t<double, 64>(acc, a + (i64)k0 * N + row0, N, nr, ut + F.ut_off + col0, N, nc, kb, sm);
#pragma unroll
  for (int bi = 0; bi < 2; bi++)
#pragma unroll
    for (int bj = 0; bj < 2; bj++)
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int i = acc_row(bi), j = acc_col<double>(bj, r);
        if (i < nr && j < nc) C[(i64)j * N + i] = cv[bi][bj][r] - acc[bi][bj][r];
      }
}

// The front's interior DOFs in pivot order: perm[int_off + j] = dofs[int_off + pi(j)], pi the
// product of the interchanges (P b)[j] = b[pi(j)].  One workgroup per front; the interchanges
// are applied in order by one thread, on LDS when the front fits (n <= 12288).
__global__ __launch_bounds__(256) void k_lu_perm(const int* __restrict__ ids, const FrontDev* __restrict__ fr,
                                                 const int* __restrict__ ipiv, const int* __restrict__ dofs,
                                                 int* __restrict__ perm) {
  const FrontDev F = fr[ids[blockIdx.x]];
  __shared__ int pi[12288];
  const int* di = dofs + F.int_off;
  int* out = perm + F.int_off;
  const int* pv = ipiv + F.int_off;
  if (F.n <= 12288) {
    for (int j = threadIdx.x; j < F.n; j += 256) pi[j] = j;
    __syncthreads();
    if (threadIdx.x == 0)
      for (int j = 0; j < F.n; j++) {
        const int p = pv[j];
        const int x = pi[j];
        pi[j] = pi[p];
        pi[p] = x;
      }
    __syncthreads();
    for (int j = threadIdx.x; j < F.n; j += 256) out[j] = di[pi[j]];
  } else {
    for (int j = threadIdx.x; j < F.n; j += 256) out[j] = j;
    __syncthreads();
    if (threadIdx.x == 0)
      for (int j = 0; j < F.n; j++) {
        const int p = pv[j];
        const int x = out[j];
        out[j] = out[p];
        out[p] = x;
      }
    __syncthreads();
    for (int j = threadIdx.x; j < F.n; j += 256) out[j] = di[out[j]];
  }
}

// Backward sweep of one LU front: X[0:n] = y (the forward result, W[int]), X[n:N] = W[bd];
// block by block from the last: z_j = y_j - U(block j, later columns) X(later), x_j = Uinv_j z_j
// (= U^{-1} (y - U12 x_sk) overall, the reference's A22^{-1} b_rd - XR b_sk); X[0:n] -> W[int].
__global__ __launch_bounds__(256) void k_lu_bwd(const int* __restrict__ ids, const FrontDev* __restrict__ fr,
                                                const double* __restrict__ A, const int* __restrict__ dofs,
                                                double* __restrict__ W, double* __restrict__ xs,
                                                const double* __restrict__ uinv) {
  const FrontDev F = fr[ids[blockIdx.x]];
  const double* a = A + F.off;
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
    // T(i, r) = sum_{c >= jb+kb} U(jb+i, c) X(c, r): wave w -> rows 16w..16w+15
    const int row = 16 * w + c16;
    const bool rin = row < kb;
    const double* ar = a + jb + (rin ? row : 0);
    V4<double> acc = vzero<double>();
    for (int c0 = jb + kb; c0 < F.N; c0 += 4 * GU) {
      double av[GU], bv[GU];
#pragma unroll
      for (int u = 0; u < GU; u++) {
        const int c = c0 + 4 * u + q;
        const int cc = min(c, F.N - 1);
        av[u] = ld_or0(ar + (i64)cc * F.N, rin && c < F.N);
        bv[u] = X[(i64)cc * RB + c16];
      }
#pragma unroll
      for (int u = 0; u < GU; u++) acc = mfma4(av[u], bv[u], acc);
    }
#pragma unroll
    for (int r = 0; r < 4; r++) {
      const int i = 16 * w + mrow<double>(q, r);
      zb[i][c16] = ld_or0(X + (i64)(jb + min(i, kb - 1)) * RB + c16, i < kb) - (i < kb ? acc[r] : 0.0);
    }
    __syncthreads();
    {  // x_j = Uinv_j z: wave w -> rows 16w..16w+15
      const double* Ui = uinv + F.uinv_off + (i64)bk * NB * NB;
      V4<double> a2 = vzero<double>();
#pragma unroll
      for (int k4 = 0; k4 < NB / 4; k4++) {
        const int c = 4 * k4 + q;
        a2 = mfma4(Ui[(i64)c * NB + 16 * w + c16], zb[c][c16], a2);
      }
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int i = 16 * w + mrow<double>(q, r);
        if (i < kb) X[(i64)(jb + i) * RB + c16] = a2[r];
      }
    }
  }
  __syncthreads();
  for (int e = t; e < F.n * RB; e += 256) W[(i64)di[e >> 4] * RB + (e & 15)] = X[e];
}

// Upward step of Apply for one LU front: u = [U U12] [x_rd; x_sk] (kept in X[0:n] for
// k_apply_down, whose X[n:N] = L21 u is the reference's XL b_rd), b_rd <- P^T L u written to the
// permuted interior DOFs (the reference's A22 (b_rd + XR b_sk), src/DE/mf2d.cpp:904-974).
__global__ __launch_bounds__(256) void k_lu_apply_up(const int* __restrict__ ids, const FrontDev* __restrict__ fr,
                                                     const double* __restrict__ A, const int* __restrict__ dofs,
                                                     double* __restrict__ W, double* __restrict__ xs, int nr,
                                                     const int* __restrict__ perm) {
  const FrontDev F = fr[ids[blockIdx.x]];
  const double* a = A + F.off;
  double* X = xs + F.x_off;
  const int* di = dofs + F.int_off;
  const int* db = dofs + F.bd_off;
  const int* pm = perm + F.int_off;
  const int t = threadIdx.x;
  gather_rhs(F, X, W, di, db);
  __syncthreads();
  // u_i = sum_{c >= i} F(i, c) X(c) into W[int] (X still holds x)
  for (int i = t; i < F.n; i += 256)
    for (int r = 0; r < nr; r++) {
      double s = 0.0;
      for (int c = i; c < F.N; c++) s += a[(i64)c * F.N + i] * X[(i64)c * RB + r];
      W[(i64)di[i] * RB + r] = s;
    }
  __syncthreads();
  for (int e = t; e < F.n * RB; e += 256) X[e] = (e & 15) < nr ? W[(i64)di[e >> 4] * RB + (e & 15)] : 0.0;
  __syncthreads();
  // (L u)_i = u_i + sum_{c < i} F(i, c) u_c to row pi(i) of b_rd
  for (int i = t; i < F.n; i += 256)
    for (int r = 0; r < nr; r++) {
      double s = X[(i64)i * RB + r];
      for (int c = 0; c < i; c++) s += a[(i64)c * F.N + i] * X[(i64)c * RB + r];
      W[(i64)pm[i] * RB + r] = s;
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

// ----------------------------------------------------------------------------- refinement
// FP64 kernels of the preconditioned conjugate-gradient refinement that turns the FP32 factor
// into an FP64-accurate solve.  Its vectors are row-major n x cw blocks (v[i*cw + c]) with cw
// the smallest power of two >= the number of right-hand sides (1..16): a solve of 4 columns
// moves 4 columns, not 16.  The FP32 preconditioner's work block stays n x 16 (the sweeps' MFMA
// width); only its first cw columns are written and read here.
//
// Reductions (dot products, maxima) run over NCH fixed chunks of rows: chunk q holds rows
// [q*ch, (q+1)*ch), ch = ceil(n / NCH); inside a chunk, row lane rl = 0..15 sums rows i with
// (i - q*ch) % 16 == rl in ascending order, the chunk's value is the sum over rl in order, and
// the result the sum over chunks in a fixed order (k_red_final).  The order depends on n alone,
// not on cw or the grid: a column's result is the same whatever other columns share the solve.
#define NCH 8192
struct ChunkLane {
  int c, rl, g, q;  // column, row lane, chunk within the workgroup, chunk
  i64 i0, i1;       // rows of the chunk
};
// Workgroup of 256 threads = 16 / cw chunks x 16 row lanes x cw columns (t = (g*16 + rl)*cw + c);
// grid NCH * cw / 16.
__device__ __forceinline__ ChunkLane chunk_lane(int n, int lcw) {
  ChunkLane L;
  const int t = threadIdx.x;
  L.c = t & ((1 << lcw) - 1);
  L.rl = (t >> lcw) & 15;
  L.g = t >> (lcw + 4);
  L.q = blockIdx.x * (16 >> lcw) + L.g;
  const i64 ch = ((i64)n + NCH - 1) / NCH;
  L.i0 = min((i64)n, (i64)L.q * ch);
  L.i1 = min((i64)n, L.i0 + ch);
  return L;
}
// Sum (MAX = false) or NaN-keeping max (MAX = true) of v over the row lanes of each (chunk,
// column) of the workgroup, rl in order, into part[(pair*NCH + q)*16 + c].
template <bool MAX>
__device__ __forceinline__ void chunk_reduce(double v, int lcw, int pair, double* __restrict__ red,
                                             double* __restrict__ part) {
  red[threadIdx.x] = v;
  __syncthreads();
  const int cw = 1 << lcw, per = 16 >> lcw;  // chunks per workgroup
  if ((int)threadIdx.x < per * cw) {
    const int g = threadIdx.x >> lcw, c = threadIdx.x & (cw - 1);
    double s = red[(g * 16) * cw + c];
    for (int rl = 1; rl < 16; rl++) {
      const double x = red[(g * 16 + rl) * cw + c];
      s = MAX ? ((x > s || x != x) ? x : s) : s + x;
    }
    part[((i64)pair * NCH + blockIdx.x * per + g) * 16 + c] = s;
  }
  __syncthreads();
}
__device__ __forceinline__ double nanmax(double m, double v) { return (v > m || v != v) ? v : m; }

// Second pass: out[pair*16 + c] = sum (or max) over the NCH chunks of part, one workgroup per
// (pair, c) output, grid npair x cw: thread t takes chunks t, t + 256, ... in order, then a
// fixed pairwise tree over the threads.
template <bool MAX>
__global__ __launch_bounds__(256) void k_red_final(const double* __restrict__ part, double* __restrict__ out) {
  const int c = blockIdx.x, pr = blockIdx.y, t = threadIdx.x;
  const double* src = part + (i64)pr * NCH * 16 + c;
  double v[NCH / 256];
#pragma unroll
  for (int j = 0; j < NCH / 256; j++) v[j] = src[(i64)(t + 256 * j) * 16];
  double s = v[0];
#pragma unroll
  for (int j = 1; j < NCH / 256; j++) s = MAX ? nanmax(s, v[j]) : s + v[j];
  __shared__ double red[256];
  red[t] = s;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if (t < h) red[t] = MAX ? nanmax(red[t], red[t + h]) : red[t] + red[t + h];
    __syncthreads();
  }
  if (t == 0) out[pr * 16 + c] = red[0];
}

// Start of the refinement: B = R = the right-hand sides (column-major in, columns >= nr zero),
// X = 0; part (pair 0) <- max |b| per column.
__global__ __launch_bounds__(256) void k_cg_init(int n, int lcw, int nr, const double* __restrict__ in, int ld,
                                                 double* __restrict__ Bv, double* __restrict__ R,
                                                 double* __restrict__ X, double* __restrict__ part) {
  __shared__ double red[256];
  const ChunkLane L = chunk_lane(n, lcw);
  const int cw = 1 << lcw;
  double mx = 0.0;
  for (i64 i = L.i0 + L.rl; i < L.i1; i += 16) {
    const double v = L.c < nr ? in[(i64)L.c * ld + i] : 0.0;
    const i64 e = i * cw + L.c;
    Bv[e] = v;
    R[e] = v;
    X[e] = 0.0;
    mx = nanmax(mx, fabs(v));
  }
  chunk_reduce<true>(mx, lcw, 0, red, part);
}

// w (the preconditioner's n x 16 FP32 work block) <- (float) r, columns < cw.
__global__ __launch_bounds__(256) void k_pc_in(int n, int lcw, const double* __restrict__ R, float* __restrict__ w) {
  const i64 e = (i64)blockIdx.x * 256 + threadIdx.x;
  if (e >= ((i64)n << lcw)) return;
  w[(e >> lcw) * RB + (e & ((1 << lcw) - 1))] = (float)R[e];
}

// z = (double) w after the preconditioner (P = z as well when P is given, the first
// direction); part (pair 0) <- r.z per column.
__global__ __launch_bounds__(256) void k_pc_out(int n, int lcw, const float* __restrict__ w,
                                                const double* __restrict__ R, double* __restrict__ Z,
                                                double* __restrict__ P, double* __restrict__ part) {
  __shared__ double red[256];
  const ChunkLane L = chunk_lane(n, lcw);
  const int cw = 1 << lcw;
  double s = 0.0;
  for (i64 i = L.i0 + L.rl; i < L.i1; i += 16) {
    const i64 e = i * cw + L.c;
    const double z = (double)w[i * RB + L.c];
    Z[e] = z;
    if (P) P[e] = z;
    s += R[e] * z;
  }
  chunk_reduce<false>(s, lcw, 0, red, part);
}

// q = A p (the CSR operator, one thread per row and column); part (pair 0) <- p.q per column.
__global__ __launch_bounds__(256) void k_spmv_dot(int n, int lcw, const int* __restrict__ rp,
                                                  const int* __restrict__ ci, const double* __restrict__ v,
                                                  const double* __restrict__ P, double* __restrict__ Q,
                                                  double* __restrict__ part) {
  __shared__ double red[256];
  const ChunkLane L = chunk_lane(n, lcw);
  const int cw = 1 << lcw;
  double s = 0.0;
  for (i64 i = L.i0 + L.rl; i < L.i1; i += 16) {
    double y = 0.0;
    const int k1 = rp[i + 1];
    for (int k = rp[i]; k < k1; k++) y += v[k] * P[(i64)ci[k] * cw + L.c];
    const i64 e = i * cw + L.c;
    Q[e] = y;
    s += P[e] * y;
  }
  chunk_reduce<false>(s, lcw, 0, red, part);
}

// x += a p, r -= a q with a = rz / pq per column (columns outside `active`, a bit mask, stay as
// they are); w <- (float) r, the next preconditioner input; part <- max |r| (pair 0) and
// max |x| (pair 1) per column.
__global__ __launch_bounds__(256) void k_cg_xr(int n, int lcw, const double* __restrict__ rz,
                                               const double* __restrict__ pq, int active,
                                               const double* __restrict__ P, const double* __restrict__ Q,
                                               double* __restrict__ X, double* __restrict__ R,
                                               float* __restrict__ w, double* __restrict__ part) {
  __shared__ double red[256];
  const ChunkLane L = chunk_lane(n, lcw);
  const int cw = 1 << lcw;
  const bool on = (active >> L.c) & 1;
  const double a = on && pq[L.c] != 0.0 ? rz[L.c] / pq[L.c] : 0.0;
  double mr = 0.0, mx = 0.0;
  for (i64 i = L.i0 + L.rl; i < L.i1; i += 16) {
    const i64 e = i * cw + L.c;
    double x = X[e], r = R[e];
    if (on) {
      x += a * P[e];
      r -= a * Q[e];
      X[e] = x;
      R[e] = r;
    }
    w[i * RB + L.c] = (float)r;
    mr = nanmax(mr, fabs(r));
    mx = nanmax(mx, fabs(x));
  }
  chunk_reduce<true>(mr, lcw, 0, red, part);
  chunk_reduce<true>(mx, lcw, 1, red, part);
}

// p = z + b p with b = rz_new / rz per column, active columns.
__global__ __launch_bounds__(256) void k_cg_p(int n, int lcw, const double* __restrict__ rz_new,
                                              const double* __restrict__ rz, int active,
                                              const double* __restrict__ Z, double* __restrict__ P) {
  const i64 e = (i64)blockIdx.x * 256 + threadIdx.x;
  if (e >= ((i64)n << lcw)) return;
  const int c = (int)(e & ((1 << lcw) - 1));
  if (!((active >> c) & 1)) return;
  const double b = rz[c] != 0.0 ? rz_new[c] / rz[c] : 0.0;
  P[e] = Z[e] + b * P[e];
}

// out (column-major, ld) <- X, columns < nr.
__global__ __launch_bounds__(256) void k_cg_out(int n, int lcw, int nr, const double* __restrict__ X,
                                                double* __restrict__ out, int ld) {
  const i64 e = (i64)blockIdx.x * 256 + threadIdx.x;
  if (e >= ((i64)n << lcw)) return;
  const int c = (int)(e & ((1 << lcw) - 1));
  if (c < nr) out[(i64)c * ld + (e >> lcw)] = X[e];
}

// ||A||_inf of a CSR operator, first pass: part[g*32] = max over block g's rows of sum_j |a_ij|.
__global__ __launch_bounds__(256) void k_rowabs(int n, const int* __restrict__ rp, const double* __restrict__ v,
                                                double* __restrict__ part) {
  const int chunk = (n + gridDim.x - 1) / gridDim.x;
  const int i0 = blockIdx.x * chunk, i1 = min(n, i0 + chunk);
  double m = 0.0;
  for (int i = i0 + threadIdx.x; i < i1; i += 256) {
    double s = 0.0;
    for (int k = rp[i]; k < rp[i + 1]; k++) s += fabs(v[k]);
    m = nanmax(m, s);
  }
  __shared__ double red[256];
  red[threadIdx.x] = m;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if (threadIdx.x < h) red[threadIdx.x] = nanmax(red[threadIdx.x], red[threadIdx.x + h]);
    __syncthreads();
  }
  if (threadIdx.x < 32) part[blockIdx.x * 32 + threadIdx.x] = threadIdx.x == 0 ? red[0] : 0.0;
}
// Second pass of k_rowabs: out[c] = max_g part[g*32 + c], c < 32.
__global__ __launch_bounds__(256) void k_max_final(int nblk, const double* __restrict__ part,
                                                   double* __restrict__ out) {
  const int c = threadIdx.x & 31, s8 = threadIdx.x >> 5;
  double m = 0.0;
  for (int g = s8; g < nblk; g += 8) m = nanmax(m, part[(i64)g * 32 + c]);
  __shared__ double red[256];
  red[threadIdx.x] = m;
  __syncthreads();
  if (threadIdx.x < 32) {
    double t = 0.0;
    for (int q = 0; q < 8; q++) t = nanmax(t, red[q * 32 + c]);
    out[c] = t;
  }
}

// ----------------------------------------------------------------------------- operator values
// Coefficient-change requests on the harness operator -div(a grad u)
// (test_include/DE/buildLaplacian.hpp:10-132, :244-391): a is the padded nodal field,
// (N_a + 2w) per axis, stored a[i][j](k) with the last index fastest; DOF id = i + j Nx (+ k Nx Ny).
struct LapGeom {
  int dim, w, N[3], P[3];  // interior and padded extents
  double hh;               // h^2
  double cf[4];            // stencil weights cf[1..w]
};

__device__ __forceinline__ i64 lap_cidx(const LapGeom& G, const int c[3]) {  // padded coords
  return G.dim == 2 ? (i64)c[0] * G.P[1] + c[1] : ((i64)c[0] * G.P[1] + c[1]) * G.P[2] + c[2];
}

// a *= factor on the padded box [lo, lo + side)^dim
__global__ __launch_bounds__(256) void k_box_scale(LapGeom G, double* __restrict__ a, int lo0, int lo1, int lo2,
                                                   int side, double factor) {
  const i64 cnt = G.dim == 2 ? (i64)side * side : (i64)side * side * side;
  for (i64 e = (i64)blockIdx.x * 256 + threadIdx.x; e < cnt; e += (i64)gridDim.x * 256) {
    const int c[3] = {lo0 + (int)(e % side), lo1 + (int)(e / side % side), lo2 + (int)(e / ((i64)side * side))};
    a[lap_cidx(G, c)] *= factor;
  }
}

// Recompute the values of the rows in the interior box [r0, r1) (per axis) in the generator's
// entry order and arithmetic: negative offsets (slowest axis first, far to near), the diagonal
// (-sum over k = w..1 and the axes of cf[k] (a1 + a2) / h^2), positive offsets (fastest axis
// first, near to far); off-diagonal value cf[k] (a_p + a_q) / 2 / h^2.  Entries whose value
// changes flag their column with the epoch.
__global__ __launch_bounds__(256) void k_lap_rows(LapGeom G, const double* __restrict__ a, const int* __restrict__ rp,
                                                  double* __restrict__ val, int b0, int b1, int b2, int e0,
                                                  int e1, int e2, int* __restrict__ flag, int epoch) {
#pragma clang fp contract(off)
  const int ext[3] = {e0, e1, e2};
  const i64 cnt = (i64)e0 * e1 * e2;
  for (i64 e = (i64)blockIdx.x * 256 + threadIdx.x; e < cnt; e += (i64)gridDim.x * 256) {
    const int p[3] = {b0 + (int)(e % ext[0]), b1 + (int)(e / ext[0] % ext[1]), b2 + (int)(e / ((i64)ext[0] * ext[1]))};
    const int row = p[0] + G.N[0] * (p[1] + G.N[1] * p[2]);
    const int stride[3] = {1, G.N[0], G.N[0] * G.N[1]};
    auto coef = [&](int ax, int s) {  // a at the padded position of p shifted by s along ax
      int c[3] = {p[0] + G.w, p[1] + G.w, p[2] + G.w};
      c[ax] += s;
      return a[lap_cidx(G, c)];
    };
    int k_pos = rp[row];
    auto put = [&](int col, double v) {
      if (val[k_pos] != v) flag[col] = epoch;
      val[k_pos++] = v;
    };
    for (int ax = G.dim - 1; ax >= 0; ax--)
      for (int k = G.w; k >= 1; k--)
        if (p[ax] - k >= 0) {
          const double ah = 0.5 * (coef(ax, -k) + coef(ax, -k + 1));
          put(row - k * stride[ax], G.cf[k] * ah / G.hh);
        }
    double diag = 0.0;
    for (int k = G.w; k >= 1; k--)
      for (int ax = 0; ax < G.dim; ax++) {
        const double a1 = 0.5 * (coef(ax, -k) + coef(ax, -k + 1));
        const double a2 = 0.5 * (coef(ax, k) + coef(ax, k - 1));
        diag = diag - G.cf[k] * (a1 + a2) / G.hh;
      }
    put(row, diag);
    for (int ax = 0; ax < G.dim; ax++)
      for (int k = 1; k <= G.w; k++)
        if (p[ax] + k < G.N[ax]) {
          const double ah = 0.5 * (coef(ax, k) + coef(ax, k - 1));
          put(row + k * stride[ax], G.cf[k] * ah / G.hh);
        }
  }
}

// DOFs of the interior box [b, b + e) flagged with the epoch, appended to list (order
// arbitrary; the count in *cnt)
__global__ __launch_bounds__(256) void k_lap_collect(LapGeom G, int b0, int b1, int b2, int e0, int e1, int e2,
                                                     const int* __restrict__ flag, int epoch, int* __restrict__ list,
                                                     int* __restrict__ cnt) {
  const i64 n = (i64)e0 * e1 * e2;
  for (i64 e = (i64)blockIdx.x * 256 + threadIdx.x; e < n; e += (i64)gridDim.x * 256) {
    const int p0 = b0 + (int)(e % e0), p1 = b1 + (int)(e / e0 % e1), p2 = b2 + (int)(e / ((i64)e0 * e1));
    const int d = p0 + G.N[0] * (p1 + G.N[1] * p2);
    if (flag[d] == epoch) list[atomicAdd(cnt, 1)] = d;
  }
}

}  // namespace umf
