// runtime.hip -- device front pool, level scheduler and the C ABI of libumf_hip.so.
//
// One umf_plan owns: the planner output (host), one device arena holding every front and
// separator matrix (planned once, 256-byte aligned), device copies of the DOF lists, assembly
// programs and CSR operator, and one HIP stream.  Factor / Update walk the tree level by level
// (leaves first), each stage one batched launch over the level's (dirty) fronts.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/umf.h"
#include "kernels.hip"
#include "plan.h"

using namespace umf;

namespace {
thread_local std::string g_err;

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

#define TOP64 3  // the mixed form's (UMF_PREC_FP32T64) FP64 levels: 0..TOP64-1

#define HIPCHK(x)                                                                        \
  do {                                                                                   \
    hipError_t e_ = (x);                                                                 \
    if (e_ != hipSuccess) return fail(UMF_EDEVICE, std::string(#x) + ": " + hipGetErrorString(e_)); \
  } while (0)

template <class T>
hipError_t upload(T** d, const std::vector<T>& h) {
  *d = nullptr;
  if (h.empty()) return hipSuccess;
  hipError_t e = hipMalloc((void**)d, sizeof(T) * h.size());
  if (e != hipSuccess) return e;
  return hipMemcpy(*d, h.data(), sizeof(T) * h.size(), hipMemcpyHostToDevice);
}
}  // namespace

// Per-kernel event timing (enabled by umf_profile_enable).
struct KStat {
  std::string name;
  double ms = 0, flops = 0, bytes = 0;
  int launches = 0;
};
struct Prof {
  bool on = false;
  std::string only;  // when non-empty, only launches of this kernel are timed
  std::vector<KStat> stats;
  std::vector<hipEvent_t> pool;
  size_t used = 0;
  struct Pending {
    int k;
    double flops, bytes;
  };
  std::vector<Pending> pend;
  int find(const char* name) {
    for (size_t i = 0; i < stats.size(); i++)
      if (stats[i].name == name) return (int)i;
    stats.push_back(KStat());
    stats.back().name = name;
    return (int)stats.size() - 1;
  }
};

namespace umf {
// Schedule of one factor/update pass: per level, the front ids to (re)factor, the separator ids
// to rebuild and the separator ids receiving Schur contributions, all uploaded in one buffer.
struct Sched {
  struct Lvl {
    int f_off, f_cnt;      // fronts of the general path (front ids == matrix ids)
    int lf_off, lf_cnt;    // fronts of the one-launch small-front kernel (leaf_front)
    int s_off, s_cnt;      // separators (separator ids)
    int sm_off, sm_cnt;    // separator matrix ids (NF + id) for zero/assemble
    int maxN, maxn, maxm;  // extents over the level's fronts
    int max_k;
    int f_grid, s_grid;    // assembly grid widths (tiles / triples per matrix)
  };
  std::vector<Lvl> lv;
  std::vector<int> ids;
};

// Level lists of all fronts / separators (solve and apply touch every front).
struct AllLists {
  std::vector<int> ids;
  std::vector<int> f_off, f_cnt, s_off, s_cnt;
};
}  // namespace umf

struct umf_plan {
  // the planner's output; serving slots share their base plan's (umf_server_create)
  std::shared_ptr<Plan> Pp = std::make_shared<Plan>();
  Plan& P = *Pp;
  umf_plan() = default;
  explicit umf_plan(std::shared_ptr<Plan> shared) : Pp(std::move(shared)) {}
  umf_plan(const umf_plan&) = delete;
  umf_plan& operator=(const umf_plan&) = delete;
  // a serving slot: its device front pool, programs and CSR pattern are the base plan's
  // (borrowed, not freed with the slot); descriptors, operator values, stream and scratch are its own
  bool is_slot = false;
  int servers = 0;  // live servers answering requests against this plan's factors (it stays immutable)
  int args[6] = {0, 0, 0, 0, 0, 0};  // dim, nx, ny, nz, width, max_side_length it was made with
  Prof prof;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  double* arena = nullptr;
  int* d_dofs = nullptr;
  MatDev* d_mats = nullptr;
  FrontDev* d_fronts = nullptr;
  SepDev* d_seps = nullptr;
  Triple* d_tri = nullptr;
  i64* d_tri_off = nullptr;
  int* d_fly_tab = nullptr;  // leaf fronts' entries by panel tile (Plan::fly_*, rest_*)
  Triple* d_fly_tri = nullptr;
  Triple* d_rest_tri = nullptr;
  i64* d_rest_off = nullptr;
  BlockOp* d_ops = nullptr;
  Tile* d_tiles = nullptr;
  i64* d_tile_off = nullptr;
  int* d_rp = nullptr;
  int* d_ci = nullptr;
  double* d_val = nullptr;
  int* d_err = nullptr;     // [0..3] error flags, [4] the queue counter of the persistent launches
  int* d_flags = nullptr;  // block flags of the blocked solve sweeps (fenced publication)
  unsigned long long* d_ll = nullptr;  // published blocks of a blocked sweep launch (flag-in-data)
  i64 ll_blocks = 0;                   // (front, block) capacity of d_ll: kLLBlock words each
  int* d_arrive = nullptr;  // per-front arrival counters of k_chol_dpanel (zero between launches)
  int n_flags = 0;
  // coefficient-change requests: per-DOF change flags (epoch-tagged), the collected list
  int* d_bflag = nullptr;
  int* d_blist = nullptr;  // count, then the DOFs
  int blist_cap = 0;
  int bepoch = 0;
  int epoch = 0;           // value the next blocked sweep launch publishes its blocks with
  bool fenced = false;     // blocked sweeps publish with agent-scope release / acquire fences
  int device = 0;          // HIP device the plan's memory and stream live on
  // LU fronts (unsymmetric operators): pivots and permuted interior DOF lists (indexed like the
  // DOF lists), inverted diagonal blocks of U, transposed U rows of the trailing updates
  bool lu = false;
  int* d_ipiv = nullptr;
  int* d_perm = nullptr;
  double* d_uinv = nullptr;
  double* d_ut = nullptr;
  int* d_ids = nullptr;  // id lists of the current schedule
  size_t ids_cap = 0;
  // tuning of the factorization's GEMM stages (umf_set_tune / umf_autotune): the outer column
  // group width of the two-level trailing update, and the trailing/Schur tile size (0: 128 for
  // FP32 fronts when that still yields 2048 tiles, else 64)
  int tune_gw2 = 1024;
  int tune_tile = 0;
  // the full factorization captured as a HIP graph (umf_set_graph): its own id list, the
  // precision it was captured in
  bool use_graph = false;
  hipGraphExec_t fgraph = nullptr;
  int fgraph_prec = -1;
  int* d_fids = nullptr;
  std::vector<int> ids_h;  // host copy, alive until the stream has consumed it
  std::vector<std::vector<int>> gen_fronts;  // per level: fronts on the general factor path
  std::map<const void*, int> resident;  // workgroups per CU of the persistent kernels
  int n_cu = 0;
  // the sweeps' level lists (fixed by the plan): built and uploaded once, then every sweep
  // launches from them without a host rebuild or a wait for the stream
  std::unique_ptr<umf::AllLists> sweep_lists;
  int* d_sids = nullptr;
  // the full factorization's schedule (fixed by the plan), its id list in d_fids
  std::unique_ptr<umf::Sched> full_sched;
  std::unique_ptr<Plan::Dirty> full_dirty;
  double* d_xs = nullptr;  // solve scratch: sum_f N_f x RB
  double* d_w = nullptr;   // solve work block n x RB, row-major
  // FP32 form: its own front pool and solve buffers, the FP64 refinement vectors
  int prec = UMF_PREC_FP64;       // precision of the next factorization
  int fact_prec = UMF_PREC_FP64;  // precision the current factors are held in
  double ir_tol = 1e-10;
  int ir_max = 50;
  float* arena32 = nullptr;
  float* d_xs32 = nullptr;
  float* d_w32 = nullptr;
  double* d_cg = nullptr;    // 5 blocks of n x RB doubles: B, X, R, P, Q/Z (n x cw row-major each)
  double* d_part = nullptr;  // reduction partials (2 x NCH x 16)
  double* d_dots = nullptr;  // 4 x 32 reduction results
  double norm_a = 0;         // ||A||_inf of the current operator
  double* d_stage = nullptr;  // device copy of host right-hand sides
  int stage_cols = 0;
  i64 xs_elems = 0;
  std::vector<FrontDev> fronts_h;
  std::vector<SepDev> seps_h;
  std::vector<MatDev> mats_h;
  std::vector<double> val_h;  // current operator values
  bool factored = false;
  bool host_only = false;  // analyzed plan without device memory
  Plan::Dirty last;
  umf_stats stats{};
  int max_tiles_per_mat = 1;
};

namespace {

// Fronts factored whole by k_leaf: assembled from the operator alone (no child blocks), one
// pivot block, at most LEAF_M boundary rows.  A property of the plan, never of the pass.
inline bool leaf_front(const Plan& P, int f) {
  const FrontInfo& F = P.fronts[f];
  return F.n >= 1 && F.n <= NB && F.m <= LEAF_M && P.tile_off[f + 1] == P.tile_off[f];
}

// small: split off the leaf_front fronts (Cholesky schedules); LU schedules keep every front on
// the general path.
Sched make_sched(const Plan& P, const Plan::Dirty& D, bool small = false) {
  Sched S;
  S.lv.assign(P.L, {});
  int NF = P.nfronts();
  std::vector<int> leaves;
  auto grid_of = [&](int mid) {
    i64 tiles = P.tile_off[mid + 1] - P.tile_off[mid];
    i64 tri = P.tri_off[mid + 1] - P.tri_off[mid];
    return (int)std::min<i64>(std::max<i64>(std::max(tiles, (tri + 1023) / 1024), 1), 1024);
  };
  for (int ell = P.L - 1; ell >= 0; ell--) {
    Sched::Lvl& L = S.lv[ell];
    L.f_off = (int)S.ids.size();
    L.maxN = L.maxn = L.maxm = 0;
    L.f_grid = L.s_grid = 1;
    leaves.clear();
    for (int b : D.boxes[ell]) {
      int f = P.box_gid[ell][b];
      if (small && leaf_front(P, f)) {
        leaves.push_back(f);
        continue;
      }
      S.ids.push_back(f);
      L.maxN = std::max(L.maxN, P.fronts[f].n + P.fronts[f].m);
      L.maxn = std::max(L.maxn, P.fronts[f].n);
      L.maxm = std::max(L.maxm, P.fronts[f].m);
      L.f_grid = std::max(L.f_grid, grid_of(f));
    }
    L.f_cnt = (int)S.ids.size() - L.f_off;
    // largest pivot block first: the 64-column steps of a level then launch over a prefix of
    // its fronts (those with n > k0), not over every front of the level
    std::stable_sort(S.ids.begin() + L.f_off, S.ids.end(),
                     [&](int a, int b) { return P.fronts[a].n > P.fronts[b].n; });
    L.lf_off = (int)S.ids.size();
    L.lf_cnt = (int)leaves.size();
    std::stable_sort(leaves.begin(), leaves.end(), [&](int a, int b) {  // k_leaf's queue: largest first
      return P.fronts[a].n + P.fronts[a].m > P.fronts[b].n + P.fronts[b].m;
    });
    S.ids.insert(S.ids.end(), leaves.begin(), leaves.end());
    L.s_off = (int)S.ids.size();
    L.max_k = 0;
    if (ell >= 1)
      for (int e : D.seps[ell]) {
        int s = P.sep_gid[ell][e];
        S.ids.push_back(s);
        L.max_k = std::max(L.max_k, P.seps[s].k);
      }
    L.s_cnt = (int)S.ids.size() - L.s_off;
    L.sm_off = (int)S.ids.size();
    if (ell >= 1)
      for (int e : D.seps[ell]) {
        S.ids.push_back(NF + P.sep_gid[ell][e]);
        L.s_grid = std::max(L.s_grid, grid_of(NF + P.sep_gid[ell][e]));
      }
    L.sm_cnt = (int)S.ids.size() - L.sm_off;
  }
  return S;
}

int ensure_ids(umf_plan* p, size_t n) {
  if (n <= p->ids_cap) return UMF_OK;
  HIPCHK(hipStreamSynchronize(p->stream));
  if (p->d_ids) (void)hipFree(p->d_ids);
  p->d_ids = nullptr;
  HIPCHK(hipMalloc((void**)&p->d_ids, sizeof(int) * std::max<size_t>(n, 1)));
  p->ids_cap = n;
  return UMF_OK;
}

// Launch on the plan stream; with profiling on, bracket it by a pair of pooled events and
// remember its algorithmic work.
// tree level of the launches that follow (-1: none), appended to kernel names when profiling
thread_local int g_level = -1;

template <class K, class... Args>
void launch_b(umf_plan* p, const char* name, double flops, double bytes, K kern, dim3 grid, dim3 block,
              Args... args) {
  Prof& pr = p->prof;
  hipEvent_t a = nullptr, b = nullptr;
  char tagged[64];
  const bool timed = pr.on && (pr.only.empty() || pr.only == name);
  if (timed && g_level >= 0) {
    snprintf(tagged, sizeof tagged, "%s@L%d", name, g_level);
    name = tagged;
  }
  if (timed) {
    while (pr.pool.size() < pr.used + 2) {
      hipEvent_t e;
      if (hipEventCreate(&e) != hipSuccess) break;
      pr.pool.push_back(e);
    }
    if (pr.pool.size() >= pr.used + 2) {
      a = pr.pool[pr.used];
      b = pr.pool[pr.used + 1];
      pr.used += 2;
      pr.pend.push_back(Prof::Pending{pr.find(name), flops, bytes});
      (void)hipEventRecord(a, p->stream);
    }
  }
  hipLaunchKernelGGL(kern, grid, block, 0, p->stream, args...);
  if (b) (void)hipEventRecord(b, p->stream);
}
template <class K, class... Args>
void launch(umf_plan* p, const char* name, double flops, double bytes, K kern, dim3 grid, Args... args) {
  launch_b(p, name, flops, bytes, kern, grid, dim3(256), args...);
}

// After the stream has been synchronized: fold the pending event pairs into the totals.
void prof_collect(umf_plan* p) {
  Prof& pr = p->prof;
  for (size_t i = 0; i < pr.pend.size(); i++) {
    float ms = 0;
    if (hipEventElapsedTime(&ms, pr.pool[2 * i], pr.pool[2 * i + 1]) != hipSuccess) ms = 0;
    KStat& k = pr.stats[pr.pend[i].k];
    k.ms += ms;
    k.launches++;
    k.flops += pr.pend[i].flops;
    k.bytes += pr.pend[i].bytes;
  }
  pr.pend.clear();
  pr.used = 0;
}

// ts x ts tiles of the lower region rows [s, rend) x cols [s, cend)
inline int tiles_of(int s, int rend, int cend, int ts = 64) {
  if (s >= cend) return 0;
  int T = (cend - s + ts - 1) / ts, RT = (rend - s + ts - 1) / ts;
  return T * (T + 1) / 2 + (RT - T) * T;
}

// k_gemm instance of a stage: the native MFMA of T, or for FP32 fronts with `split` the
// product on BF16 pieces
// STAGE 0 = group update, 1 = trailing update, 2 = Schur complement
template <class T, int STAGE, int TS>
auto gemm_fn(bool split) {
  auto pick = [](auto sp) {
    constexpr bool S = decltype(sp)::value;
    if constexpr (STAGE == 0) return k_gemm_group<T, TS, S>;
    else if constexpr (STAGE == 1) return k_gemm_trail<T, TS, S>;
    else return k_gemm_schur<T, TS, S>;
  };
  if constexpr (std::is_same<T, float>::value) {
    if (split) return pick(std::true_type{});
  }
  return pick(std::false_type{});
}
// the native-MFMA 128-tile trailing (STAGE 1) or Schur (STAGE 2) GEMM on 8 waves, used for FP64
// fronts: four waves per SIMD instead of two took the 3D 127^3 FP64 trailing update from 1604 to
// 1504 ms and the Schur update from 423 to 386 ms; FP32 fronts measured slower on 8 waves
// (728 -> 748 ms) and keep 4
template <class T, int STAGE>
auto gemm_fn8() {
  if constexpr (STAGE == 1) return k_gemm_trail<T, 128, false, 8>;
  else return k_gemm_schur<T, 128, false, 8>;
}
template <class T>
constexpr bool kGemm8 = sizeof(T) == 8;

// Workgroups per item (front, separator) of a level stage whose largest item can use `want`
// of them: enough in all to fill the chip (about 32k), no more.  Every stage loops over its
// item's work with a stride of gridDim.x, so a smaller grid only lengthens the loops; levels of
// tens of thousands of small fronts and one large one then launch one workgroup per front.
inline int per_item(int want, int cnt) { return std::max(1, std::min(want, (32768 + cnt - 1) / std::max(cnt, 1))); }

// Grid of a persistent launch of kern (256 threads) over cnt queue items: as many workgroups as
// the device holds at once, at most cnt.
unsigned persistent_grid(umf_plan* p, const void* kern, int cnt) {
  int& per_cu = p->resident[kern];
  if (per_cu == 0) {
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, 256, 0) != hipSuccess || per_cu < 1) per_cu = 1;
  }
  if (p->n_cu == 0 && hipDeviceGetAttribute(&p->n_cu, hipDeviceAttributeMultiprocessorCount, p->device) != hipSuccess)
    p->n_cu = 256;
  return (unsigned)std::max(1, std::min(cnt, per_cu * std::max(p->n_cu, 1)));
}

// The fronts of level ell that take the general path (not leaf_front), cached per plan: the
// level-wide algorithm choices of the factorization are made over them.
const std::vector<int>& general_fronts(umf_plan* p, int ell) {
  const Plan& P = p->P;
  if ((int)p->gen_fronts.size() != P.L) {
    p->gen_fronts.assign(P.L, {});
    for (int l = 0; l < P.L; l++)
      for (int f : P.box_gid[l])
        if (!leaf_front(P, f)) p->gen_fronts[l].push_back(f);
  }
  return p->gen_fronts[ell];
}

// Start of a factorization pass over schedule S: its id list to the device (unless the caller
// holds it there: ids non-null), the error flags and the arrival counters cleared (a launch
// whose wait timed out may leave a count behind, which must not reach the next factorization).
int factor_begin(umf_plan* p, const Sched& S, const int*& ids) {
  if (!ids) {
    int rc = ensure_ids(p, S.ids.size());
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(p->stream));  // the previous id list may still be in flight
    p->ids_h = S.ids;
    if (!p->ids_h.empty())
      HIPCHK(hipMemcpyAsync(p->d_ids, p->ids_h.data(), sizeof(int) * p->ids_h.size(), hipMemcpyHostToDevice,
                            p->stream));
    ids = p->d_ids;
  }
  const int NF = p->P.nfronts();
  hipLaunchKernelGGL(k_zero_ints, dim3(1), dim3(64), 0, p->stream, p->d_err, 4);
  hipLaunchKernelGGL(k_zero_ints, dim3((unsigned)std::min((NF + 255) / 256, 1024)), dim3(256), 0, p->stream,
                     p->d_arrive, NF);
  return UMF_OK;
}

// One level of the factorization over the schedule S (its id list on the device at ids), fronts
// of type T in `arena`.  a32 / lev32: in the mixed form (FP64 top levels over FP32 ones) the FP32
// pool, whose matrices (MatDev::lev >= lev32) the assembly of the lowest FP64 level reads and
// converts; otherwise null / INT_MAX.
template <class T>
int factor_level(umf_plan* p, const Sched& S, const int* ids, int ell, T* arena, bool split, const float* a32,
                 int lev32) {
  const Plan& P = p->P;
  auto lowerN = [](double N) { return N * (N + 1) / 2; };
  constexpr double E = sizeof(T);  // bytes per front entry
  const Sched::Lvl& L = S.lv[ell];
  g_level = ell;
  // algorithmic traffic of the assembly of a list of matrices: zero the lower triangle,
  // then per triple (12 B index + 8 B value read + one entry written), per copied element two entries
  auto asm_work = [&](int off, int cnt, double& zb, double& ab) {
    zb = ab = 0;
    for (int i = 0; i < cnt; i++) {
      int mid = S.ids[off + i];
      zb += E * lowerN(P.mats[mid].rows);
      ab += (20.0 + E) * (P.tri_off[mid + 1] - P.tri_off[mid]);
      for (i64 t = P.tile_off[mid]; t < P.tile_off[mid + 1]; t++) {
        const BlockOp& op = P.ops[P.tiles[t].op];
        int r = std::min(64, op.nr - P.tiles[t].r0), c = std::min(64, op.nc - P.tiles[t].c0);
        ab += 2 * E * r * c;
      }
    }
  };
  // 1. separators of this level rebuilt from their children (upper levels) or from the CSR
  //    operator (leaf level): zero + scatter
  if (L.sm_cnt > 0) {
    double zb, ab;
    asm_work(L.sm_off, L.sm_cnt, zb, ab);
    launch(p, "k_zero_lower", 0, zb, k_zero_lower<T>,
           dim3(per_item(std::min((std::max(L.max_k, 1) + 3) / 4, 256), L.sm_cnt), L.sm_cnt), ids + L.sm_off,
           (const MatDev*)p->d_mats, arena);
    launch(p, "k_assemble", 0, ab, k_assemble<T>, dim3(per_item(L.s_grid, L.sm_cnt), L.sm_cnt), ids + L.sm_off,
           (const MatDev*)p->d_mats, arena, (const i64*)p->d_tri_off, (const Triple*)p->d_tri,
           (const i64*)p->d_tile_off, (const Tile*)p->d_tiles, (const BlockOp*)p->d_ops, (const double*)p->d_val, a32, lev32);
  }
  // 2'. small leaf fronts: assembly, factorization and Schur complement in one launch
  if (L.lf_cnt > 0) {
    double fl = 0, by = 0;
    for (int i = 0; i < L.lf_cnt; i++) {
      const int f = S.ids[L.lf_off + i];
      const double n = P.fronts[f].n, m = P.fronts[f].m;
      fl += 2.0 * n * n * n / 3.0 + m * n * n + n * m * (m + 1);  // factor + inverse, panel, Schur
      by += sizeof(T) * (lowerN(n) + m * n + lowerN(m) + NB * NB) + 20.0 * (P.tri_off[f + 1] - P.tri_off[f]);
    }
    launch(p, "k_leaf", fl, by, k_leaf<T>, dim3(persistent_grid(p, (const void*)k_leaf<T>, L.lf_cnt)),
           ids + L.lf_off, L.lf_cnt, p->d_err + 4, (const FrontDev*)p->d_fronts, arena, (const i64*)p->d_tri_off,
           (const Triple*)p->d_tri, (const double*)p->d_val, p->d_err);
  }
  const int lrc = L.f_cnt == 0 ? UMF_OK : [&]() -> int {
  // 3. (decided first: it shapes the assembly) blocked partial Cholesky of the pivot columns,
  //    64 at a time.  Levels with many fronts go left-looking (each block column is formed
  //    once by an MFMA GEMM over the finished columns, then factored; no trailing
  //    read-modify-write).  Levels with few large fronts go right-looking in column groups of
  //    GW: each block updates only the rest of its group (K = 64), and the group then updates
  //    the trailing matrix once (K = GW), so the trailing read-modify-write happens once per GW
  //    columns.
  // The choice of algorithm (left- or right-looking, fused or lazy column groups) fixes the
  // order of every front's floating-point operations, so it is made from the plan's whole
  // level, never from the fronts this pass happens to factor: an Update then computes every
  // dirty front exactly as a full factorization of the new operator would, bit for bit, and a
  // front's arithmetic does not depend on what else shares its launches.
  const std::vector<int>& level_fronts = general_fronts(p, ell);
  const bool left = level_fronts.size() >= 32;
  // BF16-split GEMMs (split) only on levels whose GEMM work (~ n N^2 per front, over the
  // plan's whole level) lies mostly in fronts of 1024+ pivots: the split pays where the GEMM
  // stages are MFMA-bound (3D 127^3: 1259 -> 834 ms per step); on smaller fronts they are bound
  // by memory and latency (2D 1023^2, fronts of up to 2045 pivots: 21.2 -> 20.9 ms), where the
  // split only costs accuracy (refinement iterations)
  double wbig = 0, wall = 0;
  for (int f : level_fronts) {
    const double n = P.fronts[f].n, N = n + P.fronts[f].m, wf = n * N * N;
    wall += wf;
    if (n >= 1024) wbig += wf;
  }
  const bool sp = split && 2 * wbig >= wall && wall > 0;
  // left-looking leaf fronts take the entries of their below-diagonal tiles in the panel step
  // (fly): only the diagonal blocks and the trailing block are zeroed and scattered
  bool fly = left && !P.rest_off.empty();
  for (int i = 0; fly && i < L.f_cnt; i++) fly = P.fly_off[S.ids[L.f_off + i]] >= 0;
  // 2. front assembly
  {
    double zb, ab;
    asm_work(L.f_off, L.f_cnt, zb, ab);
    if (fly) {
      zb = ab = 0;
      for (int i = 0; i < L.f_cnt; i++) {
        const int f = S.ids[L.f_off + i];
        const double n = P.fronts[f].n, m = P.fronts[f].m;
        zb += E * (lowerN(m) + std::floor(n / NB) * lowerN(NB) + lowerN(n - NB * std::floor(n / NB)));
        ab += (20.0 + E) * (P.rest_off[f + 1] - P.rest_off[f]);
      }
      launch(p, "k_zero_lower", 0, zb, k_zero_fly<T>,
             dim3(per_item(std::min((L.maxn + NB - 1) / NB + (L.maxm + 3) / 4, 64), L.f_cnt), L.f_cnt),
             ids + L.f_off, (const FrontDev*)p->d_fronts, arena);
      launch(p, "k_assemble", 0, ab, k_assemble<T>, dim3(per_item(L.f_grid, L.f_cnt), L.f_cnt), ids + L.f_off,
             (const MatDev*)p->d_mats, arena, (const i64*)p->d_rest_off, (const Triple*)p->d_rest_tri,
             (const i64*)p->d_tile_off, (const Tile*)p->d_tiles, (const BlockOp*)p->d_ops, (const double*)p->d_val, a32, lev32);
    } else {
      launch(p, "k_zero_lower", 0, zb, k_zero_lower<T>, dim3(per_item(std::min((L.maxN + 3) / 4, 256), L.f_cnt), L.f_cnt),
             ids + L.f_off, (const MatDev*)p->d_mats, arena);
      launch(p, "k_assemble", 0, ab, k_assemble<T>, dim3(per_item(L.f_grid, L.f_cnt), L.f_cnt), ids + L.f_off,
             (const MatDev*)p->d_mats, arena, (const i64*)p->d_tri_off, (const Triple*)p->d_tri,
             (const i64*)p->d_tile_off, (const Tile*)p->d_tiles, (const BlockOp*)p->d_ops, (const double*)p->d_val, a32, lev32);
    }
  }
  const int GW = 256;
  // 128 x 128 tiles (half the operand traffic) for a stage whose largest front has 16+ of them
  // and whose fronts have 2048+ in all: 3D 95^3 FP64 trailing update 319 -> 302 ms; fronts of a
  // few 128-tiles waste the partial tiles (256-problem batch FP64 Schur 13.5 -> 27.3 ms), and
  // the 2D stages stay below 2048 tiles
  // The trailing update is blocked on two levels: a finished group of GW columns updates only
  // the rest of its outer group of GW2 columns (K = GW), and the finished outer group updates
  // the trailing matrix once with K = GW2.  At K = GW the tile GEMMs are bound by operand and
  // C traffic: with K = 1024 the 3D 95^3 FP64 trailing update goes from 43 to 52 TFLOP/s, and
  // the FP32 / BF16-split tiles of a 16k-front microbenchmark from 76 / 103 to 126 / 247;
  // GW2 = 2048 measured the same as 1024.
  const int GW2 = p->tune_gw2;
  // 128 x 128 GEMM tiles for a stage over `fronts` fronts of at most `per_front` such tiles
  auto wide = [&](i64 per_front, i64 fronts) {
    return p->tune_tile == 128 || (p->tune_tile == 0 && per_front >= 16 && per_front * fronts >= 2048);
  };
  // levels of many fronts: keep all tiles of a front on one XCD (one L2 fetches its shared
  // diagonal-block columns); levels of few fronts spread each front over the whole chip
  bool xcd = L.f_cnt >= 64;
  int act = L.f_cnt;  // fronts taking part in the current step (a prefix: n > k0)
  std::vector<int> seg_t, suf;  // per active front: panel row tiles, flops, bytes of the step
  std::vector<double> seg_f, seg_b;
  // suf[i] = max t[i..cnt); then, unless the runs would save at least half of the one
  // launch's workgroups, suf is flattened to one run (each launch costs a few microseconds)
  auto suffix_max = [&](const std::vector<int>& t, int cnt) {
    suf.assign(cnt + 1, 0);
    for (int i = cnt - 1; i >= 0; i--) suf[i] = std::max(suf[i + 1], t[i]);
    i64 one = (i64)suf[0] * cnt, runs = 0;
    for (int i = 0; i < cnt;) {
      int e = i + 1;
      while (e < cnt && 2 * suf[e] > suf[i]) e++;
      runs += (i64)suf[i] * (e - i);
      i = e;
    }
    if (2 * runs > one)
      for (int i = 0; i < cnt; i++) suf[i] = suf[0];
  };
  auto xcd_grid = [&](int per_front) { return (unsigned)(8 * ((act + 7) / 8) * per_front); };
  // GEMM launches over fewer fronts: a 2-D grid (tiles rounded up to a multiple of 8, front)
  // whose kernels take each front's tiles in band order, an eighth per XCD (xt = -1)
  auto band_grid = [](int per_front, int fronts) { return dim3((unsigned)((per_front + 7) / 8 * 8), fronts); };
  auto lower_elems = [](double s, double rend, double cend) {  // lower part of [s,rend) x [s,cend)
    const double nc = cend - s, nr = rend - s;
    return nc > 0 ? nc * nr - nc * (nc - 1) / 2 : 0.0;
  };
  // right-looking levels: a column group whose first step runs fused (k_chol_dpanel) applies
  // its block columns lazily -- each step's panel takes the GEMM over the group's earlier
  // columns and updates the group's later diagonal blocks -- instead of a group GEMM launch
  // after every step (one launch fewer on the step chain)
  bool lazy = false;
  // Left-looking levels take block columns in pairs (k_chol_diag, k_chol_diag2, k_chol_panel2)
  // where every front of the plan's level reaching k0 has two full blocks left: the panel below
  // both blocks then reads its rows' finished columns once per 128 columns, and the factors are
  // bit-identical to single steps (2D 1023^2 FP32 step 19.8 -> 19.4 ms, 3D 127^3 1231 -> 1226
  // ms; FP64, with the MFMA-phase priority of the 64 x 128 GEMM: 3D panels 296 -> 283 ms, 2D
  // step 21.7 -> 21.3 ms).
  auto paired = [&](int k0) {
    if (!left) return false;
    bool any = false;
    for (int f : level_fronts) {
      const int n = P.fronts[f].n;
      if (n <= k0) continue;
      if (n < k0 + 2 * NB) return false;
      any = true;
    }
    return any;
  };
  for (int k0 = 0; k0 < L.maxn; k0 += NB) {
    while (act > 0 && P.fronts[S.ids[L.f_off + act - 1]].n <= k0) act--;
    xcd = act >= 64;
    if (act > 0 && paired(k0)) {
      double fd = 0, fd2 = 0, bd2 = 0;
      seg_t.assign(act, 0);
      seg_f.assign(act, 0.0);
      seg_b.assign(act, 0.0);
      for (int i = 0; i < act; i++) {
        const FrontInfo& F = P.fronts[S.ids[L.f_off + i]];
        const double N = F.n + F.m, kb = NB, K = k0, nr2 = N - k0 - 2 * NB;
        fd += kb * kb * kb / 3.0 + kb * kb * K;
        // first panel tile of block k0 (64 rows), then block k0 + 64's diagonal block
        fd2 += kb * kb * kb + 2.0 * kb * kb * K + kb * kb * kb / 3.0 + kb * kb * (K + kb);
        bd2 += E * (2.0 * kb * K + kb * (K + kb) + 2.0 * kb * kb);
        seg_t[i] = (int)((nr2 + NB - 1) / NB);
        seg_f[i] = nr2 * (4.0 * kb * kb + 4.0 * kb * K);  // both blocks: GEMMs and TRSMs
        seg_b[i] = (fly ? E : 2 * E) * nr2 * 2 * kb + E * nr2 * K;
      }
      launch(p, "k_chol_diag", fd, 0, k_chol_diag<T>, dim3(act), ids + L.f_off, (const FrontDev*)p->d_fronts,
             arena, k0, 1, p->d_err);
      launch(p, "k_chol_diag", fd2, bd2, k_chol_diag2<T>, dim3(act), ids + L.f_off, (const FrontDev*)p->d_fronts,
             arena, k0, fly ? 1 : 0, (const int*)p->d_fly_tab, (const Triple*)p->d_fly_tri, (const double*)p->d_val,
             p->d_err);
      suffix_max(seg_t, act);
      for (int i = 0; i < act;) {
        const int pt = suf[i];
        int e = i;
        double f = 0, b = 0;
        while (e < act && (e == i || 2 * suf[e] > pt)) {
          f += seg_f[e];
          b += seg_b[e];
          e++;
        }
        const int c = e - i;
        const bool xc = c >= 64;
        if (pt > 0)
          launch(p, "k_chol_panel", f, b, k_chol_panel2<T>, xc ? dim3(8 * ((c + 7) / 8) * pt) : dim3(pt, c),
                 ids + L.f_off + i, (const FrontDev*)p->d_fronts, arena, k0, xc ? pt : 0, c, fly ? 1 : 0,
                 (const int*)p->d_fly_tab, (const Triple*)p->d_fly_tri, (const double*)p->d_val);
        i = e;
      }
      k0 += NB;  // the pair's second block column
      continue;
    }
    const int g0 = k0 / GW * GW, gend = g0 + GW;
    int rows = 0;
    for (int i = 0; i < act; i++) {
      const FrontInfo& F = P.fronts[S.ids[L.f_off + i]];
      if (k0 < F.n) rows = std::max(rows, F.n + F.m - k0 - std::min(NB, F.n - k0));
    }
    // right-looking steps whose panel fits one round of workgroups (2 per CU: the fused kernel
    // repeats the diagonal factorization in each) run diagonal block and panel in one launch;
    // decided over the whole level (above), whose fronts bound this pass's, so a fused launch
    // of this pass fits one round too
    const int ptd = std::max(1, (rows + NB - 1) / NB);
    i64 fit = 0;  // workgroups of the step over the whole level
    if (!left) {
      int rows_all = 0, act_all = 0;
      for (int f : level_fronts) {
        const FrontInfo& F = P.fronts[f];
        if (k0 >= F.n) continue;
        act_all++;
        rows_all = std::max(rows_all, F.n + F.m - k0 - std::min(NB, F.n - k0));
      }
      fit = (i64)std::max(1, (rows_all + NB - 1) / NB) * act_all;
    }
    if (k0 == g0) lazy = !left && fit <= 512;
    const bool fused = !left && (lazy || fit <= 512);
    const int h0 = k0 / GW2 * GW2, hend = h0 + GW2;
    const bool group_done = k0 + NB >= gend || k0 + NB >= L.maxn;
    const bool outer_done = k0 + NB >= hend || k0 + NB >= L.maxn;
    int tin = 0, ttr = 0, ttr128 = 0, tou = 0, tou128 = 0;
    double fdiag = 0, fpan = 0, bpan = 0, fin = 0, bin = 0, ftr = 0, btr = 0, fou = 0, bou = 0;
    seg_t.assign(act, 0);
    seg_f.assign(act, 0.0);
    seg_b.assign(act, 0.0);
    for (int i = 0; i < act; i++) {
      const FrontInfo& F = P.fronts[S.ids[L.f_off + i]];
      if (k0 >= F.n) continue;
      int kb = std::min(NB, F.n - k0), N = F.n + F.m;
      double K = left ? k0 : lazy ? k0 - g0 : 0;
      double s0 = k0 + kb, nr = N - s0;
      // the block's factorization (+ left-looking: its update by the K finished columns)
      fdiag += (double)kb * kb * kb / 3.0 + (double)kb * kb * K * (left ? 1 : 0);
      // TRSM (substitution count) + left-looking GEMM + (lazy groups) the later diagonal
      // blocks' update
      const double dcols = lazy ? std::max(0.0, std::min(F.n, gend) - s0) : 0.0;
      seg_t[i] = (N - k0 - kb + NB - 1) / NB;
      seg_f[i] = nr * kb * kb + 2.0 * nr * kb * K + dcols * kb * kb;
      seg_b[i] = (fly ? E : 2 * E) * nr * kb + E * nr * K;
      fpan += seg_f[i];
      bpan += seg_b[i];
      if (left) continue;
      const int ce = std::min(F.n, gend);
      if (!lazy && k0 + kb < ce) {
        tin = std::max(tin, tiles_of(k0 + kb, N, ce));
        const double e = lower_elems(s0, N, ce);
        fin += 2.0 * kb * e;
        bin += 2 * E * e + E * kb * (nr + ce - s0);
      }
      if (group_done) {  // the group's update of the rest of its outer group
        const int kg = std::min(GW, F.n - g0), sg = g0 + kg, ce = std::min(F.n, hend);
        if (sg < ce) {
          ttr = std::max(ttr, tiles_of(sg, N, ce));
          ttr128 = std::max(ttr128, tiles_of(sg, N, ce, 128));
          const double e = lower_elems(sg, N, ce);
          ftr += 2.0 * kg * e;
          btr += 2 * E * e + E * kg * (N - sg + ce - sg);
        }
      }
      if (outer_done) {  // the outer group's trailing update
        const int kg = std::min(GW2, F.n - h0), sg = h0 + kg;
        if (sg < F.n) {
          tou = std::max(tou, tiles_of(sg, N, F.n));
          tou128 = std::max(tou128, tiles_of(sg, N, F.n, 128));
          const double e = lower_elems(sg, N, F.n);
          fou += 2.0 * kg * e;
          bou += 2 * E * e + E * kg * (N - sg + F.n - sg);
        }
      }
    }
    if (fused)
      launch(p, "k_chol_dpanel", fdiag + fpan, bpan, k_chol_dpanel<T>, xcd ? dim3(xcd_grid(ptd)) : dim3(ptd, act),
             ids + L.f_off, (const FrontDev*)p->d_fronts, arena, k0, lazy ? g0 : k0, lazy ? gend : 0,
             xcd ? ptd : 0, act, p->d_arrive, p->d_err);
    else
      launch(p, "k_chol_diag", fdiag, 0, k_chol_diag<T>, dim3(act), ids + L.f_off, (const FrontDev*)p->d_fronts,
             arena, k0, left ? 1 : 0, p->d_err);
    // the panel launches over at most a few runs of the active fronts: a run ends where the
    // largest row-tile count of the fronts after it has halved, so that a level of many small
    // fronts and a few tall ones does not give every front the tallest one's grid
    suffix_max(seg_t, act);
    for (int i = 0; !fused && rows > 0 && i < act;) {
      const int pt = suf[i];
      int e = i;
      double f = 0, b = 0;
      while (e < act && (e == i || 2 * suf[e] > pt)) {
        f += seg_f[e];
        b += seg_b[e];
        e++;
      }
      const int c = e - i;
      const bool xc = c >= 64;
      if (pt > 0)
        launch(p, "k_chol_panel", f, b, k_chol_panel<T>, xc ? dim3(8 * ((c + 7) / 8) * pt) : dim3(pt, c),
               ids + L.f_off + i, (const FrontDev*)p->d_fronts, arena, k0, left ? 1 : 0, xc ? pt : 0, c,
               fly ? 1 : 0, (const int*)p->d_fly_tab, (const Triple*)p->d_fly_tri, (const double*)p->d_val);
      i = e;
    }
    if (tin > 0)
      launch(p, "k_gemm_group", fin, bin, gemm_fn<T, 0, 64>(sp), xcd ? dim3(xcd_grid(tin)) : band_grid(tin, act),
             ids + L.f_off, (const FrontDev*)p->d_fronts, arena, k0, NB, gend, xcd ? tin : -1, act);
    auto trail = [&](int t64, int t128, double fl, double by, int kstart, int kw, int ccap) {
      if (wide(t128, act) && kGemm8<T>)
        launch_b(p, "k_gemm_trail", fl, by, gemm_fn8<T, 1>(), band_grid(t128, act), dim3(512), ids + L.f_off,
                 (const FrontDev*)p->d_fronts, arena, kstart, kw, ccap, -1, act);
      else if (wide(t128, act))
        launch(p, "k_gemm_trail", fl, by, gemm_fn<T, 1, 128>(sp), band_grid(t128, act), ids + L.f_off,
               (const FrontDev*)p->d_fronts, arena, kstart, kw, ccap, -1, act);
      else
        launch(p, "k_gemm_trail", fl, by, gemm_fn<T, 1, 64>(sp), xcd ? dim3(xcd_grid(t64)) : band_grid(t64, act),
               ids + L.f_off, (const FrontDev*)p->d_fronts, arena, kstart, kw, ccap, xcd ? t64 : -1, act);
    };
    if (ttr > 0) trail(ttr, ttr128, ftr, btr, g0, GW, hend);
    if (tou > 0) trail(tou, tou128, fou, bou, h0, GW2, 0x7fffffff);
  }
  // 4. Schur complement into the trailing block (every front of the level)
  if (L.maxm > 0) {
    // per front: 64- and 128-tiles of the trailing block, flops, bytes; launched over runs
    // ending where the largest tile count after them halves, as the panel steps
    std::vector<int> t64(L.f_cnt), t128(L.f_cnt);
    seg_f.assign(L.f_cnt, 0.0);
    seg_b.assign(L.f_cnt, 0.0);
    for (int i = 0; i < L.f_cnt; i++) {
      const FrontInfo& F = P.fronts[S.ids[L.f_off + i]];
      t64[i] = tiles_of(F.n, F.n + F.m, F.n + F.m);
      t128[i] = tiles_of(F.n, F.n + F.m, F.n + F.m, 128);
      seg_f[i] = (double)F.n * F.m * (F.m + 1);
      seg_b[i] = E * F.n * F.m + 2 * E * lowerN(F.m);
    }
    suffix_max(t128, L.f_cnt);
    std::vector<int> suf128 = suf;
    suffix_max(t64, L.f_cnt);
    for (int i = 0; i < L.f_cnt;) {
      const int tiles = suf[i], tiles128 = suf128[i];
      int e = i;
      double fl = 0, by = 0;
      while (e < L.f_cnt && (e == i || 2 * suf[e] > tiles)) {
        fl += seg_f[e];
        by += seg_b[e];
        e++;
      }
      const int c = e - i;
      const bool xc = c >= 64;
      auto grid = [&](int t) { return xc ? dim3(8 * ((c + 7) / 8) * t) : band_grid(t, c); };
      if (tiles > 0) {
        if (wide(tiles128, c) && kGemm8<T>)
          launch_b(p, "k_gemm_schur", fl, by, gemm_fn8<T, 2>(), grid(tiles128), dim3(512), ids + L.f_off + i,
                   (const FrontDev*)p->d_fronts, arena, 0, 0, 0, xc ? tiles128 : -1, c);
        else if (wide(tiles128, c))
          launch(p, "k_gemm_schur", fl, by, gemm_fn<T, 2, 128>(sp), grid(tiles128), ids + L.f_off + i,
                 (const FrontDev*)p->d_fronts, arena, 0, 0, 0, xc ? tiles128 : -1, c);
        else
          launch(p, "k_gemm_schur", fl, by, gemm_fn<T, 2, 64>(sp), grid(tiles), ids + L.f_off + i,
                 (const FrontDev*)p->d_fronts, arena, 0, 0, 0, xc ? tiles : -1, c);
      }
      i = e;
    }
  }
  return UMF_OK;
  }();
  if (lrc) return lrc;
  // 5. separators of this level receive both neighbours' Schur blocks
  if (L.s_cnt > 0) {
    double by = 0;
    for (int i = 0; i < L.s_cnt; i++) by += 4 * E * lowerN(P.seps[S.ids[L.s_off + i]].k);
    launch(p, "k_sep_add", 0, by, k_sep_add<T>,
           dim3(per_item(std::min((std::max(L.max_k, 1) + 3) / 4, 256), L.s_cnt), L.s_cnt), ids + L.s_off,
           (const SepDev*)p->d_seps, (const FrontDev*)p->d_fronts, arena);
  }
  return UMF_OK;
}

// Run the factorization over the dirty schedule on the plan stream, fronts of type T in `arena`.
// ids_pre: the schedule's id list already on the device (a factorization being captured into a
// graph: no synchronization, allocation or copy may enter the stream).
template <class T>
int run_factor(umf_plan* p, const Plan::Dirty& D, T* arena, bool split, const int* ids_pre = nullptr,
               const Sched* S_pre = nullptr) {
  const Plan& P = p->P;
  Sched S_own;
  if (!S_pre) S_own = make_sched(P, D, true);
  const Sched& S = S_pre ? *S_pre : S_own;
  const int* ids = ids_pre;
  int rc = factor_begin(p, S, ids);
  if (rc) return rc;
  for (int ell = P.L - 1; ell >= 0; ell--) {
    rc = factor_level<T>(p, S, ids, ell, arena, split, nullptr, INT_MAX);
    if (rc) return rc;
  }
  g_level = -1;
  HIPCHK(hipGetLastError());
  return UMF_OK;
}

// The mixed form (UMF_PREC_FP32T64): levels TOP64 and below in FP32 (arena32), the top levels in
// FP64 (arena); the lowest FP64 level assembles the FP32 blocks of the level below, converting.
int run_factor_mixed(umf_plan* p, const Plan::Dirty& D, const int* ids_pre, const Sched* S_pre) {
  const Plan& P = p->P;
  Sched S_own;
  if (!S_pre) S_own = make_sched(P, D, true);
  const Sched& S = S_pre ? *S_pre : S_own;
  const int* ids = ids_pre;
  int rc = factor_begin(p, S, ids);
  if (rc) return rc;
  for (int ell = P.L - 1; ell >= 0; ell--) {
    rc = ell >= TOP64 ? factor_level<float>(p, S, ids, ell, p->arena32, false, nullptr, INT_MAX)
                      : factor_level<double>(p, S, ids, ell, p->arena, false, p->arena32, TOP64);
    if (rc) return rc;
  }
  g_level = -1;
  HIPCHK(hipGetLastError());
  return UMF_OK;
}

// Factor over a dirty schedule in the precision the plan is set to (the FP32 front pool is
// allocated on first use).
// LU factorization over the dirty schedule (unsymmetric operators, FP64): per level, zero and
// assemble the full separator and front matrices, then right-looking 64-column panels of every
// front (panel with partial pivoting among the interior rows, interchanges on the other columns,
// U row block, trailing update including the boundary block's Schur complement), the pivot-order
// DOF lists, and the separators' Schur accumulation.  Every front's arithmetic depends on the
// front alone, as in the Cholesky schedule.
int run_factor_lu(umf_plan* p, const Plan::Dirty& D) {
  const Plan& P = p->P;
  Sched S = make_sched(P, D);
  int rc = ensure_ids(p, S.ids.size());
  if (rc) return rc;
  HIPCHK(hipStreamSynchronize(p->stream));
  p->ids_h = S.ids;
  if (!p->ids_h.empty())
    HIPCHK(hipMemcpyAsync(p->d_ids, p->ids_h.data(), sizeof(int) * p->ids_h.size(), hipMemcpyHostToDevice,
                          p->stream));
  const int* ids = p->d_ids;
  const int NF = P.nfronts();
  hipLaunchKernelGGL(k_zero_ints, dim3(1), dim3(64), 0, p->stream, p->d_err, 4);
  double* A = p->arena;
  for (int ell = P.L - 1; ell >= 0; ell--) {
    const Sched::Lvl& L = S.lv[ell];
    g_level = ell;
    if (L.sm_cnt > 0) {
      launch(p, "k_zero_full", 0, 0, k_zero_full<double>,
             dim3(per_item(std::min(std::max(L.max_k, 1), 256), L.sm_cnt), L.sm_cnt), ids + L.sm_off,
             (const MatDev*)p->d_mats, A);
      launch(p, "k_assemble", 0, 0, k_assemble<double>, dim3(per_item(L.s_grid, L.sm_cnt), L.sm_cnt), ids + L.sm_off,
             (const MatDev*)p->d_mats, A, (const i64*)p->d_tri_off, (const Triple*)p->d_tri,
             (const i64*)p->d_tile_off, (const Tile*)p->d_tiles, (const BlockOp*)p->d_ops, (const double*)p->d_val,
           (const float*)nullptr, INT_MAX);
    }
    if (L.f_cnt == 0) continue;
    launch(p, "k_zero_full", 0, 0, k_zero_full<double>, dim3(per_item(std::min(L.maxN, 256), L.f_cnt), L.f_cnt),
           ids + L.f_off, (const MatDev*)p->d_mats, A);
    launch(p, "k_assemble", 0, 0, k_assemble<double>, dim3(per_item(L.f_grid, L.f_cnt), L.f_cnt), ids + L.f_off,
           (const MatDev*)p->d_mats, A, (const i64*)p->d_tri_off, (const Triple*)p->d_tri, (const i64*)p->d_tile_off,
           (const Tile*)p->d_tiles, (const BlockOp*)p->d_ops, (const double*)p->d_val,
           (const float*)nullptr, INT_MAX);
    int act = L.f_cnt;  // fronts sorted by pivot-block size: the prefix with n > k0 takes part
    for (int k0 = 0; k0 < L.maxn; k0 += NB) {
      while (act > 0 && P.fronts[S.ids[L.f_off + act - 1]].n <= k0) act--;
      int maxN = 0;
      double fl = 0;
      for (int i = 0; i < act; i++) {
        const FrontInfo& F = P.fronts[S.ids[L.f_off + i]];
        maxN = std::max(maxN, F.n + F.m);
        const double kb = std::min(NB, F.n - k0), r = F.n + F.m - k0 - kb;
        fl += kb * kb * kb / 3.0 + 2.0 * r * kb * kb + 2.0 * r * r * kb;  // LU form: panel, TRSM, trailing
      }
      launch(p, "k_lu_panel", 0, 0, k_lu_panel, dim3(act), ids + L.f_off, (const FrontDev*)p->d_fronts, A, k0,
             p->d_ipiv, p->d_uinv, p->d_err);
      launch(p, "k_lu_swap", 0, 0, k_lu_swap, dim3((maxN + 255) / 256, act), ids + L.f_off,
             (const FrontDev*)p->d_fronts, A, k0, (const int*)p->d_ipiv);
      const int rest = maxN - k0 - 1;  // columns right of the first panel column (an upper bound)
      if (rest > 0) {
        const int tb = (rest + NB - 1) / NB;
        launch(p, "k_lu_trsm", 0, 0, k_lu_trsm, dim3(tb, act), ids + L.f_off, (const FrontDev*)p->d_fronts, A, k0,
               p->d_ut);
        launch(p, "k_lu_gemm", fl, 0, k_lu_gemm, dim3(tb * tb, act), ids + L.f_off, (const FrontDev*)p->d_fronts,
               A, k0, (const double*)p->d_ut);
      }
    }
    launch(p, "k_lu_perm", 0, 0, k_lu_perm, dim3(L.f_cnt), ids + L.f_off, (const FrontDev*)p->d_fronts,
           (const int*)p->d_ipiv, (const int*)p->d_dofs, p->d_perm);
    if (L.s_cnt > 0)
      launch(p, "k_sep_add", 0, 0, k_sep_add_full<double>,
             dim3(per_item(std::min(std::max(L.max_k, 1), 256), L.s_cnt), L.s_cnt), ids + L.s_off,
             (const SepDev*)p->d_seps, (const FrontDev*)p->d_fronts, A);
  }
  (void)NF;
  g_level = -1;
  HIPCHK(hipGetLastError());
  return UMF_OK;
}

// S_pre: the pass's schedule when the caller holds it (the full factorization's, whose id list
// ids_pre is already on the device).
int factor_any(umf_plan* p, const Plan::Dirty& D, const int* ids_pre = nullptr, const Sched* S_pre = nullptr) {
  if (p->lu) return run_factor_lu(p, D);
  if (p->prec != UMF_PREC_FP64) {
    if (!p->arena32)
      HIPCHK(hipMalloc((void**)&p->arena32, sizeof(float) * std::max<i64>(p->P.arena_elems, 1)));
    if (p->prec == UMF_PREC_FP32T64) return run_factor_mixed(p, D, ids_pre, S_pre);
    return run_factor<float>(p, D, p->arena32, p->prec == UMF_PREC_BF16X3, ids_pre, S_pre);
  }
  return run_factor<double>(p, D, p->arena, false, ids_pre, S_pre);
}

// The full factorization's dirty set and Cholesky schedule, built once per plan, with its id
// list uploaded once to d_fids (read, never written, by every later full factorization).
int ensure_full_sched(umf_plan* p) {
  if (p->full_sched) return UMF_OK;
  std::unique_ptr<Plan::Dirty> D(new Plan::Dirty(p->P.all_dirty()));
  std::unique_ptr<Sched> S(new Sched(make_sched(p->P, *D, true)));
  if (p->d_fids) (void)hipFree(p->d_fids);
  p->d_fids = nullptr;
  HIPCHK(hipMalloc((void**)&p->d_fids, sizeof(int) * std::max<size_t>(S->ids.size(), 1)));
  if (!S->ids.empty())
    HIPCHK(hipMemcpyAsync(p->d_fids, S->ids.data(), sizeof(int) * S->ids.size(), hipMemcpyHostToDevice, p->stream));
  p->full_dirty = std::move(D);
  p->full_sched = std::move(S);
  return UMF_OK;
}

// The full factorization as one HIP graph launch: captured on first use (and again after a
// precision change) with the plan's fixed buffers and its own id list, then replayed.  Every
// input it reads (operator values, front pool) lives at a fixed address, so a replay computes
// what the eager launches would.
int factor_graph(umf_plan* p, const Plan::Dirty& D) {
  if (!p->fgraph || p->fgraph_prec != p->prec) {
    if (p->fgraph) (void)hipGraphExecDestroy(p->fgraph);
    p->fgraph = nullptr;
    int rs = ensure_full_sched(p);  // the Cholesky schedule run_factor replays
    if (rs) return rs;
    // device queries before the capture: the persistent launches' grids
    (void)persistent_grid(p, (const void*)k_leaf<double>, 1);
    (void)persistent_grid(p, (const void*)k_leaf<float>, 1);
    HIPCHK(hipStreamSynchronize(p->stream));
    if (p->prec != UMF_PREC_FP64 && !p->arena32)
      HIPCHK(hipMalloc((void**)&p->arena32, sizeof(float) * std::max<i64>(p->P.arena_elems, 1)));
    HIPCHK(hipStreamBeginCapture(p->stream, hipStreamCaptureModeThreadLocal));
    const int rc = factor_any(p, D, p->d_fids, p->full_sched.get());
    hipGraph_t g = nullptr;
    const hipError_t ec = hipStreamEndCapture(p->stream, &g);
    if (rc || ec != hipSuccess || !g) {
      if (g) (void)hipGraphDestroy(g);
      return rc ? rc : fail(UMF_EDEVICE, std::string("graph capture failed: ") + hipGetErrorString(ec));
    }
    const hipError_t ei = hipGraphInstantiate(&p->fgraph, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (ei != hipSuccess) {
      p->fgraph = nullptr;
      return fail(UMF_EDEVICE, std::string("graph instantiation failed: ") + hipGetErrorString(ei));
    }
    p->fgraph_prec = p->prec;
  }
  HIPCHK(hipGraphLaunch(p->fgraph, p->stream));
  return UMF_OK;
}

// Copy cnt front-pool entries at offset off of the current factors to the host as double; lev:
// the tree level of the matrix they belong to (the mixed form holds levels < TOP64 in FP64).
int read_pool(umf_plan* p, i64 off, size_t cnt, std::vector<double>& h, int lev) {
  h.resize(cnt);
  if (!cnt) return UMF_OK;
  if (p->fact_prec == UMF_PREC_FP64 || (p->fact_prec == UMF_PREC_FP32T64 && lev < TOP64)) {
    HIPCHK(hipMemcpy(h.data(), p->arena + off, sizeof(double) * cnt, hipMemcpyDeviceToHost));
    return UMF_OK;
  }
  std::vector<float> f(cnt);
  HIPCHK(hipMemcpy(f.data(), p->arena32 + off, sizeof(float) * cnt, hipMemcpyDeviceToHost));
  for (size_t i = 0; i < cnt; i++) h[i] = f[i];
  return UMF_OK;
}

int check_err(umf_plan* p) {
  int h[4];
  // stream-ordered (a serving slot's stream must not wait on other slots, as a null-stream
  // copy would)
  HIPCHK(hipMemcpyAsync(h, p->d_err, sizeof(h), hipMemcpyDeviceToHost, p->stream));
  HIPCHK(hipStreamSynchronize(p->stream));
  if (h[0]) {
    const FrontInfo& F = p->P.fronts[h[1]];
    char buf[256];
    if (p->lu)
      snprintf(buf, sizeof buf, "zero pivot: level %d box %d pivot column %d (front of n=%d): singular pivot block",
               F.level, F.box, h[2], F.n);
    else if (p->P.forest)
      snprintf(buf, sizeof buf,
               "non-positive pivot: problem %d level %d box %d pivot column %d (front of n=%d): operator not SPD",
               F.prob, F.level, F.box, h[2], F.n);
    else
      snprintf(buf, sizeof buf,
               "non-positive pivot: level %d box %d pivot column %d (front of n=%d): operator not SPD", F.level, F.box,
               h[2], F.n);
    g_err = buf;
    return 1 + h[2];
  }
  if (h[3]) return fail(UMF_EDEVICE, "factorization: a workgroup wait exceeded its spin limit (results invalid)");
  return UMF_OK;
}

#define DOT_BLOCKS 512

int ensure_solve_scratch(umf_plan* p) {
  const size_t nw = (size_t)p->P.n * RB, nx = (size_t)std::max<i64>(p->xs_elems, 1);
  if (p->fact_prec == UMF_PREC_FP64) {
    if (!p->d_xs) HIPCHK(hipMalloc((void**)&p->d_xs, sizeof(double) * nx));
    if (!p->d_w) HIPCHK(hipMalloc((void**)&p->d_w, sizeof(double) * nw));
    return UMF_OK;
  }
  if (p->fact_prec == UMF_PREC_FP32T64) {  // the top levels sweep in FP64
    if (!p->d_xs) HIPCHK(hipMalloc((void**)&p->d_xs, sizeof(double) * nx));
    if (!p->d_w) HIPCHK(hipMalloc((void**)&p->d_w, sizeof(double) * nw));
  }
  if (!p->d_xs32) HIPCHK(hipMalloc((void**)&p->d_xs32, sizeof(float) * nx));
  if (!p->d_w32) {  // zeroed once: columns a narrower refinement never writes hold finite values
    HIPCHK(hipMalloc((void**)&p->d_w32, sizeof(float) * nw));
    HIPCHK(hipMemsetAsync(p->d_w32, 0, sizeof(float) * nw, p->stream));
  }
  if (!p->d_cg) HIPCHK(hipMalloc((void**)&p->d_cg, sizeof(double) * 5 * nw));
  if (!p->d_part) HIPCHK(hipMalloc((void**)&p->d_part, sizeof(double) * std::max(DOT_BLOCKS * 32, 2 * NCH * 16)));
  if (!p->d_dots) HIPCHK(hipMalloc((void**)&p->d_dots, sizeof(double) * 5 * 32));
  return UMF_OK;
}

AllLists all_lists(const Plan& P) {
  AllLists A;
  int L = P.L;
  A.f_off.assign(L, 0); A.f_cnt.assign(L, 0); A.s_off.assign(L, 0); A.s_cnt.assign(L, 0);
  for (int ell = 0; ell < L; ell++) {
    A.f_off[ell] = (int)A.ids.size();
    for (int f : P.box_gid[ell]) A.ids.push_back(f);
    A.f_cnt[ell] = (int)A.ids.size() - A.f_off[ell];
    // smallest fronts first: the forward sweep launches each size class with its own kernel
    std::stable_sort(A.ids.begin() + A.f_off[ell], A.ids.end(), [&](int a, int b) {
      return P.fronts[a].n + P.fronts[a].m < P.fronts[b].n + P.fronts[b].m;
    });
    A.s_off[ell] = (int)A.ids.size();
    if (ell >= 1)
      for (int s : P.sep_gid[ell]) A.ids.push_back(s);
    A.s_cnt[ell] = (int)A.ids.size() - A.s_off[ell];
  }
  return A;
}

// The plan's sweep lists on the device, built and uploaded on first use.  They outlive every
// launch that reads them, so later sweeps neither rebuild nor wait.
int ensure_sweep_lists(umf_plan* p) {
  if (!p->sweep_lists) p->sweep_lists.reset(new AllLists(all_lists(p->P)));
  if (p->d_sids) return UMF_OK;
  const AllLists& A = *p->sweep_lists;
  HIPCHK(hipMalloc((void**)&p->d_sids, sizeof(int) * std::max<size_t>(A.ids.size(), 1)));
  if (!A.ids.empty())
    HIPCHK(hipMemcpyAsync(p->d_sids, A.ids.data(), sizeof(int) * A.ids.size(), hipMemcpyHostToDevice, p->stream));
  return UMF_OK;
}

// Epoch of the next blocked sweep launch: flags still holding an earlier epoch read as unset.
int next_epoch(umf_plan* p) {
  p->epoch = p->epoch == 0x7fffffff ? 1 : p->epoch + 1;
  return p->epoch;
}

// Levels of few (<= 16) fronts, or of at most 128 fronts of 2048+ rows (max_rows: the level's
// largest n + m), are swept by 64-row blocks, one workgroup each (k_fwd_blk / k_bwd_blk): one
// workgroup per front would leave most of the 256 CUs idle while it streams its factor (3D 127^3:
// 64 fronts of 8557 rows at 0.7 TB/s).  Levels of many fronts keep one workgroup per front
// (blocks on the 3D 127^3 leaf level, 512 fronts of 4725 rows, measured slower).
bool sweep_blocked(i64 nfronts, int max_rows) { return nfronts <= 16 || (nfronts <= 128 && max_rows >= 2048); }

// Solve (mode 0) or Apply (mode 1) of nr <= RB columns held in the row-major work block w
// (n x RB), fronts of type T in `arena`, per-front scratch xs.
template <class T>
// ids_dev: device buffer of at least all_lists(P).ids.size() ints for the id list (the no-alloc
// solve's scratch), or null for the plan's own (grown on demand).
// lo, hi, phases: the levels [lo, hi) only (hi < 0: all), the upward / forward phase (phases & 1)
// and the downward / backward phase (phases & 2); the mixed form sweeps its FP32 and FP64 levels
// in separate calls.
int sweep(umf_plan* p, int mode, int nr, T* arena, T* w, T* xs, int* ids_dev = nullptr, int lo = 0, int hi = -1,
          int phases = 3) {
  const Plan& P = p->P;
  int L = P.L;
  if (hi < 0) hi = L;
  if (!p->sweep_lists) p->sweep_lists.reset(new AllLists(all_lists(P)));
  const AllLists& A = *p->sweep_lists;
  if (!ids_dev) {
    int rc = ensure_sweep_lists(p);
    if (rc) return rc;
    ids_dev = p->d_sids;
  } else {  // caller's scratch (the no-alloc solve): the previous list may still be in flight
    HIPCHK(hipStreamSynchronize(p->stream));
    p->ids_h = A.ids;
    HIPCHK(hipMemcpyAsync(ids_dev, p->ids_h.data(), sizeof(int) * p->ids_h.size(), hipMemcpyHostToDevice,
                          p->stream));
  }
  const int* ids = ids_dev;
  int n = P.n;
  const FrontDev* fr = p->d_fronts;
  if constexpr (std::is_same<T, double>::value) {
    if (p->lu) {  // LU fronts: one workgroup per front throughout
      if (mode == 0) {
        for (int ell = L - 1; ell >= 0; ell--) {
          g_level = ell;
          launch(p, "k_solve_fwd", 0, 0, k_solve_fwd<double>, dim3(A.f_cnt[ell]), ids + A.f_off[ell], fr,
                 (const double*)arena, (const int*)p->d_dofs, w, xs, (const int*)p->d_perm);
          if (ell >= 1)
            launch(p, "k_sep_acc", 0, 0, k_sep_acc<double>, dim3(A.s_cnt[ell]), ids + A.s_off[ell],
                   (const SepDev*)p->d_seps, fr, (const int*)p->d_dofs, w, (const double*)xs);
        }
        for (int ell = 0; ell < L; ell++) {
          g_level = ell;
          launch(p, "k_lu_bwd", 0, 0, k_lu_bwd, dim3(A.f_cnt[ell]), ids + A.f_off[ell], fr, (const double*)arena,
                 (const int*)p->d_dofs, w, xs, (const double*)p->d_uinv);
        }
      } else {
        for (int ell = L - 1; ell >= 0; ell--)
          launch(p, "k_lu_apply_up", 0, 0, k_lu_apply_up, dim3(A.f_cnt[ell]), ids + A.f_off[ell], fr,
                 (const double*)arena, (const int*)p->d_dofs, w, xs, nr, (const int*)p->d_perm);
        for (int ell = 1; ell < L; ell++) {
          launch(p, "k_apply_down", 0, 0, k_apply_down<double>, dim3(A.f_cnt[ell]), ids + A.f_off[ell], fr,
                 (const double*)arena, xs, nr);
          launch(p, "k_sep_acc", 0, 0, k_sep_acc<double>, dim3(A.s_cnt[ell]), ids + A.s_off[ell],
                 (const SepDev*)p->d_seps, fr, (const int*)p->d_dofs, w, (const double*)xs);
        }
      }
      g_level = -1;
      HIPCHK(hipGetLastError());
      return UMF_OK;
    }
  }
  // per level: solve flops nr*(2n^2 + 4nm) split evenly over the two sweeps; bytes = factor read
  auto work = [&](int ell, double& fl, double& by) {
    fl = by = 0;
    for (int f : P.box_gid[ell]) {
      double nn = P.fronts[f].n, mm = P.fronts[f].m;
      fl += nr * (nn * nn + 2 * nn * mm);
      by += sizeof(T) * (nn * (nn + 1) / 2 + nn * mm + 2.0 * nr * (nn + mm));
    }
  };
  double fl, by;
  std::vector<int> maxN(L, 0);
  for (int ell = 0; ell < L; ell++)
    for (int f : P.box_gid[ell]) maxN[ell] = std::max(maxN[ell], P.fronts[f].n + P.fronts[f].m);
  auto blocked = [&](int ell) { return sweep_blocked(A.f_cnt[ell], maxN[ell]); };
  // the most interior / boundary 64-row blocks of a front of level ell; a blocked launch
  // publishes at most f_cnt x nbi blocks, which the plan's buffer holds by construction
  int blk_rc = UMF_OK;
  auto blocks = [&](int ell, int& nbi, int& nbb) {
    nbi = 1;
    nbb = 0;
    for (int f : P.box_gid[ell]) {
      const FrontInfo& F = P.fronts[f];
      nbi = std::max(nbi, (F.n + NB - 1) / NB);
      nbb = std::max(nbb, (F.m + NB - 1) / NB);
    }
    if ((i64)A.f_cnt[ell] * nbi > p->ll_blocks) blk_rc = UMF_EINVAL;
  };
  // FP64 levels of many fronts: fronts of at most 64 G rows keep the right-hand-side rows in
  // registers in the forward sweep (G = 6, 12 or 18 chunks of 64 rows), larger ones stream.
  // The level's fronts are sorted by size, so each class is a contiguous run of the list.
  auto fwd = [&](int ell) {
    work(ell, fl, by);
    g_level = ell;
    if (!blocked(ell)) {
      const int* li = ids + A.f_off[ell];
      const int cnt = A.f_cnt[ell];
      int i = 0;
      auto run_end = [&](int rows) {  // end of the run of fronts with n + m <= rows
        int e = i;
        while (e < cnt && P.fronts[A.ids[A.f_off[ell] + e]].n + P.fronts[A.ids[A.f_off[ell] + e]].m <= rows) e++;
        return e;
      };
      // work is attributed to the launches in proportion to their fronts
      auto part = [&](int c, double x) { return x * c / std::max(cnt, 1); };
      {  // fronts of at most XLR rows (FP32: XLR2) with X in LDS
        int e = run_end(XLR);
        if (e > i)
          launch(p, "k_solve_fwd", part(e - i, fl), part(e - i, by), k_solve_fwd_l<T>,
                 dim3(persistent_grid(p, (const void*)k_solve_fwd_l<T>, e - i)), li + i, e - i, p->d_err + 4, fr,
                 (const T*)arena, (const int*)p->d_dofs, w, xs);
        i = e;
        if constexpr (sizeof(T) == 4) {
          e = run_end(XLR2);
          if (e > i)
            launch(p, "k_solve_fwd", part(e - i, fl), part(e - i, by), k_solve_fwd_l<T, XLR2>,
                   dim3(persistent_grid(p, (const void*)k_solve_fwd_l<T, XLR2>, e - i)), li + i, e - i, p->d_err + 4,
                   fr, (const T*)arena, (const int*)p->d_dofs, w, xs);
          i = e;
        }
      }
      if (sizeof(T) == 8) {  // (FP32 keeps the streaming kernel: the register form measured slower there)
        int e = run_end(6 * NB);
        if (e > i)
          launch(p, "k_solve_fwd", part(e - i, fl), part(e - i, by), k_solve_fwd_r<T, 6>, dim3(e - i), li + i, fr,
                 (const T*)arena, (const int*)p->d_dofs, w, xs);
        i = e;
        e = run_end(12 * NB);
        if (e > i)
          launch(p, "k_solve_fwd", part(e - i, fl), part(e - i, by), k_solve_fwd_r<T, 12>, dim3(e - i), li + i, fr,
                 (const T*)arena, (const int*)p->d_dofs, w, xs);
        i = e;
        e = run_end(18 * NB);
        if (e > i)
          launch(p, "k_solve_fwd", part(e - i, fl), part(e - i, by), k_solve_fwd_r<T, 18>, dim3(e - i), li + i, fr,
                 (const T*)arena, (const int*)p->d_dofs, w, xs);
        i = e;
      }
      if (cnt > i)
        launch(p, "k_solve_fwd", part(cnt - i, fl), part(cnt - i, by), k_solve_fwd<T>, dim3(cnt - i), li + i, fr,
               (const T*)arena, (const int*)p->d_dofs, w, xs, (const int*)nullptr);
    } else {
      int nbi, nbb;
      blocks(ell, nbi, nbb);
      if (blk_rc) return;
      const int nf = A.f_cnt[ell];
      launch(p, "k_fwd_blk", fl, by, p->fenced ? k_fwd_blk<T, 1> : k_fwd_blk<T, 0>, dim3(nf * (nbi + nbb)),
             ids + A.f_off[ell], nf, nbi, fr, (const T*)arena, (const int*)p->d_dofs, w, xs, p->d_flags, p->d_ll,
             next_epoch(p), p->d_err);
    }
  };
  if (mode == 0) {
    if (phases & 1) {
      for (int ell = hi - 1; ell >= std::max(lo, 1); ell--) {
        fwd(ell);
        launch(p, "k_sep_acc", 0, 0, k_sep_acc<T>, dim3(A.s_cnt[ell]), ids + A.s_off[ell], (const SepDev*)p->d_seps,
               fr, (const int*)p->d_dofs, w, (const T*)xs);
      }
      if (lo == 0) fwd(0);
    }
    for (int ell = lo; (phases & 2) && ell < hi; ell++) {
      work(ell, fl, by);
      g_level = ell;
      if (!blocked(ell)) {
        // fronts of at most XLR rows (FP32: then XLR2; prefixes: the lists are sorted by size) with X
        // in LDS
        const int cnt = A.f_cnt[ell];
        auto upto = [&](int from, int rows) {
          int e = from;
          while (e < cnt && P.fronts[A.ids[A.f_off[ell] + e]].n + P.fronts[A.ids[A.f_off[ell] + e]].m <= rows) e++;
          return e;
        };
        const int e = upto(0, XLR), e2 = sizeof(T) == 4 ? upto(e, XLR2) : e;
        auto frac = [&](int a, int b) { return cnt ? (double)(b - a) / cnt : 0.0; };
        if (e > 0)
          launch(p, "k_solve_bwd", fl * frac(0, e), by * frac(0, e), k_solve_bwd_l<T>,
                 dim3(persistent_grid(p, (const void*)k_solve_bwd_l<T>, e)), ids + A.f_off[ell], e, p->d_err + 4, fr,
                 (const T*)arena, (const int*)p->d_dofs, w);
        if constexpr (sizeof(T) == 4) {
          if (e2 > e)
            launch(p, "k_solve_bwd", fl * frac(e, e2), by * frac(e, e2), k_solve_bwd_l<T, XLR2>,
                   dim3(persistent_grid(p, (const void*)k_solve_bwd_l<T, XLR2>, e2 - e)), ids + A.f_off[ell] + e,
                   e2 - e, p->d_err + 4, fr, (const T*)arena, (const int*)p->d_dofs, w);
        }
        if (cnt > e2)
          launch(p, "k_solve_bwd", fl * frac(e2, cnt), by * frac(e2, cnt), k_solve_bwd<T>, dim3(cnt - e2),
                 ids + A.f_off[ell] + e2, fr, (const T*)arena, (const int*)p->d_dofs, w, xs);
      } else {
        int nbi, nbb;
        blocks(ell, nbi, nbb);
        if (blk_rc) break;
        const int nf = A.f_cnt[ell];
        launch(p, "k_bwd_blk", fl, by, p->fenced ? k_bwd_blk<T, 1> : k_bwd_blk<T, 0>, dim3(nf * nbi),
               ids + A.f_off[ell], nf, nbi, fr, (const T*)arena, (const int*)p->d_dofs, w, xs, p->d_flags, p->d_ll,
               next_epoch(p), p->d_err);
      }
    }
  } else {
    for (int ell = hi - 1; (phases & 1) && ell >= lo; ell--)
      launch(p, "k_apply_up", 0, 0, k_apply_up<T>, dim3(A.f_cnt[ell]), ids + A.f_off[ell], fr, (const T*)arena,
             (const int*)p->d_dofs, w, xs, nr);
    for (int ell = std::max(lo, 1); (phases & 2) && ell < hi; ell++) {
      launch(p, "k_apply_down", 0, 0, k_apply_down<T>, dim3(A.f_cnt[ell]), ids + A.f_off[ell], fr,
             (const T*)arena, xs, nr);
      launch(p, "k_sep_acc", 0, 0, k_sep_acc<T>, dim3(A.s_cnt[ell]), ids + A.s_off[ell], (const SepDev*)p->d_seps, fr,
             (const int*)p->d_dofs, w, (const T*)xs);
    }
  }
  g_level = -1;
  if (blk_rc) return fail(blk_rc, "blocked sweep: more published blocks than the plan's buffer holds");
  HIPCHK(hipGetLastError());
  return UMF_OK;
}

// Sweep of the mixed form (UMF_PREC_FP32T64) on the FP32 work block w32: the FP32 levels
// [TOP64, L) in FP32, the work block converted to FP64 (d_w), the top levels in FP64 both ways,
// converted back, the FP32 levels' downward / backward phase.  Solve (mode 0) and Apply (1).
int sweep_mixed(umf_plan* p, int mode, int nr, float* w32) {
  const int L = p->P.L, top = std::min(TOP64, L);
  const i64 n16 = (i64)p->P.n * RB;
  const unsigned g = (unsigned)((n16 + 255) / 256);
  int rc = sweep<float>(p, mode, nr, p->arena32, w32, p->d_xs32, nullptr, top, L, 1);
  if (rc) return rc;
  launch(p, "k_cvt", 0, 12.0 * n16, k_cvt<double, float>, dim3(g), n16, (const float*)w32, p->d_w);
  rc = sweep<double>(p, mode, nr, p->arena, p->d_w, p->d_xs, nullptr, 0, top, 3);
  if (rc) return rc;
  launch(p, "k_cvt", 0, 12.0 * n16, k_cvt<float, double>, dim3(g), n16, (const double*)p->d_w, w32);
  return sweep<float>(p, mode, nr, p->arena32, w32, p->d_xs32, nullptr, top, L, 2);
}

// The preconditioner's sweep on the FP32 work block w (all 16 columns): the FP32 factors, or
// the mixed form's.
int precond_sweep(umf_plan* p, float* w) {
  if (p->fact_prec == UMF_PREC_FP32T64) return sweep_mixed(p, 0, RB, w);
  return sweep<float>(p, 0, RB, p->arena32, w, p->d_xs32);
}

// FP32 form of Solve for nr <= RB columns: preconditioned conjugate gradients in FP64 (FP64
// operator and residuals, FP32 factor as preconditioner), x0 = 0, each column stopped once its
// normwise backward error in the infinity norm, ||r||_inf / (||A||_inf ||x||_inf + ||b||_inf),
// is <= ir_tol.  The vectors are n x cw blocks, cw = the smallest power of two >= nr; every
// reduction runs in an order fixed by n alone (k_red_final).  Per iteration: k_spmv_dot (q = A p,
// p.q), k_cg_xr (x, r, the preconditioner's input, max |r| and |x|), the convergence check on the
// host, then the FP32 sweep, k_pc_out (z, r.z) and k_cg_p.  Returns the iteration count and the
// largest final backward error through it / berr.
int pcg(umf_plan* p, const double* in, double* out, int ld, int nr, int& it_out, double& berr_out) {
  const int n = p->P.n;
  int lcw = 0;
  while ((1 << lcw) < nr) lcw++;
  const int cw = 1 << lcw;
  const i64 nc = (i64)n << lcw, nw = (i64)n * RB;
  const unsigned ge = (unsigned)((nc + 255) / 256), gr = (unsigned)(NCH >> (4 - lcw));
  double* Bv = p->d_cg;
  double* X = Bv + nw;
  double* R = X + nw;
  double* Pd = R + nw;
  double* QZ = Pd + nw;
  double* S1 = p->d_dots + 32;  // p.q
  double* S2 = p->d_dots + 64;  // max |r| | max |x| per column
  double* rz = p->d_dots;       // r.z, current
  double* rzn = p->d_dots + 96; // r.z, next
  float* w = p->d_w32;
  const double vb = 8.0 * (double)nc;  // bytes of one n x cw FP64 vector
  const i64 nnz = (i64)p->P.ci.size();
  // ||A||_inf of the operator the plan holds
  launch(p, "k_rowabs", 0, 12.0 * nnz, k_rowabs, dim3(DOT_BLOCKS), n, (const int*)p->d_rp, (const double*)p->d_val,
         p->d_part);
  launch(p, "k_max_final", 0, 0, k_max_final, dim3(1), DOT_BLOCKS, (const double*)p->d_part, p->d_dots + 128);
  launch(p, "k_cg_init", 0, 4.0 * vb, k_cg_init, dim3(gr), n, lcw, nr, in, ld, Bv, R, X, p->d_part);
  launch(p, "k_red_final", 0, 0, k_red_final<true>, dim3(cw, 1), (const double*)p->d_part, S2);
  double h[32], na;
  HIPCHK(hipMemcpyAsync(&na, p->d_dots + 128, sizeof(double), hipMemcpyDeviceToHost, p->stream));
  HIPCHK(hipMemcpyAsync(h, S2, sizeof(h), hipMemcpyDeviceToHost, p->stream));
  // z = M^{-1} r, p = z, r.z
  launch(p, "k_pc_in", 0, 1.5 * vb, k_pc_in, dim3(ge), n, lcw, (const double*)R, w);
  int rc = precond_sweep(p, w);
  if (rc) return rc;
  launch(p, "k_pc_out", 0, 3.5 * vb, k_pc_out, dim3(gr), n, lcw, (const float*)w, (const double*)R, QZ, Pd,
         p->d_part);
  launch(p, "k_red_final", 0, 0, k_red_final<false>, dim3(cw, 1), (const double*)p->d_part, rz);
  HIPCHK(hipStreamSynchronize(p->stream));
  p->norm_a = na;
  double nb[RB];
  for (int c = 0; c < RB; c++) nb[c] = h[c];
  int active = (1 << nr) - 1;
  double berr[RB] = {0};
  int it = 0;
  while (active && it < p->ir_max) {
    it++;
    launch(p, "k_spmv_dot", 0, 12.0 * nnz + 4.0 * (double)n + 3.0 * vb, k_spmv_dot, dim3(gr), n, lcw,
           (const int*)p->d_rp, (const int*)p->d_ci, (const double*)p->d_val, (const double*)Pd, QZ, p->d_part);
    launch(p, "k_red_final", 0, 0, k_red_final<false>, dim3(cw, 1), (const double*)p->d_part, S1);
    launch(p, "k_cg_xr", 0, 6.5 * vb, k_cg_xr, dim3(gr), n, lcw, (const double*)rz, (const double*)S1, active,
           (const double*)Pd, (const double*)QZ, X, R, w, p->d_part);
    launch(p, "k_red_final", 0, 0, k_red_final<true>, dim3(cw, 2), (const double*)p->d_part, S2);
    HIPCHK(hipMemcpyAsync(h, S2, sizeof(h), hipMemcpyDeviceToHost, p->stream));
    HIPCHK(hipStreamSynchronize(p->stream));
    for (int c = 0; c < nr; c++) {
      if (!((active >> c) & 1)) continue;
      const double den = p->norm_a * h[16 + c] + nb[c];
      berr[c] = den > 0 ? h[c] / den : (h[c] > 0 ? INFINITY : 0.0);
      if (berr[c] <= p->ir_tol) active &= ~(1 << c);
    }
    if (!active) break;
    rc = precond_sweep(p, w);  // w holds (float) r from k_cg_xr
    if (rc) return rc;
    launch(p, "k_pc_out", 0, 3.5 * vb, k_pc_out, dim3(gr), n, lcw, (const float*)w, (const double*)R, QZ,
           (double*)nullptr, p->d_part);
    launch(p, "k_red_final", 0, 0, k_red_final<false>, dim3(cw, 1), (const double*)p->d_part, rzn);
    launch(p, "k_cg_p", 0, 3.0 * vb, k_cg_p, dim3(ge), n, lcw, (const double*)rzn, (const double*)rz, active,
           (const double*)QZ, Pd);
    std::swap(rz, rzn);
  }
  launch(p, "k_cg_out", 0, 2.0 * 8.0 * n * nr, k_cg_out, dim3(ge), n, lcw, nr, (const double*)X, out, ld);
  it_out = it;
  berr_out = 0;
  for (int c = 0; c < nr; c++) berr_out = std::max(berr_out, berr[c]);
  return active ? UMF_ENOCONV : UMF_OK;
}

// Shared body of the solve/apply entry points: columns in chunks of RB through the work block.
int solve_apply(umf_plan* p, int mode, const double* in, double* out, int nrhs, int ld, bool device) {
  if (!p) return fail(UMF_EINVAL, "null plan");
  if (!p->factored) return fail(UMF_ESTATE, "factor before solve/apply");
  if (nrhs < 0 || ld < p->P.n) return fail(UMF_EINVAL, "bad nrhs or ld");
  if (nrhs == 0) return UMF_OK;
  int rc = ensure_solve_scratch(p);
  if (rc) return rc;
  int n = p->P.n;
  const bool f32 = p->fact_prec != UMF_PREC_FP64;
  HIPCHK(hipEventRecord(p->ev0, p->stream));
  // host operands are staged through a device column block first
  const double* d_in = in;
  double* d_out = out;
  if (!device) {
    if (!p->d_stage || p->stage_cols < nrhs) {
      if (p->d_stage) HIPCHK(hipFree(p->d_stage));
      p->d_stage = nullptr;
      HIPCHK(hipMalloc((void**)&p->d_stage, sizeof(double) * (size_t)n * nrhs));
      p->stage_cols = nrhs;
    }
    HIPCHK(hipMemcpy2DAsync(p->d_stage, sizeof(double) * n, in, sizeof(double) * ld, sizeof(double) * n, nrhs,
                            hipMemcpyHostToDevice, p->stream));
    d_in = d_out = p->d_stage;
  }
  const int ldd = device ? ld : n;
  int iters = 0, conv = UMF_OK;
  double berr = 0;
  for (int c0 = 0; c0 < nrhs; c0 += RB) {
    int nr = std::min(RB, nrhs - c0);
    const unsigned gb = (unsigned)((n + 63) / 64), go = gb;  // 64 rows per workgroup
    const double* ci = d_in + (size_t)c0 * ldd;
    double* co = d_out + (size_t)c0 * ldd;
    if (!f32) {
      launch(p, "k_to_rows", 0, 16.0 * n * nr, k_to_rows<double>, dim3(gb), n, ci, ldd, nr, p->d_w);
      rc = sweep<double>(p, mode, nr, p->arena, p->d_w, p->d_xs);
      if (rc) return rc;
      launch(p, "k_from_rows", 0, 16.0 * n * nr, k_from_rows<double>, dim3(go), n, (const double*)p->d_w, co, ldd,
             nr);
    } else if (mode == 1) {
      launch(p, "k_to_rows", 0, 12.0 * n * nr, k_to_rows<float>, dim3(gb), n, ci, ldd, nr, p->d_w32);
      rc = p->fact_prec == UMF_PREC_FP32T64 ? sweep_mixed(p, mode, nr, p->d_w32)
                                            : sweep<float>(p, mode, nr, p->arena32, p->d_w32, p->d_xs32);
      if (rc) return rc;
      launch(p, "k_from_rows", 0, 12.0 * n * nr, k_from_rows<float>, dim3(go), n, (const float*)p->d_w32, co, ldd,
             nr);
    } else {
      int it = 0;
      double be = 0;
      rc = pcg(p, ci, co, ldd, nr, it, be);
      if (rc && rc != UMF_ENOCONV) return rc;
      if (rc == UMF_ENOCONV) conv = rc;
      iters = std::max(iters, it);
      berr = std::max(berr, be);
    }
  }
  if (!device)
    HIPCHK(hipMemcpy2DAsync(out, sizeof(double) * ld, p->d_stage, sizeof(double) * n, sizeof(double) * n, nrhs,
                            hipMemcpyDeviceToHost, p->stream));
  HIPCHK(hipEventRecord(p->ev1, p->stream));
  HIPCHK(hipStreamSynchronize(p->stream));
  prof_collect(p);
  float ms = 0;
  HIPCHK(hipEventElapsedTime(&ms, p->ev0, p->ev1));
  int serr = 0;
  HIPCHK(hipMemcpyAsync(&serr, p->d_err + 3, sizeof(int), hipMemcpyDeviceToHost, p->stream));
  HIPCHK(hipStreamSynchronize(p->stream));
  if (serr) {
    HIPCHK(hipMemsetAsync(p->d_err + 3, 0, sizeof(int), p->stream));
    HIPCHK(hipStreamSynchronize(p->stream));
    return fail(UMF_EDEVICE, "blocked solve sweep: a block wait exceeded its spin limit (results invalid)");
  }
  if (mode == 0) {
    p->stats.t_solve_ms = ms;
    double fl = 0;
    for (const FrontInfo& F : p->P.fronts) fl += 2.0 * F.n * F.n + 4.0 * F.n * F.m;
    p->stats.flops_solve = fl * nrhs;
    p->stats.ir_iters = f32 ? iters : 0;
    p->stats.ir_berr = f32 ? berr : 0.0;
    if (conv) {
      char buf[160];
      snprintf(buf, sizeof buf, "FP32 refinement: backward error %.3e after %d iterations (tolerance %.3e)", berr,
               iters, p->ir_tol);
      return fail(UMF_ENOCONV, buf);
    }
  }
  return UMF_OK;
}

// ---- the solve that allocates nothing (umf_solve_device_noalloc): its work block, per-front
// right-hand-side scratch and id list live in one file-scope device array.  Each GPU holds its
// own copy; host code reaches it only through the address hipGetSymbolAddress returns on that
// device (the array's name in host code is a host-side stand-in, not a device address).
#define UMF_SCRATCH_BYTES (size_t(64) << 20)
static __device__ __align__(256) unsigned char scratch[UMF_SCRATCH_BYTES];
std::mutex g_scratch_mu;
void* g_scratch_ptr[64] = {};  // per device index

int scratch_ptr(int dev, void** out) {
  if (dev < 0 || dev >= 64) return fail(UMF_EDEVICE, "device index out of range for the no-alloc scratch");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    int cur = 0;
    HIPCHK(hipGetDevice(&cur));
    if (cur != dev) HIPCHK(hipSetDevice(dev));
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (cur != dev) HIPCHK(hipSetDevice(cur));
    if (e != hipSuccess || !ptr)
      return fail(UMF_EDEVICE, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch_ptr[dev] = ptr;
  }
  *out = g_scratch_ptr[dev];
  return UMF_OK;
}

// Restores the plan's per-kernel timing on scope exit (the no-alloc solve runs untimed: timing
// grows an event pool).
struct ProfOff {
  Prof& pr;
  bool was;
  explicit ProfOff(Prof& p) : pr(p), was(p.on) { pr.on = false; }
  ~ProfOff() { pr.on = was; }
};

int solve_noalloc(umf_plan* p, const double* in, double* out, int nrhs, int ld) {
  if (p->host_only) return fail(UMF_ESTATE, "plan was analyzed only (no device memory)");
  if (!p->factored) return fail(UMF_ESTATE, "factor before solve/apply");
  if (p->fact_prec != UMF_PREC_FP64)
    return fail(UMF_ESTATE, "the no-alloc solve takes FP64 factors (the refined forms need work vectors of their own)");
  if (p->lu) return fail(UMF_ESTATE, "the no-alloc solve takes FP64 Cholesky factors (not LU fronts)");
  const int n = p->P.n;
  if (nrhs < 0 || ld < n) return fail(UMF_EINVAL, "bad nrhs or ld");
  if (nrhs == 0) return UMF_OK;
  auto up = [](size_t b) { return (b + 255) / 256 * 256; };
  const size_t n_ids = (size_t)p->P.nfronts() + (size_t)p->P.nseps_total();
  const size_t off_x = up(sizeof(double) * (size_t)n * RB);
  const size_t off_i = up(off_x + sizeof(double) * (size_t)std::max<i64>(p->xs_elems, 1));
  const size_t need = off_i + sizeof(int) * n_ids;
  if (need > UMF_SCRATCH_BYTES) {
    char buf[200];
    snprintf(buf, sizeof buf, "the no-alloc solve needs %.1f MiB of scratch for this plan (n=%d), above its %zu MiB",
             need / 1048576.0, n, UMF_SCRATCH_BYTES >> 20);
    return fail(UMF_EINVAL, buf);
  }
  void* base = nullptr;
  int rc = scratch_ptr(p->device, &base);
  if (rc) return rc;
  double* w = (double*)base;
  double* xs = (double*)((char*)base + off_x);
  int* ids = (int*)((char*)base + off_i);
  ProfOff off(p->prof);
  HIPCHK(hipEventRecord(p->ev0, p->stream));
  const unsigned gb = (unsigned)((n + 63) / 64);
  for (int c0 = 0; c0 < nrhs; c0 += RB) {
    const int nr = std::min(RB, nrhs - c0);
    const double* ci = in + (size_t)c0 * ld;
    double* co = out + (size_t)c0 * ld;
    launch(p, "k_to_rows", 0, 16.0 * n * nr, k_to_rows<double>, dim3(gb), n, ci, ld, nr, w);
    rc = sweep<double>(p, 0, nr, p->arena, w, xs, ids);
    if (rc) return rc;
    launch(p, "k_from_rows", 0, 16.0 * n * nr, k_from_rows<double>, dim3(gb), n, (const double*)w, co, ld, nr);
  }
  HIPCHK(hipEventRecord(p->ev1, p->stream));
  HIPCHK(hipStreamSynchronize(p->stream));
  float ms = 0;
  HIPCHK(hipEventElapsedTime(&ms, p->ev0, p->ev1));
  int serr = 0;
  HIPCHK(hipMemcpyAsync(&serr, p->d_err + 3, sizeof(int), hipMemcpyDeviceToHost, p->stream));
  HIPCHK(hipStreamSynchronize(p->stream));
  if (serr) {
    HIPCHK(hipMemsetAsync(p->d_err + 3, 0, sizeof(int), p->stream));
    HIPCHK(hipStreamSynchronize(p->stream));
    return fail(UMF_EDEVICE, "blocked solve sweep: a block wait exceeded its spin limit (results invalid)");
  }
  p->stats.t_solve_ms = ms;
  double fl = 0;
  for (const FrontInfo& F : p->P.fronts) fl += 2.0 * F.n * F.n + 4.0 * F.n * F.m;
  p->stats.flops_solve = fl * nrhs;
  p->stats.ir_iters = 0;
  p->stats.ir_berr = 0.0;
  return UMF_OK;
}

// LU buffers of a plan whose programs write full matrices: per-front offsets of the inverted U
// blocks and of the transposed U rows (in fronts_h), then the four allocations.
int alloc_lu(umf_plan* p) {
  const Plan& P = p->P;
  i64 uo = 0, to = 0;
  for (size_t f = 0; f < p->fronts_h.size(); f++) {
    FrontDev& D = p->fronts_h[f];
    D.uinv_off = uo;
    uo += (i64)((D.n + NB - 1) / NB) * NB * NB;
    D.ut_off = to;
    to += ((i64)D.N * NB + 31) / 32 * 32;
  }
  const size_t nd = std::max<size_t>(P.dofs.size(), 1);
  HIPCHK(hipMalloc((void**)&p->d_ipiv, sizeof(int) * nd));
  HIPCHK(hipMalloc((void**)&p->d_perm, sizeof(int) * nd));
  HIPCHK(hipMalloc((void**)&p->d_uinv, sizeof(double) * std::max<i64>(uo, 1)));
  HIPCHK(hipMalloc((void**)&p->d_ut, sizeof(double) * std::max<i64>(to, 1)));
  p->lu = true;
  return UMF_OK;
}

void free_lu(umf_plan* p) {
  for (void* q : {(void*)p->d_ipiv, (void*)p->d_perm, (void*)p->d_uinv, (void*)p->d_ut})
    if (q) (void)hipFree(q);
  p->d_ipiv = p->d_perm = nullptr;
  p->d_uinv = p->d_ut = nullptr;
  p->lu = false;
}

void free_plan(umf_plan* p) {
  void* ptrs[] = {p->arena, p->d_dofs, p->d_mats, p->d_fronts, p->d_seps, p->d_tri, p->d_tri_off, p->d_ops,
                  p->d_tiles, p->d_tile_off, p->d_rp, p->d_ci, p->d_val, p->d_err, p->d_flags, p->d_ids, p->d_xs, p->d_w, p->d_stage,
                  p->arena32, p->d_xs32, p->d_w32, p->d_cg, p->d_part, p->d_dots, p->d_bflag, p->d_blist,
                  p->d_fly_tab, p->d_fly_tri, p->d_rest_tri, p->d_rest_off, p->d_arrive, p->d_ipiv, p->d_perm,
                  p->d_uinv, p->d_ut, p->d_ll};
  for (void* q : ptrs)
    if (q) (void)hipFree(q);
  if (p->fgraph) (void)hipGraphExecDestroy(p->fgraph);
  if (p->d_fids) (void)hipFree(p->d_fids);
  if (p->d_sids) (void)hipFree(p->d_sids);
  for (hipEvent_t e : p->prof.pool) (void)hipEventDestroy(e);
  if (p->ev0) (void)hipEventDestroy(p->ev0);
  if (p->ev1) (void)hipEventDestroy(p->ev1);
  if (p->stream) (void)hipStreamDestroy(p->stream);
}

// Symmetric pattern and no duplicate entries (every front layout needs both); symmetric values
// (Cholesky fronts) -- or, with `unsym` non-null, asymmetric values are allowed and reported
// there (LU fronts).
std::string check_operator(int n, const int* rp, const int* ci, const double* v, bool* unsym = nullptr) {
  if (unsym) *unsym = false;
  if (rp[0] != 0) return "row_ptr[0] must be 0";
  for (int i = 0; i < n; i++)
    if (rp[i + 1] < rp[i]) return "row_ptr must be non-decreasing";
  int nnz = rp[n];
  std::vector<int> cnt(n + 1, 0);
  for (int k = 0; k < nnz; k++) {
    if (ci[k] < 0 || ci[k] >= n) return "column index out of range";
    cnt[ci[k] + 1]++;
  }
  for (int i = 0; i < n; i++) cnt[i + 1] += cnt[i];
  std::vector<int> trow(nnz);
  std::vector<double> tval(nnz);
  std::vector<int> fill(cnt.begin(), cnt.end() - 1);
  for (int i = 0; i < n; i++)
    for (int k = rp[i]; k < rp[i + 1]; k++) {
      int j = ci[k];
      trow[fill[j]] = i;
      tval[fill[j]] = v ? v[k] : 0.0;
      fill[j]++;
    }
  // row i of A (sorted copy) must equal column i of A = row i of A^T (already sorted by row)
  std::vector<std::pair<int, double>> row;
  for (int i = 0; i < n; i++) {
    row.clear();
    for (int k = rp[i]; k < rp[i + 1]; k++) row.push_back({ci[k], v ? v[k] : 0.0});
    std::sort(row.begin(), row.end(), [](const std::pair<int, double>& a, const std::pair<int, double>& b) {
      return a.first < b.first;
    });
    for (size_t q = 1; q < row.size(); q++)
      if (row[q].first == row[q - 1].first) return "duplicate entry in row " + std::to_string(i);
    if ((int)row.size() != cnt[i + 1] - cnt[i]) return "pattern is not symmetric (row " + std::to_string(i) + ")";
    for (size_t q = 0; q < row.size(); q++) {
      int k = cnt[i] + (int)q;
      if (trow[k] != row[q].first) return "pattern is not symmetric (row " + std::to_string(i) + ")";
      if (v && tval[k] != row[q].second) {
        if (!unsym) return "values are not symmetric (row " + std::to_string(i) + ")";
        *unsym = true;
      }
    }
  }
  return "";
}

}  // namespace

// ============================================================================== C ABI
namespace {
int setup_device(umf_plan* p, umf_plan** out);
}

extern "C" {

const char* umf_last_error(void) { return g_err.c_str(); }

int umf_plan_create(int dim, int nx, int ny, int nz, int width, int max_side_length, int n, const int* row_ptr,
                    const int* col_ind, const double* val, umf_plan** out) {
  if (!out || !row_ptr || !col_ind || !val) return fail(UMF_EINVAL, "null argument");
  *out = nullptr;
  if (n <= 0) return fail(UMF_EINVAL, "n must be positive");
  bool unsym = false;  // asymmetric values: LU fronts, as the reference factors every front
  std::string e = check_operator(n, row_ptr, col_ind, val, &unsym);
  if (!e.empty()) return fail(UMF_EINVAL, e);
  umf_plan* p = new umf_plan();
  bool geom = false;
  e = build_plan(p->P, dim, nx, ny, nz, width, max_side_length, n, row_ptr, col_ind, geom, unsym);
  if (!e.empty()) {
    delete p;
    return fail(geom ? UMF_EGEOM : UMF_EINVAL, e);
  }
  p->val_h.assign(val, val + row_ptr[n]);
  const int a[6] = {dim, nx, ny, nz, width, max_side_length};
  std::copy(a, a + 6, p->args);
  return setup_device(p, out);
}

}  // extern "C"

namespace {
// Device side of a planned (single or batched) plan whose values are in val_h: descriptors,
// front pool, assembly programs, CSR operator, stream.  Frees p on failure.
int setup_device(umf_plan* p, umf_plan** out) {
  Plan& P = p->P;
  // device descriptors
  std::vector<MatDev>& mats = p->mats_h;
  mats.clear();
  for (const MatDesc& M : P.mats) mats.push_back(MatDev{M.off, M.rows, M.cols, 0});
  for (int ell = 0; ell < P.L; ell++) {
    for (int f : P.box_gid[ell]) mats[f].lev = ell;
    for (int sp : P.sep_gid[ell]) mats[P.nfronts() + sp].lev = ell;
  }
  i64 xo = 0;
  for (const FrontInfo& F : P.fronts) {
    FrontDev D{};
    D.off = P.mats[&F - &P.fronts[0]].off;
    D.int_off = F.int_off;
    D.bd_off = F.bd_off;
    D.n = F.n;
    D.m = F.m;
    D.N = F.n + F.m;
    D.nsep = F.nsep;
    for (int i = 0; i < 6; i++) {
      D.sep_pos[i] = i < F.nsep ? F.sep_pos[i] : 0;
      D.sep_len[i] = i < F.nsep ? F.sep_len[i] : 0;
    }
    D.x_off = xo;
    D.dinv_off = F.dinv_off;
    D.flag_off = p->n_flags;
    p->n_flags += (F.n + NB - 1) / NB;
    D.fly_off = P.fly_off.empty() ? -1 : P.fly_off[&F - &P.fronts[0]];
    xo += (i64)D.N * RB;
    p->fronts_h.push_back(D);
  }
  p->xs_elems = xo;
  // published blocks of the largest blocked sweep launch: a blocked level's fronts times its
  // largest block count
  p->ll_blocks = 0;
  for (size_t ell = 0; ell < P.box_gid.size(); ell++) {
    i64 nbmax = 0;
    int rows = 0;
    for (int f : P.box_gid[ell]) {
      nbmax = std::max<i64>(nbmax, (P.fronts[f].n + NB - 1) / NB);
      rows = std::max(rows, P.fronts[f].n + P.fronts[f].m);
    }
    const i64 nf = (i64)P.box_gid[ell].size();
    if (sweep_blocked(nf, rows)) p->ll_blocks = std::max<i64>(p->ll_blocks, nf * nbmax);
  }
  for (const SepInfo& S : P.seps) {
    SepDev D{};
    D.off = P.mats[P.nfronts() + (&S - &P.seps[0])].off;
    D.dof_off = S.dof_off;
    D.k = S.k;
    D.box[0] = S.box[0]; D.box[1] = S.box[1];
    D.pos[0] = S.pos[0]; D.pos[1] = S.pos[1];
    p->seps_h.push_back(D);
  }
  int mt = 1;
  for (size_t m = 0; m + 1 < P.tile_off.size(); m++) mt = std::max<i64>(mt, P.tile_off[m + 1] - P.tile_off[m]);
  p->max_tiles_per_mat = mt;

  auto bail = [&](int code, const std::string& msg) {
    free_plan(p);
    delete p;
    return fail(code, msg);
  };
  hipError_t he;
#define UP(d, h)                                                         \
  if ((he = upload(&(d), h)) != hipSuccess)                              \
    return bail(he == hipErrorOutOfMemory ? UMF_ENOMEM : UMF_EDEVICE, std::string("upload: ") + hipGetErrorString(he));
  if ((he = hipGetDevice(&p->device)) != hipSuccess) return bail(UMF_EDEVICE, hipGetErrorString(he));
  if ((he = hipStreamCreate(&p->stream)) != hipSuccess) return bail(UMF_EDEVICE, hipGetErrorString(he));
  if ((he = hipEventCreate(&p->ev0)) != hipSuccess) return bail(UMF_EDEVICE, hipGetErrorString(he));
  if ((he = hipEventCreate(&p->ev1)) != hipSuccess) return bail(UMF_EDEVICE, hipGetErrorString(he));
  if ((he = hipMalloc((void**)&p->arena, sizeof(double) * std::max<i64>(P.arena_elems, 1))) != hipSuccess)
    return bail(UMF_ENOMEM, std::string("front pool of ") + std::to_string(P.arena_elems * 8.0 / 1e9) +
                                " GB: " + hipGetErrorString(he));
  UP(p->d_dofs, P.dofs);
  UP(p->d_mats, mats);
  if (P.full) {
    const int rc = alloc_lu(p);
    if (rc) {
      const std::string msg = g_err;
      free_plan(p);
      delete p;
      return fail(rc, msg);
    }
  }
  UP(p->d_fronts, p->fronts_h);
  UP(p->d_seps, p->seps_h);
  UP(p->d_tri, P.triples);
  UP(p->d_tri_off, P.tri_off);
  UP(p->d_ops, P.ops);
  UP(p->d_tiles, P.tiles);
  UP(p->d_tile_off, P.tile_off);
  if (!P.rest_off.empty()) {
    UP(p->d_fly_tab, P.fly_tab);
    UP(p->d_fly_tri, P.fly_tri);
    UP(p->d_rest_tri, P.rest_tri);
    UP(p->d_rest_off, P.rest_off);
  }
  UP(p->d_rp, P.rp);
  UP(p->d_ci, P.ci);
  UP(p->d_val, p->val_h);
#undef UP
  if ((he = hipMalloc((void**)&p->d_err, sizeof(int) * 8)) != hipSuccess) return bail(UMF_EDEVICE, hipGetErrorString(he));
  if ((he = hipMemset(p->d_err, 0, sizeof(int) * 8)) != hipSuccess) return bail(UMF_EDEVICE, hipGetErrorString(he));
  if ((he = hipMalloc((void**)&p->d_flags, sizeof(int) * std::max(p->n_flags, 1))) != hipSuccess)
    return bail(UMF_EDEVICE, hipGetErrorString(he));
  if ((he = hipMemset(p->d_flags, 0, sizeof(int) * std::max(p->n_flags, 1))) != hipSuccess)
    return bail(UMF_EDEVICE, hipGetErrorString(he));
  if ((he = hipMalloc((void**)&p->d_ll, sizeof(unsigned long long) * kLLBlock * std::max<i64>(p->ll_blocks, 1))) !=
      hipSuccess)
    return bail(UMF_EDEVICE, hipGetErrorString(he));
  if ((he = hipMemset(p->d_ll, 0, sizeof(unsigned long long) * kLLBlock * std::max<i64>(p->ll_blocks, 1))) !=
      hipSuccess)
    return bail(UMF_EDEVICE, hipGetErrorString(he));
  if ((he = hipMalloc((void**)&p->d_arrive, sizeof(int) * std::max(P.nfronts(), 1))) != hipSuccess)
    return bail(UMF_EDEVICE, hipGetErrorString(he));
  if ((he = hipMemset(p->d_arrive, 0, sizeof(int) * std::max(P.nfronts(), 1))) != hipSuccess)
    return bail(UMF_EDEVICE, hipGetErrorString(he));
  double fl = 0;
  for (int f = 0; f < P.nfronts(); f++) fl += P.front_flops(f);
  p->stats.flops_factor = fl;
  p->stats.bytes_front_pool = 8.0 * P.arena_elems;
  p->stats.n_fronts = P.nfronts();
  *out = p;
  return UMF_OK;
}
}  // namespace

extern "C" {

int umf_batch_create(int nprob, const int* desc, const int* const* row_ptr, const int* const* col_ind,
                     const double* const* val, int flags, umf_plan** out) {
  if (!out || !desc || !row_ptr || !col_ind || (!val && !(flags & 1))) return fail(UMF_EINVAL, "null argument");
  *out = nullptr;
  if (nprob < 1) return fail(UMF_EINVAL, "a batch needs at least one problem");
  std::vector<Plan> subs(nprob);
  i64 nd = 0, nnz = 0;
  for (int q = 0; q < nprob; q++) {
    const int* d = desc + 7 * q;  // dim, nx, ny, nz, width, maxSideLength, n
    const int n = d[6];
    if (!row_ptr[q] || !col_ind[q] || (!(flags & 1) && !val[q])) return fail(UMF_EINVAL, "null argument");
    if (n <= 0) return fail(UMF_EINVAL, "problem " + std::to_string(q) + ": n must be positive");
    std::string e = check_operator(n, row_ptr[q], col_ind[q], (flags & 1) ? nullptr : val[q]);
    if (!e.empty()) return fail(UMF_EINVAL, "problem " + std::to_string(q) + ": " + e);
    bool geom = false;
    e = build_plan(subs[q], d[0], d[1], d[2], d[3], d[4], d[5], n, row_ptr[q], col_ind[q], geom);
    if (!e.empty()) return fail(geom ? UMF_EGEOM : UMF_EINVAL, "problem " + std::to_string(q) + ": " + e);
    nd += n;
    nnz += row_ptr[q][n];
  }
  if (nd > 0x7fffffff || nnz > 0x7fffffff) return fail(UMF_EINVAL, "batch exceeds 2^31 unknowns or nonzeros");
  umf_plan* p = new umf_plan();
  std::vector<std::pair<const double*, int>> vals;
  for (int q = 0; q < nprob; q++) vals.push_back({(flags & 1) ? nullptr : val[q], row_ptr[q][desc[7 * q + 6]]});
  merge_plans(p->P, std::move(subs));
  p->last = p->P.all_dirty();
  if (flags & 1) {
    p->host_only = true;
    double fl = 0;
    for (int f = 0; f < p->P.nfronts(); f++) fl += p->P.front_flops(f);
    p->stats.flops_factor = fl;
    p->stats.bytes_front_pool = 8.0 * p->P.arena_elems;
    p->stats.n_fronts = p->P.nfronts();
    *out = p;
    return UMF_OK;
  }
  p->val_h.reserve(nnz);
  for (auto& v : vals) p->val_h.insert(p->val_h.end(), v.first, v.first + v.second);
  return setup_device(p, out);
}

int umf_batch_info(const umf_plan* p, int prob, int* dof0, int* n, int* levels, int* shift) {
  if (!p) return fail(UMF_EINVAL, "null plan");
  const Forest* Fo = p->P.forest.get();
  const int np = Fo ? (int)Fo->sub.size() : 1;
  if (prob < 0 || prob >= np) return fail(UMF_EINVAL, "bad problem index");
  if (dof0) *dof0 = Fo ? Fo->dof0[prob] : 0;
  if (n) *n = Fo ? Fo->sub[prob].n : p->P.n;
  if (levels) *levels = Fo ? Fo->sub[prob].L : p->P.L;
  if (shift) *shift = Fo ? Fo->shift[prob] : 0;
  return np;
}

int umf_plan_analyze(int dim, int nx, int ny, int nz, int width, int max_side_length, int n, const int* row_ptr,
                     const int* col_ind, umf_plan** out) {
  if (!out || !row_ptr || !col_ind) return fail(UMF_EINVAL, "null argument");
  *out = nullptr;
  if (n <= 0) return fail(UMF_EINVAL, "n must be positive");
  std::string e = check_operator(n, row_ptr, col_ind, nullptr);
  if (!e.empty()) return fail(UMF_EINVAL, e);
  umf_plan* p = new umf_plan();
  bool geom = false;
  e = build_plan(p->P, dim, nx, ny, nz, width, max_side_length, n, row_ptr, col_ind, geom);
  if (!e.empty()) {
    delete p;
    return fail(geom ? UMF_EGEOM : UMF_EINVAL, e);
  }
  p->host_only = true;
  p->last = p->P.all_dirty();
  double fl = 0;
  for (int f = 0; f < p->P.nfronts(); f++) fl += p->P.front_flops(f);
  p->stats.flops_factor = fl;
  p->stats.bytes_front_pool = 8.0 * p->P.arena_elems;
  p->stats.n_fronts = p->P.nfronts();
  *out = p;
  return UMF_OK;
}

int umf_analyze_update(umf_plan* p, int n_mod, const int* mod_dofs) {
  if (!p || (n_mod > 0 && !mod_dofs)) return fail(UMF_EINVAL, "null argument");
  std::vector<int> mod(mod_dofs, mod_dofs + std::max(n_mod, 0));
  Plan::Dirty D;
  std::string e;
  if (!p->P.dirty_from_dofs(mod, D, e)) return fail(UMF_EUPDATEMAP, e);
  p->last = D;
  double fl = 0;
  int nf = 0;
  for (int ell = 0; ell < p->P.L; ell++)
    for (int b : D.boxes[ell]) {
      fl += p->P.front_flops(p->P.box_gid[ell][b]);
      nf++;
    }
  p->stats.flops_update = fl;
  p->stats.n_fronts_update = nf;
  return UMF_OK;
}

int umf_plan_destroy(umf_plan* p) {
  if (!p) return UMF_OK;
  if (p->servers > 0) return fail(UMF_ESTATE, "a server still answers requests against this plan");
  free_plan(p);
  delete p;
  return UMF_OK;
}

// ---- plan files: header, CSR pattern, current values, then (factored plans) the front pool
// of the precision the factors are held in, entry by entry as the device holds it.
namespace {
const char kPlanMagic[8] = {'U', 'M', 'F', 'P', 'L', 'A', 'N', '1'};
struct PlanFileHead {
  char magic[8];
  int32_t args[6];  // dim, nx, ny, nz, width, max_side_length
  int32_t n, factored, prec, fact_prec, ir_max, elem_size;  // elem_size: 8, 4, or 0 (no pool)
  int64_t nnz, arena_elems;
  double ir_tol;
};
const size_t kChunk = size_t(1) << 26;  // bytes staged per host<->device copy
}  // namespace

int umf_plan_save(umf_plan* p, const char* path) {
  if (!p || !path) return fail(UMF_EINVAL, "null argument");
  if (p->host_only) return fail(UMF_ESTATE, "plan was analyzed only (it holds no values)");
  if (p->P.forest) return fail(UMF_EINVAL, "batch plans cannot be saved");
  const Plan& P = p->P;
  PlanFileHead h{};
  std::copy(kPlanMagic, kPlanMagic + 8, h.magic);
  std::copy(p->args, p->args + 6, h.args);
  h.n = P.n;
  h.nnz = (int64_t)P.ci.size();
  h.factored = p->factored ? 1 : 0;
  h.prec = p->prec;
  h.fact_prec = p->fact_prec;
  h.ir_max = p->ir_max;
  h.ir_tol = p->ir_tol;
  h.arena_elems = P.arena_elems;
  h.elem_size = !p->factored ? 0 : p->fact_prec != UMF_PREC_FP64 ? 4 : 8;
  // LU factors (pivots, inverted U blocks) and the mixed form's two pools are not stored: loaded
  // plans refactor
  if (p->lu || p->fact_prec == UMF_PREC_FP32T64) {
    h.factored = 0;
    h.elem_size = 0;
  }
  HIPCHK(hipStreamSynchronize(p->stream));
  std::vector<double> val((size_t)h.nnz);
  if (h.nnz) HIPCHK(hipMemcpy(val.data(), p->d_val, sizeof(double) * h.nnz, hipMemcpyDeviceToHost));
  FILE* f = std::fopen(path, "wb");
  if (!f) return fail(UMF_EINVAL, std::string("cannot open ") + path + " for writing");
  bool ok = std::fwrite(&h, sizeof h, 1, f) == 1 && std::fwrite(P.rp.data(), sizeof(int), P.rp.size(), f) == P.rp.size() &&
            std::fwrite(P.ci.data(), sizeof(int), P.ci.size(), f) == P.ci.size() &&
            std::fwrite(val.data(), sizeof(double), val.size(), f) == val.size();
  if (ok && h.elem_size) {
    const char* src = h.elem_size == 4 ? (const char*)p->arena32 : (const char*)p->arena;
    const size_t total = (size_t)h.arena_elems * h.elem_size;
    std::vector<char> buf(std::min(total, kChunk));
    for (size_t o = 0; ok && o < total; o += kChunk) {
      const size_t b = std::min(kChunk, total - o);
      if (hipMemcpy(buf.data(), src + o, b, hipMemcpyDeviceToHost) != hipSuccess) ok = false;
      ok = ok && std::fwrite(buf.data(), 1, b, f) == b;
    }
  }
  ok = (std::fclose(f) == 0) && ok;
  if (!ok) return fail(UMF_EDEVICE, std::string("writing ") + path + " failed");
  return UMF_OK;
}

int umf_plan_load(const char* path, umf_plan** out) {
  if (!path || !out) return fail(UMF_EINVAL, "null argument");
  *out = nullptr;
  FILE* f = std::fopen(path, "rb");
  if (!f) return fail(UMF_EINVAL, std::string("cannot open ") + path);
  PlanFileHead h{};
  auto bad = [&](const std::string& why) {
    std::fclose(f);
    return fail(UMF_EINVAL, std::string(path) + ": " + why);
  };
  if (std::fread(&h, sizeof h, 1, f) != 1 || !std::equal(kPlanMagic, kPlanMagic + 8, h.magic))
    return bad("not a plan file");
  auto prec_ok = [](int v) { return v >= UMF_PREC_FP64 && v <= UMF_PREC_FP32T64; };
  if (h.n <= 0 || h.nnz < 0 || (h.elem_size != 0 && h.elem_size != 4 && h.elem_size != 8) || h.arena_elems < 0 ||
      !prec_ok(h.prec) || !prec_ok(h.fact_prec) || !(h.ir_tol > 0) || !std::isfinite(h.ir_tol) || h.ir_max < 1 ||
      (h.elem_size != 0 && (h.elem_size == 8) != (h.fact_prec == UMF_PREC_FP64)))
    return bad("corrupt header");
  std::vector<int> rp((size_t)h.n + 1), ci((size_t)h.nnz);
  std::vector<double> val((size_t)h.nnz);
  if (std::fread(rp.data(), sizeof(int), rp.size(), f) != rp.size() ||
      std::fread(ci.data(), sizeof(int), ci.size(), f) != ci.size() ||
      std::fread(val.data(), sizeof(double), val.size(), f) != val.size())
    return bad("truncated");
  if (rp[h.n] != h.nnz) return bad("corrupt CSR");
  umf_plan* p = nullptr;
  int rc = umf_plan_create(h.args[0], h.args[1], h.args[2], h.args[3], h.args[4], h.args[5], h.n, rp.data(), ci.data(),
                           val.data(), &p);
  if (rc) {
    std::fclose(f);
    return rc;
  }
  auto drop = [&](int code, const std::string& why) {
    std::fclose(f);
    umf_plan_destroy(p);
    return fail(code, std::string(path) + ": " + why);
  };
  if (h.arena_elems != p->P.arena_elems) return drop(UMF_EINVAL, "front pool size differs from this build's plan");
  p->prec = h.prec;
  p->ir_tol = h.ir_tol;
  p->ir_max = h.ir_max;
  if (h.elem_size) {
    if (h.elem_size == 4 && !p->arena32 &&
        hipMalloc((void**)&p->arena32, sizeof(float) * std::max<i64>(p->P.arena_elems, 1)) != hipSuccess)
      return drop(UMF_ENOMEM, "FP32 front pool allocation failed");
    char* dst = h.elem_size == 4 ? (char*)p->arena32 : (char*)p->arena;
    const size_t total = (size_t)h.arena_elems * h.elem_size;
    std::vector<char> buf(std::min(total, kChunk));
    for (size_t o = 0; o < total; o += kChunk) {
      const size_t b = std::min(kChunk, total - o);
      if (std::fread(buf.data(), 1, b, f) != b) return drop(UMF_EINVAL, "truncated front pool");
      if (hipMemcpy(dst + o, buf.data(), b, hipMemcpyHostToDevice) != hipSuccess)
        return drop(UMF_EDEVICE, "front pool upload failed");
    }
    p->fact_prec = h.fact_prec;
    p->factored = h.factored != 0;
  }
  std::fclose(f);
  *out = p;
  return UMF_OK;
}

int umf_factor(umf_plan* p) {
  if (!p) return fail(UMF_EINVAL, "null plan");
  if (p->host_only) return fail(UMF_ESTATE, "plan was analyzed only (no device memory)");
  if (p->servers > 0) return fail(UMF_ESTATE, "a server answers requests against this plan's factors");
  int rc = ensure_full_sched(p);
  if (rc) return rc;
  const Plan::Dirty& D = *p->full_dirty;
  p->factored = false;
  p->fact_prec = p->prec;
  HIPCHK(hipEventRecord(p->ev0, p->stream));
  if (p->lu) rc = factor_any(p, D);
  else if (p->use_graph && !p->prof.on) rc = factor_graph(p, D);
  else rc = factor_any(p, D, p->d_fids, p->full_sched.get());
  if (rc) return rc;
  HIPCHK(hipEventRecord(p->ev1, p->stream));
  HIPCHK(hipStreamSynchronize(p->stream));
  prof_collect(p);
  float ms = 0;
  HIPCHK(hipEventElapsedTime(&ms, p->ev0, p->ev1));
  p->stats.t_factor_ms = ms;
  p->last = D;
  rc = check_err(p);
  p->factored = rc == UMF_OK;
  return rc;
}

}  // extern "C"

namespace {
// Shared tail of the update entry points: d_val already holds (or is about to receive, in
// stream order) the new values.
int do_update(umf_plan* p, const double* src_val, bool src_device, int n_mod, const int* mod_dofs) {
  if (p->servers > 0) return fail(UMF_ESTATE, "a server answers requests against this plan's factors");
  const Plan& P = p->P;
  Plan::Dirty D;
  std::string e;
  std::vector<int> mod(mod_dofs, mod_dofs + std::max(n_mod, 0));
  if (!P.dirty_from_dofs(mod, D, e)) return fail(UMF_EUPDATEMAP, e);
  HIPCHK(hipEventRecord(p->ev0, p->stream));
  HIPCHK(hipMemcpyAsync(p->d_val, src_val, sizeof(double) * P.ci.size(),
                        src_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, p->stream));
  double fl = 0;
  int nf = 0;
  if (D.any) {
    int rc = factor_any(p, D);
    if (rc) return rc;
    for (int ell = 0; ell < P.L; ell++)
      for (int b : D.boxes[ell]) {
        fl += P.front_flops(P.box_gid[ell][b]);
        nf++;
      }
  }
  HIPCHK(hipEventRecord(p->ev1, p->stream));
  HIPCHK(hipStreamSynchronize(p->stream));
  prof_collect(p);
  float ms = 0;
  HIPCHK(hipEventElapsedTime(&ms, p->ev0, p->ev1));
  p->stats.t_update_ms = ms;
  p->stats.flops_update = fl;
  p->stats.n_fronts_update = nf;
  p->last = D;
  int rc = check_err(p);
  p->factored = rc == UMF_OK;
  return rc;
}
}  // namespace

extern "C" {

int umf_update(umf_plan* p, const int* row_ptr, const int* col_ind, const double* val, int n_mod,
               const int* mod_dofs) {
  if (!p || !row_ptr || !col_ind || !val || (n_mod > 0 && !mod_dofs)) return fail(UMF_EINVAL, "null argument");
  if (p->host_only || !p->factored) return fail(UMF_ESTATE, "factor before update");
  const Plan& P = p->P;
  int n = P.n;
  if (row_ptr[n] != (int)P.ci.size()) return fail(UMF_EPATTERN, "nnz differs from the planned operator");
  for (int i = 0; i <= n; i++)
    if (row_ptr[i] != P.rp[i]) return fail(UMF_EPATTERN, "row_ptr differs from the planned operator");
  for (size_t k = 0; k < P.ci.size(); k++)
    if (col_ind[k] != P.ci[k]) return fail(UMF_EPATTERN, "col_ind differs from the planned operator");
  bool unsym = false;
  std::string e = check_operator(n, row_ptr, col_ind, val, &unsym);
  if (!e.empty()) return fail(UMF_EINVAL, e);
  if (unsym && !p->lu)
    return fail(UMF_EINVAL, "values are not symmetric: this plan factors with Cholesky (umf_set_lu switches it to LU "
                            "fronts)");
  p->val_h.assign(val, val + row_ptr[n]);
  return do_update(p, p->val_h.data(), false, n_mod, mod_dofs);
}

int umf_update_device(umf_plan* p, const double* d_val, int n_mod, const int* mod_dofs) {
  if (!p || !d_val || (n_mod > 0 && !mod_dofs)) return fail(UMF_EINVAL, "null argument");
  if (p->host_only || !p->factored) return fail(UMF_ESTATE, "factor before update");
  return do_update(p, d_val, true, n_mod, mod_dofs);
}

int umf_solve(umf_plan* p, const double* b, double* x, int nrhs, int ld) {
  if (!b || !x) return fail(UMF_EINVAL, "null argument");
  return solve_apply(p, 0, b, x, nrhs, ld, false);
}
int umf_apply(umf_plan* p, const double* x, double* y, int nrhs, int ld) {
  if (!x || !y) return fail(UMF_EINVAL, "null argument");
  return solve_apply(p, 1, x, y, nrhs, ld, false);
}
int umf_solve_device(umf_plan* p, const double* b, double* x, int nrhs, int ld) {
  if (!b || !x) return fail(UMF_EINVAL, "null argument");
  return solve_apply(p, 0, b, x, nrhs, ld, true);
}
int umf_solve_device_noalloc(umf_plan* p, const double* b, double* x, int nrhs, int ld) {
  if (!p || !b || !x) return fail(UMF_EINVAL, "null argument");
  return solve_noalloc(p, b, x, nrhs, ld);
}
int umf_apply_device(umf_plan* p, const double* x, double* y, int nrhs, int ld) {
  if (!x || !y) return fail(UMF_EINVAL, "null argument");
  return solve_apply(p, 1, x, y, nrhs, ld, true);
}
int umf_spmv_device(umf_plan* p, const double* x, double* y, int nrhs, int ld) {
  if (!p || !x || !y) return fail(UMF_EINVAL, "null argument");
  if (p->host_only) return fail(UMF_ESTATE, "plan was analyzed only (no device memory)");
  if (ld < p->P.n || nrhs < 0) return fail(UMF_EINVAL, "bad nrhs or ld");
  if (nrhs == 0) return UMF_OK;
  hipLaunchKernelGGL(k_spmv, dim3((p->P.n + 255) / 256), dim3(256), 0, p->stream, p->P.n, p->d_rp, p->d_ci,
                     p->d_val, x, ld, y, ld, nrhs);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(p->stream));
  return UMF_OK;
}

int umf_box_scale_device(umf_plan* p, double* d_coeff, double h, const int* lo, int side, double factor,
                         double* d_val, int* mod_dofs, int cap, int* n_mod) {
  if (!p || !d_coeff || !lo || !d_val || !n_mod || (cap > 0 && !mod_dofs)) return fail(UMF_EINVAL, "null argument");
  if (p->host_only) return fail(UMF_ESTATE, "plan was analyzed only (no device memory)");
  if (p->P.forest) return fail(UMF_EINVAL, "coefficient requests take a single-problem plan");
  const Grid& g = p->P.g;
  const int dim = g.dim, w = g.w;
  if (!(h > 0) || side < 1) return fail(UMF_EINVAL, "bad h or box side");
  for (int a = 0; a < dim; a++)
    if (lo[a] < 0 || lo[a] + side > g.N[a] + 2 * w) return fail(UMF_EINVAL, "box outside the padded field");
  LapGeom G{};
  G.dim = dim;
  G.w = w;
  G.hh = h * h;
  const double cf[4][4] = {{0, 0, 0, 0}, {0.0, -1.0, 0, 0}, {0.0, -4.0 / 3.0, 1.0 / 12.0, 0}, {0.0, -1.5, 3.0 / 20.0, -1.0 / 90.0}};
  if (w < 1 || w > 3) return fail(UMF_EINVAL, "stencil width must be 1..3");
  for (int k = 0; k < 4; k++) G.cf[k] = cf[w][k];
  for (int a = 0; a < 3; a++) {
    G.N[a] = a < dim ? g.N[a] : 1;
    G.P[a] = a < dim ? g.N[a] + 2 * w : 1;
  }
  // rows reached: interior nodes within w of the changed padded nodes (interior = padded - w),
  // columns changed: within w of those rows
  int r0[3] = {0, 0, 0}, re[3] = {1, 1, 1}, c0[3] = {0, 0, 0}, ce[3] = {1, 1, 1};
  for (int a = 0; a < dim; a++) {
    const int b = std::max(lo[a] - 2 * w, 0), e = std::min(lo[a] + side, g.N[a]);  // e > b: lo < N + 2w
    r0[a] = b;
    re[a] = e - b;
    c0[a] = std::max(b - w, 0);
    ce[a] = std::min(e + w, g.N[a]) - c0[a];
  }
  // the rows must be laid out as the generator lays them out: 1 + (valid neighbours) entries
  for (int z = r0[2]; z < r0[2] + re[2]; z++)
    for (int y = r0[1]; y < r0[1] + re[1]; y++)
      for (int x = r0[0]; x < r0[0] + re[0]; x++) {
        const int q[3] = {x, y, z};
        int cnt = 1;
        for (int a = 0; a < dim; a++) cnt += std::min(q[a], w) + std::min(g.N[a] - 1 - q[a], w);
        const int row = x + g.N[0] * (y + g.N[1] * z);
        if (p->P.rp[row + 1] - p->P.rp[row] != cnt)
          return fail(UMF_EPATTERN, "operator pattern is not the harness generator's (row " + std::to_string(row) + ")");
      }
  const int ncand = ce[0] * ce[1] * ce[2];
  if (!p->d_bflag) {
    HIPCHK(hipMalloc((void**)&p->d_bflag, sizeof(int) * std::max(p->P.n, 1)));
    HIPCHK(hipMemset(p->d_bflag, 0, sizeof(int) * std::max(p->P.n, 1)));
  }
  if (p->blist_cap < ncand) {
    if (p->d_blist) HIPCHK(hipFree(p->d_blist));
    p->d_blist = nullptr;
    HIPCHK(hipMalloc((void**)&p->d_blist, sizeof(int) * (ncand + 1)));
    p->blist_cap = ncand;
  }
  p->bepoch = p->bepoch == 0x7fffffff ? 1 : p->bepoch + 1;
  const i64 nbox = (i64)std::pow((double)side, dim) + 0.5, nrow = (i64)re[0] * re[1] * re[2];
  auto grid = [](i64 cnt) { return dim3((unsigned)std::min<i64>((cnt + 255) / 256, 1024)); };
  hipLaunchKernelGGL(k_box_scale, grid(nbox), dim3(256), 0, p->stream, G, d_coeff, lo[0], lo[1], dim == 3 ? lo[2] : 0,
                     side, factor);
  hipLaunchKernelGGL(k_lap_rows, grid(nrow), dim3(256), 0, p->stream, G, (const double*)d_coeff, (const int*)p->d_rp,
                     d_val, r0[0], r0[1], r0[2], re[0], re[1], re[2], p->d_bflag, p->bepoch);
  HIPCHK(hipMemsetAsync(p->d_blist, 0, sizeof(int), p->stream));
  hipLaunchKernelGGL(k_lap_collect, grid(ncand), dim3(256), 0, p->stream, G, c0[0], c0[1], c0[2], ce[0], ce[1], ce[2],
                     (const int*)p->d_bflag, p->bepoch, p->d_blist + 1, p->d_blist);
  HIPCHK(hipGetLastError());
  std::vector<int> out((size_t)ncand + 1);
  HIPCHK(hipMemcpyAsync(out.data(), p->d_blist, sizeof(int) * (ncand + 1), hipMemcpyDeviceToHost, p->stream));
  HIPCHK(hipStreamSynchronize(p->stream));
  const int cnt = out[0];
  std::sort(out.begin() + 1, out.begin() + 1 + cnt);
  *n_mod = cnt;
  if (cnt > cap) return fail(UMF_EINVAL, "mod_dofs capacity too small (" + std::to_string(cnt) + " needed)");
  std::copy(out.begin() + 1, out.begin() + 1 + cnt, mod_dofs);
  return UMF_OK;
}

int umf_levels(const umf_plan* p) { return p ? p->P.L : UMF_EINVAL; }
int umf_leaf_width(const umf_plan* p, int a) {
  if (p && p->P.forest) return fail(UMF_EINVAL, "tree queries take a single-problem plan");
  if (!p || a < 0 || a >= p->P.g.dim) return fail(UMF_EINVAL, "bad axis");
  return p->P.g.bw[a];
}
int umf_num_boxes(const umf_plan* p, int level) {
  if (p && p->P.forest) return fail(UMF_EINVAL, "tree queries take a single-problem plan");
  if (!p || level < 0 || level >= p->P.g.L) return fail(UMF_EINVAL, "bad level");
  return p->P.g.nboxes(level);
}
int umf_num_seps(const umf_plan* p, int level) {
  if (p && p->P.forest) return fail(UMF_EINVAL, "tree queries take a single-problem plan");
  if (!p || level < 1 || level >= p->P.g.L) return fail(UMF_EINVAL, "bad level");
  return p->P.g.nseps(level);
}
int umf_box_dofs(const umf_plan* p, int level, int box, int which, int* out) {
  if (p && p->P.forest) return fail(UMF_EINVAL, "tree queries take a single-problem plan");
  if (!p || level < 0 || level >= p->P.g.L || box < 0 || box >= p->P.g.nboxes(level))
    return fail(UMF_EINVAL, "bad box");
  const FrontInfo& F = p->P.fronts[p->P.box_gid[level][box]];
  i64 off = which == 0 ? F.int_off : F.bd_off;
  int cnt = which == 0 ? F.n : F.m;
  if (out) std::copy(p->P.dofs.begin() + off, p->P.dofs.begin() + off + cnt, out);
  return cnt;
}
int umf_sep_dofs(const umf_plan* p, int level, int sep, int* out) {
  if (p && p->P.forest) return fail(UMF_EINVAL, "tree queries take a single-problem plan");
  if (!p || level < 1 || level >= p->P.g.L || sep < 0 || sep >= p->P.g.nseps(level))
    return fail(UMF_EINVAL, "bad separator");
  const SepInfo& S = p->P.seps[p->P.sep_gid[level][sep]];
  if (out) std::copy(p->P.dofs.begin() + S.dof_off, p->P.dofs.begin() + S.dof_off + S.k, out);
  return S.k;
}
int umf_dirty(const umf_plan* p, int level, int which, int* out) {
  if (!p || level < 0 || level >= p->P.L || (int)p->last.boxes.size() != p->P.L)
    return fail(UMF_EINVAL, "bad level or nothing factored");
  const Plan& P = p->P;
  std::vector<int> v;
  if (which == 0) v = p->last.boxes[level];
  else if (which == 1) v = p->last.seps[level];
  else if (level < P.L - 1) v = p->last.seps[level];  // rebuilt from children == receive adds
  if (out) std::copy(v.begin(), v.end(), out);
  return (int)v.size();
}

int umf_get_front_block(umf_plan* p, int level, int box, int which, double* out, int* rows, int* cols) {
  if (p && p->P.forest) return fail(UMF_EINVAL, "tree queries take a single-problem plan");
  if (!p || !rows || !cols || level < 0 || level >= p->P.g.L || box < 0 || box >= p->P.g.nboxes(level))
    return fail(UMF_EINVAL, "bad block query");
  if (!p->factored) return fail(UMF_ESTATE, "factor before querying blocks");
  int f = p->P.box_gid[level][box];
  const FrontDev& F = p->fronts_h[f];
  int n = F.n, m = F.m, N = F.N;
  std::vector<double> h;
  int rc = read_pool(p, F.off, (size_t)N * N, h, level);
  if (rc) return rc;
  auto at = [&](int i, int j) { return h[(size_t)j * N + i]; };
  std::vector<double> r;
  int R = 0, Cc = 0;
  if (p->lu) {  // LU front: L\U of P A22, U12, L21, A11 - L21 U12 (full)
    const FrontInfo& FI = p->P.fronts[f];
    std::vector<int> piv(n), pi(n);
    if (n) HIPCHK(hipMemcpy(piv.data(), p->d_ipiv + FI.int_off, sizeof(int) * n, hipMemcpyDeviceToHost));
    for (int j = 0; j < n; j++) pi[j] = j;
    for (int j = 0; j < n; j++) std::swap(pi[j], pi[piv[j]]);  // (P b)[j] = b[pi[j]]
    auto Lf = [&](int i, int k) { return i == k ? 1.0 : (i > k ? at(i, k) : 0.0); };
    if (which == UMF_BLOCK_L) {
      R = Cc = n;
      r.assign((size_t)n * n, 0.0);
      for (int j = 0; j < n; j++)
        for (int i = 0; i < n; i++) r[(size_t)j * n + i] = at(i, j);
    } else if (which == UMF_BLOCK_A22) {  // A22 = P^T L U
      R = Cc = n;
      r.assign((size_t)n * n, 0.0);
      for (int c = 0; c < n; c++)
        for (int j = 0; j < n; j++) {
          double sum = 0;
          for (int k = 0; k <= std::min(j, c); k++) sum += Lf(j, k) * at(k, c);
          r[(size_t)c * n + pi[j]] = sum;
        }
    } else if (which == UMF_BLOCK_XR) {  // XR = U^{-1} U12
      R = n;
      Cc = m;
      r.assign((size_t)n * m, 0.0);
      for (int c = 0; c < m; c++) {
        double* x = &r[(size_t)c * n];
        for (int i = 0; i < n; i++) x[i] = at(i, n + c);
        for (int j = n - 1; j >= 0; j--) {
          x[j] /= at(j, j);
          for (int i = 0; i < j; i++) x[i] -= at(i, j) * x[j];
        }
      }
    } else if (which == UMF_BLOCK_XL) {  // XL = L21 L^{-1} P
      R = m;
      Cc = n;
      r.assign((size_t)m * n, 0.0);
      std::vector<double> y(n);
      for (int i = 0; i < m; i++) {
        for (int c = n - 1; c >= 0; c--) {  // y L = L21(i, :)
          double v = at(n + i, c);
          for (int k = c + 1; k < n; k++) v -= y[k] * Lf(k, c);
          y[c] = v;
        }
        for (int j = 0; j < n; j++) r[(size_t)pi[j] * m + i] = y[j];
      }
    } else if (which == UMF_BLOCK_S) {  // S = -L21 U12
      R = Cc = m;
      r.assign((size_t)m * m, 0.0);
      for (int j = 0; j < m; j++)
        for (int i = 0; i < m; i++) {
          double sum = 0;
          for (int k = 0; k < n; k++) sum += at(n + i, k) * at(k, n + j);
          r[(size_t)j * m + i] = -sum;
        }
    } else if (which == UMF_BLOCK_A11POST) {
      R = Cc = m;
      r.assign((size_t)m * m, 0.0);
      std::vector<int> run(m, -1);
      for (int q = 0; q < F.nsep; q++)
        for (int i = 0; i < F.sep_len[q]; i++) run[F.sep_pos[q] + i] = q;
      for (int j = 0; j < m; j++)
        for (int i = 0; i < m; i++) r[(size_t)j * m + i] = run[i] == run[j] ? NAN : at(n + i, n + j);
    } else {
      return fail(UMF_EINVAL, "unknown block selector");
    }
  } else if (which == UMF_BLOCK_L || which == UMF_BLOCK_A22) {
    R = Cc = n;
    std::vector<double> Lm((size_t)n * n, 0.0);
    for (int j = 0; j < n; j++)
      for (int i = j; i < n; i++) Lm[(size_t)j * n + i] = at(i, j);
    if (which == UMF_BLOCK_L) r = Lm;
    else {
      r.assign((size_t)n * n, 0.0);
      for (int j = 0; j < n; j++)
        for (int i = j; i < n; i++) {
          double s = 0;
          for (int k = 0; k <= j; k++) s += Lm[(size_t)k * n + i] * Lm[(size_t)k * n + j];
          r[(size_t)j * n + i] = s;
          r[(size_t)i * n + j] = s;
        }
    }
  } else if (which == UMF_BLOCK_XR || which == UMF_BLOCK_XL) {
    // XR = L11^{-T} W with W = L21^T (n x m)
    std::vector<double> X((size_t)n * m);
    for (int c = 0; c < m; c++) {
      double* x = &X[(size_t)c * n];
      for (int i = 0; i < n; i++) x[i] = at(n + c, i);
      for (int j = n - 1; j >= 0; j--) {
        double s = x[j];
        for (int i = j + 1; i < n; i++) s -= at(i, j) * x[i];
        x[j] = s / at(j, j);
      }
    }
    if (which == UMF_BLOCK_XR) { R = n; Cc = m; r = X; }
    else {
      R = m; Cc = n;
      r.assign((size_t)m * n, 0.0);
      for (int c = 0; c < m; c++)
        for (int i = 0; i < n; i++) r[(size_t)i * m + c] = X[(size_t)c * n + i];
    }
  } else if (which == UMF_BLOCK_S) {
    R = Cc = m;
    r.assign((size_t)m * m, 0.0);
    for (int j = 0; j < m; j++)
      for (int i = j; i < m; i++) {
        double s = 0;
        for (int k = 0; k < n; k++) s += at(n + i, k) * at(n + j, k);
        r[(size_t)j * m + i] = -s;
        r[(size_t)i * m + j] = -s;
      }
  } else if (which == UMF_BLOCK_A11POST) {
    R = Cc = m;
    r.assign((size_t)m * m, 0.0);
    std::vector<int> run(m, -1);
    for (int q = 0; q < F.nsep; q++)
      for (int i = 0; i < F.sep_len[q]; i++) run[F.sep_pos[q] + i] = q;
    for (int j = 0; j < m; j++)
      for (int i = j; i < m; i++) {
        double v = run[i] == run[j] ? NAN : at(n + i, n + j);
        r[(size_t)j * m + i] = v;
        r[(size_t)i * m + j] = v;
      }
  } else {
    return fail(UMF_EINVAL, "unknown block selector");
  }
  *rows = R;
  *cols = Cc;
  if (out && !r.empty()) std::memcpy(out, r.data(), sizeof(double) * r.size());
  return UMF_OK;
}

int umf_get_sep_block(umf_plan* p, int level, int sep, double* out, int* k) {
  if (p && p->P.forest) return fail(UMF_EINVAL, "tree queries take a single-problem plan");
  if (!p || !k || level < 1 || level >= p->P.g.L || sep < 0 || sep >= p->P.g.nseps(level))
    return fail(UMF_EINVAL, "bad separator");
  const SepDev& S = p->seps_h[p->P.sep_gid[level][sep]];
  *k = S.k;
  if (!out) return UMF_OK;
  std::vector<double> h;
  int rc = read_pool(p, S.off, (size_t)S.k * S.k, h, level);
  if (rc) return rc;
  if (p->lu) {  // full matrix
    std::copy(h.begin(), h.end(), out);
    return UMF_OK;
  }
  for (int j = 0; j < S.k; j++)
    for (int i = j; i < S.k; i++) {
      out[(size_t)j * S.k + i] = h[(size_t)j * S.k + i];
      out[(size_t)i * S.k + j] = h[(size_t)j * S.k + i];
    }
  return UMF_OK;
}

int umf_get_stats(const umf_plan* p, umf_stats* out) {
  if (!p || !out) return fail(UMF_EINVAL, "null argument");
  *out = p->stats;
  out->bytes_front_pool = (p->prec != UMF_PREC_FP64 ? 4.0 : 8.0) * p->P.arena_elems;
  return UMF_OK;
}

int umf_set_precision(umf_plan* p, int precision, double tol, int max_iter) {
  if (!p) return fail(UMF_EINVAL, "null plan");
  if (p->servers > 0) return fail(UMF_ESTATE, "a server answers requests against this plan's factors");
  if (precision != UMF_PREC_FP64 && precision != UMF_PREC_FP32 && precision != UMF_PREC_BF16X3 &&
      precision != UMF_PREC_FP32T64)
    return fail(UMF_EINVAL, "unknown precision");
  if (!(tol > 0) || max_iter < 1) return fail(UMF_EINVAL, "tolerance must be positive, max_iter >= 1");
  if (p->lu && precision != UMF_PREC_FP64) return fail(UMF_EINVAL, "LU fronts (unsymmetric operators) are FP64 only");
  if (precision != p->prec) p->factored = false;
  p->prec = precision;
  p->ir_tol = tol;
  p->ir_max = max_iter;
  return UMF_OK;
}

int umf_set_tune(umf_plan* p, int gw2, int tile) {
  if (!p) return fail(UMF_EINVAL, "null plan");
  if (gw2 < 256 || gw2 > 8192 || gw2 % 256 != 0)
    return fail(UMF_EINVAL, "outer group width must be 256..8192, a multiple of 256");
  if (tile != 0 && tile != 64 && tile != 128) return fail(UMF_EINVAL, "tile must be 0 (by size), 64 or 128");
  if (gw2 != p->tune_gw2 || tile != p->tune_tile) {
    if (p->fgraph) (void)hipGraphExecDestroy(p->fgraph);  // captured with the old settings
    p->fgraph = nullptr;
  }
  p->tune_gw2 = gw2;
  p->tune_tile = tile;
  return UMF_OK;
}

int umf_get_tune(const umf_plan* p, int* gw2, int* tile) {
  if (!p || !gw2 || !tile) return fail(UMF_EINVAL, "null argument");
  *gw2 = p->tune_gw2;
  *tile = p->tune_tile;
  return UMF_OK;
}

int umf_autotune(umf_plan* p, int reps, double* best_ms) {
  if (!p) return fail(UMF_EINVAL, "null plan");
  if (p->servers > 0) return fail(UMF_ESTATE, "a server answers requests against this plan's factors");
  if (p->host_only) return fail(UMF_ESTATE, "plan was analyzed only (no device memory)");
  reps = std::max(reps, 1);
  const bool graph = p->use_graph, prof = p->prof.on;
  const int g_in = p->tune_gw2, t_in = p->tune_tile;
  if (p->fgraph) (void)hipGraphExecDestroy(p->fgraph);  // captured with the settings in force now
  p->fgraph = nullptr;
  p->use_graph = false;  // direct launches while timing
  p->prof.on = false;
  int best_g = g_in, best_t = t_in, rc = UMF_OK;
  double best = 1e300;
  for (int g : {1024, 2048})
    for (int t : {0, 64, 128}) {
      p->tune_gw2 = g;
      p->tune_tile = t;
      double ms = 1e300;
      for (int r = 0; r < reps && rc == UMF_OK; r++) {
        rc = umf_factor(p);
        ms = std::min(ms, (double)p->stats.t_factor_ms);
      }
      if (rc) break;
      if (ms < best) {
        best = ms;
        best_g = g;
        best_t = t;
      }
    }
  p->use_graph = graph;
  p->prof.on = prof;
  if (rc) {  // a candidate failed (a non-SPD operator fails them all): the settings stay as they were
    p->tune_gw2 = g_in;
    p->tune_tile = t_in;
    return rc;
  }
  rc = umf_set_tune(p, best_g, best_t);
  if (rc) return rc;
  if (best_ms) *best_ms = best;
  return umf_factor(p);  // the factors of the chosen settings
}

int umf_set_graph(umf_plan* p, int on) {
  if (!p) return fail(UMF_EINVAL, "null plan");
  if (p->host_only) return fail(UMF_ESTATE, "plan was analyzed only (no device memory)");
  p->use_graph = on != 0;
  return UMF_OK;
}

int umf_set_lu(umf_plan* p, int on) {
  if (!p) return fail(UMF_EINVAL, "null plan");
  if (p->P.forest) return fail(UMF_EINVAL, "batches factor with Cholesky fronts");
  if (p->servers > 0) return fail(UMF_ESTATE, "a server answers requests against this plan's factors");
  const bool want = on != 0;
  if (want == p->lu) return UMF_OK;
  if (want && p->prec != UMF_PREC_FP64) return fail(UMF_EINVAL, "LU fronts (unsymmetric operators) are FP64 only");
  if (!want && !p->val_h.empty()) {
    bool unsym = false;
    const std::string e = check_operator(p->P.n, p->P.rp.data(), p->P.ci.data(), p->val_h.data(), &unsym);
    if (!e.empty() || unsym) return fail(UMF_EINVAL, "the operator's values are not symmetric: it needs LU fronts");
  }
  if (!p->host_only) HIPCHK(hipStreamSynchronize(p->stream));
  Plan& P = p->P;
  build_programs(P, want);
  p->lu = want;
  if (p->host_only) return UMF_OK;
  for (void** q : {(void**)&p->d_tri, (void**)&p->d_tri_off, (void**)&p->d_ops, (void**)&p->d_tiles,
                   (void**)&p->d_tile_off, (void**)&p->d_fly_tab, (void**)&p->d_fly_tri, (void**)&p->d_rest_tri,
                   (void**)&p->d_rest_off}) {
    if (*q) HIPCHK(hipFree(*q));
    *q = nullptr;
  }
  HIPCHK(upload(&p->d_tri, P.triples));
  HIPCHK(upload(&p->d_tri_off, P.tri_off));
  HIPCHK(upload(&p->d_ops, P.ops));
  HIPCHK(upload(&p->d_tiles, P.tiles));
  HIPCHK(upload(&p->d_tile_off, P.tile_off));
  if (!P.rest_off.empty()) {
    HIPCHK(upload(&p->d_fly_tab, P.fly_tab));
    HIPCHK(upload(&p->d_fly_tri, P.fly_tri));
    HIPCHK(upload(&p->d_rest_tri, P.rest_tri));
    HIPCHK(upload(&p->d_rest_off, P.rest_off));
  }
  for (size_t f = 0; f < p->fronts_h.size(); f++) p->fronts_h[f].fly_off = P.fly_off.empty() ? -1 : P.fly_off[f];
  free_lu(p);
  if (want) {
    const int rc = alloc_lu(p);
    if (rc) return rc;
  }
  HIPCHK(hipMemcpy(p->d_fronts, p->fronts_h.data(), sizeof(FrontDev) * p->fronts_h.size(), hipMemcpyHostToDevice));
  if (p->fgraph) (void)hipGraphExecDestroy(p->fgraph);
  p->fgraph = nullptr;
  // schedules made from the previous programs (assembly grids, the general-path fronts)
  p->full_sched.reset();
  p->full_dirty.reset();
  p->gen_fronts.clear();
  if (p->d_fids) (void)hipFree(p->d_fids);
  p->d_fids = nullptr;
  p->factored = false;
  return UMF_OK;
}

int umf_get_lu(const umf_plan* p) { return p ? (p->lu ? 1 : 0) : fail(UMF_EINVAL, "null plan"); }

int umf_set_sweep_sync(umf_plan* p, int fenced) {
  if (!p) return fail(UMF_EINVAL, "null plan");
  p->fenced = fenced != 0;
  return UMF_OK;
}

void* umf_stream(umf_plan* p) { return p ? (void*)p->stream : nullptr; }

int umf_profile_enable(umf_plan* p, int on) {
  if (!p) return fail(UMF_EINVAL, "null plan");
  p->prof.on = on != 0;
  return UMF_OK;
}
int umf_profile_select(umf_plan* p, const char* kernel) {
  if (!p) return fail(UMF_EINVAL, "null plan");
  p->prof.only = kernel ? kernel : "";
  return UMF_OK;
}
int umf_profile_reset(umf_plan* p) {
  if (!p) return fail(UMF_EINVAL, "null plan");
  p->prof.stats.clear();
  return UMF_OK;
}
int umf_profile_read(const umf_plan* p, int i, char* name, int name_len, double* ms, int* launches, double* flops,
                     double* bytes) {
  if (!p) return fail(UMF_EINVAL, "null plan");
  const std::vector<KStat>& st = p->prof.stats;
  if (i >= 0 && i < (int)st.size()) {
    if (name && name_len > 0) {
      std::strncpy(name, st[i].name.c_str(), name_len - 1);
      name[name_len - 1] = 0;
    }
    if (ms) *ms = st[i].ms;
    if (launches) *launches = st[i].launches;
    if (flops) *flops = st[i].flops;
    if (bytes) *bytes = st[i].bytes;
  }
  return (int)st.size();
}

}  // extern "C"

// ============================================================================== serving
// A server answers requests against an immutable base factorization.  Request r is a local
// change of the operator (the base operator with some values changed, and the modified DOFs of
// the reference's modDOF rule) plus right-hand sides; its answer is x = A_r^{-1} b, as if the
// base plan were Updated with A_r (the reference's use case, test_src/DE/testMF3D.cpp:137-215)
// -- but the base stays untouched and requests do not see each other's changes.
//
// A request runs in a slot: an execution context with its own stream, descriptors, operator
// values, error flags and solve scratch, sharing the base plan's planner output, programs and
// front pool.  Its dirty fronts and separators (the reference's Update recurrence) are factored
// into memory reserved for the request in the server's slot pool; the slot's descriptors point
// there for those, and at the base's factors for every other front.  The footprint of a request
// is the bytes of its dirty fronts (with their inverted diagonal blocks) and separators; a
// request is admitted when the pool has a free range that holds it.  Slots on their own streams
// run concurrently; each request's arithmetic depends on its own dirty set only (the schedule's
// algorithm choices are fixed per plan level), so its result is the same alone or among others,
// and the same as a fresh factorization of its operator.
struct umf_server {
  umf_plan* base = nullptr;
  std::vector<umf_plan*> slots;
  std::vector<double*> coeff;       // per slot: the request's coefficient field (box requests)
  std::vector<size_t> coeff_elems;
  std::vector<char> busy;           // per slot: a request is running in it
  char* pool = nullptr;
  size_t pool_bytes = 0, used = 0;
  int elem = 8;                     // bytes per front entry (8 FP64 factors, 4 FP32)
  std::mutex mu;
  std::map<size_t, size_t> free_;   // free byte ranges of the pool, start -> end
  struct Ticket {
    Plan::Dirty D;
    size_t off = 0, bytes = 0;
    std::vector<std::pair<int, i64>> fr, dinv, sep;  // id -> element offset inside the range
    bool running = false;
  };
  std::map<int, Ticket> tickets;
  int next_ticket = 1;
};

namespace {
// Layout of a request's dirty matrices inside one pool range (256-byte aligned pieces).
size_t ticket_layout(const Plan& P, const Plan::Dirty& D, int elem, umf_server::Ticket& T) {
  size_t b = 0;
  auto put = [&](i64 elems) {
    const size_t at = b;
    b += ((size_t)elems * elem + 255) / 256 * 256;
    return (i64)(at / elem);
  };
  T.fr.clear();
  T.dinv.clear();
  T.sep.clear();
  for (int ell = 0; ell < P.L && ell < (int)D.boxes.size(); ell++)
    for (int bx : D.boxes[ell]) {
      const int f = P.box_gid[ell][bx];
      const FrontInfo& F = P.fronts[f];
      const i64 N = (i64)F.n + F.m;
      T.fr.push_back({f, put(N * N)});
      T.dinv.push_back({f, put((i64)((F.n + NB - 1) / NB) * NB * NB)});
    }
  for (int ell = 1; ell < P.L && ell < (int)D.seps.size(); ell++)
    for (int e : D.seps[ell]) {
      const int sid = P.sep_gid[ell][e];
      T.sep.push_back({sid, put((i64)P.seps[sid].k * P.seps[sid].k)});
    }
  return b;
}

int check_slot(umf_server* s, int slot) {
  if (!s || !s->base) return fail(UMF_EINVAL, "null server");
  if (slot < 0 || slot >= (int)s->slots.size()) return fail(UMF_EINVAL, "bad slot index");
  return UMF_OK;
}

// Frees a slot's own device memory (its borrowed pointers are cleared first).
void free_slot(umf_plan* q) {
  q->arena = nullptr;
  q->arena32 = nullptr;
  q->d_dofs = nullptr;
  q->d_tri = nullptr;
  q->d_tri_off = nullptr;
  q->d_ops = nullptr;
  q->d_tiles = nullptr;
  q->d_tile_off = nullptr;
  q->d_rp = nullptr;
  q->d_ci = nullptr;
  q->d_fly_tab = nullptr;
  q->d_fly_tri = nullptr;
  q->d_rest_tri = nullptr;
  q->d_rest_off = nullptr;
  free_plan(q);
  delete q;
}

// A slot context over base plan b: everything a factor pass and a solve touch, with the
// plan's data borrowed and the per-call buffers its own, allocated here once.
int make_slot(umf_plan* b, umf_plan** out) {
  umf_plan* q = new umf_plan(b->Pp);
  q->is_slot = true;
  std::copy(b->args, b->args + 6, q->args);
  q->prec = q->fact_prec = b->fact_prec;
  q->ir_tol = b->ir_tol;
  q->ir_max = b->ir_max;
  q->tune_gw2 = b->tune_gw2;
  q->tune_tile = b->tune_tile;
  q->fenced = b->fenced;
  q->device = b->device;
  q->xs_elems = b->xs_elems;
  q->n_flags = b->n_flags;
  q->ll_blocks = b->ll_blocks;
  q->max_tiles_per_mat = b->max_tiles_per_mat;
  q->fronts_h = b->fronts_h;
  q->seps_h = b->seps_h;
  q->mats_h = b->mats_h;
  q->stats = b->stats;
  q->arena = b->arena;
  q->arena32 = b->arena32;
  q->d_dofs = b->d_dofs;
  q->d_tri = b->d_tri;
  q->d_tri_off = b->d_tri_off;
  q->d_ops = b->d_ops;
  q->d_tiles = b->d_tiles;
  q->d_tile_off = b->d_tile_off;
  q->d_rp = b->d_rp;
  q->d_ci = b->d_ci;
  q->d_fly_tab = b->d_fly_tab;
  q->d_fly_tri = b->d_fly_tri;
  q->d_rest_tri = b->d_rest_tri;
  q->d_rest_off = b->d_rest_off;
  const Plan& P = *b->Pp;
  const size_t nnz = P.ci.size(), NF = P.nfronts(), NS = P.nseps_total();
  auto bail = [&](hipError_t e) {
    free_slot(q);
    return fail(e == hipErrorOutOfMemory ? UMF_ENOMEM : UMF_EDEVICE,
                std::string("serving slot: ") + hipGetErrorString(e));
  };
  hipError_t e;
#define SLOTCHK(x) if ((e = (x)) != hipSuccess) return bail(e)
  // non-blocking: a slot's work never waits on the null stream's, so slots overlap
  SLOTCHK(hipStreamCreateWithFlags(&q->stream, hipStreamNonBlocking));
  SLOTCHK(hipEventCreate(&q->ev0));
  SLOTCHK(hipEventCreate(&q->ev1));
  SLOTCHK(hipMalloc((void**)&q->d_mats, sizeof(MatDev) * std::max<size_t>(q->mats_h.size(), 1)));
  SLOTCHK(hipMalloc((void**)&q->d_fronts, sizeof(FrontDev) * std::max<size_t>(NF, 1)));
  SLOTCHK(hipMalloc((void**)&q->d_seps, sizeof(SepDev) * std::max<size_t>(NS, 1)));
  SLOTCHK(hipMalloc((void**)&q->d_val, sizeof(double) * std::max<size_t>(nnz, 1)));
  SLOTCHK(hipMemcpy(q->d_val, b->d_val, sizeof(double) * nnz, hipMemcpyDeviceToDevice));
  SLOTCHK(hipMalloc((void**)&q->d_err, sizeof(int) * 8));
  SLOTCHK(hipMemset(q->d_err, 0, sizeof(int) * 8));
  SLOTCHK(hipMalloc((void**)&q->d_flags, sizeof(int) * std::max(q->n_flags, 1)));
  SLOTCHK(hipMemset(q->d_flags, 0, sizeof(int) * std::max(q->n_flags, 1)));
  SLOTCHK(hipMalloc((void**)&q->d_ll, sizeof(unsigned long long) * kLLBlock * std::max<i64>(q->ll_blocks, 1)));
  SLOTCHK(hipMemset(q->d_ll, 0, sizeof(unsigned long long) * kLLBlock * std::max<i64>(q->ll_blocks, 1)));
  SLOTCHK(hipMalloc((void**)&q->d_arrive, sizeof(int) * std::max<size_t>(NF, 1)));
  SLOTCHK(hipMemset(q->d_arrive, 0, sizeof(int) * std::max<size_t>(NF, 1)));
  q->ids_cap = NF + 2 * NS;
  SLOTCHK(hipMalloc((void**)&q->d_ids, sizeof(int) * std::max<size_t>(q->ids_cap, 1)));
  SLOTCHK(hipMalloc((void**)&q->d_bflag, sizeof(int) * std::max(P.n, 1)));
  SLOTCHK(hipMemset(q->d_bflag, 0, sizeof(int) * std::max(P.n, 1)));
  q->blist_cap = P.n;
  SLOTCHK(hipMalloc((void**)&q->d_blist, sizeof(int) * (P.n + 1)));
#undef SLOTCHK
  q->factored = true;
  int rc = ensure_solve_scratch(q);
  if (!rc) rc = ensure_sweep_lists(q);
  if (rc) {
    free_slot(q);
    return rc;
  }
  *out = q;
  return UMF_OK;
}
}  // namespace

extern "C" {

int umf_server_create(umf_plan* base, int nslots, double pool_bytes, umf_server** out) {
  if (!base || !out) return fail(UMF_EINVAL, "null argument");
  *out = nullptr;
  if (base->host_only || !base->factored) return fail(UMF_ESTATE, "the base plan must be factored");
  if (base->P.forest) return fail(UMF_EINVAL, "serving takes a single-problem plan");
  if (base->lu) return fail(UMF_EINVAL, "serving takes Cholesky fronts (symmetric operators)");
  if (base->fact_prec == UMF_PREC_FP32T64) return fail(UMF_EINVAL, "serving takes the FP64, FP32 or BF16X3 form");
  if (nslots < 1 || nslots > 32) return fail(UMF_EINVAL, "slots must be 1..32");
  if (!(pool_bytes >= 0)) return fail(UMF_EINVAL, "bad pool size");
  HIPCHK(hipStreamSynchronize(base->stream));
  umf_server* s = new umf_server();
  s->elem = base->fact_prec == UMF_PREC_FP64 ? 8 : 4;
  s->pool_bytes = (size_t)pool_bytes / 256 * 256;
  int cur = 0;
  HIPCHK(hipGetDevice(&cur));
  if (cur != base->device) HIPCHK(hipSetDevice(base->device));
  int rc = UMF_OK;
  if (s->pool_bytes) {
    const hipError_t e = hipMalloc((void**)&s->pool, s->pool_bytes);
    if (e != hipSuccess)
      rc = fail(UMF_ENOMEM, "slot pool of " + std::to_string(pool_bytes / 1e9) + " GB: " + hipGetErrorString(e));
  }
  for (int i = 0; rc == UMF_OK && i < nslots; i++) {
    umf_plan* q = nullptr;
    rc = make_slot(base, &q);
    if (rc == UMF_OK) s->slots.push_back(q);
  }
  if (cur != base->device) (void)hipSetDevice(cur);
  if (rc) {
    for (umf_plan* q : s->slots) free_slot(q);
    if (s->pool) (void)hipFree(s->pool);
    delete s;
    return rc;
  }
  s->base = base;
  s->coeff.assign(nslots, nullptr);
  s->coeff_elems.assign(nslots, 0);
  s->busy.assign(nslots, 0);
  if (s->pool_bytes) s->free_[0] = s->pool_bytes;
  base->servers++;
  *out = s;
  return UMF_OK;
}

int umf_server_destroy(umf_server* s) {
  if (!s) return UMF_OK;
  for (umf_plan* q : s->slots) {
    if (q->stream) (void)hipStreamSynchronize(q->stream);
    free_slot(q);
  }
  for (double* c : s->coeff)
    if (c) (void)hipFree(c);
  if (s->pool) (void)hipFree(s->pool);
  if (s->base) s->base->servers--;
  delete s;
  return UMF_OK;
}

int umf_server_footprint(umf_server* s, int n_mod, const int* mod_dofs, double* bytes) {
  if (!s || !s->base || !bytes || (n_mod > 0 && !mod_dofs)) return fail(UMF_EINVAL, "null argument");
  const Plan& P = s->base->P;
  std::vector<int> mod(mod_dofs, mod_dofs + std::max(n_mod, 0));
  Plan::Dirty D;
  std::string e;
  if (!P.dirty_from_dofs(mod, D, e)) return fail(UMF_EUPDATEMAP, e);
  umf_server::Ticket T;
  *bytes = (double)ticket_layout(P, D, s->elem, T);
  return UMF_OK;
}

int umf_server_admit(umf_server* s, int n_mod, const int* mod_dofs, int* ticket) {
  if (!s || !s->base || !ticket || (n_mod > 0 && !mod_dofs)) return fail(UMF_EINVAL, "null argument");
  const Plan& P = s->base->P;
  std::vector<int> mod(mod_dofs, mod_dofs + std::max(n_mod, 0));
  umf_server::Ticket T;
  std::string e;
  if (!P.dirty_from_dofs(mod, T.D, e)) return fail(UMF_EUPDATEMAP, e);
  T.bytes = ticket_layout(P, T.D, s->elem, T);
  if (T.bytes > s->pool_bytes) {
    char buf[200];
    snprintf(buf, sizeof buf, "request needs %.3f GB of slot pool, the pool holds %.3f GB", T.bytes / 1e9,
             s->pool_bytes / 1e9);
    return fail(UMF_EINVAL, buf);
  }
  std::lock_guard<std::mutex> lk(s->mu);
  if (T.bytes > 0) {
    auto it = s->free_.begin();
    for (; it != s->free_.end(); ++it)
      if (it->second - it->first >= T.bytes) break;
    if (it == s->free_.end()) return fail(UMF_EBUSY, "slot pool: no free range holds the request now");
    T.off = it->first;
    const size_t end = it->second;
    s->free_.erase(it);
    if (T.off + T.bytes < end) s->free_[T.off + T.bytes] = end;
    s->used += T.bytes;
  }
  *ticket = s->next_ticket++;
  s->tickets[*ticket] = std::move(T);
  return UMF_OK;
}

int umf_server_release(umf_server* s, int ticket) {
  if (!s || !s->base) return fail(UMF_EINVAL, "null server");
  std::lock_guard<std::mutex> lk(s->mu);
  auto it = s->tickets.find(ticket);
  if (it == s->tickets.end()) return fail(UMF_EINVAL, "unknown ticket");
  if (it->second.running) return fail(UMF_ESTATE, "the ticket's request is still running");
  const size_t a = it->second.off, bb = a + it->second.bytes;
  if (it->second.bytes > 0) {
    size_t lo = a, hi = bb;
    auto nx = s->free_.lower_bound(a);  // merge with the neighbouring free ranges
    if (nx != s->free_.end() && nx->first == bb) {
      hi = nx->second;
      s->free_.erase(nx);
    }
    auto pv = s->free_.lower_bound(a);
    if (pv != s->free_.begin()) {
      --pv;
      if (pv->second == a) {
        lo = pv->first;
        s->free_.erase(pv);
      }
    }
    s->free_[lo] = hi;
    s->used -= it->second.bytes;
  }
  s->tickets.erase(it);
  return UMF_OK;
}

int umf_server_pool(umf_server* s, double* total, double* used) {
  if (!s || !s->base) return fail(UMF_EINVAL, "null server");
  std::lock_guard<std::mutex> lk(s->mu);
  if (total) *total = (double)s->pool_bytes;
  if (used) *used = (double)s->used;
  return UMF_OK;
}

int umf_server_set_values(umf_server* s, int slot, const double* d_val) {
  int rc = check_slot(s, slot);
  if (rc) return rc;
  if (!d_val) return fail(UMF_EINVAL, "null argument");
  umf_plan* q = s->slots[slot];
  HIPCHK(hipMemcpyAsync(q->d_val, d_val, sizeof(double) * q->P.ci.size(), hipMemcpyDeviceToDevice, q->stream));
  HIPCHK(hipStreamSynchronize(q->stream));
  return UMF_OK;
}

int umf_server_box_scale(umf_server* s, int slot, const double* d_coeff, double h, const int* lo, int side,
                         double factor, int* mod_dofs, int cap, int* n_mod) {
  int rc = check_slot(s, slot);
  if (rc) return rc;
  if (!d_coeff || !lo || !n_mod || (cap > 0 && !mod_dofs)) return fail(UMF_EINVAL, "null argument");
  umf_plan* q = s->slots[slot];
  const Grid& g = q->P.g;
  size_t ce = 1;
  for (int a = 0; a < g.dim; a++) ce *= (size_t)(g.N[a] + 2 * g.w);
  if (s->coeff_elems[slot] < ce) {
    if (s->coeff[slot]) HIPCHK(hipFree(s->coeff[slot]));
    s->coeff[slot] = nullptr;
    HIPCHK(hipMalloc((void**)&s->coeff[slot], sizeof(double) * ce));
    s->coeff_elems[slot] = ce;
  }
  // the request's field and operator start from the base's
  HIPCHK(hipMemcpyAsync(s->coeff[slot], d_coeff, sizeof(double) * ce, hipMemcpyDeviceToDevice, q->stream));
  HIPCHK(hipMemcpyAsync(q->d_val, s->base->d_val, sizeof(double) * q->P.ci.size(), hipMemcpyDeviceToDevice,
                        q->stream));
  return umf_box_scale_device(q, s->coeff[slot], h, lo, side, factor, q->d_val, mod_dofs, cap, n_mod);
}

int umf_server_run(umf_server* s, int slot, int ticket, const double* d_b, double* d_x, int nrhs, int ld) {
  int rc = check_slot(s, slot);
  if (rc) return rc;
  if (!d_b || !d_x) return fail(UMF_EINVAL, "null argument");
  umf_plan* q = s->slots[slot];
  const Plan& P = q->P;
  umf_server::Ticket* T = nullptr;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->tickets.find(ticket);
    if (it == s->tickets.end()) return fail(UMF_EINVAL, "unknown ticket");
    if (it->second.running) return fail(UMF_ESTATE, "the ticket's request is already running");
    if (s->busy[slot]) return fail(UMF_ESTATE, "the slot is running another request");
    it->second.running = true;
    s->busy[slot] = 1;
    T = &it->second;  // map nodes stay put while other tickets come and go
  }
  struct Done {
    umf_server* s;
    umf_server::Ticket* T;
    int slot;
    ~Done() {
      std::lock_guard<std::mutex> lk(s->mu);
      T->running = false;
      s->busy[slot] = 0;
    }
  } done{s, T, slot};
  // descriptors: the request's dirty matrices in its pool range, every other one the base's
  q->mats_h = s->base->mats_h;
  q->fronts_h = s->base->fronts_h;
  q->seps_h = s->base->seps_h;
  const char* abase = s->elem == 8 ? (const char*)s->base->arena : (const char*)s->base->arena32;
  const i64 shift = (i64)(((intptr_t)(s->pool + T->off) - (intptr_t)abase) / s->elem);
  const int NF = P.nfronts();
  for (auto& fo : T->fr) {
    q->fronts_h[fo.first].off = shift + fo.second;
    q->mats_h[fo.first].off = shift + fo.second;
  }
  for (auto& fo : T->dinv) q->fronts_h[fo.first].dinv_off = shift + fo.second;
  for (auto& so : T->sep) {
    q->seps_h[so.first].off = shift + so.second;
    q->mats_h[NF + so.first].off = shift + so.second;
  }
  HIPCHK(hipMemcpyAsync(q->d_mats, q->mats_h.data(), sizeof(MatDev) * q->mats_h.size(), hipMemcpyHostToDevice,
                        q->stream));
  HIPCHK(hipMemcpyAsync(q->d_fronts, q->fronts_h.data(), sizeof(FrontDev) * q->fronts_h.size(),
                        hipMemcpyHostToDevice, q->stream));
  if (!q->seps_h.empty())
    HIPCHK(hipMemcpyAsync(q->d_seps, q->seps_h.data(), sizeof(SepDev) * q->seps_h.size(), hipMemcpyHostToDevice,
                          q->stream));
  // factor the request's dirty fronts (the reference's Update over its operator)
  HIPCHK(hipEventRecord(q->ev0, q->stream));
  double fl = 0;
  int nf = 0;
  if (T->D.any) {
    rc = factor_any(q, T->D);
    if (rc) return rc;
    for (int ell = 0; ell < P.L; ell++)
      for (int bx : T->D.boxes[ell]) {
        fl += P.front_flops(P.box_gid[ell][bx]);
        nf++;
      }
  }
  HIPCHK(hipEventRecord(q->ev1, q->stream));
  HIPCHK(hipStreamSynchronize(q->stream));
  float ms = 0;
  HIPCHK(hipEventElapsedTime(&ms, q->ev0, q->ev1));
  q->stats.t_update_ms = ms;
  q->stats.flops_update = fl;
  q->stats.n_fronts_update = nf;
  q->last = T->D;
  rc = check_err(q);
  if (rc) return rc;
  q->factored = true;
  return solve_apply(q, 0, d_b, d_x, nrhs, ld, true);
}

int umf_server_stats(const umf_server* s, int slot, umf_stats* out) {
  if (!s || !s->base || !out) return fail(UMF_EINVAL, "null argument");
  if (slot < 0 || slot >= (int)s->slots.size()) return fail(UMF_EINVAL, "bad slot index");
  *out = s->slots[slot]->stats;
  out->bytes_front_pool = (double)s->pool_bytes;
  return UMF_OK;
}

}  // extern "C"
