// plan.h -- host-side planner of the elimination tree and of the device assembly programs.
//
// The planner reproduces the reference's nested-dissection geometry exactly (levels, leaf
// widths, DOF lists and their order, separator numbering, center points, dirty sets), because
// those are the parity contract; everything downstream of it (front storage, assembly programs,
// batching) is this library's own layout.
#pragma once
#include <cstdint>
#include <memory>
#include <string>
#include <vector>

namespace umf {

typedef long long i64;

// Regular-grid ND geometry.  A separator is identified by its normal axis `a` and the grid
// coordinates `c` of the box on its low side; it lies between boxes c and c + e_a.
struct Grid {
  int dim = 2, N[3] = {1, 1, 1}, w = 1, L = 0, bw[3] = {0, 0, 0};

  int side(int ell) const { return 1 << ell; }
  int nboxes(int ell) const;
  int nseps(int ell) const;  // separators between level-ell boxes
  void box_coords(int box, int ell, int c[3]) const;
  int box_index(const int c[3], int ell) const;
  int sep_index(int a, const int c[3], int ell) const;
  int sep_decode(int sep, int ell, int c[3]) const;  // returns the normal axis
  int parent(int box, int ell) const;
  int children(int box, int ell, int out[8]) const;
  int nbor_seps(int box, int ell, int out[6]) const;  // x-, x+, y-, y+ (, z-, z+); -1 = none
  void sep_boxes(int sep, int ell, int out[2]) const;
  int sep_children(int sep, int ell, int out[4]) const;
  // children separators (level ell+1) of a level-ell box: interior ones (self) and the ones on
  // the box boundary (shared), each ascending.
  void child_seps(int box, int ell, std::vector<int>& self, std::vector<int>& shared) const;
  int nominal(int axis, int ell) const;  // nominal box width along axis at level ell
  void box_center(int box, int ell, std::vector<int>& out) const;
  void sep_center(int sep, int ell, std::vector<int>& out) const;
  int dof(int i, int j, int k) const { return i + j * N[0] + k * N[0] * N[1]; }
  int leaf_of_dof(int d) const;  // -1 when the reference's map would leave the leaf grid
};

// One dense column-major matrix in the device arena.
struct MatDesc {
  i64 off;   // element offset in the arena
  int rows;  // = ld
  int cols;
};

// Rectangular block copy inside the arena: dst(dr0+i, dc0+j) = src(sr0+i, sc0+j), or
// src(sr0+j, sc0+i) when `trans`; `lower` restricts a diagonal block to i >= j.
struct BlockOp {
  int src;  // matrix id
  int sr0, sc0, dr0, dc0, nr, nc;
  int flags;  // 1 = transposed source, 2 = lower triangle only
};
struct Tile {
  int op, r0, c0;  // 64 x 64 tile of an op, tile-local origin within the op
};
struct Triple {
  int row, col, k;  // dst(row, col) = csr_val[k]
};

struct FrontInfo {
  int level, box;     // level and box index in its own problem's tree
  int prob;           // problem of a batch (0 for a single problem)
  int n, m;           // pivot size, boundary size
  i64 int_off, bd_off;  // offsets into dofs[]
  int nsep;           // number of existing neighbour separators
  int sep[6];         // global separator ids of the neighbours, in bd order
  int sep_pos[6];     // their offsets within bdDOFs
  int sep_len[6];
  i64 dinv_off;  // arena offset of the inverted 64x64 diagonal blocks of L11 (ceil(n/64) of them)
};
struct SepInfo {
  int level, idx;
  int k;
  i64 dof_off;
  int box[2];   // global front ids of the two neighbour boxes (low side first)
  int pos[2];   // offset of this separator inside each neighbour's bdDOFs
};

struct Forest;

struct Plan {
  Grid g;
  int L = 0;  // levels (of the deepest tree for a batch)
  int n = 0;  // DOFs
  std::vector<int> rp, ci;  // CSR pattern

  std::vector<std::vector<int>> box_gid;  // [level][box] -> front id
  std::vector<std::vector<int>> sep_gid;  // [level][sep] -> separator id (level >= 1)
  std::vector<FrontInfo> fronts;
  std::vector<SepInfo> seps;
  std::vector<int> dofs;  // all DOF lists back to back

  // matrices: fronts 0..F-1, then separators F..F+S-1
  std::vector<MatDesc> mats;
  i64 arena_elems = 0;

  // assembly programs, per matrix: ranges into triples / tiles (ops are shared)
  std::vector<Triple> triples;
  std::vector<i64> tri_off;  // size mats+1
  std::vector<BlockOp> ops;
  std::vector<Tile> tiles;
  std::vector<i64> tile_off;  // size mats+1

  // Leaf fronts (assembled from CSR entries only) at left-looking levels skip the dense zeroing:
  // the entries below a column block's diagonal block go straight into the panel step that
  // forms that tile (fly_*), the rest (diagonal blocks, the trailing boundary block) are zeroed
  // and scattered as usual (rest_*).
  std::vector<int> fly_off;      // per front: its table in fly_tab, -1 = not such a front
  std::vector<int> fly_tab;      // per front: for column block cb, row tile t below it:
                                 // fly_tri[fly_tab[off + cb*(NT+1) + t] .. +1], NT = ceil(N/64)
  std::vector<Triple> fly_tri;   // grouped by (column block, row tile)
  std::vector<Triple> rest_tri;  // the front's other entries
  std::vector<i64> rest_off;     // per matrix (size mats+1)
  // programs write full matrices (LU fronts of an unsymmetric operator) instead of lower
  // triangles (Cholesky fronts); no fly split then
  bool full = false;

  int nfronts() const { return (int)fronts.size(); }
  int nseps_total() const { return (int)seps.size(); }

  // Dirty sets (reference recurrence): boxes[level], seps receiving Schur adds[level].
  struct Dirty {
    std::vector<std::vector<int>> boxes, seps;
    bool any = false;
  };
  Dirty all_dirty() const;
  // returns false (with err) if a DOF maps outside the leaf grid
  bool dirty_from_dofs(const std::vector<int>& mod, Dirty& out, std::string& err) const;

  // A batch of independent problems merged into one block-diagonal operator: the problems'
  // own plans, and where each one sits in the merged numbering.  Null for a single problem.
  std::shared_ptr<Forest> forest;

  double front_flops(int f) const {
    double nn = fronts[f].n, mm = fronts[f].m;
    return nn * nn * nn / 3.0 + nn * nn * mm + nn * mm * mm;
  }
};

// A batch: tree levels are aligned at the leaves-to-root depth of the deepest tree, problem p's
// level ell being merged level ell + shift[p] (shift = L_max - L_p), so one batched launch per
// stage covers every problem's fronts at that depth.
struct Forest {
  std::vector<Plan> sub;
  std::vector<int> shift;           // merged level = own level + shift
  std::vector<int> front0, sep0;    // first merged front / separator id
  std::vector<int> dof0;            // first merged DOF
  std::vector<i64> nnz0;            // first merged CSR entry
  // merged box / separator index of the problem's first box / separator at each merged level
  std::vector<std::vector<int>> box_base, sep_base;  // [problem][merged level]
};

// Merges the problems' plans (moved into the forest) into one plan of the block-diagonal batch.
void merge_plans(Plan& M, std::vector<Plan>&& subs);
// the fly_* / rest_* split of build_plan's assembly programs
void build_fly(Plan& P);

// (Re)builds the assembly programs of a planned tree, lower triangles or full matrices.
void build_programs(Plan& P, bool full);

// Builds the plan; returns an empty string on success or the error message.  `geom_error`
// is set when the failure is the reference's geometry exception.  full: programs for full
// matrices (LU fronts).
std::string build_plan(Plan& P, int dim, int nx, int ny, int nz, int width, int max_side, int n, const int* rp,
                       const int* ci, bool& geom_error, bool full = false);

}  // namespace umf
