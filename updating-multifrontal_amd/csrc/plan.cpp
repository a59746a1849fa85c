// plan.cpp -- elimination-tree planner and assembly-program generator (host C++).
//
// Geometry follows the reference bit for bit where it defines identity and order:
//   levels / leaf widths          src/DE/mf2d.cpp:23-69, src/DE/mf3d.cpp:41-92
//   separator numbering           src/DE/tree2d_common.cpp:8-35, src/DE/tree3d_common.cpp:9-71
//   neighbours / children         src/DE/tree2d_common.cpp:91-226, src/DE/tree3d_common.cpp:120-312
//   DOF lists and their order     src/DE/mf2d.cpp:73-131, :293-302, :357-383; src/DE/mf3d.cpp:96-185,
//                                 :546-565, :612-640
//   center points                 src/DE/mf2d.cpp:186-270, src/DE/mf3d.cpp:253-524
//   dirty-set recurrence          src/DE/mf2d.cpp:150-181, :579-620, :704-747
// The assembly programs restate the reference's extend-add (src/DE/mf2d.cpp:322-446) as
// contiguous separator-pair block copies plus CSR scatter triples, lower triangle only.
#include "plan.h"

#include <algorithm>
#include <cmath>
#include <map>
#include <set>
#include <stdexcept>

namespace umf {

// ------------------------------------------------------------------------------ Grid
int Grid::nboxes(int ell) const {
  int b = side(ell);
  return dim == 2 ? b * b : b * b * b;
}
int Grid::nseps(int ell) const {
  int b = side(ell);
  return dim == 2 ? 2 * b * (b - 1) : 3 * b * b * (b - 1);
}
void Grid::box_coords(int box, int ell, int c[3]) const {
  int b = side(ell);
  c[0] = box % b;
  c[1] = (box / b) % b;
  c[2] = dim == 3 ? box / (b * b) : 0;
}
int Grid::box_index(const int c[3], int ell) const {
  int b = side(ell);
  return c[0] + b * (c[1] + (dim == 3 ? b * c[2] : 0));
}
// Separator numbering: per z-slice (3D), rows of (b-1) x-normal then b y-normal entries, then
// b*b z-normal entries (tree3d_common.cpp:23-40); 2D is one slice without z-normal entries.
int Grid::sep_index(int a, const int c[3], int ell) const {
  int b = side(ell);
  int row = 2 * b - 1;
  int slice = dim == 3 ? 2 * (b - 1) * b + b * b : 0;
  int base = dim == 3 ? c[2] * slice : 0;
  if (a == 0) return base + c[1] * row + c[0];
  if (a == 1) return base + c[1] * row + c[0] + (b - 1);
  return base + 2 * (b - 1) * b + c[1] * b + c[0];
}
int Grid::sep_decode(int sep, int ell, int c[3]) const {
  int b = side(ell);
  int row = 2 * b - 1;
  c[2] = 0;
  int r = sep;
  if (dim == 3) {
    int slice = 2 * (b - 1) * b + b * b;
    c[2] = sep / slice;
    r = sep % slice;
    if (r >= 2 * (b - 1) * b) {
      r -= 2 * (b - 1) * b;
      c[1] = r / b;
      c[0] = r % b;
      return 2;
    }
  }
  c[1] = r / row;
  int x = r % row;
  if (x < b - 1) { c[0] = x; return 0; }
  c[0] = x - (b - 1);
  return 1;
}
int Grid::parent(int box, int ell) const {
  int c[3];
  box_coords(box, ell, c);
  for (int a = 0; a < 3; a++) c[a] /= 2;
  return box_index(c, ell - 1);
}
int Grid::children(int box, int ell, int out[8]) const {
  // child order of tree2d_common.cpp:126-129 ((x,y) offsets (0,0),(0,1),(1,0),(1,1)) and
  // tree3d_common.cpp:171-178
  static const int o2[4][3] = {{0, 0, 0}, {0, 1, 0}, {1, 0, 0}, {1, 1, 0}};
  static const int o3[8][3] = {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1},
                               {1, 1, 0}, {1, 0, 1}, {0, 1, 1}, {1, 1, 1}};
  int c[3];
  box_coords(box, ell, c);
  int k = dim == 2 ? 4 : 8;
  for (int i = 0; i < k; i++) {
    const int* o = dim == 2 ? o2[i] : o3[i];
    int d[3] = {2 * c[0] + o[0], 2 * c[1] + o[1], 2 * c[2] + o[2]};
    out[i] = box_index(d, ell + 1);
  }
  return k;
}
int Grid::nbor_seps(int box, int ell, int out[6]) const {
  int c[3], b = side(ell);
  box_coords(box, ell, c);
  for (int a = 0; a < dim; a++) {
    int lo[3] = {c[0], c[1], c[2]};
    lo[a] -= 1;
    out[2 * a] = c[a] > 0 ? sep_index(a, lo, ell) : -1;
    out[2 * a + 1] = c[a] < b - 1 ? sep_index(a, c, ell) : -1;
  }
  return 2 * dim;
}
void Grid::sep_boxes(int sep, int ell, int out[2]) const {
  int c[3];
  int a = sep_decode(sep, ell, c);
  out[0] = box_index(c, ell);
  c[a] += 1;
  out[1] = box_index(c, ell);
}
int Grid::sep_children(int sep, int ell, int out[4]) const {
  int c[3];
  int a = sep_decode(sep, ell, c);
  int d[3] = {2 * c[0], 2 * c[1], 2 * c[2]};
  d[a] += 1;
  if (dim == 2) {
    int t = 1 - a;
    for (int i = 0; i < 2; i++) {
      int e[3] = {d[0], d[1], 0};
      e[t] += i;
      out[i] = sep_index(a, e, ell + 1);
    }
    return 2;
  }
  // outer / inner tangential axes per normal (tree3d_common.cpp:140-154)
  int outer = a == 0 ? 1 : 0, inner = a == 2 ? 1 : 2;
  int k = 0;
  for (int i = 0; i < 2; i++)
    for (int j = 0; j < 2; j++) {
      int e[3] = {d[0], d[1], d[2]};
      e[outer] += i;
      e[inner] += j;
      out[k++] = sep_index(a, e, ell + 1);
    }
  return 4;
}
void Grid::child_seps(int box, int ell, std::vector<int>& self, std::vector<int>& shared) const {
  std::set<int> s0, s1;
  int ch[8], k = children(box, ell, ch);
  for (int i = 0; i < k; i++) {
    int ns[6], q = nbor_seps(ch[i], ell + 1, ns);
    for (int j = 0; j < q; j++) {
      if (ns[j] < 0) continue;
      int c[3];
      int a = sep_decode(ns[j], ell + 1, c);
      (c[a] % 2 == 0 ? s0 : s1).insert(ns[j]);
    }
  }
  self.assign(s0.begin(), s0.end());
  shared.assign(s1.begin(), s1.end());
}
int Grid::nominal(int axis, int ell) const {
  int d = L - 1 - ell;
  return bw[axis] * (1 << d) + ((1 << d) - 1) * w;
}
namespace {
void add_block(const Grid& g, std::set<int>& s, const int lo[3], const int hi[3]) {
  for (int k = lo[2]; k < hi[2]; k++)
    for (int j = lo[1]; j < hi[1]; j++)
      for (int i = lo[0]; i < hi[0]; i++) s.insert(g.dof(i, j, k));
}
}  // namespace
// Points where the children's separators cross: a w^d block in 2D; three w x w lines through
// the box in 3D (a line spans the box, reaching the grid end for the last box).
void Grid::box_center(int box, int ell, std::vector<int>& out) const {
  int c[3], b = side(ell);
  box_coords(box, ell, c);
  std::set<int> s;
  int cross[3], start[3], full[3];
  for (int a = 0; a < 3; a++) {
    if (a >= dim) { cross[a] = start[a] = 0; full[a] = 1; continue; }
    start[a] = c[a] * (nominal(a, ell) + w);
    cross[a] = start[a] + nominal(a, ell + 1);
    full[a] = c[a] == b - 1 ? N[a] : start[a] + nominal(a, ell);
  }
  if (dim == 2) {
    int lo[3] = {cross[0], cross[1], 0}, hi[3] = {cross[0] + w, cross[1] + w, 1};
    add_block(*this, s, lo, hi);
  } else {
    for (int t = 0; t < 3; t++) {
      int lo[3], hi[3];
      for (int a = 0; a < 3; a++) {
        lo[a] = a == t ? start[a] : cross[a];
        hi[a] = a == t ? full[a] : cross[a] + w;
      }
      add_block(*this, s, lo, hi);
    }
  }
  out.assign(s.begin(), s.end());
}
// Points of a parent separator between its children: in 2D the w x w block where it crosses
// the perpendicular child separator; in 3D the two lines along its tangential axes.
void Grid::sep_center(int sep, int ell, std::vector<int>& out) const {
  int c[3], b = side(ell);
  int an = sep_decode(sep, ell, c);
  std::set<int> s;
  int plane[3], cross[3], start[3], full[3];
  for (int a = 0; a < 3; a++) {
    if (a >= dim) { plane[a] = cross[a] = start[a] = 0; full[a] = 1; continue; }
    start[a] = c[a] * (nominal(a, ell) + w);
    plane[a] = start[a] + nominal(a, ell);
    cross[a] = start[a] + nominal(a, ell + 1);
    full[a] = c[a] == b - 1 ? N[a] : start[a] + nominal(a, ell);
  }
  if (dim == 2) {
    int t = 1 - an;
    int lo[3], hi[3];
    lo[an] = plane[an]; hi[an] = plane[an] + w;
    lo[t] = cross[t]; hi[t] = cross[t] + w;
    lo[2] = 0; hi[2] = 1;
    add_block(*this, s, lo, hi);
  } else {
    for (int t = 0; t < 3; t++) {
      if (t == an) continue;
      int lo[3], hi[3];
      for (int a = 0; a < 3; a++) {
        if (a == an) { lo[a] = plane[a]; hi[a] = plane[a] + w; }
        else if (a == t) { lo[a] = start[a]; hi[a] = full[a]; }
        else { lo[a] = cross[a]; hi[a] = cross[a] + w; }
      }
      add_block(*this, s, lo, hi);
    }
  }
  out.assign(s.begin(), s.end());
}
int Grid::leaf_of_dof(int d) const {
  int b = side(L - 1);
  int c[3] = {0, 0, 0};
  int r = d;
  c[0] = r % N[0]; r /= N[0];
  c[1] = r % N[1]; r /= N[1];
  c[2] = dim == 3 ? r : 0;
  for (int a = 0; a < dim; a++) {
    c[a] /= (bw[a] + w);
    if (c[a] >= b) return -1;
  }
  return box_index(c, L - 1);
}

// ------------------------------------------------------------------------------ dirty sets
namespace {
// Appends problem p's own dirty sets to the merged ones.
void map_dirty(const Forest& Fo, int p, const Plan::Dirty& d, Plan::Dirty& D) {
  const Plan& S = Fo.sub[p];
  for (int ell = 0; ell < S.L; ell++) {
    const int lam = ell + Fo.shift[p];
    for (int b : d.boxes[ell]) D.boxes[lam].push_back(Fo.box_base[p][lam] + b);
    for (int e : d.seps[ell]) D.seps[lam].push_back(Fo.sep_base[p][lam] + e);
  }
  D.any = D.any || d.any;
}
}  // namespace

Plan::Dirty Plan::all_dirty() const {
  Dirty D;
  if (forest) {
    D.boxes.assign(L, {});
    D.seps.assign(L, {});
    for (int p = 0; p < (int)forest->sub.size(); p++) map_dirty(*forest, p, forest->sub[p].all_dirty(), D);
    return D;
  }
  D.boxes.assign(g.L, {});
  D.seps.assign(g.L, {});
  for (int b = 0; b < g.nboxes(g.L - 1); b++) D.boxes[g.L - 1].push_back(b);
  // propagate
  for (int ell = g.L - 1; ell >= 1; ell--) {
    std::set<int> es, ps;
    for (int b : D.boxes[ell]) {
      int ns[6], q = g.nbor_seps(b, ell, ns);
      for (int j = 0; j < q; j++)
        if (ns[j] >= 0) es.insert(ns[j]);
    }
    for (int e : es) {
      int bb[2];
      g.sep_boxes(e, ell, bb);
      ps.insert(g.parent(bb[0], ell));
      ps.insert(g.parent(bb[1], ell));
    }
    D.seps[ell].assign(es.begin(), es.end());
    D.boxes[ell - 1].assign(ps.begin(), ps.end());
  }
  D.any = true;
  return D;
}

bool Plan::dirty_from_dofs(const std::vector<int>& mod, Dirty& D, std::string& err) const {
  if (forest) {  // split the DOFs by problem, each problem's own recurrence, merged
    const Forest& Fo = *forest;
    const int np = (int)Fo.sub.size();
    std::vector<std::vector<int>> per(np);
    for (int d : mod) {
      if (d < 0 || d >= n) { err = "modified DOF out of range"; return false; }
      const int p = (int)(std::upper_bound(Fo.dof0.begin(), Fo.dof0.end(), d) - Fo.dof0.begin()) - 1;
      per[p].push_back(d - Fo.dof0[p]);
    }
    D.boxes.assign(L, {});
    D.seps.assign(L, {});
    D.any = false;
    for (int p = 0; p < np; p++) {
      if (per[p].empty()) continue;
      Dirty d;
      if (!Fo.sub[p].dirty_from_dofs(per[p], d, err)) {
        err = "problem " + std::to_string(p) + ": " + err;
        return false;
      }
      map_dirty(Fo, p, d, D);
    }
    return true;
  }
  D.boxes.assign(g.L, {});
  D.seps.assign(g.L, {});
  std::set<int> leaves;
  for (int d : mod) {
    if (d < 0 || d >= n) { err = "modified DOF out of range"; return false; }
    int b = g.leaf_of_dof(d);
    if (b < 0) {
      err = "modified DOF " + std::to_string(d) +
            " maps outside the leaf box grid (uneven last box holds >= 2 extra points)";
      return false;
    }
    leaves.insert(b);
  }
  D.boxes[g.L - 1].assign(leaves.begin(), leaves.end());
  for (int ell = g.L - 1; ell >= 1; ell--) {
    std::set<int> es, ps;
    for (int b : D.boxes[ell]) {
      int ns[6], q = g.nbor_seps(b, ell, ns);
      for (int j = 0; j < q; j++)
        if (ns[j] >= 0) es.insert(ns[j]);
    }
    for (int e : es) {
      int bb[2];
      g.sep_boxes(e, ell, bb);
      ps.insert(g.parent(bb[0], ell));
      ps.insert(g.parent(bb[1], ell));
    }
    D.seps[ell].assign(es.begin(), es.end());
    D.boxes[ell - 1].assign(ps.begin(), ps.end());
  }
  D.any = !leaves.empty();
  return true;
}

// ------------------------------------------------------------------------------ build
namespace {

struct Builder {
  Plan& P;
  Grid& g;
  std::vector<std::vector<int>> box_int, box_bd;  // per front id
  std::vector<std::vector<int>> sep_dofs;         // per separator id
  std::vector<int> where;                          // scratch DOF -> position (or -1)

  Builder(Plan& p) : P(p), g(p.g) {}

  void leaf_lists() {
    int ell = g.L - 1, b = g.side(ell);
    for (int box = 0; box < g.nboxes(ell); box++) {
      int c[3];
      g.box_coords(box, ell, c);
      int lo[3], hi[3];
      for (int a = 0; a < 3; a++) {
        if (a >= g.dim) { lo[a] = 0; hi[a] = 1; continue; }
        lo[a] = c[a] * (g.bw[a] + g.w);
        hi[a] = c[a] == b - 1 ? g.N[a] : lo[a] + g.bw[a];
      }
      std::vector<int>& self = box_int[P.box_gid[ell][box]];
      for (int k = lo[2]; k < hi[2]; k++)
        for (int j = lo[1]; j < hi[1]; j++)
          for (int i = lo[0]; i < hi[0]; i++) self.push_back(g.dof(i, j, k));
      // each leaf separator is listed by one of its boxes: the x-low, y-high and z-low sides
      // (src/DE/mf2d.cpp:98-131, src/DE/mf3d.cpp:123-185)
      int ns[6];
      g.nbor_seps(box, ell, ns);
      auto fill = [&](int sep, int axis, int a0, int a1) {
        std::vector<int>& d = sep_dofs[P.sep_gid[ell][sep]];
        int l2[3] = {lo[0], lo[1], lo[2]}, h2[3] = {hi[0], hi[1], hi[2]};
        l2[axis] = a0; h2[axis] = a1;
        for (int k = l2[2]; k < h2[2]; k++)
          for (int j = l2[1]; j < h2[1]; j++)
            for (int i = l2[0]; i < h2[0]; i++) d.push_back(g.dof(i, j, k));
      };
      if (c[0] > 0) fill(ns[0], 0, lo[0] - g.w, lo[0]);
      if (c[1] < b - 1) fill(ns[3], 1, hi[1], hi[1] + g.w);
      if (g.dim == 3 && c[2] > 0) fill(ns[4], 2, lo[2] - g.w, lo[2]);
    }
  }

  void upper_sep_lists(int ell) {
    for (int e = 0; e < g.nseps(ell); e++) {
      std::vector<int>& d = sep_dofs[P.sep_gid[ell][e]];
      int ch[4], k = g.sep_children(e, ell, ch);
      for (int i = 0; i < k; i++) {
        const std::vector<int>& cd = sep_dofs[P.sep_gid[ell + 1][ch[i]]];
        d.insert(d.end(), cd.begin(), cd.end());
      }
      std::vector<int> cp;
      g.sep_center(e, ell, cp);
      d.insert(d.end(), cp.begin(), cp.end());
    }
  }

  void upper_box_int(int ell) {
    for (int box = 0; box < g.nboxes(ell); box++) {
      std::vector<int>& self = box_int[P.box_gid[ell][box]];
      std::vector<int> sc, sh;
      g.child_seps(box, ell, sc, sh);
      for (int e : sc) {
        const std::vector<int>& cd = sep_dofs[P.sep_gid[ell + 1][e]];
        self.insert(self.end(), cd.begin(), cd.end());
      }
      std::vector<int> cp;
      g.box_center(box, ell, cp);
      self.insert(self.end(), cp.begin(), cp.end());
    }
  }

  void box_bd_lists() {
    for (int f = 0; f < P.nfronts(); f++) {
      FrontInfo& F = P.fronts[f];
      int ns[6], q = g.nbor_seps(F.box, F.level, ns);
      F.nsep = 0;
      std::vector<int>& bd = box_bd[f];
      if (F.level == 0) continue;
      for (int j = 0; j < q; j++) {
        if (ns[j] < 0) continue;
        int sg = P.sep_gid[F.level][ns[j]];
        F.sep[F.nsep] = sg;
        F.sep_pos[F.nsep] = (int)bd.size();
        F.sep_len[F.nsep] = (int)sep_dofs[sg].size();
        F.nsep++;
        bd.insert(bd.end(), sep_dofs[sg].begin(), sep_dofs[sg].end());
      }
    }
  }

  // position of DOF d in front f: [int ; bd]
  void mark_front(int f, int val_or_clear) {
    const std::vector<int>& a = box_int[f];
    const std::vector<int>& b = box_bd[f];
    int n = (int)a.size();
    for (int i = 0; i < n; i++) where[a[i]] = val_or_clear ? i : -1;
    for (int i = 0; i < (int)b.size(); i++) where[b[i]] = val_or_clear ? n + i : -1;
  }

  // CSR triples of a front, lower triangle (full = false) or both (full = true, unsymmetric
  // operators): leaf fronts take every entry inside the front except the trailing diagonal
  // separator blocks (never read downstream); upper fronts take only the entries with a center
  // point in row or column (AddNewInteractions_ into A22, AddNewRowInteractions_ into A12 and
  // AddNewColInteractions_ into A21, src/DE/mf2d.cpp:369-401, :836-901).
  void front_triples(int f, std::vector<Triple>& out, bool full) {
    const FrontInfo& F = P.fronts[f];
    bool leaf = F.level == g.L - 1;
    mark_front(f, 1);
    int n = F.n;
    // separator run id of a trailing position (for the dead diagonal blocks)
    std::vector<int> run(F.m, -1);
    for (int s = 0; s < F.nsep; s++)
      for (int i = 0; i < F.sep_len[s]; i++) run[F.sep_pos[s] + i] = s;
    std::vector<char> center(n + F.m, 0);
    if (!leaf) {
      std::vector<int> cp;
      g.box_center(F.box, F.level, cp);
      // center points sit at the end of intDOFs
      for (int i = n - (int)cp.size(); i < n; i++) center[i] = 1;
      for (int s = 0; s < F.nsep; s++) {
        const SepInfo& S = P.seps[F.sep[s]];
        std::vector<int> ecp;
        g.sep_center(S.idx, S.level, ecp);
        for (int i = F.sep_len[s] - (int)ecp.size(); i < F.sep_len[s]; i++) center[n + F.sep_pos[s] + i] = 2;
      }
    }
    const std::vector<int>& a = box_int[f];
    const std::vector<int>& b = box_bd[f];
    for (int p = 0; p < n + F.m; p++) {
      int d = p < n ? a[p] : b[p - n];
      for (int k = P.rp[d]; k < P.rp[d + 1]; k++) {
        int q = where[P.ci[k]];
        if (q < 0 || (!full && q > p)) continue;
        if (leaf) {
          if (p >= n && q >= n && run[p - n] == run[q - n]) continue;
        } else {
          // A22 (both < n): a box-center endpoint; A12 (row >= n, col < n): an edge-center row;
          // A21 (row < n, col >= n): an edge-center column
          bool take = (p < n && q < n && (center[p] == 1 || center[q] == 1)) || (p >= n && q < n && center[p] == 2) ||
                      (p < n && q >= n && center[q] == 2);
          if (!take) continue;
        }
        out.push_back(Triple{p, q, k});
      }
    }
    mark_front(f, 0);
  }

  // separator A11 triples (lower, or both triangles when full): leaf = A_(DOFs, DOFs)
  // (src/DE/mf2d.cpp:595); upper = entries with a center point endpoint (AddNewInteractions_,
  // src/DE/mf2d.cpp:307)
  void sep_triples(int s, std::vector<Triple>& out, bool full) {
    const SepInfo& S = P.seps[s];
    const std::vector<int>& d = sep_dofs[s];
    bool leaf = S.level == g.L - 1;
    int ncp = 0;
    if (!leaf) {
      std::vector<int> cp;
      g.sep_center(S.idx, S.level, cp);
      ncp = (int)cp.size();
    }
    int k0 = S.k - ncp;  // first center position
    for (int i = 0; i < S.k; i++) where[d[i]] = i;
    for (int p = 0; p < S.k; p++)
      for (int k = P.rp[d[p]]; k < P.rp[d[p] + 1]; k++) {
        int q = where[P.ci[k]];
        if (q < 0 || (!full && q > p)) continue;
        if (!leaf && p < k0 && q < k0) continue;
        out.push_back(Triple{p, q, k});
      }
    for (int i = 0; i < S.k; i++) where[d[i]] = -1;
  }

  void add_op(std::vector<BlockOp>& ops, std::vector<Tile>& tiles, const BlockOp& op) {
    int id = (int)ops.size();
    ops.push_back(op);
    for (int r = 0; r < op.nr; r += 64)
      for (int c = 0; c < op.nc; c += 64) {
        if ((op.flags & 2) && c > r + 63) continue;  // fully above the diagonal
        tiles.push_back(Tile{id, r, c});
      }
  }
};

}  // namespace

// The assembly programs of every matrix (CSR triples and block-copy tiles), lower triangles only
// (Cholesky fronts of symmetric operators) or full matrices (LU fronts of unsymmetric ones), and
// for the former the fly / rest split of the leaf fronts.
void build_programs(Plan& P, bool full) {
  Builder B(P);
  Grid& g = P.g;
  const int n = P.n;
  B.box_int.assign(P.nfronts(), {});
  B.box_bd.assign(P.nfronts(), {});
  B.sep_dofs.assign(P.nseps_total(), {});
  for (int f = 0; f < P.nfronts(); f++) {
    const FrontInfo& F = P.fronts[f];
    B.box_int[f].assign(P.dofs.begin() + F.int_off, P.dofs.begin() + F.int_off + F.n);
    B.box_bd[f].assign(P.dofs.begin() + F.bd_off, P.dofs.begin() + F.bd_off + F.m);
  }
  for (int s = 0; s < P.nseps_total(); s++)
    B.sep_dofs[s].assign(P.dofs.begin() + P.seps[s].dof_off, P.dofs.begin() + P.seps[s].dof_off + P.seps[s].k);
  B.where.assign(n, -1);
  int NF = P.nfronts();
  P.triples.clear();
  P.ops.clear();
  P.tiles.clear();
  P.tri_off.assign(P.mats.size() + 1, 0);
  P.tile_off.assign(P.mats.size() + 1, 0);
  for (int mid = 0; mid < (int)P.mats.size(); mid++) {
    P.tri_off[mid] = (i64)P.triples.size();
    P.tile_off[mid] = (i64)P.tiles.size();
    if (mid < NF) {
      int f = mid;
      const FrontInfo& F = P.fronts[f];
      B.front_triples(f, P.triples, full);
      if (F.level < g.L - 1) {
        // position of every child separator (level+1) inside this front
        std::map<int, int> pos;
        std::vector<int> sc, sh;
        g.child_seps(F.box, F.level, sc, sh);
        int p = 0;
        for (int e : sc) {
          int sg = P.sep_gid[F.level + 1][e];
          pos[sg] = p;
          p += P.seps[sg].k;
        }
        for (int s = 0; s < F.nsep; s++) {
          const SepInfo& S = P.seps[F.sep[s]];
          int ch[4], k = g.sep_children(S.idx, S.level, ch);
          int q = F.n + F.sep_pos[s];
          for (int i = 0; i < k; i++) {
            int sg = P.sep_gid[S.level + 1][ch[i]];
            pos[sg] = q;
            q += P.seps[sg].k;
          }
        }
        // off-diagonal separator-pair blocks of each child's update matrix
        int ch[8], kc = g.children(F.box, F.level, ch);
        for (int i = 0; i < kc; i++) {
          int cf = P.box_gid[F.level + 1][ch[i]];
          const FrontInfo& C = P.fronts[cf];
          for (int u = 0; u < C.nsep; u++)
            for (int v = 0; v < C.nsep; v++) {
              if (u == v) continue;
              int Pu = pos.at(C.sep[u]), Pv = pos.at(C.sep[v]);
              if (!full && Pu < Pv) continue;  // upper triangle of this front
              BlockOp op{};
              op.src = cf;
              op.dr0 = Pu;
              op.dc0 = Pv;
              op.nr = C.sep_len[u];
              op.nc = C.sep_len[v];
              if (full || C.sep_pos[u] > C.sep_pos[v]) {  // full: every block as it is stored
                op.sr0 = C.n + C.sep_pos[u];
                op.sc0 = C.n + C.sep_pos[v];
                op.flags = 0;
              } else {  // stored transposed in the child's lower triangle
                op.sr0 = C.n + C.sep_pos[v];
                op.sc0 = C.n + C.sep_pos[u];
                op.flags = 1;
              }
              B.add_op(P.ops, P.tiles, op);
            }
        }
        // interior child separators: their A11 is the diagonal block of A22
        for (int e : sc) {
          int sg = P.sep_gid[F.level + 1][e];
          BlockOp op{};
          op.src = NF + sg;
          op.sr0 = op.sc0 = 0;
          op.dr0 = op.dc0 = pos.at(sg);
          op.nr = op.nc = P.seps[sg].k;
          op.flags = full ? 0 : 2;
          B.add_op(P.ops, P.tiles, op);
        }
      }
    } else {
      int s = mid - NF;
      const SepInfo& S = P.seps[s];
      B.sep_triples(s, P.triples, full);
      if (S.level < g.L - 1) {
        int ch[4], k = g.sep_children(S.idx, S.level, ch);
        int q = 0;
        for (int i = 0; i < k; i++) {
          int sg = P.sep_gid[S.level + 1][ch[i]];
          BlockOp op{};
          op.src = NF + sg;
          op.sr0 = op.sc0 = 0;
          op.dr0 = op.dc0 = q;
          op.nr = op.nc = P.seps[sg].k;
          op.flags = full ? 0 : 2;
          B.add_op(P.ops, P.tiles, op);
          q += P.seps[sg].k;
        }
      }
    }
  }
  P.tri_off[P.mats.size()] = (i64)P.triples.size();
  P.tile_off[P.mats.size()] = (i64)P.tiles.size();
  P.full = full;
  if (full) {  // no fly split: LU fronts are zeroed and scattered whole
    P.fly_off.assign(P.nfronts(), -1);
    P.fly_tab.clear();
    P.fly_tri.clear();
    P.rest_tri.clear();
    P.rest_off.clear();
  } else {
    build_fly(P);
  }
}

std::string build_plan(Plan& P, int dim, int nx, int ny, int nz, int width, int max_side, int n, const int* rp,
                       const int* ci, bool& geom_error, bool full) {
  geom_error = false;
  Grid& g = P.g;
  if (dim != 2 && dim != 3) return "dim must be 2 or 3";
  if (width < 1 || width > 3) return "separator width must be 1, 2 or 3";
  if (nx < 1 || ny < 1 || (dim == 3 && nz < 1)) return "grid dimensions must be positive";
  g.dim = dim;
  g.N[0] = nx; g.N[1] = ny; g.N[2] = dim == 3 ? nz : 1;
  g.w = width;
  long long nd = (long long)nx * ny * g.N[2];
  if (nd != n) return "n must equal the number of grid points";
  // tree depth (src/DE/mf2d.cpp:23-31)
  int top = std::max(nx, std::max(ny, dim == 3 ? nz : 0));
  g.L = (int)std::ceil(std::log2((double)(top + width) / (double)(max_side + width)));
  if (g.L <= 1) {
    geom_error = true;
    return "Based on the input parameters, the recursion depth is 1 level (i.e., do not recurse).  Try "
           "increasing the problem size or decreasing maxSideLength.";
  }
  if (g.L > 14) return "tree too deep";
  P.L = g.L;
  int b = g.side(g.L - 1);
  for (int a = 0; a < dim; a++) {
    g.bw[a] = (int)std::floor((double)(g.N[a] - (b - 1) * width) / (double)b);
    if (g.N[a] < g.bw[a] * (b - 1) + width * (b - 1)) {
      geom_error = true;
      return "You found a bug!  The math for computing the number of points per box at a leaf level broke on this "
             "input.";
    }
    if (g.bw[a] < 1) {
      geom_error = true;
      return "leaf boxes would be empty along an axis";
    }
  }
  P.n = n;
  P.rp.assign(rp, rp + n + 1);
  P.ci.assign(ci, ci + rp[n]);

  // ids: fronts level by level from the leaves up, separators likewise
  P.box_gid.assign(g.L, {});
  P.sep_gid.assign(g.L, {});
  for (int ell = g.L - 1; ell >= 0; ell--) {
    for (int bx = 0; bx < g.nboxes(ell); bx++) {
      P.box_gid[ell].push_back(P.nfronts());
      FrontInfo F{};
      F.level = ell;
      F.box = bx;
      P.fronts.push_back(F);
    }
    if (ell >= 1)
      for (int e = 0; e < g.nseps(ell); e++) {
        P.sep_gid[ell].push_back(P.nseps_total());
        SepInfo S{};
        S.level = ell;
        S.idx = e;
        P.seps.push_back(S);
      }
  }
  Builder B(P);
  B.box_int.assign(P.nfronts(), {});
  B.box_bd.assign(P.nfronts(), {});
  B.sep_dofs.assign(P.nseps_total(), {});
  B.leaf_lists();
  for (int ell = g.L - 2; ell >= 0; ell--) {
    if (ell >= 1) B.upper_sep_lists(ell);
    B.upper_box_int(ell);
  }
  B.box_bd_lists();

  // DOF arrays, sizes, separator neighbours
  P.dofs.clear();
  for (int f = 0; f < P.nfronts(); f++) {
    FrontInfo& F = P.fronts[f];
    F.n = (int)B.box_int[f].size();
    F.m = (int)B.box_bd[f].size();
    F.int_off = (i64)P.dofs.size();
    P.dofs.insert(P.dofs.end(), B.box_int[f].begin(), B.box_int[f].end());
    F.bd_off = (i64)P.dofs.size();
    P.dofs.insert(P.dofs.end(), B.box_bd[f].begin(), B.box_bd[f].end());
  }
  for (int s = 0; s < P.nseps_total(); s++) {
    SepInfo& S = P.seps[s];
    S.k = (int)B.sep_dofs[s].size();
    S.dof_off = (i64)P.dofs.size();
    P.dofs.insert(P.dofs.end(), B.sep_dofs[s].begin(), B.sep_dofs[s].end());
    int bb[2];
    g.sep_boxes(S.idx, S.level, bb);
    for (int i = 0; i < 2; i++) {
      int f = P.box_gid[S.level][bb[i]];
      S.box[i] = f;
      S.pos[i] = -1;
      for (int j = 0; j < P.fronts[f].nsep; j++)
        if (P.fronts[f].sep[j] == s) S.pos[i] = P.fronts[f].sep_pos[j];
    }
  }

  // matrices in the arena (256-byte aligned starts)
  P.mats.clear();
  i64 off = 0;
  auto place = [&](int rows, int cols) {
    MatDesc M{off, rows, cols};
    off += ((i64)rows * cols + 31) / 32 * 32;
    P.mats.push_back(M);
  };
  for (int f = 0; f < P.nfronts(); f++) place(P.fronts[f].n + P.fronts[f].m, P.fronts[f].n + P.fronts[f].m);
  for (int s = 0; s < P.nseps_total(); s++) place(P.seps[s].k, P.seps[s].k);
  for (int f = 0; f < P.nfronts(); f++) {
    P.fronts[f].dinv_off = off;
    off += (i64)((P.fronts[f].n + 63) / 64) * 64 * 64;
  }
  P.arena_elems = off;

  build_programs(P, full);
  return "";
}


// Split the entries of every front without block copies by where the left-looking
// factorization first reads them: the tile of column block cb = col/64 and rows
// [s0 + 64t, s0 + 64t + 64), s0 = the end of cb's diagonal block (row >= s0, col < n), or
// otherwise the diagonal block / trailing block, which stay zeroed-and-scattered.
void build_fly(Plan& P) {
  const int NF = P.nfronts();
  P.fly_off.assign(NF, -1);
  P.fly_tab.clear();
  P.fly_tri.clear();
  P.rest_tri.clear();
  P.rest_off.assign(P.mats.size() + 1, 0);
  for (int mid = 0; mid < (int)P.mats.size(); mid++) {
    P.rest_off[mid] = (i64)P.rest_tri.size();
    if (mid >= NF || P.tile_off[mid + 1] != P.tile_off[mid]) continue;
    const FrontInfo& F = P.fronts[mid];
    const int n = F.n, N = F.n + F.m, nbc = (n + 63) / 64, NT = (N + 63) / 64;
    if (n == 0) continue;
    std::vector<std::vector<Triple>> bins((size_t)nbc * NT);
    for (i64 t = P.tri_off[mid]; t < P.tri_off[mid + 1]; t++) {
      const Triple& tr = P.triples[t];
      const int cb = tr.col / 64;
      const int s0 = std::min(64 * (cb + 1), n);
      if (tr.col < n && tr.row >= s0)
        bins[(size_t)cb * NT + (tr.row - s0) / 64].push_back(tr);
      else
        P.rest_tri.push_back(tr);
    }
    P.fly_off[mid] = (int)P.fly_tab.size();
    for (int cb = 0; cb < nbc; cb++) {
      for (int t = 0; t < NT; t++) {
        P.fly_tab.push_back((int)P.fly_tri.size());
        for (const Triple& tr : bins[(size_t)cb * NT + t]) P.fly_tri.push_back(tr);
      }
      P.fly_tab.push_back((int)P.fly_tri.size());
    }
  }
  P.rest_off[P.mats.size()] = (i64)P.rest_tri.size();
}

// ------------------------------------------------------------------------------ batches
void merge_plans(Plan& M, std::vector<Plan>&& subs) {
  M = Plan();
  auto Fo = std::make_shared<Forest>();
  Fo->sub = std::move(subs);
  std::vector<Plan>& S = Fo->sub;
  const int np = (int)S.size();
  int Lmax = 0;
  for (const Plan& P : S) Lmax = std::max(Lmax, P.L);
  M.g = S[0].g;
  M.L = Lmax;
  int nf = 0, ns = 0, nd = 0;
  i64 nnz = 0, dsz = 0, ar = 0;
  int nops = 0;
  std::vector<i64> dofsz0(np), arena0(np);
  std::vector<int> ops0(np);
  Fo->shift.resize(np);
  Fo->front0.resize(np);
  Fo->sep0.resize(np);
  Fo->dof0.resize(np);
  Fo->nnz0.resize(np);
  for (int p = 0; p < np; p++) {
    Fo->shift[p] = Lmax - S[p].L;
    Fo->front0[p] = nf;
    Fo->sep0[p] = ns;
    Fo->dof0[p] = nd;
    Fo->nnz0[p] = nnz;
    dofsz0[p] = dsz;
    arena0[p] = ar;
    ops0[p] = nops;
    nf += S[p].nfronts();
    ns += S[p].nseps_total();
    nd += S[p].n;
    nnz += (i64)S[p].ci.size();
    dsz += (i64)S[p].dofs.size();
    ar += S[p].arena_elems;
    nops += (int)S[p].ops.size();
  }
  M.n = nd;
  M.arena_elems = ar;
  // block-diagonal CSR pattern
  M.rp.reserve(nd + 1);
  M.ci.reserve(nnz);
  for (int p = 0; p < np; p++) {
    for (int i = 0; i < S[p].n; i++) M.rp.push_back((int)(Fo->nnz0[p] + S[p].rp[i]));
    for (int c : S[p].ci) M.ci.push_back(c + Fo->dof0[p]);
  }
  M.rp.push_back((int)nnz);
  // level lists, aligned at the root depth
  M.box_gid.assign(Lmax, {});
  M.sep_gid.assign(Lmax, {});
  Fo->box_base.assign(np, std::vector<int>(Lmax, 0));
  Fo->sep_base.assign(np, std::vector<int>(Lmax, 0));
  for (int lam = 0; lam < Lmax; lam++)
    for (int p = 0; p < np; p++) {
      const int ell = lam - Fo->shift[p];
      Fo->box_base[p][lam] = (int)M.box_gid[lam].size();
      Fo->sep_base[p][lam] = (int)M.sep_gid[lam].size();
      if (ell >= 0)
        for (int f : S[p].box_gid[ell]) M.box_gid[lam].push_back(Fo->front0[p] + f);
      if (ell >= 1)
        for (int e : S[p].sep_gid[ell]) M.sep_gid[lam].push_back(Fo->sep0[p] + e);
    }
  // fronts, separators, DOF lists
  for (int p = 0; p < np; p++)
    for (FrontInfo F : S[p].fronts) {
      F.prob = p;
      F.int_off += dofsz0[p];
      F.bd_off += dofsz0[p];
      for (int q = 0; q < F.nsep; q++) F.sep[q] += Fo->sep0[p];
      F.dinv_off += arena0[p];
      M.fronts.push_back(F);
    }
  for (int p = 0; p < np; p++)
    for (SepInfo E : S[p].seps) {
      E.dof_off += dofsz0[p];
      E.box[0] += Fo->front0[p];
      E.box[1] += Fo->front0[p];
      M.seps.push_back(E);
    }
  M.dofs.reserve(dsz);
  for (int p = 0; p < np; p++)
    for (int d : S[p].dofs) M.dofs.push_back(d + Fo->dof0[p]);
  // matrices (all fronts, then all separators) and their assembly programs
  auto merged_mid = [&](int p, int mid) {
    const int NFp = S[p].nfronts();
    return mid < NFp ? Fo->front0[p] + mid : nf + Fo->sep0[p] + (mid - NFp);
  };
  std::vector<std::pair<int, int>> order;  // merged mid -> (problem, own mid)
  for (int p = 0; p < np; p++)
    for (int f = 0; f < S[p].nfronts(); f++) order.push_back({p, f});
  for (int p = 0; p < np; p++)
    for (int e = 0; e < S[p].nseps_total(); e++) order.push_back({p, S[p].nfronts() + e});
  for (int p = 0; p < np; p++)
    for (BlockOp op : S[p].ops) {
      op.src = merged_mid(p, op.src);
      M.ops.push_back(op);
    }
  M.tri_off.assign(order.size() + 1, 0);
  M.tile_off.assign(order.size() + 1, 0);
  for (size_t mid = 0; mid < order.size(); mid++) {
    const int p = order[mid].first, sm = order[mid].second;
    const Plan& P = S[p];
    MatDesc md = P.mats[sm];
    md.off += arena0[p];
    M.mats.push_back(md);
    M.tri_off[mid] = (i64)M.triples.size();
    for (i64 t = P.tri_off[sm]; t < P.tri_off[sm + 1]; t++) {
      Triple tr = P.triples[t];
      tr.k = (int)(tr.k + Fo->nnz0[p]);
      M.triples.push_back(tr);
    }
    M.tile_off[mid] = (i64)M.tiles.size();
    for (i64 t = P.tile_off[sm]; t < P.tile_off[sm + 1]; t++) {
      Tile tl = P.tiles[t];
      tl.op += ops0[p];
      M.tiles.push_back(tl);
    }
  }
  M.tri_off[order.size()] = (i64)M.triples.size();
  M.tile_off[order.size()] = (i64)M.tiles.size();
  build_fly(M);  // the merged leaf fronts' entries by panel tile (front-local rows, merged CSR)
  // the problems' own plans keep only what their dirty sets need (grid, sizes, level lists)
  for (Plan& P : S) {
    std::vector<int>().swap(P.rp);
    std::vector<int>().swap(P.ci);
    std::vector<int>().swap(P.dofs);
    std::vector<Triple>().swap(P.triples);
    std::vector<BlockOp>().swap(P.ops);
    std::vector<Tile>().swap(P.tiles);
    std::vector<MatDesc>().swap(P.mats);
  }
  M.forest = Fo;
}

}  // namespace umf
