"""CPU: the planner (csrc/plan.cpp, host C++ index arithmetic) under AddressSanitizer and
UndefinedBehaviorSanitizer.

A driver compiled with plan.cpp plans 2D and 3D grids of even and uneven sizes with stencil
widths 1-3 (the CSR pattern of the harness operator), checks that the fronts' interior DOF
lists partition the DOFs, computes dirty sets for random and out-of-range modified DOFs, and
merges plans into batches -- every access checked, any out-of-bounds read or write, overflow or
leak aborts the run.
"""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "..", "updating-multifrontal_amd", "csrc")

DRIVER = r"""
#include <cstdio>
#include <cstdlib>
#include <algorithm>
#include <random>
#include <vector>
#include "plan.h"
using namespace umf;

// CSR pattern of the width-w star stencil on an nx x ny (x nz) grid, DOF i + j nx + k nx ny
static void pattern(int dim, int nx, int ny, int nz, int w, std::vector<int>& rp, std::vector<int>& ci) {
  const int n = nx * ny * (dim == 3 ? nz : 1);
  rp.assign(1, 0);
  ci.clear();
  for (int d = 0; d < n; d++) {
    const int i = d % nx, j = (d / nx) % ny, k = dim == 3 ? d / (nx * ny) : 0;
    std::vector<int> cols;
    for (int s = -w; s <= w; s++) {
      if (s == 0) continue;
      if (i + s >= 0 && i + s < nx) cols.push_back(d + s);
      if (j + s >= 0 && j + s < ny) cols.push_back(d + s * nx);
      if (dim == 3 && k + s >= 0 && k + s < nz) cols.push_back(d + s * nx * ny);
    }
    cols.push_back(d);
    std::sort(cols.begin(), cols.end());
    ci.insert(ci.end(), cols.begin(), cols.end());
    rp.push_back((int)ci.size());
  }
}

int main() {
  struct G { int dim, nx, ny, nz, w, ms; };
  const G cases[] = {{2, 127, 127, 1, 1, 8}, {2, 63, 47, 1, 1, 4}, {2, 40, 33, 1, 2, 5}, {2, 61, 58, 1, 3, 6},
                     {2, 200, 97, 1, 1, 8}, {3, 15, 15, 15, 1, 3}, {3, 11, 14, 9, 1, 2}, {3, 23, 21, 19, 2, 4},
                     {3, 31, 31, 31, 1, 3}, {3, 20, 17, 25, 3, 5}};
  std::mt19937 rng(7);
  std::vector<Plan> subs;
  int planned = 0;
  for (const G& g : cases) {
    std::vector<int> rp, ci;
    pattern(g.dim, g.nx, g.ny, g.nz, g.w, rp, ci);
    const int n = (int)rp.size() - 1;
    Plan P;
    bool geom = false;
    std::string e = build_plan(P, g.dim, g.nx, g.ny, g.nz, g.w, g.ms, n, rp.data(), ci.data(), geom);
    if (!e.empty()) { std::printf("plan error: %s\n", e.c_str()); return 1; }
    planned++;
    std::vector<int> seen(n, 0);  // interior DOF lists partition 0..n-1
    for (const FrontInfo& F : P.fronts)
      for (int q = 0; q < F.n; q++) seen[P.dofs[F.int_off + q]]++;
    for (int d = 0; d < n; d++)
      if (seen[d] != 1) { std::printf("DOF %d in %d fronts\n", d, seen[d]); return 1; }
    for (int t = 0; t < 20; t++) {  // dirty sets of random modified DOFs
      std::vector<int> mod;
      const int cnt = 1 + (int)(rng() % 40);
      for (int q = 0; q < cnt; q++) mod.push_back((int)(rng() % n));
      std::sort(mod.begin(), mod.end());
      mod.erase(std::unique(mod.begin(), mod.end()), mod.end());
      Plan::Dirty D;
      std::string err;
      P.dirty_from_dofs(mod, D, err);  // may refuse DOFs the leaf map cannot place
    }
    Plan::Dirty D;
    std::string err;
    if (P.dirty_from_dofs({n}, D, err)) { std::printf("DOF %d out of range accepted\n", n); return 1; }
    if (P.dirty_from_dofs({-1}, D, err)) { std::printf("DOF -1 accepted\n"); return 1; }
    Plan::Dirty all = P.all_dirty();
    if (g.dim == 2 && g.w == 1) subs.push_back(std::move(P));
  }
  Plan M;
  merge_plans(M, std::move(subs));
  Plan::Dirty all = M.all_dirty();
  std::printf("planned %d, batch of %d fronts\n", planned, M.nfronts());
  return 0;
}
"""


@pytest.mark.timeout(600)
def test_planner_under_asan_ubsan(tmp_path):
    if shutil.which("g++") is None:
        pytest.skip("g++ not available")
    (tmp_path / "drv.cpp").write_text(DRIVER)
    exe = tmp_path / "drv"
    san = ["-fsanitize=address,undefined", "-fno-sanitize-recover=all", "-fno-omit-frame-pointer"]
    r = subprocess.run(["g++", "-std=c++17", "-O1", "-g", *san, "-I", CSRC, str(tmp_path / "drv.cpp"),
                        os.path.join(CSRC, "plan.cpp"), "-o", str(exe)], capture_output=True, text=True, timeout=300)
    if r.returncode != 0 and "asan" in (r.stderr or "").lower():
        pytest.skip("sanitizer runtime not available: " + r.stderr[-200:])
    assert r.returncode == 0, r.stderr[-2000:]
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=1:verify_asan_link_order=0",
               UBSAN_OPTIONS="print_stacktrace=1")
    run = subprocess.run([str(exe)], capture_output=True, text=True, timeout=300, env=env)
    assert run.returncode == 0, (run.stdout + run.stderr)[-3000:]
    assert "planned 10" in run.stdout
