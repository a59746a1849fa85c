#!/usr/bin/env python3
"""Benchmark: factor + local update + 16-RHS solve of a variable-coefficient elliptic operator,
in the FP32-MFMA + FP64-refinement form against the plain FP64 form.

Headline workload (the default): 3D 7-point -div(a grad u) on the unit cube, 127^3 interior grid
(n = 2,048,383), a = lognormal(0,1) iid per padded node (numpy PCG64, seed 3), h = 1/128,
nested dissection with maxSideLength 8 (15^3 leaf boxes, 4 levels, root front of 48,007).
One step = Factor() + Update() (a multiplied by one lognormal(0,2) draw on a random 8^3 box of
nodes, modified DOFs by the reference's modDOF rule) + Solve() of 16 N(0,1) right-hand sides.
All inputs (CSR values of both operators, right-hand sides) are resident in HBM before timing;
the update's modified-DOF list is host data, as in the reference API.  The two operators
alternate, so every step really refactors.

value = the FP32 form's algorithmic TFLOP/s over the timed steps: fronts stored and factored in
FP32 on FP32 MFMA, the solve refined in FP64 (preconditioned CG on the FP64 operator) until every
column's infinity-norm backward error ||b - Ax|| / (||A|| ||x|| + ||b||) is <= 1e-10.  The plain
FP64 form, the BF16-split form and the mixed form (FP64 top three separator levels over FP32
ones) run the same steps beside it (objects "fp64", "bf16x3_ir", "fp32t64_ir");
every form reports its backward error and its forward error against the FP64 form's solution.
TFLOP/s numerator (the same for every form and for the CPU baseline): F = sum over fronts of
n^3/3 + n^2 m + n m^2 (factor; dirty fronts for the update) + nrhs * sum (2 n^2 + 4 n m) (solve),
n / m = pivot / boundary size of each front.
"""
import argparse
import hashlib
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(ROOT, "updating-multifrontal_amd")
sys.path.insert(0, PKG)

METRIC = "factor+update+solve s & TFLOP/s; FP32-MFMA+IR vs FP64 at 1e-10 backward err"
# published MI355X dense peaks: FP64 / FP32 MFMA, BF16 MFMA; HBM3E
# (the mixed form's dominant GEMMs are its FP64 top levels': FP64 peak)
PEAK_TFLOPS = {"fp64": 78.6, "fp32": 157.3, "bf16x3": 2500.0, "fp32t64": 78.6}
PEAK_HBM_GBS = 8000.0
FORM_DTYPE = {
    "fp64": "f64",
    "fp32": "f32 fronts (FP32 MFMA) + f64 residual and refinement",
    "bf16x3": "f32 fronts, Schur/trailing GEMMs on BF16 MFMA over 2 bf16 pieces (3 products, f32 accumulate) "
              "+ f64 residual and refinement",
    "fp32t64": "f64 fronts on the top 3 separator levels, f32 fronts (FP32 MFMA) below + f64 residual and "
               "refinement",
}
ALL_FORMS = ("fp64", "fp32", "bf16x3", "fp32t64")
SOURCES = [os.path.join(PKG, "csrc", f) for f in ("kernels.hip", "runtime.hip", "plan.cpp")]


def source_hash():
    """First 16 hex digits of the SHA-256 of the library sources that decide what a kernel
    launch reads and writes."""
    h = hashlib.sha256()
    for f in SOURCES:
        with open(f, "rb") as fh:
            h.update(fh.read())
    return h.hexdigest()[:16]


def measured_traffic(kernel, workload, prec):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE
    summary of this workload and form (profiles/<kernel>_traffic_<workload>_<form>.md, calibrated
    on a plain copy), used only when the summary was measured on these exact sources (its
    'source hash' line); otherwise None.  Counters cannot be read inside the timed run."""
    import re
    path = os.path.join(ROOT, "profiles", f"{kernel}_traffic_{workload}_{prec}.md")
    try:
        with open(path) as f:
            txt = f.read()
    except OSError:
        return None
    m = re.search(r"^bytes per launch:\s*([0-9.eE+-]+)\s*$", txt, re.M)
    hs = re.search(r"^source hash:\s*([0-9a-f]+)\s*$", txt, re.M)
    if not m or not hs or hs.group(1) != source_hash():
        return None
    return float(m.group(1))


# workloads: dims, maxSideLength, the update (a box of `box` nodes per axis at `lo` (random when
# None), multiplied by `scale` (one LN(0,2) draw when None) or set to `setval`), coefficient
# field, right-hand sides, seed, the form whose rate is `value`, and the CPU baseline's sample
WORKLOADS = {
    "3d127": dict(dims=(127, 127, 127), max_side=8, box=8, scale=None, seed=3, value_form="fp32", cpu="3d63"),
    "3d95": dict(dims=(95, 95, 95), max_side=8, box=8, scale=None, seed=3, value_form="fp32", cpu="3d63"),
    "2d1023": dict(dims=(1023, 1023), max_side=16, box=32, scale=100.0, seed=2, value_form="fp32", cpu="2d511"),
    "2d127": dict(dims=(127, 127), max_side=8, box=8, lo=(60, 60), setval=10.0, coeff="uniform", rhs="uniform",
                  nrhs=1, seed=1, value_form="fp64", cpu="2d127"),
    "batch256": None,  # make_batch
    "serve95": None,   # run_serve
}
# bounded samples of the same generators for the reference's CPU run (about 10-30 s of CPU work)
CPU_SAMPLES = {
    "3d47": dict(dims=(47, 47, 47), max_side=8, box=8, scale=None, seed=3),
    "3d63": dict(dims=(63, 63, 63), max_side=8, box=8, scale=None, seed=3),
    "3d79": dict(dims=(79, 79, 79), max_side=8, box=8, scale=None, seed=3),
    "3d95": dict(dims=(95, 95, 95), max_side=8, box=8, scale=None, seed=3),
    "2d511": dict(dims=(511, 511), max_side=16, box=32, scale=100.0, seed=2),
    "2d127": WORKLOADS["2d127"],
}


def make_problem(dims, seed, box=32, scale=100.0, nrhs=16, coeff="lognormal", lo=None, setval=None, rhs="normal",
                 **_):
    """Nodal coefficients on the padded grid (lognormal(0,1) iid, or 1), and the updated field:
    a box of `box` nodes per axis at `lo` (random) multiplied by `scale` (one LN(0,2) draw when
    None) or set to `setval`.  Returns both operators' CSR values, the modified DOFs (the
    reference's modDOF rule, test_src/DE/testMF2D.cpp:143-150), right-hand sides (nrhs x n: row r
    = right-hand side r, i.e. column-major n x nrhs) and the box corner."""
    import umf
    d = len(dims)
    rng = np.random.default_rng(seed)
    shape = tuple(n + 2 for n in dims)
    co = np.exp(rng.standard_normal(shape)) if coeff == "lognormal" else np.ones(shape)
    co2 = co.copy()
    if lo is None:
        lo = [int(v) for v in rng.integers(0, [n + 2 - box for n in dims])]
    sl = tuple(slice(l, l + box) for l in lo)
    if setval is not None:
        co2[sl] = setval
    else:
        if scale is None:
            scale = float(np.exp(2.0 * rng.standard_normal()))
        co2[sl] *= scale
    h = 1.0 / (dims[0] + 1)
    rp, ci, v1 = umf.build_laplacian(dims, co, 1, h)
    _, _, v2 = umf.build_laplacian(dims, co2, 1, h)
    mod = np.unique(ci[v1 != v2]).astype(np.int32)
    n = int(np.prod(dims))
    B = rng.standard_normal((nrhs, n)) if rhs == "normal" else rng.uniform(-1.0, 1.0, (nrhs, n))
    return rp, ci, v1, v2, mod, B, tuple(lo)


def describe(spec, nrhs, levels=None, leaf=None):
    dims = spec["dims"]
    d = len(dims)
    upd = (f"a = {spec['setval']:g}" if spec.get("setval") is not None else
           f"a x {spec['scale']:g}" if spec.get("scale") is not None else "a x LN(0,2)")
    geo = f" ({levels} levels, {leaf}^{d} leaves)" if levels else ""
    return (f"{d}D {'5' if d == 2 else '7'}-point -div(a grad u), {'x'.join(map(str, dims))} grid "
            f"(n={int(np.prod(dims))}), a~{'lognormal(0,1)' if spec.get('coeff', 'lognormal') == 'lognormal' else '1'}"
            f" per node, maxSideLength {spec['max_side']}{geo}; Factor + Update ({upd} on a {spec['box']}^{d} node box) "
            f"+ {nrhs}-RHS Solve")


def make_batch(seed=4, nprob=256, nrhs=4, box=16, scale=100.0):
    """The batched workload: nprob 2D problems, N ~ U{63..511}, a ~ lognormal(0,1), maxSideLength 8,
    planned as one batch; the update multiplies a random box of one random problem by `scale`.
    Returns merged CSR arrays, both value arrays, the merged modified DOFs, right-hand sides."""
    import umf
    rng = np.random.default_rng(seed)
    Ns = rng.integers(63, 512, size=nprob)
    q_upd = int(rng.integers(0, nprob))
    problems, v2s = [], []
    mod = None
    for q, N in enumerate(Ns):
        N = int(N)
        co = np.exp(rng.standard_normal((N + 2, N + 2)))
        h = 1.0 / (N + 1)
        rp, ci, v = umf.build_laplacian((N, N), co, 1, h)
        v2 = v
        if q == q_upd:
            i0, j0 = (int(x) for x in rng.integers(0, N + 2 - box, size=2))
            co[i0:i0 + box, j0:j0 + box] *= scale
            _, _, v2 = umf.build_laplacian((N, N), co, 1, h)
            mod = np.unique(ci[v != v2]).astype(np.int64)
        problems.append((Csr(rp, ci, v), (N, N), 1, 8))
        v2s.append(v2)
    plan = umf.Batch(problems)
    mod = (mod + plan.offsets[q_upd]).astype(np.int32)
    v1 = plan._val.copy()
    v2 = np.concatenate(v2s)
    B = rng.standard_normal((nrhs, plan.n))
    desc = (f"batch of {nprob} 2D 5-point -div(a grad u) problems, N ~ U{{63..511}} (n={plan.n} in total), "
            f"a~lognormal(0,1), maxSideLength 8, one plan ({plan.levels} root-aligned levels); "
            f"Factor + Update (a x{scale:g} on a {box}^2 node box of problem {q_upd}) + {nrhs}-RHS Solve, FP64")
    return plan, plan._rp, plan._ci, v1, v2, mod, B, desc


def progress(msg):
    """One line per finished step on stderr (long 3D runs stay visibly alive)."""
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


class Csr:
    def __init__(self, rp, ci, val):
        self.rp, self.ci, self.val = rp, ci, val


def algorithmic_flops(plan, nrhs):
    fl_solve = 0.0
    for ell in range(plan.levels):
        for b in range(plan.nboxes(ell)):
            n = len(plan.box_dofs(ell, b, 0))
            m = len(plan.box_dofs(ell, b, 1))
            fl_solve += 2.0 * n * n + 4.0 * n * m
    return fl_solve * nrhs


def cpu_threads():
    """Host cores this process may use (the GPU job's CPU share), at most 16."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(n, 16))


def host_info():
    info = {"cpu_threads_available": cpu_threads()}
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    info["cpu_model"] = line.split(":", 1)[1].strip()
                    break
        with open("/proc/meminfo") as f:
            for line in f:
                if line.startswith("MemTotal"):
                    info["host_ram_gb"] = round(int(line.split()[1]) / 1e6, 1)
                    break
    except OSError:
        pass
    return info


def cpu_sample(name, nrhs):
    """The reference itself (oracle/_ref, built from its sources) on the bounded sample `name`:
    Factor + Update + nrhs single-vector Solve calls, as the reference API requires; the
    algorithmic numerator is the one the GPU line uses."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_py
    import umf
    kind = "reference" if oracle_py.have_ref() else "port"
    spec = CPU_SAMPLES[name]
    dims = spec["dims"]
    nrhs = spec.get("nrhs", nrhs)
    rp, ci, v1, v2, mod, B, _ = make_problem(**dict(spec, nrhs=nrhs))
    P = (umf.MF2D if len(dims) == 2 else umf.MF3D)(Csr(rp, ci, v1), *dims, 1, spec["max_side"], analyze_only=True)
    f_factor = P.stats()["flops_factor"]
    P.analyze_update(mod)
    f_update = P.stats()["flops_update"]
    f_solve = algorithmic_flops(P, nrhs)
    F = oracle_py.Factorization(oracle_py.Csr(rp, ci, v1), dims, 1, spec["max_side"],
                                "ref" if kind == "reference" else "orc")
    t0 = time.perf_counter()
    F.factor(trace=False)
    t1 = time.perf_counter()
    progress(f"reference CPU {name}: Factor {t1 - t0:.1f} s")
    F.update(oracle_py.Csr(rp, ci, v2), mod, trace=False)
    t2 = time.perf_counter()
    progress(f"reference CPU {name}: Update {t2 - t1:.1f} s")
    for r in range(nrhs):
        F.solve(B[r])
    t3 = time.perf_counter()
    progress(f"reference CPU {name}: {nrhs} Solve calls {t3 - t2:.1f} s")
    tot = t3 - t0
    return {"value": (f_factor + f_update + f_solve) / tot / 1e12, "unit": "TFLOP/s", "cores": 1, "kind": kind,
            "sample": describe(spec, nrhs, P.levels, P.leaf_width(0)) + "; the reference's LU form, MKL sequential",
            "t_factor_s": t1 - t0, "t_update_s": t2 - t1, "t_solve_s": t3 - t2, "seconds": tot,
            "flops": f_factor + f_update + f_solve}


def cpu_sample_threaded(name, nrhs, threads=None):
    """The sample on the reference built against threaded MKL (oracle/_ref/libumfref_mt.so) on
    all of this job's host cores, in a child process started before this one touches the GPU
    (the threaded and sequential MKL layers cannot share a process).  None when that build is
    absent or the child fails."""
    import subprocess
    lib = os.path.join(ROOT, "oracle", "_ref", "libumfref_mt.so")
    if not os.path.exists(lib):
        return None
    t = threads or cpu_threads()
    env = dict(os.environ, UMF_REF_SO=lib, MKL_NUM_THREADS=str(t), OMP_NUM_THREADS=str(t))
    try:
        # the child's progress lines (stderr) pass through; its result line is on stdout
        r = subprocess.run([sys.executable, os.path.abspath(__file__), "--cpu-sample", name, "--nrhs", str(nrhs)],
                           env=env, stdout=subprocess.PIPE, stderr=None, text=True, timeout=1800)
    except subprocess.TimeoutExpired:
        return None
    if r.returncode != 0 or not r.stdout.strip():
        progress(f"threaded CPU baseline failed (exit {r.returncode})")
        return None
    out = json.loads(r.stdout.strip().splitlines()[-1])
    out["cores"] = t
    out["sample"] = out["sample"].replace("MKL sequential", f"threaded MKL ({t} threads)")
    return out


def fold(prof_lv):
    """Per kernel, summed over tree levels (names carry '@L<level>')."""
    prof = {}
    for k, v in prof_lv.items():
        d = prof.setdefault(k.split("@")[0], {"ms": 0.0, "launches": 0, "flops": 0.0, "bytes": 0.0})
        for f in d:
            d[f] += v[f]
    return prof


class Run:
    """One plan and its inputs on the device; step() = Factor + Update + Solve."""

    def __init__(self, torch, plan, v1, v2, mod, B, nrhs, rp):
        self.torch, self.plan, self.mod, self.nrhs = torch, plan, mod, nrhs
        self.d_v = [torch.from_numpy(v1).cuda(), torch.from_numpy(v2).cuda()]
        self.d_B = torch.from_numpy(B).cuda()
        self.d_X = torch.empty_like(self.d_B)
        self.d_R = torch.empty_like(self.d_B)
        rows = np.repeat(np.arange(len(rp) - 1), np.diff(rp))
        self.norm_inf = []
        for v in (v1, v2):  # ||A||_inf of each operator
            rs = np.zeros(len(rp) - 1)
            np.add.at(rs, rows, np.abs(v))
            self.norm_inf.append(float(rs.max()))
        self.cur = 0  # operator the plan holds
        self.X64 = [None, None]

    def step(self):
        self.plan.Factor()
        nxt = 1 - self.cur
        self.plan.update_device(self.d_v[nxt].data_ptr(), self.mod)
        self.cur = nxt
        self.plan.solve_device(self.d_B.data_ptr(), self.d_X.data_ptr(), self.nrhs)

    def errors(self):
        """Infinity-norm backward error of the last solve (FP64 residual by the device SpMV) and
        its forward error against the FP64 form's solution of the same operator, max over
        columns."""
        torch = self.torch
        self.plan.spmv_device(self.d_X.data_ptr(), self.d_R.data_ptr(), self.nrhs)
        R = self.d_B - self.d_R
        den = self.norm_inf[self.cur] * self.d_X.abs().amax(dim=1) + self.d_B.abs().amax(dim=1)
        eta = float((R.abs().amax(dim=1) / den).max())
        fwd = None
        ref = self.X64[self.cur]
        if ref is not None:
            fwd = float((torch.linalg.norm(self.d_X - ref, dim=1) / torch.linalg.norm(ref, dim=1)).max())
        return eta, fwd


def run_form(run, prec, steps, warmup, args, keep_x64=False):
    """Time `steps` steps of one precision form after `warmup` untimed ones and one untimed
    profiling step (every kernel bracketed by events).  In the timed steps only the form's
    dominant kernel is bracketed by events (umf_profile_select), which gives its roofline."""
    torch, plan = run.torch, run.plan
    plan.SetPrecision(prec, tol=args.ir_tol, max_iter=100)
    plan.profile(False)
    # --graph: Factor() replays its captured HIP graph in the warmup and timed steps (captured in
    # the first warmup step); events cannot bracket launches inside a replayed graph, so the
    # roofline object then comes from the untimed profiling step's events
    graph = bool(getattr(args, "graph", False))
    plan.use_graph(graph)
    for i in range(warmup):
        run.step()
        progress(f"{prec}: warmup step {i} done")
    prof_lv = {}
    if not args.no_profile:
        plan.profile(True)
        plan.profile_select(None)
        plan.profile_reset()
        run.step()
        torch.cuda.synchronize()
        prof_lv = plan.profile_read()
        plan.profile_reset()
        plan.profile(False)
    prof = fold(prof_lv)
    dom = max(prof, key=lambda k: prof[k]["ms"]) if prof else None
    if dom and not graph:
        plan.profile(True)
        plan.profile_select(dom)
    tf = tu = ts = flops = 0.0
    iters = []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(steps):
        run.step()
        st = plan.stats()
        progress(f"{prec} step {i}: factor {st['t_factor_ms']:.1f} ms, update {st['t_update_ms']:.1f} ms, "
                 f"solve {st['t_solve_ms']:.1f} ms")
        tf += st["t_factor_ms"]
        tu += st["t_update_ms"]
        ts += st["t_solve_ms"]
        iters.append(st["ir_iters"])
        flops += st["flops_factor"] + st["flops_update"] + st["flops_solve"]
    torch.cuda.synchronize()
    secs = time.perf_counter() - t0
    plan.profile(False)
    plan.use_graph(False)
    st = plan.stats()
    roof = None
    if dom:
        if graph:
            d = prof[dom]
        else:
            d = fold(plan.profile_read()).get(dom, {"ms": 0.0, "launches": 0, "flops": 0.0, "bytes": 0.0})
        plan.profile_reset()
        launches = max(d["launches"], 1)
        # the bound is the ceiling the kernel's algorithmic intensity meets first: MFMA above the
        # ridge point (peak flop/s over 8 TB/s), HBM below it (the sweeps, small-front panels)
        peak = PEAK_TFLOPS[prec]
        ridge = peak * 1e12 / (PEAK_HBM_GBS * 1e9)
        if d["flops"] > 0 and d["ms"] > 0 and d["flops"] >= ridge * d["bytes"]:
            ach = d["flops"] / (d["ms"] * 1e-3) / 1e12
            roof = {"bound": "mfma", "achieved": ach, "peak": peak, "unit": "TFLOP/s", "frac": ach / peak}
        else:
            ach = d["bytes"] / (max(d["ms"], 1e-9) * 1e-3) / 1e9
            roof = {"bound": "hbm", "achieved": ach, "peak": PEAK_HBM_GBS, "unit": "GB/s", "frac": ach / PEAK_HBM_GBS}
        roof.update({"traffic": measured_traffic(dom, args.workload, prec), "kernel": dom,
                     "launches": d["launches"], "avg_launch_ms": d["ms"] / launches,
                     "timed_in": "the profiling step (Factor via the graph in the timed steps)" if graph else
                                 "the timed steps",
                     "algorithmic_flops_per_launch": d["flops"] / launches,
                     "algorithmic_bytes_per_launch": d["bytes"] / launches,
                     "share_of_kernel_time": d["ms"] / (1 if graph else max(steps, 1)) /
                                             max(sum(v["ms"] for v in prof.values()), 1e-9)})
    eta, fwd = run.errors()
    if keep_x64:  # the FP64 solutions of both operators, for the other forms' forward errors
        run.X64[run.cur] = run.d_X.clone()
        run.step()
        run.X64[run.cur] = run.d_X.clone()
    progress(f"{prec}: {secs / steps * 1e3:.1f} ms per step, iterations {iters}, eta_inf {eta:.2e}, fwd {fwd}")
    return {"value": flops / secs / 1e12, "unit": "TFLOP/s", "steps": steps, "warmup": warmup,
            "ms_per_step": secs / steps * 1e3, "t_factor_ms": tf / steps, "t_update_ms": tu / steps,
            "t_solve_ms": ts / steps, "flops_per_step": flops / steps, "ir_iters": iters if prec != "fp64" else None,
            "ir_tol": args.ir_tol if prec != "fp64" else None, "backward_error_inf": eta,
            "forward_error_vs_fp64": fwd, "front_pool_gb": st["bytes_front_pool"] / 1e9, "dtype": FORM_DTYPE[prec],
            "roofline": roof,
            "kernels": {k: {"ms_per_step": v["ms"], "launches_per_step": v["launches"],
                            "tflops": (v["flops"] / (v["ms"] * 1e-3) / 1e12) if v["ms"] > 0 and v["flops"] else None,
                            "gbs": (v["bytes"] / (v["ms"] * 1e-3) / 1e9) if v["ms"] > 0 and v["bytes"] else None}
                        for k, v in sorted(prof.items())},
            "kernels_by_level": {k: round(v["ms"], 4) for k, v in sorted(prof_lv.items())}}


def gemm_probe(reps=10):
    """The vendor library's GEMM (torch.matmul: hipBLASLt / rocBLAS) on this GPU, FP32 without TF32-like
    reduced precision and FP64, at m = n = 16384, k = 1024 (one outer group of the 3D trailing update)
    and at 8192^3: mean TFLOP/s of `reps` calls after 3 untimed ones.  A reference point for the
    attainable MFMA rate; no part of the factorization uses it."""
    import torch
    torch.backends.cuda.matmul.allow_tf32 = False
    out = {}
    for dt in (torch.float32, torch.float64):
        for (m, n, k) in ((16384, 16384, 1024), (8192, 8192, 8192)):
            a = torch.randn(m, k, device="cuda", dtype=dt)
            b = torch.randn(k, n, device="cuda", dtype=dt)
            c = torch.empty(m, n, device="cuda", dtype=dt)
            for _ in range(3):
                torch.matmul(a, b, out=c)
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            for _ in range(reps):
                torch.matmul(a, b, out=c)
            e1.record()
            torch.cuda.synchronize()
            ms = e0.elapsed_time(e1) / reps
            out[f"{str(dt).split('.')[-1]} {m}x{n}x{k}"] = round(2.0 * m * n * k / (ms * 1e-3) / 1e12, 1)
            del a, b, c
    print(json.dumps({"gemm_probe_tflops": out}), flush=True)


def factor_graph_timing(plan, torch, reps=2):
    """Factor() alone, direct launches against the captured HIP graph (umf_set_graph): mean ms
    of `reps` calls each after one untimed call (the graph's capture)."""
    plan.profile(False)
    res = {}
    for mode in (False, True):
        plan.use_graph(mode)
        plan.Factor()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            plan.Factor()
        torch.cuda.synchronize()
        res["graph_ms" if mode else "direct_ms"] = (time.perf_counter() - t0) / reps * 1e3
    plan.use_graph(False)
    progress(f"Factor alone: direct {res['direct_ms']:.2f} ms, graph {res['graph_ms']:.2f} ms")
    return res


def build_run(torch, name, nrhs=None):
    import umf
    spec = WORKLOADS[name]
    dims = spec["dims"]
    nrhs = nrhs or spec.get("nrhs", 16)
    t0 = time.perf_counter()
    rp, ci, v1, v2, mod, B, corner = make_problem(**dict(spec, nrhs=nrhs))
    plan = (umf.MF2D if len(dims) == 2 else umf.MF3D)(Csr(rp, ci, v1), *dims, 1, spec["max_side"])
    run = Run(torch, plan, v1, v2, mod, B, nrhs, rp)
    desc = describe(spec, nrhs, plan.levels, plan.leaf_width(0)) + f" (update box at {corner})"
    progress(f"workload {name}: n={int(np.prod(dims))}, {plan.levels} levels, front pool "
             f"{plan.stats()['bytes_front_pool'] / 1e9:.1f} GB, set up in {time.perf_counter() - t0:.1f} s")
    return run, desc, nrhs


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5, help="timed steps of the value form")
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--side-steps", type=int, default=8,
                    help="timed steps of the forms beside the value form (at most --steps)")
    ap.add_argument("--workload", choices=sorted(WORKLOADS), default="3d127")
    ap.add_argument("--nrhs", type=int, default=None, help="right-hand sides (16; 4 for the batch)")
    ap.add_argument("--forms", nargs="+", choices=sorted(FORM_DTYPE), default=list(ALL_FORMS),
                    help="precision forms to time (the value form is always timed)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-secondary", action="store_true", help="skip the 2D secondary workloads")
    ap.add_argument("--no-profile", action="store_true", help="skip per-kernel event timing")
    ap.add_argument("--no-graph", action="store_true", help="skip the Factor() direct-vs-graph comparison")
    ap.add_argument("--ir-tol", type=float, default=1e-10, help="backward error the refined forms refine to")
    ap.add_argument("--cpu-sample", default=None, help=argparse.SUPPRESS)  # child of the CPU baseline
    ap.add_argument("--cpu-only", default=None, choices=sorted(CPU_SAMPLES),
                    help="time only the reference on the host (threaded MKL) on this sample and exit")
    ap.add_argument("--gemm-probe", action="store_true",
                    help="time the vendor library's FP32 and FP64 GEMM (torch.matmul) at the trailing update's "
                         "shape on this GPU, as the attainable MFMA rate beside the published peaks, and exit")
    ap.add_argument("--copy-probe", action="store_true",
                    help="copy 1 GiB once on the GPU and exit (the HBM counters' calibration run)")
    ap.add_argument("--graph", action="store_true",
                    help="Factor() as the plan's captured HIP graph in the warmup and timed steps (the roofline "
                         "object then comes from the untimed profiling step)")
    ap.add_argument("--autotune", action="store_true",
                    help="run the plan's autotuner (GEMM tile size, outer group width) before timing")
    ap.add_argument("--rates", type=float, nargs="+", default=[2.0, 8.0, 32.0],
                    help="serve95: request arrival rates (per second)")
    ap.add_argument("--requests", type=int, default=24, help="serve95: requests per rate")
    ap.add_argument("--slots", type=int, default=4, help="serve95: request slots (concurrent requests)")
    ap.add_argument("--pool-gb", type=float, default=120.0,
                    help="serve95: slot pool for the requests' dirty fronts, FP64 (half for the FP32 forms)")
    args = ap.parse_args()
    if args.cpu_sample:  # the threaded CPU baseline's child process: no GPU work
        print(json.dumps(cpu_sample(args.cpu_sample, args.nrhs or 16)), flush=True)
        return
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU only; --gpus {args.gpus} is not supported "
              "(the sharded path is design-only)", file=sys.stderr)
        sys.exit(2)
    if args.cpu_only:
        out = cpu_sample_threaded(args.cpu_only, args.nrhs or 16)
        print(json.dumps({"cpu_only": out, "host": host_info()}), flush=True)
        return
    if args.gemm_probe:
        return gemm_probe()
    if args.copy_probe:
        import torch
        a = torch.empty(1 << 27, dtype=torch.float64, device="cuda")  # 1 GiB
        a.fill_(1.0)
        torch.cuda.synchronize()
        b = a.clone()
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes": a.numel() * 8, "sum": float(b[:16].sum())}), flush=True)
        return
    if args.workload == "serve95":
        return run_serve(args)
    if args.workload == "batch256":
        return run_batch(args)

    spec = WORKLOADS[args.workload]
    value_form = spec["value_form"]
    forms = [f for f in ALL_FORMS if f in args.forms or f == value_form]
    cpu_mt = None
    if not args.no_cpu_baseline:
        progress(f"CPU baseline: reference on sample {spec['cpu']} ({cpu_threads()} threads)")
        cpu_mt = cpu_sample_threaded(spec["cpu"], args.nrhs or spec.get("nrhs", 16))
    import torch
    torch.cuda.set_device(0)
    run, wdesc, nrhs = build_run(torch, args.workload, args.nrhs)
    plan = run.plan
    tuned = None
    if args.autotune:
        plan.SetPrecision(value_form, tol=args.ir_tol, max_iter=100)
        plan.Factor()
        tuned = plan.autotune(reps=2)
        progress(f"autotune: {tuned}")
    side = max(1, min(args.side_steps, args.steps))
    res = {}
    # FP64 first: its solutions of both operators are the other forms' forward-error reference
    for f in forms:
        is_value = f == value_form
        res[f] = run_form(run, f, args.steps if is_value else side, args.warmup if is_value else 1, args,
                          keep_x64=(f == "fp64"))
    main_ = res[value_form]
    plan.SetPrecision(value_form, tol=args.ir_tol, max_iter=100)
    fgraph = None if args.no_graph else factor_graph_timing(plan, torch)
    out = {
        "metric": METRIC,
        "value": main_["value"],
        "unit": "TFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": main_["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64-refinement" if value_form != "fp64" else "f64",
        "data": "synthetic",
        "config": {"workload": wdesc, "nrhs": nrhs, "precision": value_form, "ir_tol": args.ir_tol,
                   "tune": plan.tune(), "autotuned": tuned is not None},
        "t_factor_ms": main_["t_factor_ms"],
        "t_update_ms": main_["t_update_ms"],
        "t_solve_ms": main_["t_solve_ms"],
        "flops_per_step": main_["flops_per_step"],
        "backward_error_inf": main_["backward_error_inf"],
        "forward_error_vs_fp64": main_["forward_error_vs_fp64"],
        "roofline": main_["roofline"],
        "cpu_baseline": None,
        "factor_graph": fgraph,
    }
    for f in forms:
        if f == value_form:
            out[f] = {k: main_[k] for k in ("kernels", "kernels_by_level", "ir_iters", "front_pool_gb")}
            continue
        key = f if f == "fp64" else f + "_ir"
        out[key] = res[f]
    if "fp64" in res:  # wall time of the FP64 form over each form's (same work)
        out["speedup_vs_fp64"] = {f: res["fp64"]["ms_per_step"] / res[f]["ms_per_step"] for f in res}
    out["host"] = host_info()
    if cpu_mt is not None:
        out["cpu_baseline"] = cpu_mt
        out["gpu_over_cpu"] = out["value"] / cpu_mt["value"]
    if args.workload == "3d127" and not args.no_secondary:
        del run, plan
        torch.cuda.empty_cache()
        out["secondary"] = secondary(torch, args)
    print(json.dumps(out), flush=True)


def secondary(torch, args):
    """The 2D configurations beside the 3D headline: 2D 1023^2 (FP64 and FP32 + refinement) and
    2D 127^2 (the plain FP64 form against the reference's own CPU path, one core, MKL
    sequential, at full size)."""
    out = {}
    run, desc, nrhs = build_run(torch, "2d1023")
    a = argparse.Namespace(**vars(args))
    a.workload = "2d1023"
    r64 = run_form(run, "fp64", 10, 2, a, keep_x64=True)
    r32 = run_form(run, "fp32", 10, 2, a)
    out["2d1023"] = {"workload": desc, "fp64": {k: r64[k] for k in ("value", "ms_per_step", "backward_error_inf")},
                     "fp32_ir": {k: r32[k] for k in ("value", "ms_per_step", "backward_error_inf",
                                                      "forward_error_vs_fp64", "ir_iters")},
                     "speedup_fp32_vs_fp64": r64["ms_per_step"] / r32["ms_per_step"]}
    del run
    run, desc, nrhs = build_run(torch, "2d127")
    a.workload = "2d127"
    r64 = run_form(run, "fp64", 20, 3, a)
    ent = {"workload": desc, "fp64_gpu": {k: r64[k] for k in ("value", "ms_per_step", "t_factor_ms", "t_update_ms",
                                                                "t_solve_ms", "backward_error_inf")}}
    del run
    if not args.no_cpu_baseline:
        cpu1 = cpu_sample("2d127", nrhs)
        ent["reference_cpu_1core"] = cpu1
        ent["gpu_over_cpu_1core"] = r64["value"] / cpu1["value"]
    out["2d127"] = ent
    return out


def run_batch(args):
    """The batched workload (make_batch): every form on one plan of 256 problems."""
    import torch
    torch.cuda.set_device(0)
    nrhs = args.nrhs or 4
    plan, rp, ci, v1, v2, mod, B, wdesc = make_batch(nrhs=nrhs)
    run = Run(torch, plan, v1, v2, mod, B, nrhs, rp)
    tuned = None
    if args.autotune:
        tuned = plan.autotune(reps=2)
    res = {}
    for f in ALL_FORMS:
        if f in args.forms:
            res[f] = run_form(run, f, args.steps, args.warmup, args, keep_x64=(f == "fp64"))
    main_ = res.get("fp32") or next(iter(res.values()))
    out = {"metric": METRIC, "value": main_["value"], "unit": "TFLOP/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": main_["ms_per_step"], "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f32+f64-refinement", "data": "synthetic",
           "config": {"workload": wdesc, "nrhs": nrhs, "tune": plan.tune(), "autotuned": tuned is not None},
           "forms": res, "cpu_baseline": None}
    if "fp64" in res:
        out["speedup_vs_fp64"] = {f: res["fp64"]["ms_per_step"] / res[f]["ms_per_step"] for f in res}
    print(json.dumps(out), flush=True)


def run_serve(args):
    """Serving on a resident 3D 95^3 factorization (a ~ lognormal(0,1), maxSideLength 8, seed 5):
    Poisson arrivals of requests, each a coefficient change on a random box of side U{4..16} by
    one LN(0,2) factor plus U{1..32} N(0,1) right-hand sides, answered against the unchanged base
    (updating-multifrontal_amd/serve.py: slots on their own streams, admission by the dirty
    fronts' footprint in the slot pool, the real clock).  Latency = completion - arrival."""
    import torch
    import umf
    from serve import Request, SlotServer
    torch.cuda.set_device(0)
    N, ms, seed = 95, 8, 5
    dims = (N, N, N)
    n = N ** 3
    h = 1.0 / (N + 1)
    rng = np.random.default_rng(seed)
    coeff = np.exp(rng.standard_normal((N + 2,) * 3))
    rp, ci, v = umf.build_laplacian(dims, coeff, 1, h)
    plan = umf.MF3D(Csr(rp, ci, v), *dims, 1, ms)
    d_coeff = torch.from_numpy(coeff).cuda()
    out = {"metric": "update+solve request latency (p50/p99) and requests/s, continuous batching", "unit": "ms",
           "n_gpus": 1, "higher_is_better": False, "data": "synthetic",
           "config": {"workload": f"3D 7-point -div(a grad u), 95^3 grid (n={n}), a~lognormal(0,1), maxSideLength "
                                  f"{ms}; requests: a x LN(0,2) on a U{{4..16}}-side node box + U{{1..32}} N(0,1) "
                                  f"right-hand sides, Poisson arrivals, each answered against the unchanged base",
                      "requests_per_rate": args.requests, "slots": args.slots},
           "forms": []}
    for prec in ("fp64", "fp32", "bf16x3"):
        if prec not in args.forms:
            continue
        plan.SetPrecision(prec, tol=args.ir_tol, max_iter=100)
        plan.Factor()
        pool = args.pool_gb * 1e9 * (1.0 if prec == "fp64" else 0.5)
        S = SlotServer(plan, torch, coeff, d_coeff, v, rp, ci, h, slots=args.slots, pool_bytes=pool)

        def make(count, rate):
            reqs, t = [], 0.0
            for _ in range(count):
                t += float(rng.exponential(1.0 / rate)) if rate else 0.0
                side = int(rng.integers(4, 17))
                lo = [int(x) for x in rng.integers(0, N + 2 - side, size=3)]
                reqs.append(Request(rng.standard_normal((int(rng.integers(1, 33)), n)), t,
                                    box=(lo, side, float(np.exp(2.0 * rng.standard_normal())))))
            return reqs

        S.run(make(args.slots, 0), realtime=False)  # warm-up
        form = {"precision": prec, "pool_gb": pool / 1e9, "rates": []}
        for rate in args.rates:
            reqs = make(args.requests, rate)
            lat = S.run(reqs) * 1e3
            span = max(r.done for r in reqs) - min(r.arrival for r in reqs)
            tr = S.inflight_trace
            ent = {"rate_per_s": rate, "p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
                   "mean_ms": float(lat.mean()), "throughput_per_s": len(reqs) / span,
                   "queue_wait_ms_mean": float(np.mean([(r.admitted - r.arrival) * 1e3 for r in reqs])),
                   "gpu_update_ms_mean": float(np.mean([r.t_update_ms for r in reqs])),
                   "gpu_solve_ms_mean": float(np.mean([r.t_solve_ms for r in reqs])),
                   "max_in_flight": int(max(c for _, c in tr)),
                   "footprint_gb_mean": float(np.mean([r.footprint for r in reqs]) / 1e9),
                   "rhs_total": int(sum(len(r.rhs) for r in reqs))}
            if prec != "fp64":
                ent["ir_iters_max"] = int(max(r.ir_iters for r in reqs))
            progress(f"{prec} rate {rate}/s: p50 {ent['p50_ms']:.0f} ms, p99 {ent['p99_ms']:.0f} ms, "
                     f"{ent['throughput_per_s']:.2f} req/s, in flight <= {ent['max_in_flight']}")
            form["rates"].append(ent)
        # accuracy of the last request: infinity-norm backward error against its own operator
        r = reqs[-1]
        lo, side, f = r.box
        c2 = coeff.copy()
        c2[tuple(slice(l, l + side) for l in lo)] *= f
        _, _, v2 = umf.build_laplacian(dims, c2, 1, h)
        X = r.x.cpu().numpy()
        rows = np.repeat(np.arange(n), np.diff(rp))
        rs = np.zeros(n)
        np.add.at(rs, rows, np.abs(v2))
        import scipy.sparse as sp
        A2 = sp.csr_matrix((v2, ci, rp), shape=(n, n))
        R = r.rhs - (A2 @ X.T).T
        form["backward_error_inf_last"] = float(max(np.abs(R[j]).max() / (rs.max() * np.abs(X[j]).max() +
                                                                        np.abs(r.rhs[j]).max())
                                                    for j in range(len(R))))
        S.close()
        out["forms"].append(form)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
