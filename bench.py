#!/usr/bin/env python3
"""Benchmark: factor + local update + 16-RHS solve of the 2D variable-coefficient operator.

Workload (the 2D 1023^2 configuration): 5-point -div(a grad u) on the unit square, 1023 x 1023
interior grid (n = 1,046,529), a = lognormal(0,1) iid per padded node (numpy PCG64, seed 2),
h = 1/1024, nested dissection with maxSideLength 16 (31 x 31 leaf boxes, 6 levels), FP64
plain form.  One step = Factor() + Update() (a multiplied by 100 on a random 32 x 32 block of
nodes, modified DOFs by the reference's modDOF rule) + Solve() of 16 N(0,1) right-hand sides.
All inputs (CSR values of both operators, right-hand sides) are resident in HBM before timing;
the update's modified-DOF list is host data, as in the reference API.

value = algorithmic TFLOP/s of the whole step: (F_factor + F_update + F_solve) / time, with
F = sum over fronts of n^3/3 + n^2 m + n m^2 (factor, dirty fronts for update) and
nrhs * sum (2 n^2 + 4 n m) (solve), n / m = pivot / boundary size of each front.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(ROOT, "updating-multifrontal_amd")
sys.path.insert(0, PKG)

METRIC = "factor+update+solve s & TFLOP/s; FP32-MFMA+IR vs FP64 at 1e-10 backward err"
PEAK_FP64_TFLOPS = 78.6
PEAK_HBM_GBS = 8000.0


def measured_traffic(kernel, workload):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 --pmc FETCH_SIZE/WRITE_SIZE
    summary of this workload (profiles/<kernel>_traffic_<workload>.md, calibrated on a plain
    copy), or None when there is none: counters cannot be read inside the timed run."""
    import re
    path = os.path.join(ROOT, "profiles", f"{kernel}_traffic_{workload}.md")
    try:
        with open(path) as f:
            m = re.search(r"^bytes per launch:\s*([0-9.eE+-]+)\s*$", f.read(), re.M)
        return float(m.group(1)) if m else None
    except OSError:
        return None


# workloads: name -> (dims, maxSideLength, update box side, update scale or None for one LN(0,2)
# factor, seed); the first is the headline, the 3D ones are scale runs of the same step
WORKLOADS = {
    "2d1023": ((1023, 1023), 16, 32, 100.0, 2),
    "3d95": ((95, 95, 95), 8, 8, None, 3),
    "3d127": ((127, 127, 127), 8, 8, None, 3),
    "batch256": None,  # make_batch
    "serve95": None,   # run_serve
}


def make_problem(N, seed, box=32, scale=100.0, nrhs=16, dims=None):
    """Lognormal(0,1) nodal coefficients on the padded grid, the update multiplies a random
    box of nodes by `scale` (or by one LN(0,2) draw when scale is None)."""
    import umf
    dims = dims or (N, N)
    d = len(dims)
    rng = np.random.default_rng(seed)
    co = np.exp(rng.standard_normal(tuple(n + 2 for n in dims)))
    co2 = co.copy()
    lo = [int(v) for v in rng.integers(0, [n + 2 - box for n in dims])]
    if scale is None:
        scale = float(np.exp(2.0 * rng.standard_normal()))
    co2[tuple(slice(l, l + box) for l in lo)] *= scale
    h = 1.0 / (dims[0] + 1)
    rp, ci, v1 = umf.build_laplacian(dims, co, 1, h)
    _, _, v2 = umf.build_laplacian(dims, co2, 1, h)
    mod = np.unique(ci[v1 != v2]).astype(np.int32)  # modDOF rule (testMF2D.cpp:143-150)
    B = rng.standard_normal((nrhs, int(np.prod(dims))))  # row r = right-hand side r: column-major n x nrhs
    return rp, ci, v1, v2, mod, B, tuple(lo)


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


def cpu_baseline_threaded(nrhs=16):
    """The same sample on the reference built against threaded MKL (oracle/_ref/libumfref_mt.so)
    on all of this job's host cores.  It runs in a child process started before this one touches
    the GPU: the threaded and sequential MKL layers cannot share a process.  None when that build
    is absent or the child fails."""
    import subprocess
    lib = os.path.join(ROOT, "oracle", "_ref", "libumfref_mt.so")
    if not os.path.exists(lib):
        return None
    t = cpu_threads()
    env = dict(os.environ, UMF_REF_SO=lib, MKL_NUM_THREADS=str(t), OMP_NUM_THREADS=str(t))
    try:
        r = subprocess.run([sys.executable, os.path.abspath(__file__), "--cpu-sample", "--nrhs", str(nrhs)],
                           env=env, capture_output=True, text=True, timeout=300)
    except subprocess.TimeoutExpired:
        return None
    if r.returncode != 0 or not r.stdout.strip():
        progress(f"threaded CPU baseline failed: {r.stderr.strip()[-300:]}")
        return None
    out = json.loads(r.stdout.strip().splitlines()[-1])
    out["cores"] = t
    out["sample"] = out["sample"].replace("MKL sequential", f"threaded MKL ({t} threads)")
    return out


def cpu_baseline(N=511, seed=2, nrhs=16, max_side=16):
    """The reference itself (oracle/_ref, MKL sequential, 1 thread) on a bounded sample of the
    same workload: the same generator at 511^2 (5 levels, the same 31^2 leaf fronts),
    Factor + Update + nrhs single-vector Solve calls as the reference API requires."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_py
    import umf
    kind = "reference" if oracle_py.have_ref() else "port"
    rp, ci, v1, v2, mod, B, _ = make_problem(N, seed, nrhs=nrhs)
    A1, A2 = Csr(rp, ci, v1), Csr(rp, ci, v2)
    P = umf.MF2D(A1, N, N, 1, max_side, analyze_only=True)
    f_factor = P.stats()["flops_factor"]
    P.analyze_update(mod)
    f_update = P.stats()["flops_update"]
    f_solve = algorithmic_flops(P, nrhs)
    Ac1 = oracle_py.Csr(rp, ci, v1)
    Ac2 = oracle_py.Csr(rp, ci, v2)
    F = oracle_py.Factorization(Ac1, (N, N), 1, max_side, "ref" if kind == "reference" else "orc")
    t0 = time.perf_counter()
    F.factor(trace=False)
    t1 = time.perf_counter()
    F.update(Ac2, mod, trace=False)
    t2 = time.perf_counter()
    for r in range(nrhs):
        F.solve(B[r])
    t3 = time.perf_counter()
    tot = t3 - t0
    return {
        "value": (f_factor + f_update + f_solve) / tot / 1e12,
        "unit": "TFLOP/s",
        "cores": 1,
        "kind": kind,
        "sample": f"2D {N}x{N} lognormal(0,1), maxSideLength {max_side}: Factor + Update (x100 on 32x32) + "
                  f"{nrhs} single-RHS Solve calls; reference LU form via MKL sequential",
        "t_factor_s": t1 - t0,
        "t_update_s": t2 - t1,
        "t_solve_s": t3 - t2,
        "seconds": tot,
    }


def backward_error(plan, d_B, d_X, d_R, vals, B, nrhs):
    """max_j ||b_j - A x_j|| / (||A||_F ||x_j|| + ||b_j||), residual by the FP64 SpMV."""
    plan.spmv_device(d_X.data_ptr(), d_R.data_ptr(), nrhs)
    R = (d_B - d_R).cpu().numpy()
    X = d_X.cpu().numpy()
    normA = float(np.linalg.norm(vals))
    return max(np.linalg.norm(R[j]) / (normA * np.linalg.norm(X[j]) + np.linalg.norm(B[j])) for j in range(nrhs))


FORM_DTYPE = {
    "fp32": "f32 fronts (FP32 MFMA) + f64 refinement",
    "bf16x3": "f32 fronts, Schur/trailing GEMMs on BF16 MFMA over 2 bf16 pieces (3 products, f32 accumulate) "
              "+ f64 refinement",
}


def low_form(prec, plan, step, state, d_v, d_B, d_X, d_R, vals, B, args, nrhs):
    """The same step in a lower-precision form (FP32 fronts; GEMMs on FP32 MFMA or on BF16
    pieces) with the Solve refined in FP64 to backward error <= --ir-tol; same algorithmic flop
    numerator as the FP64 form."""
    import torch
    plan.SetPrecision(prec, tol=args.ir_tol, max_iter=100)
    plan.profile(False)
    for _ in range(max(args.warmup, 1)):
        step()
    kern = {}
    if not args.no_profile:  # untimed step with per-kernel events
        plan.profile(True)
        plan.profile_select(None)
        plan.profile_reset()
        step()
        torch.cuda.synchronize()
        for k, v in plan.profile_read().items():
            d = kern.setdefault(k.split("@")[0], {"ms": 0.0, "launches": 0, "flops": 0.0})
            for f in d:
                d[f] += v[f]
        plan.profile_reset()
        plan.profile(False)
    tf = tu = ts = flops = 0.0
    iters = []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(args.steps):
        step()
        st = plan.stats()
        progress(f"step {i}: factor {st['t_factor_ms']:.1f} ms, update {st['t_update_ms']:.1f} ms, "
                 f"solve {st['t_solve_ms']:.1f} ms")
        tf += st["t_factor_ms"]
        tu += st["t_update_ms"]
        ts += st["t_solve_ms"]
        iters.append(st["ir_iters"])
        flops += st["flops_factor"] + st["flops_update"] + st["flops_solve"]
    torch.cuda.synchronize()
    secs = time.perf_counter() - t0
    K = args.steps
    st = plan.stats()
    be = backward_error(plan, d_B, d_X, d_R, vals[state["cur"]], B, nrhs)
    plan.SetPrecision("fp64")
    progress(f"{prec} form: {secs / K * 1e3:.1f} ms per step, refinement iterations {iters}, berr {be:.1e}")
    return {"value": flops / secs / 1e12, "unit": "TFLOP/s", "ms_per_step": secs / K * 1e3,
            "t_factor_ms": tf / K, "t_update_ms": tu / K, "t_solve_ms": ts / K,
            "ir_iters": iters, "ir_tol": args.ir_tol, "backward_error": be, "ir_berr": st["ir_berr"],
            "front_pool_gb": st["bytes_front_pool"] / 1e9, "dtype": FORM_DTYPE[prec],
            "kernels_ms_per_step": {k: round(v["ms"], 3) for k, v in sorted(kern.items())},
            "kernels_tflops": {k: round(v["flops"] / (v["ms"] * 1e-3) / 1e12, 2)
                               for k, v in sorted(kern.items()) if v["ms"] > 0 and v["flops"]}}


def factor_graph_timing(plan, torch, reps=5):
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


def run_serve(args):
    """Serving: Poisson arrivals of coefficient-change + solve requests on one 3D 95^3
    factorization (a ~ lognormal(0,1), maxSideLength 8, seed 5); a request multiplies a box of
    side U{4..16} at a uniform position by one LN(0,2) factor and brings U{1..32} N(0,1)
    right-hand sides.  The server batches every waiting request into one Update + one Solve
    (updating-multifrontal_amd/serve.py); latency = completion - arrival on a virtual clock
    advanced by the measured GPU service time."""
    import torch
    import umf
    from serve import Request, UpdateServer
    torch.cuda.set_device(0)
    N, ms, seed = 95, 8, 5
    dims = (N, N, N)
    n = N ** 3
    h = 1.0 / (N + 1)
    rng = np.random.default_rng(seed)
    coeff = np.exp(rng.standard_normal((N + 2,) * 3))
    rp, ci, v = umf.build_laplacian(dims, coeff, 1, h)
    plan = umf.MF3D(Csr(rp, ci, v), *dims, 1, ms)
    out = {"metric": "update+solve request latency, continuous batching", "unit": "ms", "n_gpus": 1,
           "higher_is_better": False, "data": "synthetic",
           "config": {"workload": f"3D 7-point -div(a grad u), 95^3 grid (n={n}), a~lognormal(0,1), maxSideLength {ms}; "
                                  f"requests: a x LN(0,2) on a U{{4..16}}-side node box (applied on the device) + "
                                  f"U{{1..32}} N(0,1) RHS, Poisson arrivals", "requests_per_rate": args.requests},
           "forms": []}
    # one operator state throughout: the FP32 pass continues from the FP64 pass's last round;
    # the coefficient field and the operator values live on the device, where each request's
    # box change is applied inside its round
    d_val = torch.from_numpy(v).cuda()
    d_coeff = torch.from_numpy(coeff.copy()).cuda()
    for prec in ("fp64", "fp32", "bf16x3"):
        plan.SetPrecision(prec, tol=args.ir_tol)
        plan.Factor()
        srv = UpdateServer(plan, d_val, torch, d_coeff=d_coeff, h=h)
        form = {"precision": prec, "rates": []}
        for rate in args.rates:
            reqs = []
            t = 0.0
            for _ in range(args.requests):
                t += float(rng.exponential(1.0 / rate))
                side = int(rng.integers(4, 17))
                lo = [int(x) for x in rng.integers(0, N + 2 - side, size=3)]
                fac = float(np.exp(2.0 * rng.standard_normal()))
                k = int(rng.integers(1, 33))
                reqs.append(Request(None, None, None, rng.standard_normal((k, n)), t, box=(lo, side, fac)))
            srv.rounds = []
            lat = srv.run(reqs) * 1e3
            # accuracy of the last round: every right-hand side of its last request
            r = reqs[-1]
            X = r.x.contiguous()
            Bd = r.rhs_d
            R = torch.empty_like(X)
            plan.spmv_device(X.data_ptr(), R.data_ptr(), X.shape[0])
            na = float(torch.linalg.norm(d_val))
            be = max(float(torch.linalg.norm(Bd[j] - R[j]) / (na * torch.linalg.norm(X[j]) + torch.linalg.norm(Bd[j])))
                     for j in range(X.shape[0]))
            rounds = srv.rounds
            span = max(q.done for q in reqs) - min(q.arrival for q in reqs)
            ent = {"rate_per_s": rate, "p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
                   "mean_ms": float(lat.mean()), "rounds": len(rounds),
                   "requests_per_round": float(np.mean([x[0] for x in rounds])),
                   "rhs_per_round": float(np.mean([x[1] for x in rounds])),
                   "round_ms_mean": float(np.mean([x[2] for x in rounds]) * 1e3),
                   "throughput_per_s": len(reqs) / span, "backward_error_last": be}
            progress(f"{prec} rate {rate}/s: p50 {ent['p50_ms']:.0f} ms, p99 {ent['p99_ms']:.0f} ms, "
                     f"{len(rounds)} rounds, berr {be:.1e}")
            form["rates"].append(ent)
        out["forms"].append(form)
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", choices=sorted(WORKLOADS), default="2d1023")
    ap.add_argument("--nrhs", type=int, default=None, help="right-hand sides (16; 4 for the batch)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-profile", action="store_true", help="skip per-kernel event timing")
    ap.add_argument("--no-fp32", action="store_true", help="skip the lower-precision + FP64-refinement forms")
    ap.add_argument("--forms", nargs="+", choices=sorted(FORM_DTYPE), default=["fp32", "bf16x3"],
                    help="lower-precision forms measured beside FP64 (each refined in FP64)")
    ap.add_argument("--ir-tol", type=float, default=1e-10, help="backward error the FP32 form refines to")
    ap.add_argument("--cpu-sample", action="store_true", help=argparse.SUPPRESS)  # child of the CPU baseline
    ap.add_argument("--autotune", action="store_true",
                    help="run the plan's autotuner (GEMM tile size, outer group width) before timing")
    ap.add_argument("--rates", type=float, nargs="+", default=[2.0, 8.0, 32.0],
                    help="serve95: request arrival rates (per second)")
    ap.add_argument("--requests", type=int, default=24, help="serve95: requests per rate")
    args = ap.parse_args()
    if args.cpu_sample:  # the threaded CPU baseline's child process: no GPU work
        print(json.dumps(cpu_baseline(nrhs=args.nrhs or 16)), flush=True)
        return
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU only; --gpus {args.gpus} is not supported "
              "(the sharded path is design-only)", file=sys.stderr)
        sys.exit(2)

    if args.workload == "serve95":
        return run_serve(args)
    cpu_mt = None
    if not args.no_cpu_baseline and args.workload == "2d1023":
        progress(f"threaded CPU baseline ({cpu_threads()} threads)")
        cpu_mt = cpu_baseline_threaded(nrhs=args.nrhs or 16)
    import torch
    import umf
    torch.cuda.set_device(0)
    if args.workload == "batch256":
        nrhs = args.nrhs or 4
        plan, rp, ci, v1, v2, mod, B, wdesc = make_batch(nrhs=nrhs)
        n = plan.n
    else:
        dims, max_side, box, scale, seed = WORKLOADS[args.workload]
        nrhs = args.nrhs or 16
        rp, ci, v1, v2, mod, B, corner = make_problem(dims[0], seed, box=box, scale=scale, nrhs=nrhs, dims=dims)
        n = int(np.prod(dims))
        plan = (umf.MF2D if len(dims) == 2 else umf.MF3D)(Csr(rp, ci, v1), *dims, 1, max_side)
        wdesc = (f"{len(dims)}D {'5' if len(dims) == 2 else '7'}-point -div(a grad u), "
                 f"{'x'.join(map(str, dims))} grid (n={n}), a~lognormal(0,1) per node, "
                 f"maxSideLength {max_side} ({plan.levels} levels, {plan.leaf_width(0)}^{len(dims)} leaves), "
                 f"Factor + Update (a x{'100' if scale else 'LN(0,2)'} on {box}^{len(dims)} nodes at {corner}) "
                 f"+ {nrhs}-RHS Solve, FP64")
    d_v = [torch.from_numpy(v1).cuda(), torch.from_numpy(v2).cuda()]
    d_B = torch.from_numpy(B).cuda()
    d_X = torch.empty_like(d_B)
    d_R = torch.empty_like(d_B)

    state = {"cur": 0}  # index of the operator the plan currently holds

    def step():
        plan.Factor()
        nxt = 1 - state["cur"]
        plan.update_device(d_v[nxt].data_ptr(), mod)
        state["cur"] = nxt
        plan.solve_device(d_B.data_ptr(), d_X.data_ptr(), nrhs)

    progress(f"workload {args.workload}: n={n}, {plan.levels} levels, front pool "
             f"{plan.stats()['bytes_front_pool'] / 1e9:.1f} GB")
    tuned = None
    if args.autotune:
        tuned = plan.autotune(reps=2)
        progress(f"autotune: {tuned}")
    for i in range(args.warmup):
        step()
        progress(f"warmup step {i} done")
    # untimed profiling step: every kernel bracketed by events -> per-kernel table, dominant kernel
    prof_lv = {}
    if not args.no_profile:
        plan.profile(True)
        plan.profile_select(None)
        plan.profile_reset()
        step()
        torch.cuda.synchronize()
        prof_lv = plan.profile_read()
        plan.profile_reset()
    prof = {}  # per kernel, summed over tree levels (names carry "@L<level>")
    for k, v in prof_lv.items():
        d = prof.setdefault(k.split("@")[0], {"ms": 0.0, "launches": 0, "flops": 0.0, "bytes": 0.0})
        for f in d:
            d[f] += v[f]
    dom = max(prof, key=lambda k: prof[k]["ms"]) if prof else None
    if dom:
        plan.profile_select(dom)  # the timed region carries events around this kernel only
    tf = tu = ts = 0.0
    flops = 0.0
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(args.steps):
        step()
        st = plan.stats()
        progress(f"step {i}: factor {st['t_factor_ms']:.1f} ms, update {st['t_update_ms']:.1f} ms, "
                 f"solve {st['t_solve_ms']:.1f} ms")
        tf += st["t_factor_ms"]
        tu += st["t_update_ms"]
        ts += st["t_solve_ms"]
        flops += st["flops_factor"] + st["flops_update"] + st["flops_solve"]
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    plan.profile(False)
    secs = t1 - t0
    K = args.steps

    # accuracy of the last step: normwise backward error max_j ||b - A x|| / (||A||_F ||x|| + ||b||)
    berr = backward_error(plan, d_B, d_X, d_R, (v1, v2)[state["cur"]], B, nrhs)

    roof = None
    if dom:
        d = {"ms": 0.0, "launches": 0, "flops": 0.0, "bytes": 0.0}  # the dominant kernel in the timed region
        for k, v in plan.profile_read().items():
            if k.split("@")[0] == dom:
                for f in d:
                    d[f] += v[f]
        avg_ms = d["ms"] / max(d["launches"], 1)
        if d["flops"] > 0:
            ach = d["flops"] / (d["ms"] * 1e-3) / 1e12
            roof = {"bound": "mfma", "achieved": ach, "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                    "frac": ach / PEAK_FP64_TFLOPS,
                    "traffic": measured_traffic(dom, args.workload)}
        else:
            ach = d["bytes"] / (d["ms"] * 1e-3) / 1e9
            roof = {"bound": "hbm", "achieved": ach, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                    "frac": ach / PEAK_HBM_GBS, "traffic": None}
        roof.update({"kernel": dom, "launches": d["launches"], "avg_launch_ms": avg_ms,
                     "share_of_kernel_time": d["ms"] / K / sum(v["ms"] for v in prof.values())})

    st = plan.stats()
    out = {
        "metric": METRIC,
        "value": flops / secs / 1e12,
        "unit": "TFLOP/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": args.warmup,
        "ms_per_step": secs / K * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": wdesc,
            "nrhs": nrhs,
            "precision": "fp64",
            "tune": plan.tune(),
            "autotuned": tuned is not None,
        },
        "t_factor_ms": tf / K,
        "t_update_ms": tu / K,
        "t_solve_ms": ts / K,
        "flops_per_step": flops / K,
        "fronts": st["n_fronts"],
        "fronts_update": st["n_fronts_update"],
        "front_pool_gb": st["bytes_front_pool"] / 1e9,
        "backward_error": berr,
        "roofline": roof,
        "kernels": {k: {"ms_per_step": v["ms"], "launches_per_step": v["launches"],
                        "tflops": (v["flops"] / (v["ms"] * 1e-3) / 1e12) if v["ms"] > 0 and v["flops"] else None,
                        "gbs": (v["bytes"] / (v["ms"] * 1e-3) / 1e9) if v["ms"] > 0 and v["bytes"] else None}
                    for k, v in prof.items()},
        "kernels_by_level": {k: round(v["ms"], 4) for k, v in sorted(prof_lv.items())},
        "cpu_baseline": None,
    }
    out["factor_graph"] = factor_graph_timing(plan, torch)
    if not args.no_fp32:
        for prec in args.forms:
            out[prec + "_ir"] = low_form(prec, plan, step, state, d_v, d_B, d_X, d_R, (v1, v2), B, args, nrhs)
    if not args.no_cpu_baseline and args.workload == "2d1023":
        one = cpu_baseline(nrhs=nrhs)
        if cpu_mt is not None:  # all host cores (threaded MKL), with the 1-thread run beside it
            cpu_mt["one_core"] = {k: one[k] for k in ("value", "unit", "cores", "kind", "seconds")}
            out["cpu_baseline"] = cpu_mt
        else:
            out["cpu_baseline"] = one
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
