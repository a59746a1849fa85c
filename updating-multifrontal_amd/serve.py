"""Continuous-batching request server over one resident multifrontal factorization.

A request changes the coefficient field on a box of nodes and asks for the solution of k
right-hand sides with that operator -- the reference's Update use case (a local coefficient
change, Update with the modified DOFs, then Solve; test_src/DE/testMF3D.cpp:137-215) -- but
each request is answered against the unchanged base factorization: requests never see each
other's changes, and the answer is what Update + Solve on the base would give, bit for bit.

The library's server (umf.Server, umf_server_* in include/umf.h) holds the base plan immutable
and runs requests in slots, each with its own HIP stream.  A request's dirty fronts (the
reference's Update recurrence over its modified DOFs) are factored into a range of the slot
pool; its footprint is their bytes.  This loop is the scheduler around it, on the real clock:

- requests arrive at their arrival times and queue in arrival order;
- the head of the queue is admitted as soon as a slot is idle and the pool has a free range for
  its footprint (umf_server_admit); it waits otherwise, and so do the requests behind it;
- an admitted request runs in its slot on a worker thread: its box change is applied on the
  device (umf_server_box_scale), its dirty fronts are factored and its right-hand sides solved
  (umf_server_run); then its pool range is released and the next request can be admitted.

Requests that arrive while others are on the GPU are therefore admitted as slots and memory
free up, and their work overlaps on the device (slots on separate streams, driven by separate
host threads; ctypes releases the GIL during the library calls).
"""
import queue
import threading
import time

import numpy as np

from umf import Server, build_laplacian


class Request:
    """A coefficient box change box = (lo, side, factor) on the padded nodal field, and k
    right-hand sides (rhs: k x n, row r = right-hand side r), arriving at `arrival` seconds.
    d_b / d_x: optional device buffers (objects with data_ptr()) holding the right-hand sides
    and receiving the solution; without them the server makes torch tensors."""

    def __init__(self, rhs, arrival=0.0, box=None, d_b=None, d_x=None):
        self.rhs, self.arrival, self.box = rhs, arrival, box
        self.d_b, self.d_x = d_b, d_x
        self.mod = None
        self.x = None         # k x n solution (device tensor)
        self.admitted = None  # seconds, on the server's clock
        self.done = None
        self.slot = None
        self.footprint = None
        self.t_update_ms = None
        self.t_solve_ms = None
        self.ir_iters = None


def box_mods(dims, coeff, vals, lo, side, factor, rp, ci, h, width=1):
    """Modified DOFs of a box change -- the columns of the operator entries whose value changes,
    the reference's rule (test_src/DE/testMF3D.cpp:160-166) -- without changing `coeff` or
    `vals`: the operator is rebuilt on the window of rows the change can reach (interior nodes
    within 2w of the box, the padded field around them included, so the window's entries are
    exact) and compared with `vals`."""
    d = len(dims)
    w = width
    w0 = [max(l - 2 * w, 0) for l in lo]
    w1 = [min(l + side, n) for l, n in zip(lo, dims)]  # exclusive interior index
    if any(b <= a for a, b in zip(w0, w1)):
        return np.zeros(0, np.int32)
    wdims = tuple(b - a for a, b in zip(w0, w1))
    sub = np.array(coeff[tuple(slice(a, b + 2 * w) for a, b in zip(w0, w1))])
    box = tuple(slice(max(l - a, 0), max(l + side - a, 0)) for l, a in zip(lo, w0))
    sub[box] *= factor
    lrp, lci, lval = build_laplacian(wdims, sub, w, h)
    loc = np.indices(wdims).reshape(d, -1)
    lid = np.ravel_multi_index(loc, wdims, order="F")
    gid = np.ravel_multi_index(loc + np.array(w0)[:, None], dims, order="F")
    g_of_l = np.empty(len(lid), np.int64)
    g_of_l[lid] = gid
    rows = np.repeat(np.arange(len(lrp) - 1), np.diff(lrp))
    grow, gcol = g_of_l[rows], g_of_l[lci]
    pos = np.empty(len(lci), np.int64)
    for r in np.unique(grow):  # columns ascend within a CSR row
        sel = grow == r
        pos[sel] = rp[r] + np.searchsorted(ci[rp[r]:rp[r + 1]], gcol[sel])
    changed = vals[pos] != lval
    return np.unique(gcol[changed]).astype(np.int32)


class SlotServer:
    """Serves Requests against `plan` (an MF2D / MF3D factored with the values `vals`, built by
    the harness generator from the padded nodal field `coeff` with grid spacing h).  The field
    lives on the device (`d_coeff`, any object with data_ptr()) for the slots' box changes.
    torch (or None when every request brings its own device buffers) makes the right-hand-side
    and solution buffers, each slot's on a stream of its own."""

    def __init__(self, plan, torch, coeff, d_coeff, vals, rp, ci, h, slots=4, pool_bytes=16e9):
        self.plan, self.torch = plan, torch
        self.coeff, self.d_coeff, self.vals, self.rp, self.ci, self.h = coeff, d_coeff, vals, rp, ci, h
        self.dims = plan.dims
        self.nslots = int(slots)
        self.srv = Server(plan, self.nslots, pool_bytes)
        self.streams = [torch.cuda.Stream() for _ in range(self.nslots)] if torch is not None else None

    def close(self):
        self.srv.close()

    def mods(self, r):
        if r.mod is None:
            lo, side, f = r.box
            r.mod = box_mods(self.dims, self.coeff, self.vals, lo, side, f, self.rp, self.ci, self.h,
                             self.plan.width)
        return r.mod

    def _work(self, slot, job, clock, done):
        """Worker thread of one slot: runs the requests handed to it."""
        torch = self.torch
        while True:
            item = job.get()
            if item is None:
                return
            r, ticket = item
            try:
                if r.d_b is not None:
                    B, X = r.d_b, r.d_x
                else:
                    with torch.cuda.stream(self.streams[slot]):
                        B = torch.from_numpy(np.ascontiguousarray(r.rhs)).cuda()
                        X = torch.empty_like(B)
                        self.streams[slot].synchronize()
                lo, side, f = r.box
                dmod = self.srv.box_scale(slot, self.d_coeff.data_ptr(), self.h, lo, side, f)
                if not np.array_equal(dmod, r.mod):
                    raise RuntimeError("device and host modified-DOF sets differ")
                self.srv.run(slot, ticket, B.data_ptr(), X.data_ptr(), len(r.rhs))
                st = self.srv.stats(slot)
                r.t_update_ms, r.t_solve_ms, r.ir_iters = st["t_update_ms"], st["t_solve_ms"], st["ir_iters"]
                r.x = X
                r.error = None
            except Exception as e:  # reported to the dispatcher
                r.error = e
            finally:
                self.srv.release(ticket)
                r.done = clock()
                done.put(slot)

    def run(self, reqs, realtime=True):
        """Serve `reqs`: each becomes available at its arrival time (seconds after the start, on
        the real clock; realtime=False makes them all available at once, in arrival order).
        Sets r.admitted / r.done / r.x; returns the latencies done - arrival in seconds."""
        t0 = time.perf_counter()
        clock = lambda: time.perf_counter() - t0
        done = queue.Queue()
        jobs = [queue.Queue() for _ in range(self.nslots)]
        workers = [threading.Thread(target=self._work, args=(s, jobs[s], clock, done), daemon=True)
                   for s in range(self.nslots)]
        for w in workers:
            w.start()
        order = sorted(reqs, key=lambda r: r.arrival)
        for r in order:  # host-side admission data: modified DOFs and footprint
            self.mods(r)
            r.footprint = self.srv.footprint(r.mod)
        idle = list(range(self.nslots))
        waiting, nxt, finished = [], 0, 0
        self.inflight_trace = []
        try:
            while finished < len(order):
                now = clock()
                while nxt < len(order) and (not realtime or order[nxt].arrival <= now):
                    waiting.append(order[nxt])
                    nxt += 1
                # admit from the head of the queue while a slot is idle and the pool holds it
                while waiting and idle:
                    ticket = self.srv.admit(waiting[0].mod)
                    if ticket is None:
                        break
                    r = waiting.pop(0)
                    r.slot = idle.pop(0)
                    r.admitted = clock()
                    jobs[r.slot].put((r, ticket))
                self.inflight_trace.append((clock(), self.nslots - len(idle)))
                # wait for a completion, or until the next arrival
                timeout = None
                if nxt < len(order) and realtime:
                    timeout = max(order[nxt].arrival - clock(), 0.0)
                try:
                    s = done.get(timeout=timeout if timeout is not None else 60.0)
                except queue.Empty:
                    if timeout is None:
                        raise RuntimeError("no request finished within 60 s")
                    continue
                idle.append(s)
                finished += 1
                while True:  # collect every completion already signalled
                    try:
                        idle.append(done.get_nowait())
                        finished += 1
                    except queue.Empty:
                        break
        finally:
            for j in jobs:
                j.put(None)
            for w in workers:
                w.join()
        for r in order:
            if getattr(r, "error", None) is not None:
                raise r.error
        if realtime:
            return np.array([r.done - r.arrival for r in reqs])
        return np.array([r.done for r in reqs])
