"""Continuous-batching request server over one multifrontal factorization.

A request changes the coefficient field on a box of nodes (the reference's Update use case,
test_src/DE/testMF3D.cpp:150-180: a local coefficient change, then Update with the modified
DOFs) and asks for the solution of k right-hand sides with the updated operator.  The server
keeps the operator values (and, for box requests, the nodal coefficient field) resident on the
GPU.  Whenever it becomes free it takes every request that has arrived: their changes are
applied in arrival order (device scatter of given values, or the box scaling and the operator
rows it reaches recomputed on the device), one Update runs over the union of their modified
DOFs (the dirty sets of all boxes at once, so the shared ancestors up to the root are
refactored once per round, not once per request), and one multi-RHS Solve answers all their
right-hand sides.  Each request therefore sees the operator
with every change of its round applied.
"""
import time

import numpy as np

from umf import build_laplacian


class Request:
    """A value change -- CSR positions and new values with their modified DOFs, or a coefficient
    box change box = (lo, side, factor) done on the device by the server -- and right-hand
    sides."""

    def __init__(self, pos, val, mod, rhs, arrival=0.0, box=None):
        self.pos, self.val, self.mod, self.rhs, self.arrival = pos, val, mod, rhs, arrival
        self.box = box
        self.done = None
        self.x = None


def box_change(dims, coeff, vals, lo, side, factor, rp, ci, h):
    """Apply `factor` to the nodal coefficients of the box [lo, lo + side) (padded indices) and
    to the CSR values `vals`, both in place, and return (CSR positions, new values, modified
    DOFs) of the operator entries that change; the modified DOFs follow the reference's rule,
    the columns of changed entries (test_src/DE/testMF3D.cpp:160-166).  Only rows within one
    node of the box change: the operator is rebuilt on that window from the padded coefficients
    around it, whose diagonals and internal entries are then exact (entries leaving the window
    join two unchanged nodes)."""
    d = len(dims)
    coeff[tuple(slice(l, l + side) for l in lo)] *= factor
    # interior-node window: padded box widened by 1, as interior indices (padded index - 1)
    w0 = [max(l - 2, 0) for l in lo]
    w1 = [min(l + side - 1 + 1, n) for l, n in zip(lo, dims)]  # exclusive
    wdims = tuple(b - a for a, b in zip(w0, w1))
    sub = coeff[tuple(slice(a, b + 2) for a, b in zip(w0, w1))]
    lrp, lci, lval = build_laplacian(wdims, sub, 1, h)
    # local -> global DOF ids
    loc = np.indices(wdims).reshape(d, -1)  # [axis, local dof] with axis 0 fastest in id
    lid = np.ravel_multi_index(loc, wdims, order="F")
    gcoord = loc + np.array(w0)[:, None]
    gid = np.ravel_multi_index(gcoord, dims, order="F")
    g_of_l = np.empty(len(lid), np.int64)
    g_of_l[lid] = gid
    rows = np.repeat(np.arange(len(lrp) - 1), np.diff(lrp))
    grow, gcol = g_of_l[rows], g_of_l[lci]
    pos = np.empty(len(lci), np.int64)
    for r in np.unique(grow):  # columns ascend within a CSR row
        sel = grow == r
        pos[sel] = rp[r] + np.searchsorted(ci[rp[r]:rp[r + 1]], gcol[sel])
    changed = vals[pos] != lval
    vals[pos] = lval
    return pos[changed], lval[changed], np.unique(gcol[changed]).astype(np.int32)


class UpdateServer:
    """Serves Requests on `plan` (an MF2D/MF3D already factored with the values in `d_val`, a
    torch device tensor kept in sync with the plan).  Box requests need the padded nodal
    coefficient field on the device (`d_coeff`) and the generator's grid spacing `h`: the
    server scales it and recomputes the reached operator values there
    (MF.box_scale_device), so no operator is rebuilt on the host."""

    def __init__(self, plan, d_val, torch, d_coeff=None, h=None):
        self.plan, self.d_val, self.torch = plan, d_val, torch
        self.d_coeff, self.h = d_coeff, h
        self.rounds = []  # (requests in the round, right-hand sides, seconds)

    def serve_round(self, reqs):
        torch = self.torch
        t0 = time.perf_counter()
        mods = []
        for r in reqs:  # arrival order: a later change of the same entry wins
            if r.box is not None:
                lo, side, f = r.box
                r.mod = self.plan.box_scale_device(self.d_coeff.data_ptr(), self.h, lo, side, f,
                                                   self.d_val.data_ptr())
            else:
                self.d_val[r.pos_d] = r.val_d
            mods.append(r.mod)
        mod = np.unique(np.concatenate(mods))
        self.plan.update_device(self.d_val.data_ptr(), mod)
        k = [r.rhs_d.shape[0] for r in reqs]
        B = torch.cat([r.rhs_d for r in reqs], dim=0)  # (sum k) x n, row = right-hand side
        X = torch.empty_like(B)
        self.plan.solve_device(B.data_ptr(), X.data_ptr(), B.shape[0])
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        o = 0
        for r, kk in zip(reqs, k):
            r.x = X[o:o + kk]
            o += kk
        self.rounds.append((len(reqs), B.shape[0], dt))
        return dt

    def run(self, reqs):
        """Virtual-clock simulation: requests arrive at r.arrival (seconds); a round starts when
        the server is free and at least one request is waiting, takes every waiting request, and
        lasts its measured GPU time.  Sets r.done; returns the latencies done - arrival."""
        torch = self.torch
        for r in reqs:
            if r.box is None:
                r.pos_d = torch.from_numpy(r.pos).cuda()
                r.val_d = torch.from_numpy(r.val).cuda()
            r.rhs_d = torch.from_numpy(np.ascontiguousarray(r.rhs)).cuda()
        order = sorted(range(len(reqs)), key=lambda i: reqs[i].arrival)
        t, i = 0.0, 0
        while i < len(order):
            t = max(t, reqs[order[i]].arrival)
            batch = []
            while i < len(order) and reqs[order[i]].arrival <= t:
                batch.append(reqs[order[i]])
                i += 1
            t += self.serve_round(batch)
            for r in batch:
                r.done = t
        return np.array([r.done - r.arrival for r in reqs])
