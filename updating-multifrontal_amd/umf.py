"""Python host mirror of the reference's multifrontal interface, over libumf_hip.so (ctypes).

`MF2D` / `MF3D` keep the reference's constructor arguments and method names
(include/DE/mf.hpp:67-80, :173-188): ``Factor()``, ``Update(A, modifiedDOFs)``, ``Solve(b)``,
``Apply(x)``.  ``Solve``/``Apply`` also accept an n x nrhs array.  Errors follow the
reference: geometry problems raise RuntimeError with the reference's message; a non-SPD pivot
(Cholesky fronts) or a zero pivot (LU fronts, for operators with asymmetric values) raises
ArithmeticError (the reference asserts on LAPACK info != 0).

The compute path is the HIP library only: if libumf_hip.so is missing this module raises at
import time rather than falling back to anything else.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libumf_hip.so")

if not os.path.exists(LIB_PATH):
    raise ImportError(f"{LIB_PATH} is not built (run __graft_entry__.build() or make -C csrc)")

_lib = C.CDLL(LIB_PATH)

UMF_OK = 0
UMF_EINVAL, UMF_EGEOM, UMF_EDEVICE, UMF_ENOMEM, UMF_EPATTERN, UMF_EUPDATEMAP, UMF_ESTATE, UMF_ENOCONV, UMF_EBUSY = \
    range(-1, -10, -1)
PRECISIONS = {"fp64": 0, "fp32": 1, "bf16x3": 2, "fp32t64": 3}  # UMF_PREC_*


class RefinementError(ArithmeticError):
    """FP32 form: the refinement stopped above the requested backward error (the returned
    solution is the last iterate)."""
BLOCKS = {"A22": 0, "XL": 1, "XR": 2, "S": 3, "A11post": 4, "L": 5}

_ip = C.POINTER(C.c_int)
_dp = C.POINTER(C.c_double)
_vp = C.c_void_p


class Stats(C.Structure):
    _fields_ = [("t_factor_ms", C.c_double), ("t_update_ms", C.c_double), ("t_solve_ms", C.c_double),
                ("flops_factor", C.c_double), ("flops_update", C.c_double), ("flops_solve", C.c_double),
                ("bytes_front_pool", C.c_double), ("n_fronts", C.c_int), ("n_fronts_update", C.c_int),
                ("ir_iters", C.c_int), ("ir_berr", C.c_double)]


def _sig(name, res, args):
    f = getattr(_lib, name)
    f.restype = res
    f.argtypes = args
    return f


_create = _sig("umf_plan_create", C.c_int, [C.c_int] * 7 + [_ip, _ip, _dp, C.POINTER(_vp)])
_analyze = _sig("umf_plan_analyze", C.c_int, [C.c_int] * 7 + [_ip, _ip, C.POINTER(_vp)])
_analyze_update = _sig("umf_analyze_update", C.c_int, [_vp, C.c_int, _ip])
_destroy = _sig("umf_plan_destroy", C.c_int, [_vp])
_factor = _sig("umf_factor", C.c_int, [_vp])
_update = _sig("umf_update", C.c_int, [_vp, _ip, _ip, _dp, C.c_int, _ip])
_update_dev = _sig("umf_update_device", C.c_int, [_vp, _vp, C.c_int, _ip])
_prof_enable = _sig("umf_profile_enable", C.c_int, [_vp, C.c_int])
_prof_reset = _sig("umf_profile_reset", C.c_int, [_vp])
_prof_select = _sig("umf_profile_select", C.c_int, [_vp, C.c_char_p])
_prof_read = _sig("umf_profile_read", C.c_int, [_vp, C.c_int, C.c_char_p, C.c_int, _dp, _ip, _dp, _dp])
_solve = _sig("umf_solve", C.c_int, [_vp, _dp, _dp, C.c_int, C.c_int])
_apply = _sig("umf_apply", C.c_int, [_vp, _dp, _dp, C.c_int, C.c_int])
_solve_dev = _sig("umf_solve_device", C.c_int, [_vp, _vp, _vp, C.c_int, C.c_int])
_solve_noalloc = _sig("umf_solve_device_noalloc", C.c_int, [_vp, _vp, _vp, C.c_int, C.c_int])
_apply_dev = _sig("umf_apply_device", C.c_int, [_vp, _vp, _vp, C.c_int, C.c_int])
_spmv_dev = _sig("umf_spmv_device", C.c_int, [_vp, _vp, _vp, C.c_int, C.c_int])
_save = _sig("umf_plan_save", C.c_int, [_vp, C.c_char_p])
_load = _sig("umf_plan_load", C.c_int, [C.c_char_p, C.POINTER(_vp)])
_box_scale_dev = _sig("umf_box_scale_device", C.c_int,
                      [_vp, _vp, C.c_double, _ip, C.c_int, C.c_double, _vp, _ip, C.c_int, _ip])
_last_error = _sig("umf_last_error", C.c_char_p, [])
_levels = _sig("umf_levels", C.c_int, [_vp])
_leaf_width = _sig("umf_leaf_width", C.c_int, [_vp, C.c_int])
_num_boxes = _sig("umf_num_boxes", C.c_int, [_vp, C.c_int])
_num_seps = _sig("umf_num_seps", C.c_int, [_vp, C.c_int])
_box_dofs = _sig("umf_box_dofs", C.c_int, [_vp, C.c_int, C.c_int, C.c_int, _ip])
_sep_dofs = _sig("umf_sep_dofs", C.c_int, [_vp, C.c_int, C.c_int, _ip])
_dirty = _sig("umf_dirty", C.c_int, [_vp, C.c_int, C.c_int, _ip])
_front_block = _sig("umf_get_front_block", C.c_int, [_vp, C.c_int, C.c_int, C.c_int, _dp, _ip, _ip])
_sep_block = _sig("umf_get_sep_block", C.c_int, [_vp, C.c_int, C.c_int, _dp, _ip])
_stats = _sig("umf_get_stats", C.c_int, [_vp, C.POINTER(Stats)])
_set_prec = _sig("umf_set_precision", C.c_int, [_vp, C.c_int, C.c_double, C.c_int])
_set_graph = _sig("umf_set_graph", C.c_int, [_vp, C.c_int])
_set_sweep_sync = _sig("umf_set_sweep_sync", C.c_int, [_vp, C.c_int])
_set_lu = _sig("umf_set_lu", C.c_int, [_vp, C.c_int])
_get_lu = _sig("umf_get_lu", C.c_int, [_vp])
_set_tune = _sig("umf_set_tune", C.c_int, [_vp, C.c_int, C.c_int])
_get_tune = _sig("umf_get_tune", C.c_int, [_vp, _ip, _ip])
_autotune = _sig("umf_autotune", C.c_int, [_vp, C.c_int, _dp])
_batch_create = _sig("umf_batch_create", C.c_int, [C.c_int, _ip, C.POINTER(_ip), C.POINTER(_ip), C.POINTER(_dp),
                                                   C.c_int, C.POINTER(_vp)])
_batch_info = _sig("umf_batch_info", C.c_int, [_vp, C.c_int, _ip, _ip, _ip, _ip])
_srv_create = _sig("umf_server_create", C.c_int, [_vp, C.c_int, C.c_double, C.POINTER(_vp)])
_srv_destroy = _sig("umf_server_destroy", C.c_int, [_vp])
_srv_footprint = _sig("umf_server_footprint", C.c_int, [_vp, C.c_int, _ip, _dp])
_srv_admit = _sig("umf_server_admit", C.c_int, [_vp, C.c_int, _ip, _ip])
_srv_release = _sig("umf_server_release", C.c_int, [_vp, C.c_int])
_srv_pool = _sig("umf_server_pool", C.c_int, [_vp, _dp, _dp])
_srv_set_values = _sig("umf_server_set_values", C.c_int, [_vp, C.c_int, _vp])
_srv_box_scale = _sig("umf_server_box_scale", C.c_int, [_vp, C.c_int, _vp, C.c_double, _ip, C.c_int, C.c_double, _ip,
                                                        C.c_int, _ip])
_srv_run = _sig("umf_server_run", C.c_int, [_vp, C.c_int, C.c_int, _vp, _vp, C.c_int, C.c_int])
_srv_stats = _sig("umf_server_stats", C.c_int, [_vp, C.c_int, C.c_void_p])

EXPORTED = ["umf_plan_create", "umf_plan_analyze", "umf_analyze_update", "umf_plan_destroy", "umf_factor", "umf_update",
            "umf_update_device", "umf_profile_enable", "umf_profile_select", "umf_profile_reset", "umf_profile_read", "umf_solve", "umf_apply",
            "umf_solve_device", "umf_solve_device_noalloc", "umf_apply_device", "umf_spmv_device", "umf_last_error", "umf_levels",
            "umf_leaf_width", "umf_num_boxes", "umf_num_seps", "umf_box_dofs", "umf_sep_dofs", "umf_dirty",
            "umf_get_front_block", "umf_get_sep_block", "umf_get_stats", "umf_stream", "umf_set_precision", "umf_set_graph",
            "umf_set_tune", "umf_get_tune", "umf_autotune", "umf_set_sweep_sync", "umf_set_lu", "umf_get_lu",
            "umf_batch_create", "umf_batch_info", "umf_box_scale_device", "umf_plan_save", "umf_plan_load",
            "umf_server_create", "umf_server_destroy", "umf_server_footprint", "umf_server_admit",
            "umf_server_release", "umf_server_pool", "umf_server_set_values", "umf_server_box_scale",
            "umf_server_run", "umf_server_stats"]


def _i(a):
    return a.ctypes.data_as(_ip)


def _d(a):
    return a.ctypes.data_as(_dp)


def _check(rc):
    if rc == UMF_OK:
        return
    msg = _last_error().decode()
    if rc > 0:
        raise ArithmeticError(msg)
    if rc == UMF_ENOCONV:
        raise RefinementError(msg)
    if rc == UMF_EGEOM:
        raise RuntimeError(msg)
    if rc in (UMF_EINVAL, UMF_EPATTERN, UMF_EUPDATEMAP):
        raise ValueError(msg)
    if rc == UMF_ENOMEM:
        raise MemoryError(msg)
    raise RuntimeError(f"umf error {rc}: {msg}")


def _csr(A):
    """Accepts an object with rp/ci/val (int32/int32/float64) or a scipy CSR matrix."""
    if hasattr(A, "indptr"):
        rp, ci, val = A.indptr, A.indices, A.data
    else:
        rp, ci, val = A.rp, A.ci, A.val
    return (np.ascontiguousarray(rp, dtype=np.int32), np.ascontiguousarray(ci, dtype=np.int32),
            np.ascontiguousarray(val, dtype=np.float64))


class _MF:
    def __init__(self, A, dims, width, maxSideLength, analyze_only=False):
        rp, ci, val = _csr(A)
        self.n = len(rp) - 1
        self.dims = tuple(dims)
        self.width = width
        h = _vp()
        nz = dims[2] if len(dims) == 3 else 1
        if analyze_only:  # host-side symbolic analysis only, no device memory
            _check(_analyze(len(dims), dims[0], dims[1], nz, width, maxSideLength, self.n, _i(rp), _i(ci),
                            C.byref(h)))
        else:
            _check(_create(len(dims), dims[0], dims[1], nz, width, maxSideLength, self.n, _i(rp), _i(ci),
                           _d(val), C.byref(h)))
        self._h = h

    def save(self, path):
        """Write the plan (geometry, current operator values, factors if factored) to `path`
        (umf_plan_save); load_plan(path) restores it without refactoring."""
        _check(_save(self._h, os.fsencode(path)))

    def analyze_update(self, modifiedDOFs):
        """Dirty sets an Update with these DOFs would factor (then read with dirty())."""
        mod = np.ascontiguousarray(sorted(set(int(d) for d in modifiedDOFs)), dtype=np.int32)
        _check(_analyze_update(self._h, len(mod), _i(mod) if len(mod) else None))

    def __del__(self, _destroy=_destroy):  # bound now: module globals may be gone at interpreter exit
        h = getattr(self, "_h", None)
        if h:
            _destroy(h)
            self._h = None

    # ---- precision form (no reference counterpart)
    def SetPrecision(self, precision="fp64", tol=1e-10, max_iter=50):
        """'fp64' (default, the reference's arithmetic) or 'fp32': FP32 fronts on FP32 MFMA, Solve
        refined in FP64 (preconditioned CG) to backward error <= tol; 'bf16x3': the fp32 form
        with the Schur and trailing-update GEMMs on BF16 MFMA over two bfloat16 pieces of each
        FP32 operand (3 piece products, FP32 accumulation); 'fp32t64': the top three separator
        levels in FP64, the levels below in FP32, Solve refined as 'fp32'.  Factor() again
        afterwards."""
        _check(_set_prec(self._h, PRECISIONS[precision], float(tol), int(max_iter)))

    def use_graph(self, on=True):
        """Run Factor() as one captured HIP graph (captured at the next Factor, then replayed;
        bit-identical to the direct launches)."""
        _check(_set_graph(self._h, 1 if on else 0))

    def set_lu(self, on=True):
        """LU fronts (partial pivoting within each front's interior rows, as the reference
        factors every front) instead of Cholesky; an operator with asymmetric values gets them
        at construction.  Factor() again afterwards."""
        _check(_set_lu(self._h, 1 if on else 0))

    @property
    def lu(self):
        """True when the fronts are factored by LU (unsymmetric operators)."""
        return _get_lu(self._h) == 1

    def sweep_sync(self, fenced):
        """Blocked solve sweeps: publish blocks with agent-scope release / acquire fences
        (fenced=True) or through write-through atomics alone (False, the default); the results
        are identical."""
        _check(_set_sweep_sync(self._h, 1 if fenced else 0))

    def set_tune(self, gw2=1024, tile=0):
        """Outer column-group width of the trailing update and GEMM tile size (0: by size)."""
        _check(_set_tune(self._h, int(gw2), int(tile)))

    def tune(self):
        g, t = C.c_int(), C.c_int()
        _check(_get_tune(self._h, C.byref(g), C.byref(t)))
        return {"gw2": g.value, "tile": t.value}

    def autotune(self, reps=2):
        """Time every candidate setting on this plan's factorization, keep the fastest (and
        refactor with it); returns the chosen settings and their factorization time."""
        ms = C.c_double()
        _check(_autotune(self._h, int(reps), C.byref(ms)))
        out = self.tune()
        out["factor_ms"] = ms.value
        return out

    # ---- the reference's public API
    def Factor(self):
        _check(_factor(self._h))

    def Update(self, A, modifiedDOFs):
        rp, ci, val = _csr(A)
        mod = np.ascontiguousarray(sorted(set(int(d) for d in modifiedDOFs)), dtype=np.int32)
        _check(_update(self._h, _i(rp), _i(ci), _d(val), len(mod), _i(mod) if len(mod) else None))

    def _sweep(self, fn, b):
        b = np.asarray(b, dtype=np.float64)
        vec = b.ndim == 1
        B = np.asfortranarray(b.reshape(self.n, -1))
        X = np.empty_like(B, order="F")
        _check(fn(self._h, _d(B), _d(X), B.shape[1], self.n))
        return X[:, 0].copy() if vec else X

    def Solve(self, b):
        return self._sweep(_solve, b)

    def Apply(self, x):
        return self._sweep(_apply, x)

    # ---- device-resident variants (raw device pointers), column-major n x nrhs
    def update_device(self, d_val, modifiedDOFs):
        """Update with new CSR values already in device memory (same pattern as planned)."""
        mod = np.ascontiguousarray(sorted(set(int(d) for d in modifiedDOFs)), dtype=np.int32)
        _check(_update_dev(self._h, d_val, len(mod), _i(mod) if len(mod) else None))

    def solve_device(self, d_b, d_x, nrhs, ld=None):
        _check(_solve_dev(self._h, d_b, d_x, nrhs, ld or self.n))

    def solve_device_noalloc(self, d_b, d_x, nrhs, ld=None):
        """solve_device through the entry point that allocates nothing (FP64 factors; the plan's
        solve must fit the library's 64 MiB per-GPU scratch array)."""
        _check(_solve_noalloc(self._h, d_b, d_x, nrhs, ld or self.n))

    def apply_device(self, d_x, d_y, nrhs, ld=None):
        _check(_apply_dev(self._h, d_x, d_y, nrhs, ld or self.n))

    def box_scale_device(self, d_coeff, h, lo, side, factor, d_val):
        """Coefficient-change request on the device (umf_box_scale_device): scale the padded
        nodal field (device pointer d_coeff) by `factor` on the box [lo, lo + side) of padded
        nodes, recompute the reached operator values in d_val (device pointer), and return the
        modified DOFs (ascending int32 array) for update_device."""
        lo = np.ascontiguousarray(lo, dtype=np.int32)
        cap = int(np.prod([min(side + 4 * self.width, n) for n in self.dims]))
        out = np.empty(max(cap, 1), np.int32)
        cnt = C.c_int()
        _check(_box_scale_dev(self._h, d_coeff, float(h), _i(lo), int(side), float(factor), d_val, _i(out), cap,
                              C.byref(cnt)))
        return out[:cnt.value].copy()

    def spmv_device(self, d_x, d_y, nrhs, ld=None):
        _check(_spmv_dev(self._h, d_x, d_y, nrhs, ld or self.n))

    # ---- structure / parity queries
    @property
    def levels(self):
        return _levels(self._h)

    def leaf_width(self, axis):
        return _leaf_width(self._h, axis)

    def nboxes(self, ell):
        return _num_boxes(self._h, ell)

    def nseps(self, ell):
        return _num_seps(self._h, ell)

    def _ints(self, fn, *args):
        k = fn(self._h, *args, None)
        if k < 0:
            _check(k)
        out = np.empty(max(k, 1), np.int32)
        fn(self._h, *args, _i(out))
        return out[:k].copy()

    def box_dofs(self, ell, box, which):
        return self._ints(_box_dofs, ell, box, which)

    def sep_dofs(self, ell, sep):
        return self._ints(_sep_dofs, ell, sep)

    def dirty(self, ell, which):
        return self._ints(_dirty, ell, which)

    def front_block(self, ell, box, which):
        r = C.c_int()
        c = C.c_int()
        _check(_front_block(self._h, ell, box, BLOCKS[which], None, C.byref(r), C.byref(c)))
        out = np.empty(max(r.value * c.value, 1))
        _check(_front_block(self._h, ell, box, BLOCKS[which], _d(out), C.byref(r), C.byref(c)))
        return out[: r.value * c.value].reshape(c.value, r.value).T.copy()

    def sep_block(self, ell, sep):
        k = C.c_int()
        _check(_sep_block(self._h, ell, sep, None, C.byref(k)))
        out = np.empty(max(k.value * k.value, 1))
        _check(_sep_block(self._h, ell, sep, _d(out), C.byref(k)))
        return out[: k.value * k.value].reshape(k.value, k.value).T.copy()

    def profile(self, on=True):
        _check(_prof_enable(self._h, 1 if on else 0))

    def profile_select(self, kernel=None):
        """Time only launches of `kernel` (base name); None = every kernel."""
        _check(_prof_select(self._h, (kernel or "").encode()))

    def profile_reset(self):
        _check(_prof_reset(self._h))

    def profile_read(self):
        """{kernel: {ms, launches, flops, bytes}} accumulated while profiling was on."""
        k = _prof_read(self._h, -1, None, 0, None, None, None, None)
        out = {}
        for i in range(k):
            name = C.create_string_buffer(64)
            ms = C.c_double(); la = C.c_int(); fl = C.c_double(); by = C.c_double()
            _prof_read(self._h, i, name, 64, C.byref(ms), C.byref(la), C.byref(fl), C.byref(by))
            out[name.value.decode()] = {"ms": ms.value, "launches": la.value, "flops": fl.value, "bytes": by.value}
        return out

    def stats(self):
        s = Stats()
        _check(_stats(self._h, C.byref(s)))
        return {f: getattr(s, f) for f, _ in Stats._fields_}


class MF2D(_MF):
    """MF2D(A, Nx, Ny, width=1, maxSideLength=4) -- include/DE/mf.hpp:67."""

    def __init__(self, A, Nx, Ny, width=1, maxSideLength=4, analyze_only=False):
        super().__init__(A, (Nx, Ny), width, maxSideLength, analyze_only)


class MF3D(_MF):
    """MF3D(A, Nx, Ny, Nz, width=1, maxSideLength=4) -- include/DE/mf.hpp:173."""

    def __init__(self, A, Nx, Ny, Nz, width=1, maxSideLength=4, analyze_only=False):
        super().__init__(A, (Nx, Ny, Nz), width, maxSideLength, analyze_only)


def load_plan(path):
    """An MF2D / MF3D restored from a file written by save() (umf_plan_load): factored if it
    was saved factored, in the same precision form."""
    import struct
    with open(path, "rb") as f:
        head = f.read(8 + 4 * 12)
    if len(head) < 8 + 4 * 12 or head[:8] != b"UMFPLAN1":
        raise ValueError(f"{path}: not a plan file")
    dim, nx, ny, nz, width, _maxs, n = struct.unpack("<7i", head[8:36])
    h = _vp()
    _check(_load(os.fsencode(path), C.byref(h)))
    cls = MF2D if dim == 2 else MF3D
    G = cls.__new__(cls)
    G.n, G.dims, G.width, G._h = n, ((nx, ny) if dim == 2 else (nx, ny, nz)), width, h
    return G


class CSR:
    """Host CSR operator (int32 row pointers / column indices, float64 values), the form every
    constructor here accepts."""

    def __init__(self, rp, ci, val, m, n):
        self.rp = np.ascontiguousarray(rp, dtype=np.int32)
        self.ci = np.ascontiguousarray(ci, dtype=np.int32)
        self.val = np.ascontiguousarray(val, dtype=np.float64)
        self.m, self.n = int(m), int(n)


def read_csr(path):
    """The reference's CSR text file (src/LinAlg/sparse.cpp:5-44): whitespace-separated
    "m n nnz", nnz values, nnz column indices, m+1 row pointers."""
    with open(path) as f:
        tok = f.read().split()
    if len(tok) < 3:
        raise ValueError("malformed CSR file header")
    m, n, nnz = int(tok[0]), int(tok[1]), int(tok[2])
    if m < 0 or n < 0 or nnz < 0 or len(tok) < 3 + 2 * nnz + m + 1:
        raise ValueError("truncated CSR file")
    val = np.array(tok[3:3 + nnz], dtype=np.float64)
    ci = np.array(tok[3 + nnz:3 + 2 * nnz], dtype=np.int64)
    rp = np.array(tok[3 + 2 * nnz:3 + 2 * nnz + m + 1], dtype=np.int64)
    nz = int(rp[-1])  # the reference takes nnz from the last row pointer
    return CSR(rp, ci[:nz], val[:nz], m, n)


def write_csr(path, A, m=None, n=None):
    """Write A (CSR object or scipy CSR) in the same text format, values to full precision."""
    rp, ci, val = _csr(A)
    m = len(rp) - 1 if m is None else m
    n = (getattr(A, "n", None) or (A.shape[1] if hasattr(A, "shape") else m)) if n is None else n
    with open(path, "w") as f:
        f.write(f"{m} {n} {int(rp[-1])}\n")
        f.write(" ".join(repr(float(v)) for v in val[:rp[-1]]) + "\n")
        f.write(" ".join(str(int(c)) for c in ci[:rp[-1]]) + "\n")
        f.write(" ".join(str(int(r)) for r in rp) + "\n")


def build_laplacian(dims, coeff, width=1, h=1.0):
    """Variable-coefficient -div(a grad u) operator on the padded nodal field `coeff`
    ((N+2w)^d, indexed [i][j][k]), face coefficient = mean of the two nodal values, Dirichlet
    boundary through the padding -- the reference harness operator
    (test_include/DE/buildLaplacian.hpp:10-132, :244-391).  Vectorized numpy; returns
    (row_ptr, col_ind, val) with ascending columns per row."""
    coeff = np.asarray(coeff, dtype=np.float64)
    dim = len(dims)
    cf = {1: [0.0, -1.0], 2: [0.0, -4.0 / 3.0, 1.0 / 12.0], 3: [0.0, -1.5, 3.0 / 20.0, -1.0 / 90.0]}[width]
    w = width
    N = tuple(dims)
    n = int(np.prod(N))
    hh = h * h
    # DOF id = i + j*Nx (+ k*Nx*Ny); grids indexed [i, j(, k)]
    idx = np.arange(n).reshape(N[::-1]).transpose()  # idx[i,j(,k)]
    inner = tuple(slice(w, w + N[a]) for a in range(dim))
    diag = np.zeros(N)

    def shifted(a, s):
        sl = list(inner)
        sl[a] = slice(w + s, w + s + N[a])
        return coeff[tuple(sl)]

    # entries in ascending column order: for axes from slowest (last) to fastest, negative
    # offsets far-to-near, then the diagonal, then positive offsets near-to-far
    order_neg = [(a, -k) for a in reversed(range(dim)) for k in range(w, 0, -1)]
    order_pos = [(a, k) for a in range(dim) for k in range(1, w + 1)]
    # the diagonal sums per k (far to near) in the harness order x, y (, z) inside each k
    for k in range(w, 0, -1):
        for a in range(dim):
            a1 = 0.5 * (shifted(a, -k) + shifted(a, -k + 1))
            a2 = 0.5 * (shifted(a, k) + shifted(a, k - 1))
            diag = diag - cf[k] * (a1 + a2) / hh
    entries = []
    for (a, s) in order_neg + [(None, 0)] + order_pos:
        if a is None:
            valid = np.ones(N, bool)
            entries.append((idx, diag, valid))
            continue
        k = abs(s)
        if s < 0:
            ah = 0.5 * (shifted(a, -k) + shifted(a, -k + 1))
        else:
            ah = 0.5 * (shifted(a, k) + shifted(a, k - 1))
        v = cf[k] * ah / hh
        pos = np.indices(N)[a] + s
        valid = (pos >= 0) & (pos < N[a])
        nb = idx + s * (1 if a == 0 else int(np.prod(N[:a])))
        entries.append((nb, v, valid))
    # assemble row by row in the entry order above (already ascending in column)
    flat = [(e[0].transpose().ravel(), e[1].transpose().ravel(), e[2].transpose().ravel()) for e in entries]
    cnt = sum(f[2].astype(np.int32) for f in flat)
    rp = np.zeros(n + 1, np.int32)
    np.cumsum(cnt, out=rp[1:])
    ci = np.empty(rp[-1], np.int32)
    val = np.empty(rp[-1])
    fill = rp[:-1].copy()
    for c_, v_, m_ in flat:
        rows = np.nonzero(m_)[0]
        ci[fill[rows]] = c_[rows]
        val[fill[rows]] = v_[rows]
        fill[rows] += 1
    return rp, ci, val


class Batch(_MF):
    """Batch(problems) -- many independent problems as one plan of the block-diagonal operator
    diag(A_0, A_1, ...); `problems` is a list of (A, dims, width, maxSideLength).  Every tree
    level of every problem runs in the same batched launches.  Solve / Apply take the stacked
    n_total x nrhs block (rows offsets[q] : offsets[q] + sizes[q] belong to problem q);
    Update(q, A_q, modifiedDOFs) changes one problem.  Per-tree structure queries take
    single-problem plans (MF2D / MF3D)."""

    def __init__(self, problems, analyze_only=False):
        desc, rps, cis, vals = [], [], [], []
        for A, dims, width, max_side in problems:
            rp, ci, val = _csr(A)
            nz = dims[2] if len(dims) == 3 else 1
            desc += [len(dims), dims[0], dims[1], nz, width, max_side, len(rp) - 1]
            rps.append(rp)
            cis.append(ci)
            vals.append(val)
        q = len(problems)
        d = np.ascontiguousarray(desc, dtype=np.int32)
        rpa = (_ip * q)(*[_i(a) for a in rps])
        cia = (_ip * q)(*[_i(a) for a in cis])
        vla = (_dp * q)(*[_d(a) for a in vals])
        h = _vp()
        _check(_batch_create(q, _i(d), rpa, cia, None if analyze_only else vla, 1 if analyze_only else 0, C.byref(h)))
        self._h = h
        self.nprob = q
        info = [self.problem(i) for i in range(q)]
        self.offsets = np.array([o for o, _, _, _ in info])
        self.sizes = np.array([n for _, n, _, _ in info])
        self.n = int(self.sizes.sum())
        self.dims = None
        # host copy of the merged block-diagonal operator (Update replaces one problem's slice)
        self._nnz0 = np.cumsum([0] + [len(c) for c in cis])
        self._rp = np.concatenate([r[:-1] + self._nnz0[i] for i, r in enumerate(rps)] + [self._nnz0[-1:]]).astype(np.int32)
        self._ci = np.concatenate([c + self.offsets[i] for i, c in enumerate(cis)]).astype(np.int32)
        self._val = np.concatenate(vals)

    def problem(self, q):
        """(first merged DOF, size, levels, level shift) of problem q."""
        o, n, l, sh = C.c_int(), C.c_int(), C.c_int(), C.c_int()
        k = _batch_info(self._h, q, C.byref(o), C.byref(n), C.byref(l), C.byref(sh))
        if k < 0:
            _check(k)
        return o.value, n.value, l.value, sh.value

    def Update(self, q, A, modifiedDOFs):
        """New values for problem q (same pattern); modifiedDOFs are problem-local."""
        _, _, val = _csr(A)
        self._val[self._nnz0[q]:self._nnz0[q + 1]] = val
        mod = np.ascontiguousarray(sorted(set(int(d) + int(self.offsets[q]) for d in modifiedDOFs)), dtype=np.int32)
        _check(_update(self._h, _i(self._rp), _i(self._ci), _d(self._val), len(mod), _i(mod) if len(mod) else None))

    def split(self, X):
        """The per-problem row blocks of a stacked n_total (x nrhs) array."""
        return [X[o:o + n] for o, n in zip(self.offsets, self.sizes)]


class Server:
    """Requests against an immutable factored plan (umf_server_*): each request is the base
    operator with a local change plus right-hand sides, answered as Update + Solve on the base
    would answer it, without changing the base.  `slots` execution contexts run requests
    concurrently (one host thread per slot); a request's dirty fronts live in a slot pool of
    `pool_bytes`, where it is admitted by footprint.  While the server lives the base plan cannot
    be refactored or updated; close() the server first."""

    def __init__(self, plan, slots=4, pool_bytes=8e9):
        h = _vp()
        _check(_srv_create(plan._h, int(slots), float(pool_bytes), C.byref(h)))
        self._h, self.plan, self.slots = h, plan, int(slots)

    @staticmethod
    def _mods(mods):
        m = np.ascontiguousarray(sorted(set(int(d) for d in mods)), dtype=np.int32)
        return m, (_i(m) if len(m) else None)

    def footprint(self, mods):
        """Slot-pool bytes a request with these modified DOFs needs."""
        m, p = self._mods(mods)
        b = C.c_double()
        _check(_srv_footprint(self._h, len(m), p, C.byref(b)))
        return b.value

    def admit(self, mods):
        """Reserve the request's footprint; a ticket, or None while the pool has no room."""
        m, p = self._mods(mods)
        t = C.c_int()
        rc = _srv_admit(self._h, len(m), p, C.byref(t))
        if rc == UMF_EBUSY:
            return None
        _check(rc)
        return t.value

    def release(self, ticket):
        _check(_srv_release(self._h, int(ticket)))

    def pool(self):
        """(total, used) bytes of the slot pool."""
        a, b = C.c_double(), C.c_double()
        _check(_srv_pool(self._h, C.byref(a), C.byref(b)))
        return a.value, b.value

    def set_values(self, slot, d_val):
        """The request's full CSR values (device pointer) as the operator of `slot`."""
        _check(_srv_set_values(self._h, int(slot), d_val))

    def box_scale(self, slot, d_coeff, h, lo, side, factor):
        """The base operator with the coefficient box change (lo, side, factor) as the operator
        of `slot`, computed on the device; returns the modified DOFs."""
        lo = np.ascontiguousarray(lo, dtype=np.int32)
        n = self.plan.n
        out = np.empty(max(n, 1), np.int32)
        cnt = C.c_int()
        _check(_srv_box_scale(self._h, int(slot), d_coeff, float(h), _i(lo), int(side), float(factor), _i(out), n,
                              C.byref(cnt)))
        return out[:cnt.value].copy()

    def run(self, slot, ticket, d_b, d_x, nrhs, ld=None):
        """Factor the ticket's dirty fronts with the slot's operator and solve (device pointers,
        column-major n x nrhs).  Releases the GIL while it runs (ctypes)."""
        _check(_srv_run(self._h, int(slot), int(ticket), d_b, d_x, int(nrhs), int(ld or self.plan.n)))

    def stats(self, slot):
        st = Stats()
        _check(_srv_stats(self._h, int(slot), C.byref(st)))
        return {f: getattr(st, f) for f, _ in Stats._fields_}

    def close(self, _destroy=None):
        if getattr(self, "_h", None):
            (_destroy or _srv_destroy)(self._h)
            self._h = None

    def __del__(self, _destroy=_srv_destroy):  # bound now: module globals may be gone at interpreter exit
        self.close(_destroy)
