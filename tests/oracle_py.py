"""ctypes front-end to the CPU oracle (oracle/liboracle.so) and to the reference harness
(oracle/_ref/libumfref.so).  Test infrastructure: used by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg only.

Both libraries expose the same shape of C ABI (prefix ``orc_`` / ``ref_``); `Factorization` wraps
either one behind one Python interface mirroring the reference's MF2D/MF3D
(include/DE/mf.hpp:67-80, :173-188).
"""
import ctypes as C
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "liboracle.so")
# UMF_REF_SO selects another build of the harness (the threaded-MKL one for the all-cores CPU
# baseline, which runs in a process of its own: the two MKL threading layers do not mix)
REF_SO = os.environ.get("UMF_REF_SO") or os.path.join(ROOT, "oracle", "_ref", "libumfref.so")

_libs = {}

P_I = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
P_D = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")


def _load(kind):
    if kind in _libs:
        return _libs[kind]
    path = ORACLE_SO if kind == "orc" else REF_SO
    if not os.path.exists(path):
        return None
    # the harness links MKL, whose CPU-specific kernels resolve core symbols globally
    lib = C.CDLL(path, mode=C.RTLD_GLOBAL)
    p = kind + "_"
    vp = C.c_void_p
    ip = C.POINTER(C.c_int)
    dp = C.POINTER(C.c_double)
    sig = {
        "last_error": (C.c_char_p, []),
        "create": (vp, [C.c_int] * 7 + [ip, ip, dp]),
        "destroy": (None, [vp]),
        "levels": (C.c_int, [vp]),
        "leaf_width": (C.c_int, [vp, C.c_int]),
        "nboxes": (C.c_int, [vp, C.c_int]),
        "nseps": (C.c_int, [vp, C.c_int]),
        "box_dofs": (C.c_int, [vp, C.c_int, C.c_int, C.c_int, ip]),
        "sep_dofs": (C.c_int, [vp, C.c_int, C.c_int, ip]),
        "trace": (C.c_int, [vp, C.c_int, C.c_int, ip]),
        "box_block": (C.c_int, [vp, C.c_int, C.c_int, C.c_int, dp, ip, ip]),
        "sep_block": (C.c_int, [vp, C.c_int, C.c_int, dp, ip]),
        "solve": (None, [vp, dp, dp]),
        "apply": (None, [vp, dp, dp]),
    }
    if kind == "orc":
        sig["factor"] = (C.c_int, [vp])
        sig["update"] = (C.c_int, [vp, dp, C.c_int, ip])
        sig["laplacian_2d"] = (C.c_int, [C.c_int, C.c_int, dp, C.c_int, C.c_double, C.c_double, ip, ip, dp])
        sig["laplacian_3d"] = (C.c_int, [C.c_int, C.c_int, C.c_int, dp, C.c_int, C.c_double, C.c_double,
                                         C.c_double, ip, ip, dp])
        sig["moddofs"] = (C.c_int, [C.c_int, ip, dp, dp, ip])
        sig["spmv"] = (None, [C.c_int, ip, ip, dp, dp, dp])
    else:
        sig["factor"] = (C.c_int, [vp, C.c_int])
        sig["update"] = (C.c_int, [vp, ip, ip, dp, C.c_int, ip, C.c_int])
        sig["laplacian"] = (C.c_int, [C.c_int] * 4 + [dp, C.c_int, C.c_double, ip, ip, dp])
    for name, (res, args) in sig.items():
        f = getattr(lib, p + name)
        f.restype = res
        f.argtypes = args
    _libs[kind] = lib
    return lib


def have_ref():
    return _load("ref") is not None


def _ip(a):
    return a.ctypes.data_as(C.POINTER(C.c_int))


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


class Csr:
    """CSR operator with int32 indices, as the reference's Sparse (include/LinAlg/sparse.hpp:33-83)."""

    def __init__(self, rp, ci, val):
        self.rp = np.ascontiguousarray(rp, dtype=np.int32)
        self.ci = np.ascontiguousarray(ci, dtype=np.int32)
        self.val = np.ascontiguousarray(val, dtype=np.float64)
        self.n = len(self.rp) - 1

    def to_scipy(self):
        import scipy.sparse as sp
        return sp.csr_matrix((self.val, self.ci, self.rp), shape=(self.n, self.n))

    def matvec(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.empty(self.n)
        lib = _load("orc")
        lib.orc_spmv(self.n, _ip(self.rp), _ip(self.ci), _dp(self.val), _dp(x), _dp(y))
        return y


def laplacian(dims, coeff, width=1, h=1.0, impl="orc"):
    """Variable-coefficient -div(a grad u) operator on the padded nodal field `coeff`
    (shape (Nx+2w, Ny+2w[, Nz+2w]), indexed [i][j][k]) -- buildLaplacian
    (test_include/DE/buildLaplacian.hpp:10, :244)."""
    coeff = np.ascontiguousarray(coeff, dtype=np.float64)
    dim = len(dims)
    n = int(np.prod(dims))
    if impl == "ref":
        lib = _load("ref")
        nx, ny = dims[0], dims[1]
        nz = dims[2] if dim == 3 else 1
        nnz = lib.ref_laplacian(dim, nx, ny, nz, _dp(coeff), width, h, None, None, None)
        rp = np.empty(n + 1, np.int32); ci = np.empty(nnz, np.int32); v = np.empty(nnz)
        lib.ref_laplacian(dim, nx, ny, nz, _dp(coeff), width, h, _ip(rp), _ip(ci), _dp(v))
        return Csr(rp, ci, v)
    lib = _load("orc")
    if dim == 2:
        args = (dims[0], dims[1], _dp(coeff), width, h, h)
        nnz = lib.orc_laplacian_2d(*args, None, None, None)
        rp = np.empty(n + 1, np.int32); ci = np.empty(nnz, np.int32); v = np.empty(nnz)
        lib.orc_laplacian_2d(*args, _ip(rp), _ip(ci), _dp(v))
    else:
        args = (dims[0], dims[1], dims[2], _dp(coeff), width, h, h, h)
        nnz = lib.orc_laplacian_3d(*args, None, None, None)
        rp = np.empty(n + 1, np.int32); ci = np.empty(nnz, np.int32); v = np.empty(nnz)
        lib.orc_laplacian_3d(*args, _ip(rp), _ip(ci), _dp(v))
    return Csr(rp, ci, v)


def moddofs(A1, A2):
    """modDOF rule: columns of CSR entries that changed (test_src/DE/testMF2D.cpp:143-150)."""
    lib = _load("orc")
    out = np.empty(len(A1.ci), np.int32)
    k = lib.orc_moddofs(len(A1.ci), _ip(A1.ci), _dp(A1.val), _dp(A2.val), _ip(out))
    return out[:k].copy()


class Factorization:
    """MF2D/MF3D held by the oracle (impl='orc') or by the reference harness (impl='ref')."""

    def __init__(self, A, dims, width=1, max_side=4, impl="orc"):
        self.kind = impl
        self.lib = _load(impl)
        if self.lib is None:
            raise RuntimeError(f"{impl} library not built")
        self.dims = tuple(dims)
        self.dim = len(dims)
        self.A = A
        nz = dims[2] if self.dim == 3 else 1
        self._f = lambda name: getattr(self.lib, f"{impl}_{name}")
        self.h = self._f("create")(self.dim, dims[0], dims[1], nz, width, max_side, A.n,
                                   _ip(A.rp), _ip(A.ci), _dp(A.val))
        if not self.h:
            raise RuntimeError(self._f("last_error")().decode())
        self.L = self._f("levels")(self.h)

    def __del__(self):
        if getattr(self, "h", None):
            self._f("destroy")(self.h)
            self.h = None

    def leaf_width(self, a):
        return self._f("leaf_width")(self.h, a)

    def nboxes(self, ell):
        return self._f("nboxes")(self.h, ell)

    def nseps(self, ell):
        return self._f("nseps")(self.h, ell)

    def _ints(self, fn, *args):
        k = fn(self.h, *args, None)
        out = np.empty(max(k, 1), np.int32)
        fn(self.h, *args, _ip(out))
        return out[:k].copy()

    def box_dofs(self, ell, b, which):
        return self._ints(self._f("box_dofs"), ell, b, which)

    def sep_dofs(self, ell, e):
        return self._ints(self._f("sep_dofs"), ell, e)

    def trace(self, ell, which):
        return self._ints(self._f("trace"), ell, which)

    def factor(self, trace=True):
        if self.kind == "ref":
            return self._f("factor")(self.h, 1 if trace else 0)
        return self._f("factor")(self.h)

    def update(self, A2, mod, trace=True):
        mod = np.ascontiguousarray(mod, dtype=np.int32)
        self.A = A2
        if self.kind == "ref":
            return self._f("update")(self.h, _ip(A2.rp), _ip(A2.ci), _dp(A2.val), len(mod), _ip(mod),
                                     1 if trace else 0)
        r = self._f("update")(self.h, _dp(A2.val), len(mod), _ip(mod))
        if r < 0:
            raise ValueError(self._f("last_error")().decode())
        return r

    BLOCKS = {"A22": 0, "XL": 1, "XR": 2, "S": 3, "A11post": 4}

    def box_block(self, ell, b, which):
        m = C.c_int(); n = C.c_int()
        self._f("box_block")(self.h, ell, b, self.BLOCKS[which], None, C.byref(m), C.byref(n))
        out = np.zeros(max(m.value * n.value, 1))
        self._f("box_block")(self.h, ell, b, self.BLOCKS[which], _dp(out), C.byref(m), C.byref(n))
        return out[: m.value * n.value].reshape(n.value, m.value).T.copy()

    def sep_block(self, ell, e):
        k = C.c_int()
        self._f("sep_block")(self.h, ell, e, None, C.byref(k))
        out = np.zeros(max(k.value * k.value, 1))
        self._f("sep_block")(self.h, ell, e, _dp(out), C.byref(k))
        return out[: k.value * k.value].reshape(k.value, k.value).T.copy()

    def solve(self, b):
        b = np.ascontiguousarray(b, dtype=np.float64)
        x = np.empty_like(b)
        self._f("solve")(self.h, _dp(b), _dp(x))
        return x

    def apply(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.empty_like(x)
        self._f("apply")(self.h, _dp(x), _dp(y))
        return y


# ------------------------------------------------------------------ synthetic inputs
def coeff_field(dims, kind="uniform", seed=0, width=1, sigma=1.0, value=1.0):
    """Padded nodal coefficient field, shape (N+2w)^d.  kind: 'uniform' (constant `value`),
    'lognormal' (exp(sigma*z), z ~ N(0,1) iid from numpy PCG64 seeded by `seed`),
    'rand' (1 + U[0,1), the reference's 3D random field shape)."""
    shape = tuple(d + 2 * width for d in dims)
    if kind == "uniform":
        return np.full(shape, float(value))
    rng = np.random.default_rng(seed)
    if kind == "lognormal":
        return np.exp(sigma * rng.standard_normal(shape))
    if kind == "rand":
        return 1.0 + rng.random(shape)
    raise ValueError(kind)
