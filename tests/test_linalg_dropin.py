"""CPU: the drop-in LinAlg header's free functions and CSR text format against the reference.

The drop-in header (updating-multifrontal_amd/include/LinAlg/linalg.hpp) defines the reference's
LU, LUSolve, LUSolveVec, LUSolveT, GEMM and transpose (include/LinAlg/dense.hpp:167-227,
:281-310) and its CSR text reader / writer (src/LinAlg/sparse.cpp:5-44).  A caller compiled
against the header (a small shared library built here with g++) is compared with the reference
itself, built from its sources into oracle/_ref by oracle/Makefile:

- LU: the same row interchanges (ipiv, exact) and L / U within rounding of MKL's blocked dgetrf;
- the three solves and GEMM (all four transpose combinations, beta = 0 and beta != 0, and the
  reference's no-op when a dimension is 0) within rounding;
- transpose exactly;
- files written by umf.write_csr and by Sparse::write read by the reference's Sparse(filename)
  into bit-identical arrays.
"""
import ctypes as C
import os
import shutil
import subprocess

import numpy as np
import pytest

import umf
from oracle_py import _load, coeff_field, laplacian

HERE = os.path.dirname(os.path.abspath(__file__))
INC = os.path.join(HERE, "..", "updating-multifrontal_amd", "include")

SRC = r"""
#include <cstring>
#include <stdexcept>
#include <string>
#include "LinAlg/linalg.hpp"
using namespace LinAlg;
#define API extern "C" __attribute__((visibility("default")))
static Dense<double> mk(int m, int n, const double* a) {
  Dense<double> A(m, n);
  std::memcpy(A.data(), a, sizeof(double) * (size_t)m * n);
  return A;
}
API int dr_lu(int n, double* a, int* ipiv) {
  Dense<double> A = mk(n, n, a);
  try { LU(A); } catch (const std::runtime_error&) { return 1; }
  std::memcpy(a, A.data(), sizeof(double) * (size_t)n * n);
  for (int i = 0; i < n; i++) ipiv[i] = A.ipiv()[i];
  return 0;
}
API void dr_lusolve(int n, const double* lu, const int* ipiv, int nrhs, double* b, int mode) {
  Dense<double> A = mk(n, n, lu);
  A.ipiv().assign(ipiv, ipiv + n);
  if (mode == 2) {
    Vec<double> v(n);
    std::memcpy(v.data(), b, sizeof(double) * n);
    LUSolveVec(A, v);
    std::memcpy(b, v.data(), sizeof(double) * n);
    return;
  }
  Dense<double> B = mode == 0 ? mk(n, nrhs, b) : mk(nrhs, n, b);
  if (mode == 0) LUSolve(A, B); else LUSolveT(A, B);
  std::memcpy(b, B.data(), sizeof(double) * (size_t)n * nrhs);
}
API void dr_gemm(double alpha, int am, int an, const double* a, char ta, int bm, int bn, const double* b, char tb,
             double beta, int cm, int cn, double* c) {
  Dense<double> A = mk(am, an, a), B = mk(bm, bn, b), Cm = mk(cm, cn, c);
  GEMM(alpha, A, ta, B, tb, beta, Cm);
  std::memcpy(c, Cm.data(), sizeof(double) * (size_t)cm * cn);
}
API void dr_transpose(int m, int n, const double* a, double* at) {
  Dense<double> A = mk(m, n, a), AT;
  transpose(A, AT);
  std::memcpy(at, AT.data(), sizeof(double) * (size_t)m * n);
}
API void dr_write_csr(const char* path, int n, const int* rp, const int* ci, const double* v) {
  std::vector<double> val(v, v + rp[n]);
  IdxVec col(ci, ci + rp[n]), row(rp, rp + n + 1);
  Sparse<double>(val, col, row, n, n).write(path);
}
"""

pytestmark = pytest.mark.ref

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


def d(a):
    return a.ctypes.data_as(_dp)


def i(a):
    return a.ctypes.data_as(_ip)


@pytest.fixture(scope="module")
def libs(tmp_path_factory):
    if shutil.which("g++") is None:
        pytest.skip("g++ not available")
    ref = _load("ref")
    if ref is None:
        pytest.skip("reference harness not built")
    dd = tmp_path_factory.mktemp("dropin")
    (dd / "dropin.cpp").write_text(SRC)
    so = dd / "libdropin.so"
    # hidden visibility: the reference harness (loaded RTLD_GLOBAL) exports the same C++ names
    # (LinAlg::LU, ...) for its own Dense layout; the drop-in's must bind inside this library
    subprocess.run(["g++", "-std=c++17", "-O2", "-fPIC", "-shared", "-fvisibility=hidden",
                    "-fvisibility-inlines-hidden", "-Wl,-Bsymbolic", "-I", INC, str(dd / "dropin.cpp"), "-o",
                    str(so)], check=True, capture_output=True, timeout=120)
    mine = C.CDLL(str(so))
    for lib, p in ((mine, "dr_"), (ref, "ref_dense_")):
        getattr(lib, p + "lusolve").restype = None
        getattr(lib, p + "gemm").argtypes = [C.c_double, C.c_int, C.c_int, _dp, C.c_char, C.c_int, C.c_int, _dp,
                                             C.c_char, C.c_double, C.c_int, C.c_int, _dp]
        getattr(lib, p + "gemm").restype = None
        getattr(lib, p + "transpose").restype = None
    ref.ref_read_csr.argtypes = [C.c_char_p, _ip, _ip, _ip, _ip, _dp]
    ref.ref_read_csr.restype = C.c_int
    mine.dr_write_csr.restype = None
    return mine, ref


def colmajor(a):
    return np.asfortranarray(a).ravel(order="F").copy()


@pytest.mark.parametrize("n", [1, 5, 37, 128])
def test_lu_and_solves(libs, n):
    mine, ref = libs
    rng = np.random.default_rng(n)
    A = rng.standard_normal((n, n)) + 0.1 * np.eye(n)
    a1, a2 = colmajor(A), colmajor(A)
    p1, p2 = np.zeros(n, np.int32), np.zeros(n, np.int32)
    assert mine.dr_lu(n, d(a1), i(p1)) == 0
    ref.ref_dense_lu(n, d(a2), i(p2))
    assert np.array_equal(p1, p2)  # the same pivot rows
    scale = np.abs(A).max()
    assert np.abs(a1 - a2).max() <= 1e-12 * scale * n
    # P A = L U holds for the drop-in's factors
    LU = a1.reshape(n, n, order="F")
    L, U = np.tril(LU, -1) + np.eye(n), np.triu(LU)
    PA = A.copy()
    for k in range(n):
        PA[[k, p1[k] - 1]] = PA[[p1[k] - 1, k]]
    assert np.abs(PA - L @ U).max() <= 1e-13 * scale * n
    for mode, shape in ((0, (n, 3)), (1, (4, n)), (2, (n, 1))):
        B = rng.standard_normal(shape)
        b1, b2 = colmajor(B), colmajor(B)
        nrhs = shape[1] if mode == 0 else shape[0]
        mine.dr_lusolve(n, d(a1), i(p1), nrhs, d(b1), mode)
        ref.ref_dense_lusolve(n, d(a2), i(p2), nrhs, d(b2), mode)
        X = b1.reshape(shape, order="F")
        resid = (A @ X - B) if mode != 1 else (X @ A - B)  # LUSolveT solves X A = B
        assert np.abs(resid).max() <= 1e-11 * max(1.0, np.abs(B).max()) * n
        assert np.abs(b1 - b2).max() <= 1e-9 * max(1.0, np.abs(b2).max())


def test_lu_singular_raises(libs):
    mine, _ = libs
    a = np.zeros(9)
    assert mine.dr_lu(3, d(a), i(np.zeros(3, np.int32))) == 1  # the reference asserts (info != 0)


@pytest.mark.parametrize("ta,tb", [("N", "N"), ("T", "N"), ("N", "T"), ("t", "t")])
@pytest.mark.parametrize("beta", [0.0, -0.5])
def test_gemm(libs, ta, tb, beta):
    mine, ref = libs
    rng = np.random.default_rng(7)
    m, n, k = 9, 6, 13
    A = rng.standard_normal((m, k) if ta == "N" else (k, m))
    B = rng.standard_normal((k, n) if tb == "N" else (n, k))
    Cm = rng.standard_normal((m, n))
    c1, c2 = colmajor(Cm), colmajor(Cm)
    args = lambda c: (1.5, A.shape[0], A.shape[1], d(colmajor(A)), ta.encode(), B.shape[0], B.shape[1],
                      d(colmajor(B)), tb.encode(), beta, m, n, d(c))
    mine.dr_gemm(*args(c1))
    ref.ref_dense_gemm(*args(c2))
    opA = A if ta == "N" else A.T
    opB = B if tb == "N" else B.T
    want = 1.5 * opA @ opB + beta * Cm
    assert np.abs(c1 - colmajor(want)).max() <= 1e-13 * np.abs(want).max() * k
    assert np.abs(c1 - c2).max() <= 1e-13 * np.abs(want).max() * k


def test_gemm_empty_is_noop(libs):
    mine, ref = libs
    Cm = np.arange(6.0)
    c1, c2 = Cm.copy(), Cm.copy()
    z = np.zeros(1)
    for lib, c in ((mine.dr_gemm, c1), (ref.ref_dense_gemm, c2)):
        lib(1.0, 3, 0, d(z), b"N", 0, 2, d(z), b"N", 0.0, 3, 2, d(c))  # k = 0: C untouched
    assert np.array_equal(c1, Cm) and np.array_equal(c2, Cm)


def test_transpose(libs):
    mine, ref = libs
    A = np.random.default_rng(1).standard_normal((5, 8))
    t1, t2 = np.zeros(40), np.zeros(40)
    mine.dr_transpose(5, 8, d(colmajor(A)), d(t1))
    ref.ref_dense_transpose(5, 8, d(colmajor(A)), d(t2))
    assert np.array_equal(t1, t2) and np.array_equal(t1, colmajor(A.T))


def read_with_reference(ref, path):
    m, n = C.c_int(), C.c_int()
    nnz = ref.ref_read_csr(path.encode(), C.byref(m), C.byref(n), None, None, None)
    assert nnz >= 0
    rp = np.zeros(m.value + 1, np.int32)
    ci = np.zeros(nnz, np.int32)
    val = np.zeros(nnz)
    ref.ref_read_csr(path.encode(), C.byref(m), C.byref(n), i(rp), i(ci), d(val))
    return m.value, n.value, rp, ci, val


@pytest.mark.parametrize("dims,kind", [((23, 17), "lognormal"), ((5, 6, 7), "rand")])
def test_csr_files_read_by_reference(libs, tmp_path, dims, kind):
    """Both writers' files through the reference's own Sparse(filename)."""
    mine, ref = libs
    A = laplacian(dims, coeff_field(dims, kind, seed=3))
    p1, p2 = str(tmp_path / "py.dat"), str(tmp_path / "cpp.dat")
    umf.write_csr(p1, A)
    mine.dr_write_csr(p2.encode(), A.n, i(A.rp), i(A.ci), d(A.val))
    for p in (p1, p2):
        m, n, rp, ci, val = read_with_reference(ref, p)
        assert (m, n) == (A.n, A.n)
        assert np.array_equal(rp, A.rp) and np.array_equal(ci, A.ci)
        assert np.array_equal(val, A.val)  # 17 significant digits: exact


def test_reference_written_format_read_by_dropin(libs, tmp_path):
    """The reference reads what umf.read_csr reads: a hand-written file in the documented layout
    (m n nnz / values / columns / row pointers, any whitespace) gives the same arrays on both."""
    _, ref = libs
    p = tmp_path / "hand.dat"
    p.write_text("3 4\n5\n1.5 -2 3e-3 4 5\n0 2 1 3 0\n0 2\n3 5\n")
    m, n, rp, ci, val = read_with_reference(ref, str(p))
    B = umf.read_csr(str(p))
    assert (m, n) == (B.m, B.n) == (3, 4)
    assert np.array_equal(rp, B.rp) and np.array_equal(ci, B.ci) and np.array_equal(val, B.val)
