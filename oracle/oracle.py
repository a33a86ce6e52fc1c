"""ctypes wrapper over the CPU oracle (oracle/kmeans_oracle.cpp).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the checker.  The product path
(onedal_amd/) never imports this module.
"""

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libkmeans_oracle.so")
FAST_PATH = os.path.join(HERE, "lib", "libkmeans_cpu_fast.so")

# Order constants of the engine's FP64 reductions (must equal the ones in
# onedal_amd/csrc/kmeans_config.hpp; the tests check that they do).
REDUCE_CHUNK = 256
OBJ_CHUNK = 4096


def reduce_groups(d):
    """Interleaved partial sums per chunk used by the engine for a given d."""
    return max(1, 64 // d) if d <= 32 else 1


_lib = None
_fast = None


def build():
    subprocess.check_call(["make", "-s", "-C", HERE])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = ctypes.CDLL(LIB_PATH)
        _declare(_lib)
    return _lib


def _declare(L):
    P = ctypes.c_void_p
    I64 = ctypes.c_int64
    D = ctypes.c_double
    I = ctypes.c_int
    for sfx in ("f32", "f64"):
        f = getattr(L, "oracle_init_centroids_" + sfx)
        f.argtypes = [P, I64, I64, I64, I, ctypes.c_uint64, P]
        f.restype = None
    L.oracle_init_mix.argtypes = [ctypes.c_uint64]
    L.oracle_init_mix.restype = ctypes.c_uint64
    for sfx in ("f32", "f64"):
        f = getattr(L, "oracle_daal_lloyd_" + sfx)
        f.argtypes = [P, I64, I64, P, I64, I64, D, I, I64, I, P, P, P, P]
        f.restype = I
        f = getattr(L, "oracle_canonical_lloyd_" + sfx)
        f.argtypes = [P, I64, I64, P, I64, I64, D, I64, I64, I64, I, P, P, P, P, P]
        f.restype = I
        f = getattr(L, "oracle_canonical_assign_" + sfx)
        f.argtypes = [P, I64, I64, P, I64, P, P, I]
        f.restype = None
        f = getattr(L, "oracle_canonical_objective_" + sfx)
        f.argtypes = [P, I64, I64, P, P, I64, I]
        f.restype = D
        f = getattr(L, "oracle_canonical_sums_" + sfx)
        f.argtypes = [P, I64, I64, I64, P, I64, I64, P, P]
        f.restype = None
        f = getattr(L, "oracle_fp64_assign_" + sfx)
        f.argtypes = [P, I64, I64, P, I64, P, P, P, I]
        f.restype = None
        f = getattr(L, "oracle_csr_lloyd_" + sfx)
        f.argtypes = [P, P, P, I64, I64, I64, P, I64, I64, D, I64, I64, I64, I, P, P, P, P]
        f.restype = I
        f = getattr(L, "oracle_csr_assign_" + sfx)
        f.argtypes = [P, P, P, I64, I64, I64, P, I64, I, P, P]
        f.restype = None
        f = getattr(L, "oracle_csr_objective_" + sfx)
        f.argtypes = [P, I64, I64]
        f.restype = D


def init_centroids(x, k, method, seed):
    """The engine's kmeans_init methods restated (0 dense, 1 random_dense,
    2 plus_plus_dense)."""
    x = np.ascontiguousarray(x)
    n, d = x.shape
    c = np.empty((k, d), x.dtype)
    f = getattr(lib(), "oracle_init_centroids_" + _sfx(x))
    f(_ptr(x), n, d, k, int(method), ctypes.c_uint64(seed & (2**64 - 1)), _ptr(c))
    return c


def _sfx(x):
    if x.dtype == np.float32:
        return "f32"
    if x.dtype == np.float64:
        return "f64"
    raise TypeError("oracle supports float32/float64, got %s" % x.dtype)


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _threads(nthreads):
    return int(nthreads if nthreads else min(8, os.cpu_count() or 1))


def daal_lloyd(x, c0, max_iter, thr=0.0, stop_rule=0, block_size=512, nthreads=None):
    """Reference (DAAL) CPU Lloyd restatement. stop_rule: 0 DAAL shift, 1 GPU objective, 2 fixed."""
    x = np.ascontiguousarray(x)
    c0 = np.ascontiguousarray(c0, dtype=x.dtype)
    n, d = x.shape
    k = c0.shape[0]
    c = np.empty_like(c0)
    labels = np.empty(n, np.int32)
    obj = ctypes.c_double()
    iters = ctypes.c_int64()
    st = getattr(lib(), "oracle_daal_lloyd_" + _sfx(x))(
        _ptr(x), n, d, _ptr(c0), k, int(max_iter), float(thr), int(stop_rule), int(block_size),
        _threads(nthreads), _ptr(c), _ptr(labels), ctypes.byref(obj), ctypes.byref(iters))
    if st != 0:
        raise RuntimeError("oracle: number of clusters is too large for the candidates")
    return c, labels, obj.value, iters.value


def canonical_lloyd(x, c0, max_iter, thr=0.0, nthreads=None):
    """The engine's arithmetic contract restated on the CPU (bit-exact target)."""
    x = np.ascontiguousarray(x)
    c0 = np.ascontiguousarray(c0, dtype=x.dtype)
    n, d = x.shape
    k = c0.shape[0]
    c = np.empty_like(c0)
    labels = np.empty(n, np.int32)
    obj = ctypes.c_double()
    iters = ctypes.c_int64()
    iter_obj = np.zeros(max(1, int(max_iter)), np.float64)
    st = getattr(lib(), "oracle_canonical_lloyd_" + _sfx(x))(
        _ptr(x), n, d, _ptr(c0), k, int(max_iter), float(thr), REDUCE_CHUNK, reduce_groups(d),
        OBJ_CHUNK, _threads(nthreads), _ptr(c), _ptr(labels), ctypes.byref(obj),
        ctypes.byref(iters), _ptr(iter_obj))
    if st != 0:
        raise RuntimeError("oracle: more empty clusters than points")
    return c, labels, obj.value, iters.value, iter_obj[: iters.value]


def canonical_assign(x, c, nthreads=None):
    x = np.ascontiguousarray(x)
    c = np.ascontiguousarray(c, dtype=x.dtype)
    n, d = x.shape
    labels = np.empty(n, np.int32)
    best = np.empty(n, x.dtype)
    getattr(lib(), "oracle_canonical_assign_" + _sfx(x))(
        _ptr(x), n, d, _ptr(c), c.shape[0], _ptr(labels), _ptr(best), _threads(nthreads))
    return labels, best


def canonical_objective(x, c, labels, nthreads=None):
    x = np.ascontiguousarray(x)
    c = np.ascontiguousarray(c, dtype=x.dtype)
    labels = np.ascontiguousarray(labels, dtype=np.int32)
    n, d = x.shape
    return getattr(lib(), "oracle_canonical_objective_" + _sfx(x))(
        _ptr(x), n, d, _ptr(c), _ptr(labels), OBJ_CHUNK, _threads(nthreads))


def canonical_sums(x, labels, k):
    x = np.ascontiguousarray(x)
    labels = np.ascontiguousarray(labels, dtype=np.int32)
    n, d = x.shape
    s1 = np.empty((k, d), np.float64)
    counts = np.empty(k, np.int64)
    getattr(lib(), "oracle_canonical_sums_" + _sfx(x))(
        _ptr(x), n, d, k, _ptr(labels), REDUCE_CHUNK, reduce_groups(d), _ptr(s1), _ptr(counts))
    return s1, counts


def fp64_assign(x, c, nthreads=None):
    """Labels from exact-in-double distances plus the best and second-best distance."""
    x = np.ascontiguousarray(x)
    c = np.ascontiguousarray(c, dtype=x.dtype)
    n, d = x.shape
    labels = np.empty(n, np.int32)
    best = np.empty(n, np.float64)
    second = np.empty(n, np.float64)
    getattr(lib(), "oracle_fp64_assign_" + _sfx(x))(
        _ptr(x), n, d, _ptr(c), c.shape[0], _ptr(labels), _ptr(best), _ptr(second),
        _threads(nthreads))
    return labels, best, second


class Csr:
    """CSR rows as csr_table holds them: values, column indices and row offsets
    (int64), zero- or one-based (base 0 or 1), n rows of d columns."""

    def __init__(self, values, cols, offs, d, base=0):
        self.values = np.ascontiguousarray(values)
        self.cols = np.ascontiguousarray(cols, dtype=np.int64)
        self.offs = np.ascontiguousarray(offs, dtype=np.int64)
        self.n = int(self.offs.shape[0] - 1)
        self.d = int(d)
        self.base = int(base)
        if self.values.dtype not in (np.float32, np.float64):
            raise TypeError("CSR values must be float32 or float64")

    @classmethod
    def from_dense(cls, x, base=0):
        x = np.asarray(x)
        mask = x != 0
        offs = np.zeros(x.shape[0] + 1, np.int64)
        offs[1:] = np.cumsum(mask.sum(axis=1))
        rows, cols = np.nonzero(mask)
        return cls(x[rows, cols], cols.astype(np.int64) + base, offs + base, x.shape[1], base)

    def to_dense(self):
        x = np.zeros((self.n, self.d), self.values.dtype)
        for i in range(self.n):
            for p in range(self.offs[i] - self.base, self.offs[i + 1] - self.base):
                x[i, self.cols[p] - self.base] += self.values[p]
        return x


def csr_lloyd(x, c0, max_iter, thr=0.0, nthreads=None):
    """The engine's CSR arithmetic contract (method lloyd_csr) restated on the CPU."""
    c0 = np.ascontiguousarray(c0, dtype=x.values.dtype)
    k = c0.shape[0]
    c = np.empty_like(c0)
    labels = np.empty(x.n, np.int32)
    obj = ctypes.c_double()
    iters = ctypes.c_int64()
    st = getattr(lib(), "oracle_csr_lloyd_" + _sfx(x.values))(
        _ptr(x.values), _ptr(x.cols), _ptr(x.offs), x.base, x.n, x.d, _ptr(c0), k, int(max_iter),
        float(thr), REDUCE_CHUNK, reduce_groups(x.d), OBJ_CHUNK, _threads(nthreads), _ptr(c),
        _ptr(labels), ctypes.byref(obj), ctypes.byref(iters))
    if st != 0:
        raise RuntimeError("oracle: more empty clusters than points")
    return c, labels, obj.value, iters.value


def csr_assign(x, c, nthreads=None):
    """Labels and closest distances (expanded form over the stored entries)."""
    c = np.ascontiguousarray(c, dtype=x.values.dtype)
    labels = np.empty(x.n, np.int32)
    closest = np.empty(x.n, x.values.dtype)
    getattr(lib(), "oracle_csr_assign_" + _sfx(x.values))(
        _ptr(x.values), _ptr(x.cols), _ptr(x.offs), x.base, x.n, x.d, _ptr(c), c.shape[0],
        _threads(nthreads), _ptr(labels), _ptr(closest))
    return labels, closest


def csr_objective(closest):
    closest = np.ascontiguousarray(closest)
    return getattr(lib(), "oracle_csr_objective_" + _sfx(closest))(
        _ptr(closest), closest.shape[0], OBJ_CHUNK)


def _fast_lib():
    """The CPU baseline library (the reference's CPU algorithm with a vectorised
    product kernel; oracle/kmeans_cpu_fast.cpp)."""
    global _fast
    if _fast is None:
        if not os.path.exists(FAST_PATH):
            build()
        L = ctypes.CDLL(FAST_PATH)
        P, I64 = ctypes.c_void_p, ctypes.c_int64
        L.cpu_fast_isa.argtypes = []
        L.cpu_fast_isa.restype = ctypes.c_int
        L.cpu_fast_lloyd_f32.argtypes = [P, I64, I64, P, I64, I64, I64, ctypes.c_int, ctypes.c_int,
                                         P, P, ctypes.POINTER(ctypes.c_double)]
        L.cpu_fast_lloyd_f32.restype = ctypes.c_int
        _fast = L
    return _fast


def cpu_fast_isa():
    """Vector kernel the CPU baseline runs here: 'avx512', 'avx2' or 'portable'."""
    return {2: "avx512", 1: "avx2", 0: "portable"}[_fast_lib().cpu_fast_isa()]


def cpu_fast_lloyd(x, c0, iters, nthreads=None, block_rows=512, final=True):
    """CPU baseline: `iters` Lloyd iterations of the reference's CPU algorithm
    (float32) with a vectorised X.C^T kernel, then (final) the final assignment.
    Returns (centroids, labels, exact objective; 0 without the final pass)."""
    x = np.ascontiguousarray(x, dtype=np.float32)
    c0 = np.ascontiguousarray(c0, dtype=np.float32)
    n, d = x.shape
    k = c0.shape[0]
    c = np.empty_like(c0)
    labels = np.empty(n, np.int32)
    obj = ctypes.c_double()
    st = _fast_lib().cpu_fast_lloyd_f32(_ptr(x), n, d, _ptr(c0), k, int(iters), int(block_rows),
                                        _threads(nthreads), 1 if final else 0, _ptr(c), _ptr(labels),
                                        ctypes.byref(obj))
    if st != 0:
        raise RuntimeError("cpu baseline: more empty clusters than candidates")
    return c, labels, obj.value
