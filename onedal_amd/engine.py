"""ctypes binding of libkmeans_hip.so (include/onedal_amd/kmeans_hip.h).

The Python layer is plumbing for tests and benchmarks: every computation runs
in the HIP library.  Importing works without a GPU; creating an Engine needs
the built library and a visible device and raises otherwise.
"""

import ctypes
import os

import numpy as np

from .build import LIB_PATH

KM_OK = 0
KM_DOMAIN_ERROR = 1
KM_INVALID_ARGUMENT = 2
KM_DEVICE_BAD_ALLOC = 3
KM_HOST_BAD_ALLOC = 4
KM_UNIMPLEMENTED = 5
KM_INTERNAL_ERROR = 6
KM_PENDING = 7

KM_FLAG_EXACT_OBJECTIVE = 0x1
KM_FLAG_BF16_REFINE = 0x2
KM_FLAG_DEVICE_POINTERS = 0x4
KM_FLAG_GRAPH = 0x8

KM_INIT_DENSE, KM_INIT_RANDOM_DENSE, KM_INIT_PLUS_PLUS_DENSE, KM_INIT_PARALLEL_PLUS_DENSE = 0, 1, 2, 3

KM_OPT_ASSIGN_PATH = 1
KM_OPT_TIMING = 2
KM_OPT_BATCH_TILE = 3
KM_BATCH_TILE_VARIANTS = 8
KM_BATCH_TILE_BUCKETS = 42

PATH_AUTO, PATH_GENERIC, PATH_MFMA_F32 = 0, 1, 2

EXPORTED_SYMBOLS = [
    "km_context_create", "km_context_destroy", "km_last_error", "km_context_set_option",
    "km_get_stats", "km_lloyd_train_f32", "km_lloyd_train_f64", "km_lloyd_infer_f32",
    "km_lloyd_infer_f64", "km_lloyd_infer_f32_noalloc", "km_device_alloc", "km_device_free", "km_memcpy",
    "km_init_centroids_f32", "km_init_centroids_f64", "km_init_centroids_csr_f32",
    "km_init_centroids_csr_f64",
    "km_lloyd_train_csr_f32", "km_lloyd_train_csr_f64", "km_lloyd_infer_csr_f32",
    "km_lloyd_infer_csr_f64", "km_measure_bf16_mfma_peak",
    "km_batch_train_f32", "km_batch_autotune", "km_batch_step_f32",
    "km_server_create",
    "km_server_destroy", "km_server_last_error", "km_server_submit_fit_f32",
    "km_server_submit_infer_f32", "km_server_wait", "km_server_get_stats",
]


class KmBatchInstance(ctypes.Structure):
    """km_batch_instance_f32 (include/onedal_amd/kmeans_hip.h)."""
    _fields_ = [("X", ctypes.c_void_p), ("C0", ctypes.c_void_p), ("C_out", ctypes.c_void_p),
                ("labels_out", ctypes.c_void_p), ("n", ctypes.c_int64), ("d", ctypes.c_int64),
                ("k", ctypes.c_int64)]


class KmServerConfig(ctypes.Structure):
    """km_server_config (include/onedal_amd/kmeans_hip.h)."""
    _fields_ = [("device", ctypes.c_int32), ("max_batch_jobs", ctypes.c_int32),
                ("hbm_budget_bytes", ctypes.c_int64), ("large_job_work", ctypes.c_int64),
                ("autotune", ctypes.c_int32), ("reserved", ctypes.c_int32)]


class KmJobResult(ctypes.Structure):
    """km_job_result."""
    _fields_ = [("iterations", ctypes.c_int64), ("objective", ctypes.c_double),
                ("queue_ms", ctypes.c_double), ("latency_ms", ctypes.c_double),
                ("lane", ctypes.c_int32), ("reserved", ctypes.c_int32)]


class KmServerStats(ctypes.Structure):
    """km_server_stats."""
    _fields_ = [(name, ctypes.c_int64) for name in (
        "submitted", "completed", "failed", "queued", "steps", "step_jobs", "max_step_jobs",
        "large_jobs", "bytes_admitted", "peak_bytes_admitted")]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


class KmStats(ctypes.Structure):
    _fields_ = [
        ("assign_launches", ctypes.c_int64),
        ("assign_kernel_ms", ctypes.c_double),
        ("iterations_ms", ctypes.c_double),
        ("refined_candidate_rows", ctypes.c_int64),
        ("refined_full_rows", ctypes.c_int64),
        ("empty_cluster_events", ctypes.c_int64),
        ("assign_path", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
        ("refined_collect_rows", ctypes.c_int64),
    ]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_ if name != "reserved"}


class EngineError(RuntimeError):
    def __init__(self, status, message):
        super().__init__("km status %d: %s" % (status, message))
        self.status = status
        self.message = message


class DomainError(EngineError):
    pass


class InvalidArgument(EngineError):
    pass


class Unimplemented(EngineError):
    pass


_ERRORS = {KM_DOMAIN_ERROR: DomainError, KM_INVALID_ARGUMENT: InvalidArgument,
           KM_UNIMPLEMENTED: Unimplemented}

_lib = None


def load_library():
    """Load libkmeans_hip.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError("libkmeans_hip.so is not built (%s); run python -m onedal_amd.build"
                           % LIB_PATH)
    # PyTorch-ROCm bundles its own HIP runtime under the same soname as the one the
    # engine links.  Loaded first, the engine would map the system runtime and torch
    # would later map its own copy, which then finds no device.  Importing torch
    # first (when it is installed) leaves one HIP runtime in the process.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    L = ctypes.CDLL(LIB_PATH)
    P, I64, D, U32, I = ctypes.c_void_p, ctypes.c_int64, ctypes.c_double, ctypes.c_uint32, ctypes.c_int
    L.km_context_create.argtypes = [I, ctypes.POINTER(P)]
    L.km_context_create.restype = I
    L.km_context_destroy.argtypes = [P]
    L.km_context_destroy.restype = I
    L.km_last_error.argtypes = [P]
    L.km_last_error.restype = ctypes.c_char_p
    L.km_context_set_option.argtypes = [P, I, I64]
    L.km_context_set_option.restype = I
    L.km_get_stats.argtypes = [P, ctypes.POINTER(KmStats)]
    L.km_get_stats.restype = I
    for sfx in ("f32", "f64"):
        f = getattr(L, "km_lloyd_train_" + sfx)
        f.argtypes = [P, P, I64, I64, P, I64, I64, D, U32, P, P, P, P]
        f.restype = I
        f = getattr(L, "km_lloyd_infer_" + sfx)
        f.argtypes = [P, P, I64, I64, P, I64, U32, P, P]
        f.restype = I
        f = getattr(L, "km_lloyd_train_csr_" + sfx)
        f.argtypes = [P, P, P, P, I64, I64, I64, ctypes.c_int32, P, I64, I64, D, U32, P, P, P, P]
        f.restype = I
        f = getattr(L, "km_lloyd_infer_csr_" + sfx)
        f.argtypes = [P, P, P, P, I64, I64, I64, ctypes.c_int32, P, I64, U32, P, P]
        f.restype = I
    L.km_lloyd_infer_f32_noalloc.argtypes = [P, P, I64, I64, P, I64, U32, P, P]
    L.km_lloyd_infer_f32_noalloc.restype = I
    L.km_batch_train_f32.argtypes = [P, ctypes.POINTER(KmBatchInstance), I64, I64, U32,
                                     ctypes.POINTER(D), ctypes.POINTER(I64)]
    L.km_batch_train_f32.restype = I
    for sfx in ("f32", "f64"):
        f = getattr(L, "km_init_centroids_" + sfx)
        f.argtypes = [P, P, I64, I64, I64, ctypes.c_int32, ctypes.c_uint64, U32, P]
        f.restype = I
        f = getattr(L, "km_init_centroids_csr_" + sfx)
        f.argtypes = [P, P, P, P, I64, I64, I64, ctypes.c_int32, I64, ctypes.c_int32,
                      ctypes.c_uint64, U32, P]
        f.restype = I
    L.km_measure_bf16_mfma_peak.argtypes = [P, I64, ctypes.POINTER(D), ctypes.POINTER(D)]
    L.km_measure_bf16_mfma_peak.restype = I
    L.km_batch_autotune.argtypes = [P, ctypes.POINTER(ctypes.c_int32), I64]
    L.km_batch_autotune.restype = I
    L.km_batch_step_f32.argtypes = [P, ctypes.POINTER(KmBatchInstance), I64, I64,
                                    ctypes.POINTER(KmBatchInstance), I64, U32, ctypes.POINTER(D)]
    L.km_batch_step_f32.restype = I
    L.km_server_create.argtypes = [ctypes.POINTER(KmServerConfig), ctypes.POINTER(P)]
    L.km_server_create.restype = I
    L.km_server_destroy.argtypes = [P]
    L.km_server_destroy.restype = I
    L.km_server_last_error.argtypes = []
    L.km_server_last_error.restype = ctypes.c_char_p
    L.km_server_submit_fit_f32.argtypes = [P, P, I64, I64, P, I64, I64, U32, P, P,
                                           ctypes.POINTER(I64)]
    L.km_server_submit_fit_f32.restype = I
    L.km_server_submit_infer_f32.argtypes = [P, P, I64, I64, P, I64, U32, P, ctypes.POINTER(I64)]
    L.km_server_submit_infer_f32.restype = I
    L.km_server_wait.argtypes = [P, I64, I64, ctypes.POINTER(KmJobResult)]
    L.km_server_wait.restype = I
    L.km_server_get_stats.argtypes = [P, ctypes.POINTER(KmServerStats)]
    L.km_server_get_stats.restype = I
    _lib = L
    return L


def _sfx(dtype):
    dtype = np.dtype(dtype)
    if dtype == np.float32:
        return "f32"
    if dtype == np.float64:
        return "f64"
    raise TypeError("k-means supports float32 and float64 data, got %s" % dtype)


class _Dev:
    """Device-pointer view of a torch CUDA tensor (duck-typed, no torch import)."""

    def __init__(self, t):
        self.ptr = t.data_ptr()
        self.shape = tuple(t.shape)
        self.dtype = np.dtype(str(t.dtype).replace("torch.", ""))


def _is_torch_cuda(a):
    return hasattr(a, "data_ptr") and getattr(a, "is_cuda", False)


def _check_csr(values, cols, offs):
    """True for device arrays (all three CUDA tensors: contiguous, int64 indices),
    False for host arrays; array lengths must agree (values and column indices of
    one length, at least one row offset)."""
    on_dev = [_is_torch_cuda(a) for a in (values, cols, offs)]
    if any(on_dev) and not all(on_dev):
        raise TypeError("CSR arrays must all be host arrays or all CUDA tensors")
    if len(values.shape) != 1 or len(cols.shape) != 1 or len(offs.shape) != 1:
        raise ValueError("CSR arrays must be one-dimensional")
    if int(values.shape[0]) != int(cols.shape[0]):
        raise ValueError("values and column indices must have the same length")
    if int(offs.shape[0]) < 2:
        raise ValueError("row offsets need at least two entries")
    if all(on_dev):
        import torch

        if cols.dtype != torch.int64 or offs.dtype != torch.int64:
            raise TypeError("device column indices and row offsets must be int64 tensors")
        if not (values.is_contiguous() and cols.is_contiguous() and offs.is_contiguous()):
            raise ValueError("device CSR arrays must be contiguous")
        return True
    return False


def _check_centroids(c, d, dtype, dev, device=None):
    """Centroid arrays are read as k x d elements of the data's type: require a
    2-D (k, d) array of that dtype (host) or a contiguous CUDA tensor of that
    dtype on the data's device (dev), before any pointer reaches the engine."""
    if dev != _is_torch_cuda(c):
        raise TypeError("centroids must be a CUDA tensor exactly when the data is one")
    if not dev:
        c = np.asarray(c)
    if len(c.shape) != 2 or int(c.shape[1]) != int(d):
        raise ValueError("centroids must have shape (k, %d), got %s" % (int(d), tuple(c.shape)))
    if dev:
        if str(c.dtype).replace("torch.", "") != np.dtype(dtype).name:
            raise TypeError("centroids must have the data's type %s, got %s" % (np.dtype(dtype).name, c.dtype))
        if not c.is_contiguous():
            raise ValueError("device centroids must be contiguous")
        if device is not None and c.device != device:
            raise ValueError("centroids must be on the data's device %s, got %s" % (device, c.device))


def _check_device_rows(x):
    """Device data: a contiguous 2-D float32 or float64 CUDA tensor."""
    if len(x.shape) != 2:
        raise ValueError("data must be two-dimensional, got shape %s" % (tuple(x.shape),))
    if not x.is_contiguous():
        raise ValueError("device data must be contiguous")
    _sfx(str(x.dtype).replace("torch.", ""))


class Engine:
    """One HIP device context (km_ctx) and its stream."""

    def __init__(self, device=0):
        self.lib = load_library()
        ctx = ctypes.c_void_p()
        st = self.lib.km_context_create(int(device), ctypes.byref(ctx))
        if st != KM_OK:
            raise EngineError(st, "km_context_create(%d) failed (no usable HIP device?)" % device)
        self.ctx = ctx

    def close(self):
        if self.ctx:
            self.lib.km_context_destroy(self.ctx)
            self.ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _check(self, st):
        if st != KM_OK:
            msg = self.lib.km_last_error(self.ctx).decode()
            raise _ERRORS.get(st, EngineError)(st, msg)

    def set_option(self, option, value):
        self._check(self.lib.km_context_set_option(self.ctx, int(option), int(value)))

    def measure_bf16_mfma_peak(self, iters=20000):
        """(TFLOP/s, shader GHz) of a pure BF16 MFMA loop on this device."""
        tf = ctypes.c_double()
        ghz = ctypes.c_double()
        self._check(self.lib.km_measure_bf16_mfma_peak(self.ctx, int(iters), ctypes.byref(tf),
                                                       ctypes.byref(ghz)))
        return tf.value, ghz.value

    def stats(self):
        s = KmStats()
        self._check(self.lib.km_get_stats(self.ctx, ctypes.byref(s)))
        return s.as_dict()

    # -- host (numpy) entry points -------------------------------------------------
    def train(self, x, c0, max_iter, accuracy_threshold=0.0, bf16=False):
        """Host arrays in, host arrays out: (centroids, labels, objective, iterations)."""
        if _is_torch_cuda(x):
            return self.train_device(x, c0, max_iter, accuracy_threshold, bf16)
        x = np.ascontiguousarray(x)
        sfx = _sfx(x.dtype)
        n, d = x.shape
        _check_centroids(c0, d, x.dtype, False)
        c0 = np.ascontiguousarray(c0, dtype=x.dtype)
        k = c0.shape[0]
        c = np.empty((k, d), x.dtype)
        labels = np.empty(n, np.int32)
        obj = ctypes.c_double()
        iters = ctypes.c_int64()
        flags = KM_FLAG_BF16_REFINE if bf16 else 0
        st = getattr(self.lib, "km_lloyd_train_" + sfx)(
            self.ctx, x.ctypes.data, n, d, c0.ctypes.data, k, int(max_iter),
            float(accuracy_threshold), flags, c.ctypes.data, labels.ctypes.data,
            ctypes.byref(obj), ctypes.byref(iters))
        self._check(st)
        return c, labels, obj.value, iters.value

    def infer(self, x, c, objective=False, bf16=False):
        if _is_torch_cuda(x):
            return self.infer_device(x, c, objective, bf16)
        x = np.ascontiguousarray(x)
        sfx = _sfx(x.dtype)
        n, d = x.shape
        _check_centroids(c, d, x.dtype, False)
        c = np.ascontiguousarray(c, dtype=x.dtype)
        labels = np.empty(n, np.int32)
        obj = ctypes.c_double()
        flags = (KM_FLAG_EXACT_OBJECTIVE if objective else 0) | (KM_FLAG_BF16_REFINE if bf16 else 0)
        st = getattr(self.lib, "km_lloyd_infer_" + sfx)(
            self.ctx, x.ctypes.data, n, d, c.ctypes.data, c.shape[0], flags, labels.ctypes.data,
            ctypes.byref(obj))
        self._check(st)
        return labels, (obj.value if objective else None)

    # -- CSR rows (method lloyd_csr) ------------------------------------------------
    def train_csr(self, values, cols, offs, d, c0, max_iter, accuracy_threshold=0.0, base=0):
        """CSR arrays in (host numpy, or CUDA tensors on this device: then the
        outputs are CUDA tensors too): (centroids, labels, objective, iterations)."""
        dev = _check_csr(values, cols, offs)
        _check_centroids(c0, d, str(values.dtype).replace("torch.", ""), dev,
                         values.device if dev else None)
        n = int(offs.shape[0]) - 1
        nnz = int(values.shape[0])
        k = int(c0.shape[0])
        obj = ctypes.c_double()
        iters = ctypes.c_int64()
        if dev:
            import torch

            sfx = _sfx(str(values.dtype).replace("torch.", ""))
            c = torch.empty((k, d), dtype=values.dtype, device=values.device)
            labels = torch.empty(n, dtype=torch.int32, device=values.device)
            ptrs = [values.data_ptr(), cols.data_ptr(), offs.data_ptr(), c0.data_ptr(),
                    c.data_ptr(), labels.data_ptr()]
            flags = KM_FLAG_DEVICE_POINTERS
        else:
            values = np.ascontiguousarray(values)
            sfx = _sfx(values.dtype)
            cols = np.ascontiguousarray(cols, dtype=np.int64)
            offs = np.ascontiguousarray(offs, dtype=np.int64)
            c0 = np.ascontiguousarray(c0, dtype=values.dtype)
            c = np.empty((k, d), values.dtype)
            labels = np.empty(n, np.int32)
            ptrs = [values.ctypes.data, cols.ctypes.data, offs.ctypes.data, c0.ctypes.data,
                    c.ctypes.data, labels.ctypes.data]
            flags = 0
        st = getattr(self.lib, "km_lloyd_train_csr_" + sfx)(
            self.ctx, ptrs[0], ptrs[1], ptrs[2], n, int(d), nnz, int(base), ptrs[3], k,
            int(max_iter), float(accuracy_threshold), flags, ptrs[4], ptrs[5], ctypes.byref(obj),
            ctypes.byref(iters))
        self._check(st)
        return c, labels, obj.value, iters.value

    def infer_csr(self, values, cols, offs, d, c, objective=False, base=0):
        """Labels (and the objective if asked) of CSR rows against centroids c."""
        dev = _check_csr(values, cols, offs)
        _check_centroids(c, d, str(values.dtype).replace("torch.", ""), dev,
                         values.device if dev else None)
        n = int(offs.shape[0]) - 1
        nnz = int(values.shape[0])
        k = int(c.shape[0])
        obj = ctypes.c_double()
        flags = KM_FLAG_EXACT_OBJECTIVE if objective else 0
        if dev:
            import torch

            sfx = _sfx(str(values.dtype).replace("torch.", ""))
            labels = torch.empty(n, dtype=torch.int32, device=values.device)
            ptrs = [values.data_ptr(), cols.data_ptr(), offs.data_ptr(), c.data_ptr(),
                    labels.data_ptr()]
            flags |= KM_FLAG_DEVICE_POINTERS
        else:
            values = np.ascontiguousarray(values)
            sfx = _sfx(values.dtype)
            cols = np.ascontiguousarray(cols, dtype=np.int64)
            offs = np.ascontiguousarray(offs, dtype=np.int64)
            c = np.ascontiguousarray(c, dtype=values.dtype)
            labels = np.empty(n, np.int32)
            ptrs = [values.ctypes.data, cols.ctypes.data, offs.ctypes.data, c.ctypes.data,
                    labels.ctypes.data]
        st = getattr(self.lib, "km_lloyd_infer_csr_" + sfx)(
            self.ctx, ptrs[0], ptrs[1], ptrs[2], n, int(d), nnz, int(base), ptrs[3], k, flags,
            ptrs[4], ctypes.byref(obj))
        self._check(st)
        return labels, (obj.value if objective else None)

    # -- device (torch CUDA tensor) entry points -----------------------------------
    def train_device(self, x, c0, max_iter, accuracy_threshold=0.0, bf16=False, out=None):
        """x, c0: contiguous CUDA tensors on this context's device.  Returns device
        tensors (centroids, labels) plus objective and iteration count."""
        import torch

        _check_device_rows(x)
        sfx = _sfx(str(x.dtype).replace("torch.", ""))
        n, d = x.shape
        _check_centroids(c0, d, str(x.dtype).replace("torch.", ""), True, x.device)
        k = c0.shape[0]
        if out is None:
            c = torch.empty((k, d), dtype=x.dtype, device=x.device)
            labels = torch.empty(n, dtype=torch.int32, device=x.device)
        else:
            c, labels = out
        obj = ctypes.c_double()
        iters = ctypes.c_int64()
        flags = KM_FLAG_DEVICE_POINTERS | (KM_FLAG_BF16_REFINE if bf16 else 0)
        st = getattr(self.lib, "km_lloyd_train_" + sfx)(
            self.ctx, x.data_ptr(), n, d, c0.data_ptr(), k, int(max_iter),
            float(accuracy_threshold), flags, c.data_ptr(), labels.data_ptr(),
            ctypes.byref(obj), ctypes.byref(iters))
        self._check(st)
        return c, labels, obj.value, iters.value

    def train_batch_device(self, xs, c0s, max_iter, graph=False, outs=None):
        """Batched training of independent instances: xs[i] (n_i x d_i) and c0s[i]
        (k_i x d_i) contiguous float32 CUDA tensors.  Returns lists of device
        centroids and labels and host lists of objectives and iteration counts."""
        import torch

        count = len(xs)
        arr = (KmBatchInstance * count)()
        cs, ls = [], []
        for i, (x, c0) in enumerate(zip(xs, c0s)):
            _check_device_rows(x)
            if x.dtype != torch.float32:
                raise TypeError("batched training takes float32 data")
            n, d = x.shape
            _check_centroids(c0, d, "float32", True, x.device)
            k = c0.shape[0]
            if outs is None:
                c = torch.empty((k, d), dtype=torch.float32, device=x.device)
                lab = torch.empty(n, dtype=torch.int32, device=x.device)
            else:
                c, lab = outs[0][i], outs[1][i]
            cs.append(c)
            ls.append(lab)
            arr[i] = KmBatchInstance(x.data_ptr(), c0.data_ptr(), c.data_ptr(), lab.data_ptr(),
                                     n, d, k)
        obj = (ctypes.c_double * count)()
        its = (ctypes.c_int64 * count)()
        flags = KM_FLAG_DEVICE_POINTERS | (KM_FLAG_GRAPH if graph else 0)
        st = self.lib.km_batch_train_f32(self.ctx, arr, count, int(max_iter), flags, obj, its)
        self._check(st)
        return cs, ls, list(obj), list(its)

    def init_centroids(self, x, k, method, seed=777):
        """Initial centroids (KM_INIT_* method) of numpy rows or a CUDA tensor
        (then the result is a CUDA tensor)."""
        if _is_torch_cuda(x):
            import torch

            assert x.is_contiguous()
            c = torch.empty((k, x.shape[1]), dtype=x.dtype, device=x.device)
            sfx = _sfx(str(x.dtype).replace("torch.", ""))
            st = getattr(self.lib, "km_init_centroids_" + sfx)(
                self.ctx, x.data_ptr(), x.shape[0], x.shape[1], int(k), int(method),
                ctypes.c_uint64(seed & (2**64 - 1)), KM_FLAG_DEVICE_POINTERS, c.data_ptr())
            self._check(st)
            return c
        x = np.ascontiguousarray(x)
        sfx = _sfx(x.dtype)
        c = np.empty((k, x.shape[1]), x.dtype)
        st = getattr(self.lib, "km_init_centroids_" + sfx)(
            self.ctx, x.ctypes.data, x.shape[0], x.shape[1], int(k), int(method),
            ctypes.c_uint64(seed & (2**64 - 1)), 0, c.ctypes.data)
        self._check(st)
        return c

    def init_centroids_csr(self, values, cols, offs, d, k, method, seed=777, base=0):
        """Initial centroids (KM_INIT_* method) of CSR rows, equal to init_centroids
        on their dense expansion; host arrays in, numpy out, or CUDA tensors in,
        a CUDA tensor out."""
        dev = _check_csr(values, cols, offs)
        n = int(offs.shape[0]) - 1
        nnz = int(values.shape[0])
        if dev:
            import torch

            sfx = _sfx(str(values.dtype).replace("torch.", ""))
            c = torch.empty((k, d), dtype=values.dtype, device=values.device)
            ptrs = [values.data_ptr(), cols.data_ptr(), offs.data_ptr(), c.data_ptr()]
            flags = KM_FLAG_DEVICE_POINTERS
        else:
            values = np.ascontiguousarray(values)
            sfx = _sfx(values.dtype)
            cols = np.ascontiguousarray(cols, dtype=np.int64)
            offs = np.ascontiguousarray(offs, dtype=np.int64)
            c = np.empty((k, d), values.dtype)
            ptrs = [values.ctypes.data, cols.ctypes.data, offs.ctypes.data, c.ctypes.data]
            flags = 0
        st = getattr(self.lib, "km_init_centroids_csr_" + sfx)(
            self.ctx, ptrs[0], ptrs[1], ptrs[2], n, int(d), nnz, int(base), int(k), int(method),
            ctypes.c_uint64(seed & (2**64 - 1)), flags, ptrs[3])
        self._check(st)
        return c

    def batch_autotune(self):
        """Tune the batched assignment's tile per (d, k) bucket; returns the chosen
        variant of each of the KM_BATCH_TILE_BUCKETS buckets."""
        table = (ctypes.c_int32 * KM_BATCH_TILE_BUCKETS)()
        self._check(self.lib.km_batch_autotune(self.ctx, table, KM_BATCH_TILE_BUCKETS))
        return list(table)

    def batch_step_device(self, iterated, iterations, finished):
        """km_batch_step_f32.  iterated: (X, C0, C_out) tensors; finished: (X, C,
        C_out or None, labels_out) tensors.  Returns the finished set's objectives."""
        ni, nf = len(iterated), len(finished)
        ia = (KmBatchInstance * max(ni, 1))()
        fa = (KmBatchInstance * max(nf, 1))()
        for i, (x, c0, cout) in enumerate(iterated):
            ia[i] = KmBatchInstance(x.data_ptr(), c0.data_ptr(), cout.data_ptr(), None,
                                    x.shape[0], x.shape[1], c0.shape[0])
        for i, (x, c, cout, lab) in enumerate(finished):
            fa[i] = KmBatchInstance(x.data_ptr(), c.data_ptr(),
                                    cout.data_ptr() if cout is not None else None, lab.data_ptr(),
                                    x.shape[0], x.shape[1], c.shape[0])
        obj = (ctypes.c_double * max(nf, 1))()
        st = self.lib.km_batch_step_f32(self.ctx, ia if ni else None, ni, int(iterations),
                                        fa if nf else None, nf, 0, obj)
        self._check(st)
        return list(obj)[:nf]

    def infer_device(self, x, c, objective=False, bf16=False):
        import torch

        _check_device_rows(x)
        sfx = _sfx(str(x.dtype).replace("torch.", ""))
        n, d = x.shape
        _check_centroids(c, d, str(x.dtype).replace("torch.", ""), True, x.device)
        labels = torch.empty(n, dtype=torch.int32, device=x.device)
        obj = ctypes.c_double()
        flags = KM_FLAG_DEVICE_POINTERS | (KM_FLAG_EXACT_OBJECTIVE if objective else 0) | (
            KM_FLAG_BF16_REFINE if bf16 else 0)
        st = getattr(self.lib, "km_lloyd_infer_" + sfx)(
            self.ctx, x.data_ptr(), n, d, c.data_ptr(), c.shape[0], flags, labels.data_ptr(),
            ctypes.byref(obj))
        self._check(st)
        return labels, (obj.value if objective else None)

    def infer_device_noalloc(self, x, c, labels, objective=False, bf16=False):
        """km_lloyd_infer_f32_noalloc: labels of the float32 device rows x against
        the device centroids c, written into the caller's int32 device tensor
        `labels` (n entries); the library allocates nothing during the call."""
        import torch

        _check_device_rows(x)
        if x.dtype != torch.float32:
            raise TypeError("km_lloyd_infer_f32_noalloc takes float32 rows")
        n, d = x.shape
        _check_centroids(c, d, "float32", True, x.device)
        if (not isinstance(labels, torch.Tensor) or labels.dtype != torch.int32 or
                labels.device != x.device or labels.numel() < n or not labels.is_contiguous()):
            raise ValueError("labels must be a contiguous int32 tensor of at least n entries "
                             "on the rows' device")
        obj = ctypes.c_double()
        flags = KM_FLAG_DEVICE_POINTERS | (KM_FLAG_EXACT_OBJECTIVE if objective else 0) | (
            KM_FLAG_BF16_REFINE if bf16 else 0)
        st = self.lib.km_lloyd_infer_f32_noalloc(
            self.ctx, x.data_ptr(), n, d, c.data_ptr(), c.shape[0], flags, labels.data_ptr(),
            ctypes.byref(obj))
        self._check(st)
        return obj.value if objective else None


def _raise_server(st):
    msg = load_library().km_server_last_error().decode()
    raise _ERRORS.get(st, EngineError)(st, msg)


class Server:
    """A km_server: fit and infer jobs admitted by an HBM budget and run with
    continuous batching on one device.

    numpy inputs are host jobs (copied in at admission, results copied back);
    contiguous float32 CUDA tensors are device jobs used in place.  The buffers of
    a job are held here until its wait() returns."""

    def __init__(self, budget_bytes, device=0, max_batch_jobs=0, large_job_work=0, autotune=False):
        self.lib = load_library()
        cfg = KmServerConfig(int(device), int(max_batch_jobs), int(budget_bytes), int(large_job_work),
                             1 if autotune else 0, 0)
        h = ctypes.c_void_p()
        st = self.lib.km_server_create(ctypes.byref(cfg), ctypes.byref(h))
        if st != KM_OK:
            _raise_server(st)
        self.handle = h
        self._held = {}

    def close(self):
        if self.handle:
            self.lib.km_server_destroy(self.handle)
            self.handle = None
            self._held.clear()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @staticmethod
    def _prepare(x, c, k_rows, want_centroids):
        """(flags, X ptr, C ptr, C_out ptr, labels ptr, keep-alive, outputs)."""
        if _is_torch_cuda(x):
            import torch

            assert x.dtype == torch.float32 and x.is_contiguous() and c.is_contiguous()
            n, d = x.shape
            labels = torch.empty(n, dtype=torch.int32, device=x.device)
            cout = torch.empty((k_rows, d), dtype=torch.float32, device=x.device) if want_centroids else None
            keep = (x, c, labels, cout)
            return (KM_FLAG_DEVICE_POINTERS, x.data_ptr(), c.data_ptr(),
                    cout.data_ptr() if cout is not None else None, labels.data_ptr(), keep, (cout, labels))
        x = np.ascontiguousarray(x, dtype=np.float32)
        c = np.ascontiguousarray(c, dtype=np.float32)
        n, d = x.shape
        labels = np.empty(n, np.int32)
        cout = np.empty((k_rows, d), np.float32) if want_centroids else None
        keep = (x, c, labels, cout)
        return (0, x.ctypes.data, c.ctypes.data, cout.ctypes.data if cout is not None else None,
                labels.ctypes.data, keep, (cout, labels))

    def submit_fit(self, x, c0, max_iter, bf16=False):
        k = c0.shape[0]
        flags, px, pc, pco, pl, keep, outs = self._prepare(x, c0, k, True)
        flags |= KM_FLAG_BF16_REFINE if bf16 else 0
        jid = ctypes.c_int64()
        st = self.lib.km_server_submit_fit_f32(self.handle, px, x.shape[0], x.shape[1], pc, k,
                                               int(max_iter), flags, pco, pl, ctypes.byref(jid))
        if st != KM_OK:
            _raise_server(st)
        self._held[jid.value] = (keep, outs)
        return jid.value

    def submit_infer(self, x, c, bf16=False):
        k = c.shape[0]
        flags, px, pc, _, pl, keep, outs = self._prepare(x, c, k, False)
        flags |= KM_FLAG_BF16_REFINE if bf16 else 0
        jid = ctypes.c_int64()
        st = self.lib.km_server_submit_infer_f32(self.handle, px, x.shape[0], x.shape[1], pc, k, flags,
                                                 pl, ctypes.byref(jid))
        if st != KM_OK:
            _raise_server(st)
        self._held[jid.value] = (keep, outs)
        return jid.value

    def wait(self, job, timeout_ms=-1):
        """Result dict (centroids, labels, objective, iterations, queue_ms, latency_ms,
        lane), or None if the job is still running after timeout_ms."""
        r = KmJobResult()
        st = self.lib.km_server_wait(self.handle, int(job), int(timeout_ms), ctypes.byref(r))
        if st == KM_PENDING:
            return None
        keep, outs = self._held.pop(job, (None, (None, None)))
        if st != KM_OK:
            _raise_server(st)
        return {"centroids": outs[0], "labels": outs[1], "objective": r.objective,
                "iterations": r.iterations, "queue_ms": r.queue_ms, "latency_ms": r.latency_ms,
                "lane": r.lane}

    def measure_bf16_mfma_peak(self, iters=20000):
        """(TFLOP/s, shader GHz) of a pure BF16 MFMA loop on this device."""
        tf = ctypes.c_double()
        ghz = ctypes.c_double()
        self._check(self.lib.km_measure_bf16_mfma_peak(self.ctx, int(iters), ctypes.byref(tf),
                                                       ctypes.byref(ghz)))
        return tf.value, ghz.value

    def stats(self):
        s = KmServerStats()
        st = self.lib.km_server_get_stats(self.handle, ctypes.byref(s))
        if st != KM_OK:
            _raise_server(st)
        return s.as_dict()
