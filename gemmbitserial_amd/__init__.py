"""Python binding of libgbs_hip.so, the MI355X bit-serial GEMM C ABI.

The product interface is the C ABI (include/gbs_hip.h) and the C++ drop-in
header include/gemmbitserial.hpp.  This module is a thin ctypes layer over
the same C ABI for tests and benchmarks; it names things after the
reference's API (BitSerialMatrix, importRegular, gemmBitSerial, sumRows,
exportRegular: gemmbitserial.hpp:19-479, arch-generic.hpp:331-345).

Loading fails loudly when the library has not been built: there is no CPU
fallback.
"""
import ctypes as C
import os

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "libgbs_hip.so")

# gbs_dtype (the floating types: threshold import only)
I8, U8, U8_RAW, I16, U16, I32, U32, I64, U64, F32, F64 = range(11)
_NP_DT = {np.dtype(np.int8): I8, np.dtype(np.uint8): U8,
          np.dtype(np.int16): I16, np.dtype(np.uint16): U16,
          np.dtype(np.int32): I32, np.dtype(np.uint32): U32,
          np.dtype(np.int64): I64, np.dtype(np.uint64): U64,
          np.dtype(np.float32): F32, np.dtype(np.float64): F64}
# gbs_algo
AUTO, POPCOUNT, MFMA, MFMA_FP4, MFMA_FP6, MFMA_BITS, MFMA_PLANES = 0, 1, 2, 3, 4, 5, 6
ALGOS = {"auto": AUTO, "popcount": POPCOUNT, "mfma": MFMA, "mfma_fp4": MFMA_FP4,
         "mfma_fp6": MFMA_FP6, "mfma_bits": MFMA_BITS, "mfma_planes": MFMA_PLANES}

EXPORTED = [
    "gbs_init", "gbs_last_error", "gbs_version", "gbs_host_alloc", "gbs_host_free",
    "gbs_matrix_create", "gbs_matrix_destroy", "gbs_matrix_device_planes",
    "gbs_matrix_mark_modified",
    "gbs_import", "gbs_import_quantize", "gbs_import_quantize_typed", "gbs_planes_from_host", "gbs_planes_to_host",
    "gbs_export", "gbs_sum_rows", "gbs_workspace_create", "gbs_workspace_destroy",
    "gbs_gemm", "gbs_gemm_async", "gbs_workspace_reserve", "gbs_gemm_select",
    "gbs_stream", "gbs_stream_sync", "gbs_event_create", "gbs_event_destroy",
    "gbs_event_record", "gbs_event_elapsed_ms",
    "gbs_batch_create", "gbs_batch_destroy", "gbs_batch_run", "gbs_batch_run_graph",
    "gbs_batch_run_per_call", "gbs_batch_info", "gbs_batch_plan_of",
    "gbs_batch_tune_classes", "gbs_batch_tune_class_of", "gbs_batch_tune_shapes",
    "gbs_batch_tune_set", "gbs_batch_tune_get", "gbs_batch_tune_load",
    "gbs_batch_set_planes", "gbs_batch_get_planes",
    "gbs_tune_classes", "gbs_tune_class_of", "gbs_tune_set", "gbs_tune_load",
    "gbs_clock_ns", "gbs_server_create", "gbs_server_destroy", "gbs_server_submit",
    "gbs_server_step", "gbs_server_take", "gbs_server_stats", "gbs_server_layer_input",
    "gbs_server_finished",
    "gbs_gemm_async_config", "gbs_gemm_async_noalloc",
    "gbs_conv2d_shape", "gbs_import_im2col", "gbs_matrix_save", "gbs_matrix_load",
]


class GbsError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"gbs error {code}: {msg}")
        self.code = code


_lib = None
_u64, _p, _i = C.c_uint64, C.c_void_p, C.c_int


def lib():
    """The loaded library with argtypes set.  Raises if it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is not built; run python -m gemmbitserial_amd.build")
    L = C.CDLL(LIB_PATH)
    sig = {
        "gbs_init": [_i], "gbs_version": [], "gbs_host_alloc": [C.c_size_t, _p],
        "gbs_host_free": [_p],
        "gbs_matrix_create": [_u64, _u64, _u64, _i, _u64, _u64, _p],
        "gbs_matrix_destroy": [_p], "gbs_matrix_device_planes": [_p, _p],
        "gbs_matrix_mark_modified": [_p],
        "gbs_import": [_p, _p, _i, _i, _i, _p],
        "gbs_import_quantize": [_p, _p, _p, _i, _i, _i, _p],
        "gbs_import_quantize_typed": [_p, _p, _p, _i, _i, _i, _i, _p],
        "gbs_planes_from_host": [_p, _p], "gbs_planes_to_host": [_p, _p],
        "gbs_export": [_p, _p, _i], "gbs_sum_rows": [_p, _p, _i],
        "gbs_workspace_create": [_p], "gbs_workspace_destroy": [_p],
        "gbs_gemm": [_p, _p, _p, _p, _i, _i], "gbs_gemm_async": [_p, _p, _p, _p, _i, _p],
        "gbs_workspace_reserve": [_p, _p, _p], "gbs_gemm_select": [_p, _p, _p],
        "gbs_stream": [_p], "gbs_stream_sync": [_p], "gbs_event_create": [_p],
        "gbs_event_destroy": [_p], "gbs_event_record": [_p, _p],
        "gbs_event_elapsed_ms": [_p, _p, _p],
        "gbs_batch_create": [_p, _u64, _p], "gbs_batch_destroy": [_p],
        "gbs_batch_run": [_p, _p], "gbs_batch_run_graph": [_p, _p],
        "gbs_batch_run_per_call": [_p, _p], "gbs_batch_info": [_p, _p, _p, _p],
        "gbs_batch_plan_of": [_p, _u64, _p, _p, _p, _p],
        "gbs_batch_tune_classes": [_p], "gbs_batch_tune_class_of": [_u64, _u64, _p],
        "gbs_batch_tune_shapes": [_i, _p, _i, _p], "gbs_batch_tune_set": [_i, _i, _i, _i],
        "gbs_batch_tune_get": [_i, _i, _p, _p], "gbs_batch_tune_load": [C.c_char_p],
        "gbs_batch_set_planes": [_i], "gbs_batch_get_planes": [_p],
        "gbs_tune_classes": [_p, _p], "gbs_tune_class_of": [_u64, _u64, _u64, _p],
        "gbs_tune_set": [_p, _p, _i], "gbs_tune_load": [C.c_char_p],
        "gbs_gemm_async_config": [_p, _p, _p, _p, _i, _i, _i, _p],
        "gbs_gemm_async_noalloc": [_p, _p, _p, _i, _p],
        "gbs_clock_ns": [_p], "gbs_server_create": [_p, _p, _i, _i, _p, _p],
        "gbs_server_destroy": [_p], "gbs_server_submit": [_p, _u64, _p, _i, C.c_int64],
        "gbs_server_step": [_p, _p],
        "gbs_server_take": [_p, _u64, _p, _u64, _p, _p, _p, _p],
        "gbs_server_stats": [_p, _p], "gbs_server_layer_input": [_p, _i, _p, _p],
        "gbs_server_finished": [_p, _p, _i, _p],
        "gbs_conv2d_shape": [_p, _p, _p, _p, _p],
        "gbs_import_im2col": [_p, _p, _i, _p, _i, _p],
        "gbs_matrix_save": [_p, C.c_char_p], "gbs_matrix_load": [C.c_char_p, _p],
    }
    for name, args in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = _i
    L.gbs_last_error.argtypes = []
    L.gbs_last_error.restype = C.c_char_p
    _lib = L
    return L


def check(rc):
    if rc != 0:
        raise GbsError(rc, lib().gbs_last_error().decode())
    return rc


def _ptr(a):
    return C.c_void_p(a.ctypes.data)


def init(device=0):
    check(lib().gbs_init(device))


class Conv2d(C.Structure):
    """gbs_conv2d: NCHW input, kernel, strides, zero padding, dilations."""
    _fields_ = [(n, C.c_uint32) for n in ("batch", "channels", "height", "width", "kh", "kw",
                                           "stride_h", "stride_w", "pad_h", "pad_w",
                                           "dil_h", "dil_w")]


def conv2d(batch, channels, height, width, kh, kw, stride=1, pad=0, dil=1):
    sh, sw = (stride, stride) if isinstance(stride, int) else stride
    ph, pw = (pad, pad) if isinstance(pad, int) else pad
    dh, dw = (dil, dil) if isinstance(dil, int) else dil
    return Conv2d(batch, channels, height, width, kh, kw, sh, sw, ph, pw, dh, dw)


def conv2d_shape(g):
    """(rows, cols, out_h, out_w) of the im2col operand of geometry g."""
    r, c, oh, ow = C.c_uint64(), C.c_uint64(), C.c_uint32(), C.c_uint32()
    check(lib().gbs_conv2d_shape(C.byref(g), C.byref(r), C.byref(c), C.byref(oh), C.byref(ow)))
    return r.value, c.value, oh.value, ow.value


PLANE_FILE_MAGIC = b"GBSPLN01"


class DeviceMatrix:
    """A gbs_matrix: bit-planes resident in HBM, reference layout
    [nbits][nrows_a][ncols_a/64] (BitSerialMatrix, gemmbitserial.hpp:19-50)."""

    def __init__(self, nbits, nrows, ncols, issigned, nrows_a=None, ncols_a=None,
                 _handle=None):
        self.nbits, self.nrows, self.ncols = int(nbits), int(nrows), int(ncols)
        self.issigned = bool(issigned)
        self.nrows_a = self.nrows if nrows_a is None else int(nrows_a)
        self.ncols_a = ((self.ncols + 63) // 64) * 64 if ncols_a is None else int(ncols_a)
        if _handle is not None:
            self.h = _handle
            return
        h = C.c_void_p()
        check(lib().gbs_matrix_create(self.nbits, self.nrows, self.ncols,
                                      int(self.issigned), self.nrows_a, self.ncols_a,
                                      C.byref(h)))
        self.h = h

    @classmethod
    def load(cls, path):
        """A matrix from a packed-matrix file (gbs_matrix_save's format)."""
        h = C.c_void_p()
        check(lib().gbs_matrix_load(os.fsencode(path), C.byref(h)))
        with open(path, "rb") as f:
            head = np.frombuffer(f.read(64)[8:], dtype="<u8")
        nbits, nrows, ncols, nrows_a, ncols_a, issigned = (int(x) for x in head[:6])
        return cls(nbits, nrows, ncols, issigned, nrows_a, ncols_a, _handle=h)

    def save(self, path):
        check(lib().gbs_matrix_save(self.h, os.fsencode(path)))

    def import_im2col(self, x, g, want_planes=False):
        """Pack the im2col matrix of NCHW tensor x (geometry g) into the planes."""
        x = np.ascontiguousarray(x)
        if x.size != g.batch * g.channels * g.height * g.width:
            raise ValueError("tensor size does not match the geometry")
        out = np.zeros(self.words, np.uint64) if want_planes else None
        check(lib().gbs_import_im2col(self.h, _ptr(x), _NP_DT[x.dtype], C.byref(g), 0,
                                      _ptr(out) if out is not None else None))
        return out

    @property
    def words(self):
        return self.nbits * self.nrows_a * (self.ncols_a // 64)

    def close(self):
        if self.h:
            check(lib().gbs_matrix_destroy(self.h))
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def device_planes(self):
        p = C.c_void_p()
        check(lib().gbs_matrix_device_planes(self.h, C.byref(p)))
        return p.value

    def import_regular(self, src, col_major=False, raw_u8=False, want_planes=False):
        """importRegular on the GPU (gemmbitserial.hpp:212-275); src is
        nrows x ncols (or ncols x nrows when col_major).  Returns the host
        copy of the planes when want_planes."""
        src = np.ascontiguousarray(src)
        dt = _NP_DT[src.dtype]
        if raw_u8:
            assert dt == U8
            dt = U8_RAW
        if src.size != self.nrows * self.ncols:
            raise ValueError("source size does not match the matrix")
        out = np.zeros(self.words, np.uint64) if want_planes else None
        check(lib().gbs_import(self.h, _ptr(src), dt, int(col_major), 0,
                               _ptr(out) if out is not None else None))
        return out

    def import_quantize(self, src, thresholds, col_major=False, want_planes=False):
        """importRegularAndQuantize<T> on the GPU (gemmbitserial.hpp:281-308),
        T = src.dtype; thresholds is nThres x nrows of the same type."""
        src = np.ascontiguousarray(src)
        thr = np.ascontiguousarray(thresholds, dtype=src.dtype)
        if thr.ndim == 1:
            thr = thr.reshape(0 if thr.size == 0 else 1, -1)
        if src.size != self.nrows * self.ncols:
            raise ValueError("source size does not match the matrix")
        out = np.zeros(self.words, np.uint64) if want_planes else None
        check(lib().gbs_import_quantize_typed(self.h, _ptr(src), _ptr(thr), _NP_DT[src.dtype],
                                              thr.shape[0], int(col_major), 0,
                                              _ptr(out) if out is not None else None))
        return out

    def planes(self):
        out = np.zeros(self.words, np.uint64)
        check(lib().gbs_planes_to_host(self.h, _ptr(out)))
        return out

    def set_planes(self, words):
        w = np.ascontiguousarray(words, dtype=np.uint64)
        assert w.size == self.words
        check(lib().gbs_planes_from_host(self.h, _ptr(w)))

    def export_regular(self):
        out = np.zeros((self.nrows, self.ncols), np.int32)
        check(lib().gbs_export(self.h, _ptr(out), 0))
        return out

    def sum_rows(self):
        out = np.zeros(self.nrows, np.int32)
        check(lib().gbs_sum_rows(self.h, _ptr(out), 0))
        return out


def from_values(vals, nbits, issigned, raw_u8=False, nrows_a=None, ncols_a=None):
    vals = np.ascontiguousarray(vals)
    m = DeviceMatrix(nbits, vals.shape[0], vals.shape[1], issigned, nrows_a, ncols_a)
    m.import_regular(vals, raw_u8=raw_u8)
    return m


def gemm(lhs, rhs, algo="auto"):
    """gemmBitSerial: res[n, m] = sum_k lhs[m,k] * rhs[n,k] (N x M int32)."""
    res = np.zeros((rhs.nrows, lhs.nrows), np.int32)
    check(lib().gbs_gemm(None, lhs.h, rhs.h, _ptr(res), 0, ALGOS[algo]))
    return res


def select(lhs, rhs):
    a = C.c_int()
    check(lib().gbs_gemm_select(lhs.h, rhs.h, C.byref(a)))
    return {v: k for k, v in ALGOS.items()}[a.value]


class DeviceBuffer:
    """Raw device memory for results that stay in HBM (hipMalloc through
    libamdhip64, plumbing for timing loops)."""

    _hip = None

    @classmethod
    def hip(cls):
        if cls._hip is None:
            rocm = os.environ.get("ROCM_PATH", "/opt/rocm")
            h = C.CDLL(os.path.join(rocm, "lib", "libamdhip64.so"))
            h.hipMalloc.argtypes = [_p, C.c_size_t]
            h.hipFree.argtypes = [_p]
            h.hipMemcpy.argtypes = [_p, _p, C.c_size_t, _i]
            h.hipMemset.argtypes = [_p, _i, C.c_size_t]
            h.hipDeviceSynchronize.argtypes = []
            cls._hip = h
        return cls._hip

    def __init__(self, nbytes):
        self.nbytes = int(nbytes)
        p = C.c_void_p()
        rc = self.hip().hipMalloc(C.byref(p), max(self.nbytes, 8))
        if rc != 0:
            raise GbsError(2, f"hipMalloc({nbytes}) failed: {rc}")
        self.ptr = p.value

    # The copies below run on the null stream, which the library's
    # non-blocking streams do not order against: each waits for the device
    # first (work already enqueued anywhere) and returns when it is done.
    def to_host(self, dtype, shape):
        out = np.zeros(shape, dtype)
        rc = self.hip().hipDeviceSynchronize()
        if rc == 0:
            rc = self.hip().hipMemcpy(_ptr(out), C.c_void_p(self.ptr), out.nbytes, 2)
        if rc != 0:
            raise GbsError(3, f"hipMemcpy D2H failed: {rc}")
        return out

    def copy_from_host(self, a):
        a = np.ascontiguousarray(a)
        if a.nbytes > self.nbytes:
            raise ValueError("array larger than the buffer")
        rc = self.hip().hipDeviceSynchronize()
        if rc == 0:
            rc = self.hip().hipMemcpy(C.c_void_p(self.ptr), _ptr(a), a.nbytes, 1)
        if rc == 0:
            rc = self.hip().hipDeviceSynchronize()
        if rc != 0:
            raise GbsError(3, f"hipMemcpy H2D failed: {rc}")

    def copy_from(self, other):
        rc = self.hip().hipDeviceSynchronize()
        if rc == 0:
            rc = self.hip().hipMemcpy(C.c_void_p(self.ptr), C.c_void_p(other.ptr),
                                      min(self.nbytes, other.nbytes), 3)
        if rc == 0:
            rc = self.hip().hipDeviceSynchronize()
        if rc != 0:
            raise GbsError(3, f"hipMemcpy D2D failed: {rc}")

    def fill_zero(self):
        rc = self.hip().hipDeviceSynchronize()
        if rc == 0:
            rc = self.hip().hipMemset(C.c_void_p(self.ptr), 0, max(self.nbytes, 8))
        if rc == 0:
            rc = self.hip().hipDeviceSynchronize()
        if rc != 0:
            raise GbsError(3, f"hipMemset failed: {rc}")

    def close(self):
        if self.ptr:
            self.hip().hipFree(C.c_void_p(self.ptr))
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Workspace:
    def __init__(self):
        h = C.c_void_p()
        check(lib().gbs_workspace_create(C.byref(h)))
        self.h = h

    def reserve(self, lhs, rhs):
        check(lib().gbs_workspace_reserve(self.h, lhs.h, rhs.h))

    def gemm_async(self, lhs, rhs, d_res_ptr, algo="auto", stream=None):
        check(lib().gbs_gemm_async(self.h, lhs.h, rhs.h, C.c_void_p(d_res_ptr),
                                   ALGOS[algo], stream))

    def close(self):
        if self.h:
            check(lib().gbs_workspace_destroy(self.h))
            self.h = None


def stream():
    s = C.c_void_p()
    check(lib().gbs_stream(C.byref(s)))
    return s


def sync(s=None):
    check(lib().gbs_stream_sync(s))


class Event:
    def __init__(self):
        e = C.c_void_p()
        check(lib().gbs_event_create(C.byref(e)))
        self.e = e

    def record(self, s=None):
        check(lib().gbs_event_record(self.e, s))

    def elapsed_ms(self, end):
        ms = C.c_float()
        check(lib().gbs_event_elapsed_ms(self.e, end.e, C.byref(ms)))
        return ms.value


class _Problem(C.Structure):
    _fields_ = [("lhs", _p), ("rhs", _p), ("d_res", _p)]


class Batch:
    """gbs_batch: many products in one persistent launch.  problems is a
    list of (lhs DeviceMatrix, rhs DeviceMatrix, device result pointer)."""

    def __init__(self, problems):
        self._keep = problems
        arr = (_Problem * max(len(problems), 1))()
        for i, (l, r, d) in enumerate(problems):
            arr[i] = _Problem(l.h, r.h, d)
        h = C.c_void_p()
        check(lib().gbs_batch_create(arr, len(problems), C.byref(h)))
        self.h = h

    def run(self, stream=None):
        check(lib().gbs_batch_run(self.h, stream))

    def run_graph(self, stream=None):
        check(lib().gbs_batch_run_graph(self.h, stream))

    def run_per_call(self, stream=None):
        check(lib().gbs_batch_run_per_call(self.h, stream))

    def info(self):
        t, a, b = C.c_uint64(), C.c_uint64(), C.c_uint64()
        check(lib().gbs_batch_info(self.h, C.byref(t), C.byref(a), C.byref(b)))
        return {"tiles": t.value, "persistent": a.value, "separate": b.value}

    def plan_of(self, i):
        """(group, shape, splits, K piece units) the batch planned for
        problem i: group 0/1/2 = int8/FP4/FP6 copies, BATCH_PLANES = read
        from the planes (shape: its tile form; piece: super-steps of 1024
        values), -1 for a popcount product."""
        f, s = C.c_int(), C.c_int()
        sp, pc = C.c_uint32(), C.c_uint32()
        check(lib().gbs_batch_plan_of(self.h, i, C.byref(f), C.byref(s), C.byref(sp),
                                      C.byref(pc)))
        return f.value, s.value, sp.value, pc.value

    def close(self):
        if self.h:
            check(lib().gbs_batch_destroy(self.h))
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


BATCH_PLANES = 3


PLANES_NONE, PLANES_SKINNY, PLANES_ALL = 0, 1, 2


def batch_set_planes(mode):
    """Which products batches planned from now on read from their bit-planes
    (the planes group): PLANES_NONE (the default: every product from value
    copies), PLANES_SKINNY (products of at most 16 rhs rows that do not fit
    FP4 codes on both sides) or PLANES_ALL (every product it holds)."""
    check(lib().gbs_batch_set_planes(int(mode)))


def batch_get_planes():
    e = C.c_int()
    check(lib().gbs_batch_get_planes(C.byref(e)))
    return e.value


def batch_tune_classes():
    n = C.c_int()
    check(lib().gbs_batch_tune_classes(C.byref(n)))
    return n.value


def batch_tune_class_of(N, K):
    c = C.c_int()
    check(lib().gbs_batch_tune_class_of(N, K, C.byref(c)))
    return c.value


def batch_tune_shapes(fmt):
    buf, n = (C.c_int * 64)(), C.c_int()
    check(lib().gbs_batch_tune_shapes(fmt, buf, 64, C.byref(n)))
    return list(buf[:n.value])


def batch_tune_set(fmt, cls, shape, piece_units=16):
    check(lib().gbs_batch_tune_set(fmt, cls, shape, piece_units))


def batch_tune_get(fmt, cls):
    s, p = C.c_int(), C.c_int()
    check(lib().gbs_batch_tune_get(fmt, cls, C.byref(s), C.byref(p)))
    return s.value, p.value


def batch_tune_load(path):
    check(lib().gbs_batch_tune_load(path.encode()))


def tune_classes():
    n, s = C.c_int(), C.c_int()
    check(lib().gbs_tune_classes(C.byref(n), C.byref(s)))
    return n.value, s.value


def tune_class_of(M, N, K):
    c = C.c_int()
    check(lib().gbs_tune_class_of(M, N, K, C.byref(c)))
    return c.value


def tune_set(shapes, splits=None):
    if shapes is None:
        check(lib().gbs_tune_set(None, None, 0))
        return
    a = (C.c_int * len(shapes))(*shapes)
    b = (C.c_int * len(shapes))(*(splits or [0] * len(shapes)))
    check(lib().gbs_tune_set(a, b, len(shapes)))


def tune_load(path):
    check(lib().gbs_tune_load(path.encode()))


class _ServerConfig(C.Structure):
    _fields_ = [("max_requests", C.c_uint32), ("max_rows", C.c_uint32),
                ("mem_budget", C.c_uint64)]


class _ServerCounters(C.Structure):
    _fields_ = [("steps", C.c_uint64), ("done_requests", C.c_uint64),
                ("done_rows", C.c_uint64), ("ops", C.c_double),
                ("row_capacity", C.c_uint32), ("bytes_reserved", C.c_uint64),
                ("waiting", C.c_uint32), ("in_flight", C.c_uint32),
                ("layer_steps", C.c_uint64), ("hbm_bytes", C.c_double)]


def clock_ns():
    t = C.c_int64()
    check(lib().gbs_clock_ns(C.byref(t)))
    return t.value


class Server:
    """gbs_server: continuous batching over resident bit-serial layers.
    weights: list of DeviceMatrix (h_{l+1} x h_l); shifts: requant shift per
    layer but the last."""

    def __init__(self, weights, shifts, act_bits=4, max_requests=64, max_rows=64,
                 mem_budget=0):
        L = lib()
        self.weights = list(weights)
        self.h = [self.weights[0].ncols] + [w.nrows for w in self.weights]
        wa = (C.c_void_p * len(self.weights))(*[w.h.value for w in self.weights])
        sh = (C.c_uint32 * max(1, len(shifts)))(*(list(shifts) or [0]))
        cfg = _ServerConfig(max_requests, max_rows, mem_budget)
        h = C.c_void_p()
        check(L.gbs_server_create(wa, sh, len(self.weights), act_bits, C.byref(cfg), C.byref(h)))
        self.s = h
        self.act_bits = act_bits

    def submit(self, rid, x, arrival_ns=0):
        x = np.ascontiguousarray(x, dtype=np.uint8)
        assert x.shape[1] == self.h[0]
        check(lib().gbs_server_submit(self.s, rid, _ptr(x), x.shape[0], int(arrival_ns)))

    def step(self):
        n = C.c_uint32()
        check(lib().gbs_server_step(self.s, C.byref(n)))
        return n.value

    def take(self, rid, max_rows=64, out=None):
        """(rows, arrival ns, done ns) of a finished request, or None.  The
        rows are written into out (max_rows x h_L int32) when given, else a
        new array; the returned rows are a view of it."""
        if out is None:
            out = np.empty((max_rows, self.h[-1]), np.int32)
        ready, rows = C.c_int(), C.c_uint32()
        arr, done = C.c_int64(), C.c_int64()
        check(lib().gbs_server_take(self.s, rid, _ptr(out), out.size, C.byref(ready),
                                    C.byref(rows), C.byref(arr), C.byref(done)))
        if not ready.value:
            return None
        return out[: rows.value], arr.value, done.value

    def finished(self, cap=1024):
        """Ids of finished requests not yet taken."""
        ids, n = (C.c_uint64 * cap)(), C.c_uint32()
        check(lib().gbs_server_finished(self.s, ids, cap, C.byref(n)))
        return list(ids[:n.value])

    def layer_input(self, layer):
        """(plane words [act_bits][capacity][words per row], rows) of the
        bit-serial input the cohort at `layer` reads next step."""
        m, rows = C.c_void_p(), C.c_uint32()
        check(lib().gbs_server_layer_input(self.s, layer, C.byref(m), C.byref(rows)))
        cap = self.stats()["row_capacity"]
        wpr = (self.h[layer] + 63) // 64
        out = np.zeros(self.act_bits * cap * wpr, np.uint64)
        check(lib().gbs_planes_to_host(m, _ptr(out)))
        return out.reshape(self.act_bits, cap, wpr), rows.value

    def stats(self):
        st = _ServerCounters()
        check(lib().gbs_server_stats(self.s, C.byref(st)))
        return {f: getattr(st, f) for f, _ in _ServerCounters._fields_}

    def close(self):
        if self.s:
            check(lib().gbs_server_destroy(self.s))
            self.s = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
