"""ctypes front end for the CPU checkers (TEST INFRASTRUCTURE ONLY).

Loads ``oracle/liborc.so`` (the C restatement in ``gbs_oracle.c``) and, when
it has been built, ``oracle/_ref/libgbs_ref.so`` (the reference compiled from
its own sources by ``oracle/Makefile``).  Only tests/, ``__graft_entry__.smoke``
and bench.py's cpu_baseline leg import this module; the product library never
does.

Plane arrays use the reference layout ``[nbits][nrows_a][ncols_a/64]`` of
uint64 words (gemmbitserial.hpp:44-50): column c is bit ``c & 63`` of word
``c >> 6``.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORC_SO = os.path.join(HERE, "liborc.so")
REF_SO = os.path.join(HERE, "_ref", "libgbs_ref.so")

# element-type codes shared with include/gbs_hip.h (gbs_dtype)
I8, U8, U8_RAW, I16, U16, I32, U32, I64, U64, F32, F64 = range(11)
_NP_DT = {np.dtype(np.int8): I8, np.dtype(np.uint8): U8,
          np.dtype(np.int16): I16, np.dtype(np.uint16): U16,
          np.dtype(np.int32): I32, np.dtype(np.uint32): U32,
          np.dtype(np.int64): I64, np.dtype(np.uint64): U64,
          np.dtype(np.float32): F32, np.dtype(np.float64): F64}


def dtype_code(arr, raw_u8=False):
    """gbs_dtype for a numpy array; raw_u8 selects the reference's
    non-template ``importRegular(uint8_t*)`` overload (gemmbitserial.hpp:249)."""
    code = _NP_DT[arr.dtype]
    if raw_u8:
        if code != U8:
            raise ValueError("raw_u8 needs a uint8 array")
        return U8_RAW
    return code


_u64 = C.c_uint64
_p = C.c_void_p


class OrcMat(C.Structure):
    _fields_ = [("planes", _p), ("nbits", _u64), ("nrows", _u64),
                ("ncols", _u64), ("nrows_a", _u64), ("ncols_a", _u64),
                ("issigned", C.c_int32)]


def _ptr(a):
    return C.c_void_p(a.ctypes.data) if a is not None else C.c_void_p(0)


_orc = None
_ref = None


def orc():
    global _orc
    if _orc is None:
        if not os.path.exists(ORC_SO):
            raise RuntimeError("oracle/liborc.so is not built (run make -C oracle)")
        lib = C.CDLL(ORC_SO)
        lib.orc_align_to.restype = _u64
        lib.orc_align_to.argtypes = [_u64, _u64]
        lib.orc_finetune_block_size.restype = _u64
        lib.orc_finetune_block_size.argtypes = [_u64, _u64, _u64]
        lib.orc_alloc_geometry.argtypes = [_u64, _u64, _u64, _p]
        lib.orc_import.restype = C.c_int
        lib.orc_import.argtypes = [_p, C.c_int, _u64, C.c_int, _u64, _u64, _u64,
                                   _u64, C.c_int, _p]
        lib.orc_import_and_quantize.restype = C.c_int
        lib.orc_import_and_quantize.argtypes = [_p, _p, C.c_int, C.c_int, _u64, _u64,
                                                _u64, _u64, _u64, C.c_int, _p]
        lib.orc_export.argtypes = [_p, _u64, C.c_int, _u64, _u64, _u64, _u64, _p]
        lib.orc_sum_rows.argtypes = [_p, _u64, C.c_int, _u64, _u64, _u64, _u64, _p]
        lib.orc_gemm.restype = C.c_int
        lib.orc_gemm.argtypes = [C.POINTER(OrcMat), C.POINTER(OrcMat), _p]
        lib.orc_naive_gemm_i64.argtypes = [_p, _p, _p, _u64, _u64, _u64]
        _orc = lib
    return _orc


def ref_available():
    return os.path.exists(REF_SO)


def ref():
    global _ref
    if _ref is None:
        if not ref_available():
            raise RuntimeError("oracle/_ref/libgbs_ref.so is not built")
        lib = C.CDLL(REF_SO)
        lib.gbsref_import.restype = C.c_int
        lib.gbsref_import.argtypes = [_p, C.c_int, _u64, C.c_int, _u64, _u64, _u64,
                                      _u64, C.c_int, _p, _p]
        lib.gbsref_import_quantize.restype = C.c_int
        lib.gbsref_import_quantize.argtypes = [_p, _p, C.c_int, C.c_int, _u64, _u64, _u64,
                                               _u64, _u64, C.c_int, _p, _p]
        lib.gbsref_export.argtypes = [_p, _u64, C.c_int, _u64, _u64, _u64, _u64, _p]
        lib.gbsref_sum_rows.argtypes = [_p, _u64, C.c_int, _u64, _u64, _u64, _u64, _p]
        lib.gbsref_alloc_geometry.argtypes = [_u64, _u64, _u64, _p]
        lib.gbsref_gemm.restype = C.c_int
        lib.gbsref_gemm.argtypes = [_u64, _u64, _u64, _u64, _u64, C.c_int, C.c_int,
                                    _p, C.c_int, _p, C.c_int, _p]
        lib.gbsref_bench_gemm.restype = C.c_double
        lib.gbsref_bench_gemm.argtypes = [_u64, _u64, _u64, _u64, _u64, C.c_int,
                                          C.c_int, _p, C.c_int, _p, C.c_int,
                                          C.c_double, C.c_int, _p, _p]
        _ref = lib
    return _ref


# ----------------------------------------------------------------- geometry
def alloc_geometry(M, K, N):
    """(lhsBlock, rhsBlock, lhs.nrows_a, lhs.ncols_a, rhs.nrows_a, rhs.ncols_a)
    as allocGEMMContext_generic would choose them."""
    out = np.zeros(6, np.uint64)
    orc().orc_alloc_geometry(M, K, N, _ptr(out))
    return tuple(int(x) for x in out)


def ref_alloc_geometry(M, K, N):
    out = np.zeros(6, np.uint64)
    ref().gbsref_alloc_geometry(M, K, N, _ptr(out))
    return tuple(int(x) for x in out)


def align_to(x, a):
    return (x + a - 1) // a * a if x % a else x


# ------------------------------------------------------------------- import
class Planes:
    """A bit-serial matrix held on the host in the reference layout."""

    def __init__(self, planes, nbits, issigned, nrows, ncols, nrows_a, ncols_a):
        self.planes = np.ascontiguousarray(planes, dtype=np.uint64)
        self.nbits, self.issigned = int(nbits), bool(issigned)
        self.nrows, self.ncols = int(nrows), int(ncols)
        self.nrows_a, self.ncols_a = int(nrows_a), int(ncols_a)

    def struct(self):
        return OrcMat(self.planes.ctypes.data, self.nbits, self.nrows, self.ncols,
                      self.nrows_a, self.ncols_a, int(self.issigned))


def import_regular(src, nbits, issigned, nrows_a=None, ncols_a=None,
                   col_major=False, raw_u8=False, nrows=None, ncols=None):
    """orc_import on a 2-D array (rows x cols, or cols x rows when
    col_major).  Returns Planes."""
    src = np.ascontiguousarray(src)
    if nrows is None:
        nrows, ncols = (src.shape[1], src.shape[0]) if col_major else src.shape
    nrows_a = nrows if nrows_a is None else nrows_a
    ncols_a = align_to(ncols, 64) if ncols_a is None else ncols_a
    planes = np.zeros(nbits * nrows_a * (ncols_a // 64), np.uint64)
    rc = orc().orc_import(_ptr(src), dtype_code(src, raw_u8), nbits, int(issigned),
                          nrows, ncols, nrows_a, ncols_a, int(col_major),
                          _ptr(planes))
    if rc != 0:
        raise ValueError("orc_import rejected the arguments")
    return Planes(planes, nbits, issigned, nrows, ncols, nrows_a, ncols_a)


def ref_import_regular(src, nbits, issigned, rowalign=1, colalign=64,
                       col_major=False, raw_u8=False):
    src = np.ascontiguousarray(src)
    nrows, ncols = (src.shape[1], src.shape[0]) if col_major else src.shape
    geom = np.zeros(2, np.uint64)
    lib = ref()
    lib.gbsref_import(_ptr(src), dtype_code(src, raw_u8), nbits, int(issigned),
                      nrows, ncols, rowalign, colalign, int(col_major), None,
                      _ptr(geom))
    nra, nca = int(geom[0]), int(geom[1])
    planes = np.zeros(nbits * nra * (nca // 64), np.uint64)
    lib.gbsref_import(_ptr(src), dtype_code(src, raw_u8), nbits, int(issigned),
                      nrows, ncols, rowalign, colalign, int(col_major),
                      _ptr(planes), _ptr(geom))
    return Planes(planes, nbits, issigned, nrows, ncols, nra, nca)


def import_and_quantize(src, thresholds, nbits, nrows_a=None, ncols_a=None,
                        col_major=False):
    """importRegularAndQuantize<T> (gemmbitserial.hpp:281-308) with T the
    arrays' dtype; thresholds is nThres x nrows (thr[t, r]), same dtype."""
    src = np.ascontiguousarray(src)
    thr = np.ascontiguousarray(thresholds, dtype=src.dtype)
    if thr.ndim == 1:
        thr = thr.reshape(0 if thr.size == 0 else 1, -1)
    nrows, ncols = (src.shape[1], src.shape[0]) if col_major else src.shape
    nrows_a = nrows if nrows_a is None else nrows_a
    ncols_a = align_to(ncols, 64) if ncols_a is None else ncols_a
    planes = np.zeros(nbits * nrows_a * (ncols_a // 64), np.uint64)
    rc = orc().orc_import_and_quantize(_ptr(src), _ptr(thr), dtype_code(src), thr.shape[0],
                                       nbits, nrows, ncols, nrows_a, ncols_a,
                                       int(col_major), _ptr(planes))
    if rc != 0:
        raise ValueError("orc_import_and_quantize rejected the arguments")
    return Planes(planes, nbits, False, nrows, ncols, nrows_a, ncols_a)


def ref_import_and_quantize(src, thresholds, nbits, rowalign=1, colalign=64,
                            col_major=False):
    """The compiled reference's importRegularAndQuantize<T> on the same input."""
    lib = ref()
    src = np.ascontiguousarray(src)
    thr = np.ascontiguousarray(thresholds, dtype=src.dtype)
    if thr.ndim == 1:
        thr = thr.reshape(0 if thr.size == 0 else 1, -1)
    nrows, ncols = (src.shape[1], src.shape[0]) if col_major else src.shape
    geom = np.zeros(2, np.uint64)
    args = (dtype_code(src), thr.shape[0], nbits, nrows, ncols, rowalign, colalign,
            int(col_major))
    lib.gbsref_import_quantize(_ptr(src), _ptr(thr), *args, None, _ptr(geom))
    nra, nca = int(geom[0]), int(geom[1])
    planes = np.zeros(nbits * nra * (nca // 64), np.uint64)
    if lib.gbsref_import_quantize(_ptr(src), _ptr(thr), *args, _ptr(planes), _ptr(geom)):
        raise ValueError("gbsref_import_quantize rejected the dtype")
    return Planes(planes, nbits, False, nrows, ncols, nra, nca)


# ---------------------------------------------------------- export / sumrows
def export_regular(p):
    out = np.zeros((p.nrows, p.ncols), np.int32)
    orc().orc_export(_ptr(p.planes), p.nbits, int(p.issigned), p.nrows, p.ncols,
                     p.nrows_a, p.ncols_a, _ptr(out))
    return out


def ref_export_regular(p):
    out = np.zeros((p.nrows, p.ncols), np.int32)
    ref().gbsref_export(_ptr(p.planes), p.nbits, int(p.issigned), p.nrows, p.ncols,
                        p.nrows_a, p.ncols_a, _ptr(out))
    return out


def sum_rows(p):
    out = np.zeros(p.nrows, np.int32)
    orc().orc_sum_rows(_ptr(p.planes), p.nbits, int(p.issigned), p.nrows, p.ncols,
                       p.nrows_a, p.ncols_a, _ptr(out))
    return out


def ref_sum_rows(p):
    out = np.zeros(p.nrows, np.int32)
    ref().gbsref_sum_rows(_ptr(p.planes), p.nbits, int(p.issigned), p.nrows,
                          p.ncols, p.nrows_a, p.ncols_a, _ptr(out))
    return out


# --------------------------------------------------------------------- gemm
def gemm(lhs, rhs):
    """orc_gemm: transposed result res[n, m] (N x M int32)."""
    res = np.zeros((rhs.nrows, lhs.nrows), np.int32)
    a, b = lhs.struct(), rhs.struct()
    rc = orc().orc_gemm(C.byref(a), C.byref(b), _ptr(res))
    if rc != 0:
        raise ValueError("orc_gemm: lhs.ncols != rhs.ncols")
    return res


def ref_gemm(lhs_vals, rhs_vals, lb, rb, ls, rs, raw_u8=False):
    """The reference flow end to end on regular matrices (M x K and N x K)."""
    lhs_vals = np.ascontiguousarray(lhs_vals)
    rhs_vals = np.ascontiguousarray(rhs_vals)
    M, K = lhs_vals.shape
    N = rhs_vals.shape[0]
    res = np.zeros((N, M), np.int32)
    rc = ref().gbsref_gemm(M, K, N, lb, rb, int(ls), int(rs), _ptr(lhs_vals),
                           dtype_code(lhs_vals, raw_u8), _ptr(rhs_vals),
                           dtype_code(rhs_vals, raw_u8), _ptr(res))
    if rc != 0:
        raise ValueError("gbsref_gemm failed")
    return res


def ref_bench_gemm(lhs_vals, rhs_vals, lb, rb, ls, rs, min_secs=1.0,
                   max_reps=1000):
    """Reference CPU timing (benchmark.cpp:93-111 shape).  Returns
    (ns_per_call, reps, res)."""
    lhs_vals = np.ascontiguousarray(lhs_vals)
    rhs_vals = np.ascontiguousarray(rhs_vals)
    M, K = lhs_vals.shape
    N = rhs_vals.shape[0]
    res = np.zeros((N, M), np.int32)
    reps = C.c_int(0)
    ns = ref().gbsref_bench_gemm(M, K, N, lb, rb, int(ls), int(rs),
                                 _ptr(lhs_vals), dtype_code(lhs_vals),
                                 _ptr(rhs_vals), dtype_code(rhs_vals),
                                 min_secs, max_reps, C.byref(reps), _ptr(res))
    return ns, reps.value, res


def naive_gemm_i64(lhs_vals, rhs_vals):
    """testhelpers.hpp:34-44 in int64: res[n, m] = sum_k lhs[m,k]*rhs[n,k]."""
    l32 = np.ascontiguousarray(lhs_vals, dtype=np.int32)
    r32 = np.ascontiguousarray(rhs_vals, dtype=np.int32)
    M, K = l32.shape
    N = r32.shape[0]
    res = np.zeros((N, M), np.int64)
    orc().orc_naive_gemm_i64(_ptr(l32), _ptr(r32), _ptr(res), M, K, N)
    return res


def blas_gemm_exact(lhs_vals, rhs_vals):
    """rhs @ lhs.T through float64 BLAS.  Every partial sum of small-integer
    products is an integer far below 2^53, so the product is exact; used as a
    fast second naive check at full sizes.  Returns int64 (N x M)."""
    l = np.asarray(lhs_vals, dtype=np.float64)
    r = np.asarray(rhs_vals, dtype=np.float64)
    bound = float(np.abs(l).max(initial=0)) * float(np.abs(r).max(initial=0)) * l.shape[1]
    if bound >= 2.0 ** 53:
        raise ValueError("float64 product would not be exact")
    return (r @ l.T).astype(np.int64)


def wrap_i32(x):
    """int64 -> int32 modulo 2^32 (the reference's int32 accumulators)."""
    return np.asarray(x, dtype=np.int64).astype(np.uint32).view(np.int32)


def values_of(vals, nbits, issigned, raw_u8=False):
    """The integer each input element denotes after import: the bipolar
    value, or the two's-complement / unsigned value of the low nbits
    (gemmbitserial.hpp:169-186, 312-333)."""
    v = np.asarray(vals).astype(np.int64)
    if nbits == 1 and issigned:
        if raw_u8:
            return np.where((v & 1) == 1, 1, -1)
        return np.where(v > 0, 1, -1)
    code = v & ((1 << nbits) - 1)
    if issigned:
        code = np.where(code >= (1 << (nbits - 1)), code - (1 << nbits), code)
    return code
