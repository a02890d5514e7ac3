// ref_shim.cpp -- C entry points around the reference gemmbitserial, compiled
// from the reference's own header where it lies (oracle/Makefile passes
// -I<reference checkout>).  TEST INFRASTRUCTURE ONLY: the result,
// oracle/_ref/libgbs_ref.so, is loaded by tests/ (to pin the oracle
// restatement) and by bench.py's cpu_baseline leg (to time the reference CPU
// path).  Nothing in the product links it.
//
// The reference API calls used here: BitSerialMatrix::alloc/importRegular/
// exportRegular (gemmbitserial.hpp:24-333), allocGEMMContext/gemmBitSerial/
// sumRows/deallocGEMMContext (gemmbitserial.hpp:415-479, arch-generic.hpp),
// and the timing loop shape of benchmark/benchmark.cpp:93-111.
#include <chrono>
#include <cstring>
#include <cstdint>
#include "gemmbitserial.hpp"

using namespace gemmbitserial;

namespace {
enum { I8 = 0, U8 = 1, U8_RAW = 2, I16 = 3, U16 = 4, I32 = 5, U32 = 6,
       I64 = 7, U64 = 8, F32 = 9, F64 = 10 };

// dispatch importRegular on the element type; U8 goes through the template
// (T = const uint8_t), U8_RAW through the non-template uint8_t* overload
int do_import(BitSerialMatrix &m, const void *src, int dt, bool cm) {
  switch (dt) {
    case I8: m.importRegular((const int8_t *)src, cm); return 0;
    case U8: m.importRegular((const uint8_t *)src, cm); return 0;
    case U8_RAW: m.importRegular((uint8_t *)const_cast<void *>(src), cm); return 0;
    case I16: m.importRegular((const int16_t *)src, cm); return 0;
    case U16: m.importRegular((const uint16_t *)src, cm); return 0;
    case I32: m.importRegular((const int32_t *)src, cm); return 0;
    case U32: m.importRegular((const uint32_t *)src, cm); return 0;
    case I64: m.importRegular((const int64_t *)src, cm); return 0;
    case U64: m.importRegular((const uint64_t *)src, cm); return 0;
  }
  return -1;
}
template <typename T>
void quantize_into(BitSerialMatrix &m, const void *src, const void *thr, int nthres,
                   bool cm) {
  m.importRegularAndQuantize((T *)const_cast<void *>(src), (T *)const_cast<void *>(thr),
                             nthres, cm);
}

int do_import_quantize(BitSerialMatrix &m, const void *src, const void *thr, int dt,
                       int nthres, bool cm) {
  switch (dt) {
    case I8: quantize_into<int8_t>(m, src, thr, nthres, cm); return 0;
    case U8: case U8_RAW: quantize_into<uint8_t>(m, src, thr, nthres, cm); return 0;
    case I16: quantize_into<int16_t>(m, src, thr, nthres, cm); return 0;
    case U16: quantize_into<uint16_t>(m, src, thr, nthres, cm); return 0;
    case I32: quantize_into<int32_t>(m, src, thr, nthres, cm); return 0;
    case U32: quantize_into<uint32_t>(m, src, thr, nthres, cm); return 0;
    case I64: quantize_into<int64_t>(m, src, thr, nthres, cm); return 0;
    case U64: quantize_into<uint64_t>(m, src, thr, nthres, cm); return 0;
    case F32: quantize_into<float>(m, src, thr, nthres, cm); return 0;
    case F64: quantize_into<double>(m, src, thr, nthres, cm); return 0;
  }
  return -1;
}
}  // namespace

extern "C" {

// Import a regular matrix into a standalone BitSerialMatrix and copy out the
// plane words.  geom[0..1] receives nrows_a, ncols_a.  planes must hold
// nbits*nrows_a*ncols_a/64 words (query with planes == nullptr first).
int gbsref_import(const void *src, int dt, uint64_t nbits, int issigned,
                  uint64_t nrows, uint64_t ncols, uint64_t rowalign,
                  uint64_t colalign, int col_major, uint64_t *planes,
                  uint64_t geom[2]) {
  BitSerialMatrix m = BitSerialMatrix::alloc(nbits, nrows, ncols,
                                             issigned != 0, rowalign, colalign);
  geom[0] = m.nrows_a;
  geom[1] = m.ncols_a;
  int rc = 0;
  if (planes) {
    rc = do_import(m, src, dt, col_major != 0);
    if (rc == 0)
      std::memcpy(planes, m.data, m.nbits * m.wordsPerBitplane() * 8);
  }
  BitSerialMatrix::dealloc(m);
  return rc;
}

// BitSerialMatrix::importRegularAndQuantize<T> (gemmbitserial.hpp:281-308) on
// an unsigned standalone matrix; src and thresholds of element type dt.
int gbsref_import_quantize(const void *src, const void *thr, int dt, int nthres,
                           uint64_t nbits, uint64_t nrows, uint64_t ncols,
                           uint64_t rowalign, uint64_t colalign, int col_major,
                           uint64_t *planes, uint64_t geom[2]) {
  BitSerialMatrix m = BitSerialMatrix::alloc(nbits, nrows, ncols, false, rowalign,
                                             colalign);
  geom[0] = m.nrows_a;
  geom[1] = m.ncols_a;
  int rc = 0;
  if (planes) {
    rc = do_import_quantize(m, src, thr, dt, nthres, col_major != 0);
    if (rc == 0) std::memcpy(planes, m.data, m.nbits * m.wordsPerBitplane() * 8);
  }
  BitSerialMatrix::dealloc(m);
  return rc;
}

int gbsref_export(const uint64_t *planes, uint64_t nbits, int issigned,
                  uint64_t nrows, uint64_t ncols, uint64_t rowalign,
                  uint64_t colalign, int32_t *out) {
  BitSerialMatrix m = BitSerialMatrix::alloc(nbits, nrows, ncols,
                                             issigned != 0, rowalign, colalign);
  std::memcpy(m.data, planes, m.nbits * m.wordsPerBitplane() * 8);
  m.exportRegular(out);
  BitSerialMatrix::dealloc(m);
  return 0;
}

int gbsref_sum_rows(const uint64_t *planes, uint64_t nbits, int issigned,
                    uint64_t nrows, uint64_t ncols, uint64_t rowalign,
                    uint64_t colalign, int32_t *out) {
  BitSerialMatrix m = BitSerialMatrix::alloc(nbits, nrows, ncols,
                                             issigned != 0, rowalign, colalign);
  std::memcpy(m.data, planes, m.nbits * m.wordsPerBitplane() * 8);
  sumRows(m, out);
  BitSerialMatrix::dealloc(m);
  return 0;
}

// out = {lhsBlock, rhsBlock, lhs.nrows_a, lhs.ncols_a, rhs.nrows_a, rhs.ncols_a}
int gbsref_alloc_geometry(uint64_t lhsRows, uint64_t depth, uint64_t rhsRows,
                          uint64_t out[6]) {
  GEMMContext ctx = allocGEMMContext(lhsRows, depth, rhsRows, 1, 1, false, false);
  out[0] = ctx.lhsBlock;
  out[1] = ctx.rhsBlock;
  out[2] = ctx.lhs.nrows_a;
  out[3] = ctx.lhs.ncols_a;
  out[4] = ctx.rhs.nrows_a;
  out[5] = ctx.rhs.ncols_a;
  deallocGEMMContext(ctx);
  return 0;
}

// Full reference flow: allocGEMMContext, importRegular x2, gemmBitSerial.
// res receives the transposed result (rhsRows x lhsRows).  When
// lhs_planes/rhs_planes are non-null, the imported plane words are copied
// out too (sizes as allocated by the context).
int gbsref_gemm(uint64_t M, uint64_t K, uint64_t N, uint64_t lb, uint64_t rb,
                int ls, int rs, const void *lhs, int dtl, const void *rhs,
                int dtr, int32_t *res) {
  GEMMContext ctx = allocGEMMContext(M, K, N, lb, rb, ls != 0, rs != 0);
  int rc = do_import(ctx.lhs, lhs, dtl, false);
  if (rc == 0) rc = do_import(ctx.rhs, rhs, dtr, false);
  if (rc == 0) {
    gemmBitSerial(ctx);
    std::memcpy(res, ctx.res, M * N * sizeof(int32_t));
  }
  deallocGEMMContext(ctx);
  return rc;
}

// Time gemmBitSerial the way benchmark.cpp:93-111 does: import outside the
// timed loop, first iteration dropped as warm-up, std::chrono, single
// thread.  Runs until min_secs have passed after warm-up or max_reps timed
// iterations were done (whichever first, at least one).  Returns ns per call.
double gbsref_bench_gemm(uint64_t M, uint64_t K, uint64_t N, uint64_t lb,
                         uint64_t rb, int ls, int rs, const void *lhs, int dtl,
                         const void *rhs, int dtr, double min_secs,
                         int max_reps, int *reps_out, int32_t *res) {
  GEMMContext ctx = allocGEMMContext(M, K, N, lb, rb, ls != 0, rs != 0);
  do_import(ctx.lhs, lhs, dtl, false);
  do_import(ctx.rhs, rhs, dtr, false);
  gemmBitSerial(ctx);  // warm-up, not timed
  int reps = 0;
  auto start = std::chrono::high_resolution_clock::now();
  auto end = start;
  do {
    gemmBitSerial(ctx);
    reps++;
    end = std::chrono::high_resolution_clock::now();
  } while (reps < max_reps &&
           std::chrono::duration<double>(end - start).count() < min_secs);
  double ns = std::chrono::duration<double, std::nano>(end - start).count() / reps;
  if (res) std::memcpy(res, ctx.res, M * N * sizeof(int32_t));
  if (reps_out) *reps_out = reps;
  deallocGEMMContext(ctx);
  return ns;
}

}  // extern "C"
