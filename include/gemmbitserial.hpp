// gemmbitserial.hpp -- drop-in replacement for the header-only gemmbitserial
// API, running the bit-serial GEMM on an AMD MI355X through libgbs_hip.so.
//
// Same namespace, class names, member names and argument order as the
// reference (gemmbitserial.hpp:8-479, arch-generic.hpp:331-345):
//
//   GEMMContext ctx = allocGEMMContext(M, K, N, lhsBits, rhsBits, lhsSigned, rhsSigned);
//   ctx.lhs.importRegular(lhs);        // M x K, row-major
//   ctx.rhs.importRegular(rhs);        // N x K ("transposed" rhs)
//   gemmBitSerial(ctx);                // ctx.res[n*M + m], N x M int32
//   deallocGEMMContext(ctx);
//
// What changes underneath:
//   - importRegular packs on the GPU (one thread per plane word or quarter
//     word, plane bits gathered by byte dot products) into the device planes;
//   - gemmBitSerial runs one kernel over all plane pairs (a matrix-core
//     engine where it holds the product exactly, else the popcount engine)
//     and copies the N x M result into `res`;
//   - `data` and `res` are pinned host memory (gbs_host_alloc).
//
// Coherence.  BitSerialMatrix and GEMMContext are copied by value, as in the
// reference; copies share `data` and a small control block holding the
// device planes.  The host mirror `data` is filled lazily: after an import
// the packed words stay on the device until a host accessor (get, set,
// unset, word, rowptr, bitplaneptr, printHex, import32As4x8) or
// syncToHost() needs them, so an import followed by gemmBitSerial moves no
// planes over PCIe.  Host-side mutators mark the host copy newer;
// gemmBitSerial, sumRows and exportRegular upload it first.  Code that reads
// `data` directly calls syncToHost() first; code that writes it directly
// calls markHostDirty() (or syncToDevice()) afterwards.
//
// Errors.  The reference asserts; here a failed device call throws
// gemmbitserial::Error with the library's message.
#pragma once

#include <math.h>
#include <stdint.h>
#include <string.h>

#include <cassert>
#include <iostream>
#include <stdexcept>
#include <string>
#include <type_traits>

#include "gbs_hip.h"

namespace gemmbitserial {

class Error : public std::runtime_error {
 public:
  Error(const std::string &what, int code)
      : std::runtime_error(what), code(code) {}
  int code;
};

namespace detail {

inline void check(int rc, const char *what) {
  if (rc != GBS_OK)
    throw Error(std::string("gemmbitserial: ") + what + ": " + gbs_last_error(), rc);
}

// Element type -> gbs_dtype for the template importRegular<T>.
template <typename T>
inline gbs_dtype dtype_of() {
  typedef typename std::remove_cv<T>::type U;
  static_assert(std::is_integral<U>::value,
                "importRegular needs an integral element type");
  static_assert(sizeof(U) <= 8, "element type wider than 64 bits");
  return sizeof(U) == 1 ? (std::is_signed<U>::value ? GBS_I8 : GBS_U8)
       : sizeof(U) == 2 ? (std::is_signed<U>::value ? GBS_I16 : GBS_U16)
       : sizeof(U) == 4 ? (std::is_signed<U>::value ? GBS_I32 : GBS_U32)
                        : (std::is_signed<U>::value ? GBS_I64 : GBS_U64);
}

// Element type -> gbs_dtype for importRegularAndQuantize<T>, which also
// takes floating element types.
template <typename U>
inline gbs_dtype quantize_dtype_impl(std::true_type) {
  static_assert(sizeof(U) == 4 || sizeof(U) == 8, "float or double");
  return sizeof(U) == 4 ? GBS_F32 : GBS_F64;
}
template <typename U>
inline gbs_dtype quantize_dtype_impl(std::false_type) {
  return dtype_of<U>();
}
template <typename T>
inline gbs_dtype quantize_dtype_of() {
  typedef typename std::remove_cv<T>::type U;
  static_assert(std::is_arithmetic<U>::value,
                "importRegularAndQuantize needs an arithmetic element type");
  return quantize_dtype_impl<U>(std::is_floating_point<U>());
}

// Shared by every copy of one BitSerialMatrix.
struct Shared {
  gbs_matrix *dev;
  bool host_dirty;     // host `data` was changed after the last device write
  bool device_newer;   // the device planes were written after `data` was filled
};

}  // namespace detail

// Utility function to increment-and-align "in" to "af"
// (gemmbitserial.hpp:11-17).
inline uint64_t alignTo(uint64_t in, uint64_t af) {
  return (in % af != 0) ? in + (af - (in % af)) : in;
}

class BitSerialMatrix {
 public:
  // Allocate a bit-serial matrix: pinned host mirror + device planes, both
  // zero (gemmbitserial.hpp:24-35).
  static BitSerialMatrix alloc(uint64_t nbits, uint64_t nrows, uint64_t ncols,
                               bool issigned, uint64_t rowalign = 1,
                               uint64_t colalign = 64) {
    BitSerialMatrix bsm;
    bsm.nbits = nbits;
    bsm.nrows = nrows;
    bsm.ncols = ncols;
    bsm.nrows_a = alignTo(nrows, rowalign);
    bsm.ncols_a = alignTo(ncols, colalign);
    bsm.issigned = issigned;
    bsm.dev = new detail::Shared();
    bsm.dev->dev = nullptr;
    bsm.dev->host_dirty = false;
    bsm.dev->device_newer = false;
    detail::check(gbs_matrix_create(nbits, nrows, ncols, issigned, bsm.nrows_a,
                                    bsm.ncols_a, &bsm.dev->dev),
                  "BitSerialMatrix::alloc");
    void *p = nullptr;
    detail::check(gbs_host_alloc(bsm.nbits * bsm.wordsPerBitplane() * sizeof(uint64_t), &p),
                  "BitSerialMatrix::alloc (host mirror)");
    bsm.data = static_cast<uint64_t *>(p);
    memset(bsm.data, 0, bsm.nbits * bsm.wordsPerBitplane() * sizeof(uint64_t));
    return bsm;
  }

  // gemmbitserial.hpp:38-40
  static void dealloc(BitSerialMatrix bsm) {
    gbs_host_free(bsm.data);
    if (bsm.dev) {
      gbs_matrix_destroy(bsm.dev->dev);
      delete bsm.dev;
    }
  }

 public:
  bool issigned;        // whether highest order bit pos is negative
  uint64_t nbits;       // bits of precision
  uint64_t nrows;       // number of real (actual) rows
  uint64_t ncols;       // number of real (actual) columns
  uint64_t nrows_a;     // number of allocated rows
  uint64_t ncols_a;     // number of allocated columns
  uint64_t *data;       // host mirror, layout [nbits][nrows_a][ncols_a/64]
  detail::Shared *dev;  // device planes (shared by copies)

  void printSummary() {
    std::cout << "BitSerialMatrix" << std::endl;
    std::cout << "Bits of precision: " << nbits << " signed: " << issigned << std::endl;
    std::cout << "Actual size: " << nrows << " x " << ncols << std::endl;
    std::cout << "Allocated size: " << nrows_a << " x " << ncols_a << std::endl;
  }

  void printHex() {
    syncToHost();
    for (uint64_t i = 0; i < nbits; i++) {
      std::cout << "Bit " << i << ":" << std::endl;
      for (uint64_t j = 0; j < nrows_a; j++) {
        for (uint64_t k = 0; k < ncols_a / 64; k++)
          std::cout << std::hex << data[i * wordsPerBitplane() + j * wordsPerRow() + k]
                    << " " << std::dec;
        std::cout << std::endl;
      }
      std::cout << std::endl;
    }
  }

  inline bool isBipolar() const { return nbits == 1 && issigned; }
  inline uint64_t wordsPerRow() const { return ncols_a / 64; }
  inline uint64_t wordsPerBitplane() const { return nrows_a * wordsPerRow(); }

  inline bool get(uint64_t bit, uint64_t row, uint64_t col) {
    if (dev && dev->device_newer) syncToHost();
    return ((data[bit * wordsPerBitplane() + row * wordsPerRow() + (col >> 6)] >>
             bitpos(col)) & 1) == 1;
  }
  inline void clearAll() {
    memset(data, 0, nbits * wordsPerBitplane() * sizeof(uint64_t));
    if (dev) dev->device_newer = false;  // every word overwritten
    markHostDirty();
  }
  inline void set(uint64_t bit, uint64_t row, uint64_t col) {
    word(bit, row, col) |= (uint64_t)1 << bitpos(col);
  }
  inline void unset(uint64_t bit, uint64_t row, uint64_t col) {
    word(bit, row, col) &= ~((uint64_t)1 << bitpos(col));
  }
  // Writable reference into the host mirror; the host copy becomes newer.
  inline uint64_t &word(uint64_t bit, uint64_t row, uint64_t col) {
    syncToHost();
    markHostDirty();
    return data[bit * wordsPerBitplane() + row * wordsPerRow() + (col >> 6)];
  }
  inline uint64_t *rowptr(uint64_t bit, uint64_t row) {
    syncToHost();
    markHostDirty();
    return &data[bit * wordsPerBitplane() + row * wordsPerRow()];
  }
  inline uint64_t *bitplaneptr(uint64_t bit) {
    syncToHost();
    markHostDirty();
    return &data[bit * wordsPerBitplane()];
  }
  uint64_t bitpos(uint64_t col) { return col & 63; }

  // Host-side mirror changed outside the accessors above.
  inline void markHostDirty() {
    if (dev) dev->host_dirty = true;
  }
  // Upload the host mirror if it is newer than the device planes.
  inline void syncToDevice() {
    if (dev && dev->host_dirty) {
      detail::check(gbs_planes_from_host(dev->dev, data), "syncToDevice");
      dev->host_dirty = false;
    }
  }
  // Fill the host mirror if the device planes are newer (after an import).
  inline void syncToHost() {
    if (dev && dev->device_newer) {
      detail::check(gbs_planes_to_host(dev->dev, data), "syncToHost");
      dev->device_newer = false;
    }
  }

  // Same result as importRegular (gemmbitserial.hpp:135-164); packed on the GPU.
  template <typename T>
  void importRegular_naive(T *matrix, bool readColMajor = false) {
    importFrom(matrix, detail::dtype_of<T>(), readColMajor);
  }

  // The code an element maps to (gemmbitserial.hpp:169-186).
  template <typename T>
  inline uint8_t quantize(T currentElem) {
    if (isBipolar()) return currentElem > 0 ? 1 : 0;
    if (issigned && currentElem < 0)
      return (uint8_t)((uint8_t)(1 << (nbits - 1)) +
                       (uint8_t)(currentElem + (1 << (nbits - 1))));
    return (uint8_t)currentElem;
  }

  // Set the bits of four packed codes at row r, columns c..c+3 of the host
  // mirror (gemmbitserial.hpp:192-207).
  inline void import32As4x8(const uint32_t igroup, const uint64_t r, const uint64_t c) {
    for (uint64_t i = 0; i < 4; i++) {
      const uint32_t code = (igroup >> (8 * i)) & 0xff;
      for (uint64_t b = 0; b < nbits; b++)
        if (code & (1u << b)) set(b, r, c + i);
    }
  }

  // Pack a regular matrix (row-major, or column-major when readColMajor) on
  // the GPU (gemmbitserial.hpp:212-245).
  template <typename T>
  void importRegular(T *matrix, bool readColMajor = false) {
    importFrom(matrix, detail::dtype_of<T>(), readColMajor);
  }

  // uint8_t overload: raw low bits, no quantize (gemmbitserial.hpp:249-275).
  void importRegular(uint8_t *matrix, bool readColMajor = false) {
    importFrom(matrix, GBS_U8_RAW, readColMajor);
  }

  // Threshold quantization for unsigned matrices; thresholds[t*nrows + r]
  // ascending per row, compared in T (gemmbitserial.hpp:281-308).
  template <typename T>
  void importRegularAndQuantize(T *matrix, T *thresholds, int nThres,
                                bool readColMajor = false) {
    if (issigned)
      throw Error("gemmbitserial: importRegularAndQuantize needs an unsigned matrix",
                  GBS_EINVAL);
    detail::check(gbs_import_quantize_typed(dev->dev, matrix, thresholds,
                                            detail::quantize_dtype_of<T>(), nThres,
                                            readColMajor, 0, nullptr),
                  "importRegularAndQuantize");
    dev->host_dirty = false;
    dev->device_newer = true;
  }

  // Unpack into a row-major nrows x ncols matrix (gemmbitserial.hpp:312-333).
  template <typename T>
  void exportRegular(T *matrix) {
    syncToDevice();
    int32_t *tmp = new int32_t[nrows * ncols + 1];
    int rc = gbs_export(dev->dev, tmp, 0);
    if (rc == GBS_OK)
      for (uint64_t i = 0; i < nrows * ncols; i++) matrix[i] = (T)tmp[i];
    delete[] tmp;
    detail::check(rc, "exportRegular");
  }

 private:
  void importFrom(const void *matrix, gbs_dtype dt, bool readColMajor) {
    detail::check(gbs_import(dev->dev, matrix, dt, readColMajor, 0, nullptr), "importRegular");
    dev->host_dirty = false;
    dev->device_newer = true;
  }
};

// Cache-block sizing of the reference (gemmbitserial.hpp:336-377).  The GPU
// kernels do not use the blocks; they fix nrows_a / ncols_a so that the plane
// words in `data` are laid out exactly as the reference lays them out.
static void computeBlockSize(float lhsMult, float rhsMult, float cacheBits, float dBits,
                             uint64_t &lhsBlock, uint64_t &rhsBlock) {
  const float a = sizeof(int32_t) * lhsMult * rhsMult;
  const float b = dBits * (lhsMult + rhsMult);
  const float c = -cacheBits;
  const float discr = sqrt(b * b - 4 * a * c);
  assert(discr > 0);
  const int64_t x0 = floor((-b + discr) / (2 * a));
  const int64_t x1 = floor((-b - discr) / (2 * a));
  const int64_t x = x0 > x1 ? x0 : x1;
  if (x > 0) {
    lhsBlock = lhsMult * x;
    rhsBlock = rhsMult * x;
  } else {
    lhsBlock = lhsMult;
    rhsBlock = rhsMult;
  }
}

static uint64_t finetuneBlockSize(uint64_t rows, uint64_t bs_max, uint64_t bs_div) {
  uint64_t best = bs_max;
  uint64_t min_pen = alignTo(rows, best) - rows;
  for (uint64_t cand = bs_max; cand > bs_div; cand -= bs_div) {
    if (cand % bs_div == 0) {
      const uint64_t pen = alignTo(rows, cand) - rows;
      if (pen < min_pen) {
        best = cand;
        min_pen = pen;
      }
    }
  }
  return best;
}

class GEMMContext {
 public:
  BitSerialMatrix lhs, rhs;
  uint64_t lhsBlock, rhsBlock;
  int32_t *res;  // pinned host result, rhs.nrows x lhs.nrows

  void printSummary() {
    std::cout << "GEMMContext" << std::endl;
    std::cout << "LHS: ";
    lhs.printSummary();
    std::cout << "Block size: " << lhsBlock << std::endl;
    std::cout << "RHS: ";
    rhs.printSummary();
    std::cout << "Block size: " << rhsBlock << std::endl;
    float actual_ops = 2 * lhs.nrows * lhs.ncols * rhs.nrows;
    float alloc_ops = 2 * lhs.nrows_a * lhs.ncols_a * rhs.nrows_a;
    std::cout << "Actual ops: " << actual_ops << std::endl;
    std::cout << "Allocated ops: " << alloc_ops << std::endl;
    std::cout << "Actual op percentage: " << 100 * actual_ops / alloc_ops << std::endl;
  }

  inline bool isBipolarTimesRegular() const { return lhs.isBipolar() != rhs.isBipolar(); }
  inline bool isBipolarTimesBipolar() const { return lhs.isBipolar() && rhs.isBipolar(); }
  inline bool isMatrixVector() const { return rhs.nrows == 1; }
};

// Block sizes and padded sizes allocGEMMContext_base chooses; a pure
// function of the shape, so the layout can be checked without a device.
struct ContextGeometry {
  uint64_t lhsBlock, rhsBlock;
  uint64_t lhs_nrows_a, lhs_ncols_a, rhs_nrows_a, rhs_ncols_a;
};

// The blocking steps of allocGEMMContext_base (gemmbitserial.hpp:415-447).
static inline ContextGeometry contextGeometry_base(const uint64_t lhsRows, const uint64_t depth,
                                            const uint64_t rhsRows,
                                            const uint64_t regblock_lhs,
                                            const uint64_t regblock_d,
                                            const uint64_t regblock_rhs,
                                            const uint64_t cacheBits) {
  ContextGeometry g;
  const uint64_t depth_al = alignTo(depth, regblock_d * 64);
  computeBlockSize(regblock_lhs, regblock_rhs, cacheBits, depth_al, g.lhsBlock, g.rhsBlock);
  if (g.lhsBlock > lhsRows || g.rhsBlock > rhsRows) {
    g.lhsBlock = alignTo(lhsRows, regblock_lhs);
    g.rhsBlock = alignTo(rhsRows, regblock_rhs);
  } else {
    if ((alignTo(lhsRows, g.lhsBlock) - lhsRows) > 0.1 * lhsRows)
      g.lhsBlock = finetuneBlockSize(lhsRows, g.lhsBlock, regblock_lhs);
    if ((alignTo(rhsRows, g.rhsBlock) - rhsRows) > 0.1 * rhsRows)
      g.rhsBlock = finetuneBlockSize(rhsRows, g.rhsBlock, regblock_rhs);
  }
  // zero-row operands: keep alignment well defined (the reference divides by 0)
  const uint64_t la = g.lhsBlock ? g.lhsBlock : 1, ra = g.rhsBlock ? g.rhsBlock : 1;
  g.lhs_nrows_a = alignTo(lhsRows, la);
  g.rhs_nrows_a = alignTo(rhsRows, ra);
  g.lhs_ncols_a = g.rhs_ncols_a = alignTo(depth, regblock_d * 64);
  return g;
}

// gemmbitserial.hpp:415-452
static GEMMContext allocGEMMContext_base(const uint64_t lhsRows, const uint64_t depth,
                                         const uint64_t rhsRows, const uint64_t lhsBits,
                                         const uint64_t rhsBits, const bool lhsSigned,
                                         const bool rhsSigned, const uint64_t regblock_lhs,
                                         const uint64_t regblock_d,
                                         const uint64_t regblock_rhs,
                                         const uint64_t cacheBits) {
  GEMMContext ret;
  const ContextGeometry g = contextGeometry_base(lhsRows, depth, rhsRows, regblock_lhs,
                                                 regblock_d, regblock_rhs, cacheBits);
  ret.lhsBlock = g.lhsBlock;
  ret.rhsBlock = g.rhsBlock;
  const uint64_t la = g.lhsBlock ? g.lhsBlock : 1, ra = g.rhsBlock ? g.rhsBlock : 1;
  ret.lhs = BitSerialMatrix::alloc(lhsBits, lhsRows, depth, lhsSigned, la, regblock_d * 64);
  ret.rhs = BitSerialMatrix::alloc(rhsBits, rhsRows, depth, rhsSigned, ra, regblock_d * 64);
  void *p = nullptr;
  detail::check(gbs_host_alloc(lhsRows * rhsRows * sizeof(int32_t), &p), "allocGEMMContext");
  ret.res = static_cast<int32_t *>(p);
  return ret;
}

// The generic register blocking of arch-generic.hpp:78-100 (2x1x2, or 1x4x1
// for matrix-vector), as allocGEMMContext applies it.
static inline ContextGeometry contextGeometry(uint64_t lhsRows, uint64_t depth, uint64_t rhsRows) {
  const uint64_t cacheBits = 32 * 1024 * 8;
  if (rhsRows == 1) return contextGeometry_base(lhsRows, depth, rhsRows, 1, 4, 1, cacheBits);
  return contextGeometry_base(lhsRows, depth, rhsRows, 2, 1, 2, cacheBits);
}

// The generic register blocking of arch-generic.hpp:78-100 (2x1x2, or 1x4x1
// for matrix-vector), which fixes the padded sizes.
static GEMMContext allocGEMMContext(uint64_t lhsRows, uint64_t depth, uint64_t rhsRows,
                                    uint64_t lhsBits, uint64_t rhsBits, bool lhsSigned,
                                    bool rhsSigned) {
  const uint64_t cacheBits = 32 * 1024 * 8;
  if (rhsRows == 1)
    return allocGEMMContext_base(lhsRows, depth, rhsRows, lhsBits, rhsBits, lhsSigned,
                                 rhsSigned, 1, 4, 1, cacheBits);
  return allocGEMMContext_base(lhsRows, depth, rhsRows, lhsBits, rhsBits, lhsSigned,
                               rhsSigned, 2, 1, 2, cacheBits);
}

// README.md spelling of the allocator.
static inline GEMMContext allocateGEMMContext(uint64_t lhsRows, uint64_t depth,
                                              uint64_t rhsRows, uint64_t lhsBits,
                                              uint64_t rhsBits, bool lhsSigned,
                                              bool rhsSigned) {
  return allocGEMMContext(lhsRows, depth, rhsRows, lhsBits, rhsBits, lhsSigned, rhsSigned);
}

// gemmbitserial.hpp:454-458
static void deallocGEMMContext(GEMMContext ctx) {
  gbs_host_free(ctx.res);
  BitSerialMatrix::dealloc(ctx.lhs);
  BitSerialMatrix::dealloc(ctx.rhs);
}

// res = lhs * rhs^T on the GPU (arch-generic.hpp:331-345 and everything it
// dispatches to).  Synchronous: ctx.res holds the result on return.
static void gemmBitSerial(GEMMContext ctx) {
  ctx.lhs.syncToDevice();
  ctx.rhs.syncToDevice();
  detail::check(gbs_gemm(nullptr, ctx.lhs.dev->dev, ctx.rhs.dev->dev, ctx.res, 0,
                         GBS_ALGO_AUTO),
                "gemmBitSerial");
}

// Row sums of the values (arch-generic.hpp:6-37).
static void sumRows(BitSerialMatrix m, int32_t *row_sums) {
  m.syncToDevice();
  detail::check(gbs_sum_rows(m.dev->dev, row_sums, 0), "sumRows");
}

}  // namespace gemmbitserial
