/*
 * gbs_oracle.c -- CPU restatement of gemmbitserial's bit-serial path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker for the HIP library:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it, and only to check or time against.  The product library
 * (libgbs_hip.so, include/gemmbitserial.hpp) never links or calls it.
 *
 * Every function restates one piece of the reference, cited file:line
 * (paths relative to the reference checkout):
 *
 *   orc_align_to              gemmbitserial.hpp:11-17
 *   orc_compute_block_size    gemmbitserial.hpp:343-360
 *   orc_finetune_block_size   gemmbitserial.hpp:364-377
 *   orc_alloc_geometry        gemmbitserial.hpp:415-452 + arch-generic.hpp:78-100
 *   orc_quantize (static)     gemmbitserial.hpp:169-186 (template path),
 *                             gemmbitserial.hpp:249-275 (raw uint8_t* overload)
 *   orc_import                gemmbitserial.hpp:212-275 (import32As4x8 :192-207)
 *   orc_import_and_quantize   gemmbitserial.hpp:281-308
 *   orc_export                gemmbitserial.hpp:312-333
 *   orc_sum_rows              arch-generic.hpp:6-37
 *   orc_gemm                  arch-generic.hpp:39-76 (accumulator init),
 *                             :169-193 (plane-pair binary GEMMs),
 *                             :239-262 (bipolar XOR path),
 *                             :267-329 (GEMV paths), :331-345 (dispatch)
 *   orc_naive_gemm_i64        test/testhelpers.hpp:34-44 (same index
 *                             convention, widened to int64)
 *
 * Pinning (what checks each function against something other than itself):
 *   - tests/test_oracle.py compares orc_import/orc_export/orc_sum_rows/
 *     orc_gemm/orc_alloc_geometry word-for-word with the reference compiled
 *     from its own sources (oracle/_ref, built by oracle/Makefile), on seeded
 *     inputs including out-of-range values, bipolar 0s, the raw uint8 overload
 *     and column-major reads;
 *   - orc_gemm is checked against orc_naive_gemm_i64 and a float64 BLAS
 *     product (exact for these magnitudes), and against the MNIST
 *     known-answer vector held by the reference's tests (test/mnistdata.h);
 *   - orc_export(orc_import(x)) == x for in-range x (round trip), and
 *     bipolar identities a.b = K - 2 popcount(a^b).
 *
 * Arithmetic: int32 accumulation wraps modulo 2^32 exactly like the
 * reference on x86; it is done in uint32_t here to keep C defined.
 */
#include <stdint.h>
#include <string.h>
#include <stdlib.h>
#include <math.h>

/* element types accepted by orc_import; the values match gbs_dtype in
 * include/gbs_hip.h */
enum {
  ORC_I8 = 0, ORC_U8 = 1, ORC_U8_RAW = 2, ORC_I16 = 3, ORC_U16 = 4,
  ORC_I32 = 5, ORC_U32 = 6, ORC_I64 = 7, ORC_U64 = 8,
  ORC_F32 = 9, ORC_F64 = 10  /* orc_import_and_quantize only */
};

static int dtype_size(int dt) {
  switch (dt) {
    case ORC_I8: case ORC_U8: case ORC_U8_RAW: return 1;
    case ORC_I16: case ORC_U16: return 2;
    case ORC_I32: case ORC_U32: return 4;
    case ORC_I64: case ORC_U64: return 8;
  }
  return 0;
}

/* read element i of a typed buffer as a mathematical integer */
static int64_t load_elem(const void *src, int dt, uint64_t i) {
  switch (dt) {
    case ORC_I8: return ((const int8_t *)src)[i];
    case ORC_U8: case ORC_U8_RAW: return ((const uint8_t *)src)[i];
    case ORC_I16: return ((const int16_t *)src)[i];
    case ORC_U16: return ((const uint16_t *)src)[i];
    case ORC_I32: return ((const int32_t *)src)[i];
    case ORC_U32: return ((const uint32_t *)src)[i];
    case ORC_I64: return ((const int64_t *)src)[i];
    case ORC_U64: return (int64_t)((const uint64_t *)src)[i];
  }
  return 0;
}

static int dtype_unsigned(int dt) {
  return dt == ORC_U8 || dt == ORC_U8_RAW || dt == ORC_U16 || dt == ORC_U32 ||
         dt == ORC_U64;
}

/* gemmbitserial.hpp:11-17 */
uint64_t orc_align_to(uint64_t in, uint64_t af) {
  if (in % af != 0) return in + (af - (in % af));
  return in;
}

/* gemmbitserial.hpp:343-360 -- float quadratic for the cache-block rows */
void orc_compute_block_size(float lhsMult, float rhsMult, float cacheBits,
                            float dBits, uint64_t *lhsBlock,
                            uint64_t *rhsBlock) {
  float a = sizeof(int32_t) * lhsMult * rhsMult;
  float b = dBits * (lhsMult + rhsMult);
  float c = -cacheBits;
  float discr = sqrt(b * b - 4 * a * c);
  int64_t x0 = floor((-b + discr) / (2 * a));
  int64_t x1 = floor((-b - discr) / (2 * a));
  int64_t x = x0 > x1 ? x0 : x1;
  if (x > 0) {
    *lhsBlock = lhsMult * x;
    *rhsBlock = rhsMult * x;
  } else {
    *lhsBlock = lhsMult;
    *rhsBlock = rhsMult;
  }
}

/* gemmbitserial.hpp:364-377 */
uint64_t orc_finetune_block_size(uint64_t rows, uint64_t bs_max,
                                 uint64_t bs_div) {
  uint64_t best = bs_max;
  uint64_t min_pen = orc_align_to(rows, best) - rows;
  for (uint64_t cand = bs_max; cand > bs_div; cand = cand - bs_div) {
    if (cand % bs_div == 0) {
      uint64_t pen = orc_align_to(rows, cand) - rows;
      if (pen < min_pen) {
        best = cand;
        min_pen = pen;
      }
    }
  }
  return best;
}

/* gemmbitserial.hpp:415-452 with the generic register blocking of
 * arch-generic.hpp:78-100.  out[] = {lhsBlock, rhsBlock, lhs.nrows_a,
 * lhs.ncols_a, rhs.nrows_a, rhs.ncols_a}. */
void orc_alloc_geometry(uint64_t lhsRows, uint64_t depth, uint64_t rhsRows,
                        uint64_t out[6]) {
  uint64_t rl = 2, rd = 1, rr = 2;
  const uint64_t cacheBits = 32 * 1024 * 8;
  if (rhsRows == 1) { rl = 1; rd = 4; rr = 1; }
  uint64_t depth_al = orc_align_to(depth, rd * 64);
  uint64_t lb, rb;
  orc_compute_block_size((float)rl, (float)rr, (float)cacheBits,
                         (float)depth_al, &lb, &rb);
  if (lb > lhsRows || rb > rhsRows) {
    lb = orc_align_to(lhsRows, rl);
    rb = orc_align_to(rhsRows, rr);
  } else {
    if ((orc_align_to(lhsRows, lb) - lhsRows) > 0.1 * lhsRows)
      lb = orc_finetune_block_size(lhsRows, lb, rl);
    if ((orc_align_to(rhsRows, rb) - rhsRows) > 0.1 * rhsRows)
      rb = orc_finetune_block_size(rhsRows, rb, rr);
  }
  out[0] = lb;
  out[1] = rb;
  out[2] = orc_align_to(lhsRows, lb);
  out[3] = orc_align_to(depth, rd * 64);
  out[4] = orc_align_to(rhsRows, rb);
  out[5] = orc_align_to(depth, rd * 64);
}

/* gemmbitserial.hpp:169-186 (template) and :257-263 (raw uint8 overload).
 * Returns the uint8 code whose low nbits become the planes.  For integral
 * element types the non-bipolar branch reduces to "elem modulo 256" (the
 * signed branch adds 2^(n-1) twice, i.e. 2^n, before truncating). */
static uint8_t orc_quantize(int64_t elem, int dt, uint64_t nbits,
                            int issigned) {
  int bipolar = (nbits == 1) && issigned;
  if (dt == ORC_U8_RAW) return (uint8_t)elem;
  if (bipolar) {
    if (dtype_unsigned(dt)) return ((uint64_t)elem) > 0 ? 1 : 0;
    return elem > 0 ? 1 : 0;
  }
  if (issigned && !dtype_unsigned(dt) && elem < 0) {
    uint8_t ret = 0;
    ret += (uint8_t)(1 << (nbits - 1));
    ret += (uint8_t)(elem + (1 << (nbits - 1)));
    return ret;
  }
  return (uint8_t)elem;
}

/* gemmbitserial.hpp:212-275: clearAll, then for each real element set plane
 * bit b of word (row, col>>6) at position col&63 when code bit b is 1.  The
 * 4-column movmask path (import32As4x8, :192-207) and the tail path set the
 * same bits; only bits b < nbits of the code are examined.
 * planes layout: [nbits][nrows_a][ncols_a/64]. */
int orc_import(const void *src, int dt, uint64_t nbits, int issigned,
               uint64_t nrows, uint64_t ncols, uint64_t nrows_a,
               uint64_t ncols_a, int read_col_major, uint64_t *planes) {
  if (dtype_size(dt) == 0 || nbits < 1 || nbits > 8) return -1;
  const uint64_t wpr = ncols_a / 64;
  const uint64_t wpp = nrows_a * wpr;
  memset(planes, 0, nbits * wpp * sizeof(uint64_t));
  for (uint64_t r = 0; r < nrows; r++) {
    for (uint64_t c = 0; c < ncols; c++) {
      uint64_t idx = read_col_major ? c * nrows + r : r * ncols + c;
      uint8_t code = orc_quantize(load_elem(src, dt, idx), dt, nbits,
                                  issigned);
      for (uint64_t b = 0; b < nbits; b++) {
        if (code & (1u << b))
          planes[b * wpp + r * wpr + (c >> 6)] |= (uint64_t)1 << (c & 63);
      }
    }
  }
  return 0;
}

/* a[i] <= b[j], compared in the element type dt (gemmbitserial.hpp:292 compares
 * two T values) */
static int le_typed(const void *a, uint64_t i, const void *b, uint64_t j, int dt) {
  switch (dt) {
    case ORC_I8: return ((const int8_t *)a)[i] <= ((const int8_t *)b)[j];
    case ORC_U8: case ORC_U8_RAW: return ((const uint8_t *)a)[i] <= ((const uint8_t *)b)[j];
    case ORC_I16: return ((const int16_t *)a)[i] <= ((const int16_t *)b)[j];
    case ORC_U16: return ((const uint16_t *)a)[i] <= ((const uint16_t *)b)[j];
    case ORC_I32: return ((const int32_t *)a)[i] <= ((const int32_t *)b)[j];
    case ORC_U32: return ((const uint32_t *)a)[i] <= ((const uint32_t *)b)[j];
    case ORC_I64: return ((const int64_t *)a)[i] <= ((const int64_t *)b)[j];
    case ORC_U64: return ((const uint64_t *)a)[i] <= ((const uint64_t *)b)[j];
    case ORC_F32: return ((const float *)a)[i] <= ((const float *)b)[j];
    case ORC_F64: return ((const double *)a)[i] <= ((const double *)b)[j];
  }
  return 0;
}

/* (uint8_t) of element i of type dt (gemmbitserial.hpp:298): integers modulo
 * 256; floating values truncated toward zero, then the low 8 bits of the
 * int32 (C leaves a float outside (-1, 256) undefined; x86-64 code converts
 * through int32, giving 0 beyond +-2^31 and for NaN, which this follows). */
static uint8_t u8_typed(const void *a, uint64_t i, int dt) {
  double f;
  switch (dt) {
    case ORC_F32: f = ((const float *)a)[i]; break;
    case ORC_F64: f = ((const double *)a)[i]; break;
    default: return (uint8_t)load_elem(a, dt, i);
  }
  if (!(f > -2147483648.0 && f < 2147483648.0)) return 0;
  return (uint8_t)(int32_t)f;
}

/* gemmbitserial.hpp:281-308: unsigned-only threshold quantization of an
 * nrows x ncols matrix of element type dt (row-major, or col-major when
 * read_col_major); thresholds[t * nrows + r] of the same type for
 * t < nThres, sorted ascending per row.  The element becomes the index of
 * the first threshold it does not exceed, nThres when it exceeds them all,
 * and stays itself when nThres == 0 (the loop at :291 never runs); the low
 * nbits of its uint8_t conversion are set. */
int orc_import_and_quantize(const void *src, const void *thresholds, int dt,
                            int nThres, uint64_t nbits, uint64_t nrows,
                            uint64_t ncols, uint64_t nrows_a,
                            uint64_t ncols_a, int read_col_major,
                            uint64_t *planes) {
  if (nbits < 1 || nbits > 8 || nThres < 0 || dt < ORC_I8 || dt > ORC_F64) return -1;
  const uint64_t wpr = ncols_a / 64;
  const uint64_t wpp = nrows_a * wpr;
  memset(planes, 0, nbits * wpp * sizeof(uint64_t));
  for (uint64_t r = 0; r < nrows; r++) {
    for (uint64_t c = 0; c < ncols; c++) {
      const uint64_t idx = read_col_major ? c * nrows + r : r * ncols + c;
      uint8_t code = nThres == 0 ? u8_typed(src, idx, dt) : (uint8_t)nThres;
      for (int t = 0; t < nThres; t++) {
        if (le_typed(src, idx, thresholds, (uint64_t)t * nrows + r, dt)) {
          code = (uint8_t)t;
          break;
        }
      }
      for (uint64_t b = 0; b < nbits; b++)
        if (code & (1u << b))
          planes[b * wpp + r * wpr + (c >> 6)] |= (uint64_t)1 << (c & 63);
    }
  }
  return 0;
}

static inline int getbit(const uint64_t *planes, uint64_t wpp, uint64_t wpr,
                         uint64_t b, uint64_t r, uint64_t c) {
  return (planes[b * wpp + r * wpr + (c >> 6)] >> (c & 63)) & 1;
}

/* gemmbitserial.hpp:312-333: bipolar -> +-1; otherwise sum of 2^b with the
 * top plane weighted -2^(n-1) when signed.  Always row-major output. */
void orc_export(const uint64_t *planes, uint64_t nbits, int issigned,
                uint64_t nrows, uint64_t ncols, uint64_t nrows_a,
                uint64_t ncols_a, int32_t *out) {
  const uint64_t wpr = ncols_a / 64;
  const uint64_t wpp = nrows_a * wpr;
  int bipolar = (nbits == 1) && issigned;
  for (uint64_t r = 0; r < nrows; r++) {
    for (uint64_t c = 0; c < ncols; c++) {
      if (bipolar) {
        out[r * ncols + c] = getbit(planes, wpp, wpr, 0, r, c) ? 1 : -1;
      } else {
        int32_t v = 0;
        for (uint64_t b = 0; b < nbits; b++) {
          if (getbit(planes, wpp, wpr, b, r, c)) {
            if (b == nbits - 1 && issigned) v -= 1 << b;
            else v += 1 << b;
          }
        }
        out[r * ncols + c] = v;
      }
    }
  }
}

/* arch-generic.hpp:6-37 */
void orc_sum_rows(const uint64_t *planes, uint64_t nbits, int issigned,
                  uint64_t nrows, uint64_t ncols, uint64_t nrows_a,
                  uint64_t ncols_a, int32_t *row_sums) {
  const uint64_t wpr = ncols_a / 64;
  const uint64_t wpp = nrows_a * wpr;
  int bipolar = (nbits == 1) && issigned;
  for (uint64_t r = 0; r < nrows; r++) {
    uint32_t acc = 0;
    if (bipolar) {
      for (uint64_t w = 0; w < wpr; w++)
        acc += (uint32_t)__builtin_popcountll(planes[r * wpr + w]);
      row_sums[r] = (int32_t)(2u * acc - (uint32_t)ncols);
    } else {
      for (uint64_t b = 0; b < nbits; b++) {
        uint32_t bit_acc = 0;
        for (uint64_t w = 0; w < wpr; w++)
          bit_acc += (uint32_t)__builtin_popcountll(planes[b * wpp + r * wpr + w]);
        bit_acc = bit_acc << b;
        if (issigned && b == nbits - 1) bit_acc = 0u - bit_acc;
        acc += bit_acc;
      }
      row_sums[r] = (int32_t)acc;
    }
  }
}

/* plane-matrix descriptor passed from Python */
typedef struct {
  const uint64_t *planes;
  uint64_t nbits, nrows, ncols, nrows_a, ncols_a;
  int32_t issigned;
} orc_mat;

static int is_bipolar(const orc_mat *m) { return m->nbits == 1 && m->issigned; }

/* arch-generic.hpp:39-76: zero, or -rowsum(regular) for bipolar x regular
 * (lhs bipolar: indexed by the output's rhs row; rhs bipolar: by lhs row). */
static void prepare_accumulators(const orc_mat *l, const orc_mat *r,
                                 uint32_t *res) {
  int lb = is_bipolar(l), rb = is_bipolar(r);
  uint64_t M = l->nrows, N = r->nrows;
  if (lb ^ rb) {
    const orc_mat *reg = lb ? r : l;
    int32_t *rs = (int32_t *)malloc((reg->nrows ? reg->nrows : 1) * sizeof(int32_t));
    orc_sum_rows(reg->planes, reg->nbits, reg->issigned, reg->nrows,
                 reg->ncols, reg->nrows_a, reg->ncols_a, rs);
    for (uint64_t n = 0; n < N; n++)
      for (uint64_t m = 0; m < M; m++)
        res[n * M + m] = (uint32_t)(-(int64_t)(lb ? rs[n] : rs[m]));
    free(rs);
  } else {
    memset(res, 0, M * N * sizeof(uint32_t));
  }
}

/* arch-generic.hpp:331-345 dispatch.  The GEMV paths (:267-329) and the
 * plane-pair path (:169-193) compute the same sums in the same modulus; the
 * restatement follows the plane-pair form for all non-bipolar-x-bipolar
 * cases and the XOR form (:239-262, :306-329) for bipolar x bipolar.
 * res is N x M row-major: res[n*M + m]. */
int orc_gemm(const orc_mat *l, const orc_mat *r, int32_t *res_out) {
  if (l->ncols != r->ncols) return -1;
  uint64_t M = l->nrows, N = r->nrows;
  uint64_t lw = l->ncols_a / 64, rw = r->ncols_a / 64;
  uint64_t dw = lw < rw ? lw : rw;
  uint64_t lwpp = l->nrows_a * lw, rwpp = r->nrows_a * rw;
  uint32_t *res = (uint32_t *)res_out;
  prepare_accumulators(l, r, res);
  if (is_bipolar(l) && is_bipolar(r)) {
    for (uint64_t n = 0; n < N; n++)
      for (uint64_t m = 0; m < M; m++) {
        uint32_t pc = 0;
        for (uint64_t k = 0; k < dw; k++)
          pc += (uint32_t)__builtin_popcountll(l->planes[m * lw + k] ^
                                               r->planes[n * rw + k]);
        res[n * M + m] += (uint32_t)l->ncols - 2u * pc;
      }
    return 0;
  }
  int bpreg = is_bipolar(l) ^ is_bipolar(r);
  for (uint64_t lb = 0; lb < l->nbits; lb++) {
    int neg_l = l->issigned && !is_bipolar(l) && lb == l->nbits - 1;
    for (uint64_t rb = 0; rb < r->nbits; rb++) {
      int neg_r = r->issigned && !is_bipolar(r) && rb == r->nbits - 1;
      uint32_t alpha = 1u << (lb + rb);
      if (neg_l ^ neg_r) alpha = 0u - alpha;
      if (bpreg) alpha = 2u * alpha;
      const uint64_t *A = l->planes + lb * lwpp;
      const uint64_t *B = r->planes + rb * rwpp;
      for (uint64_t n = 0; n < N; n++)
        for (uint64_t m = 0; m < M; m++) {
          uint32_t acc = 0;
          for (uint64_t k = 0; k < dw; k++)
            acc += (uint32_t)__builtin_popcountll(A[m * lw + k] & B[n * rw + k]);
          res[n * M + m] += acc * alpha;
        }
    }
  }
  return 0;
}

/* test/testhelpers.hpp:34-44 index convention (res[k*rows+i] with rhs given
 * as cols x depth), accumulated in int64 so it cannot wrap. */
void orc_naive_gemm_i64(const int32_t *lhs, const int32_t *rhs, int64_t *res,
                        uint64_t rows, uint64_t depth, uint64_t cols) {
  for (uint64_t k = 0; k < cols; k++)
    for (uint64_t i = 0; i < rows; i++) {
      int64_t acc = 0;
      for (uint64_t j = 0; j < depth; j++)
        acc += (int64_t)lhs[i * depth + j] * (int64_t)rhs[k * depth + j];
      res[k * rows + i] = acc;
    }
}
