// gbs_mfma.hip -- exact matrix-core engine for the bit-serial GEMM on gfx950.
//
// A matrix's planes are expanded once per change into a value copy
// ([align(rows,256)][align(K,256) values], zero padded) and the product is
// a plain GEMM on the matrix cores:
//
//   res[n*M + m] = sum_k val(rhs[n,k]) * val(lhs[m,k])      (int32)
//
// Three copy formats:
//  * int8 (one byte per value): bipolar +-1, signed two's complement of up
//    to 8 bits, unsigned codes of up to 7 bits; int32 accumulation, exact
//    whenever K * max|val(lhs)| * max|val(rhs)| < 2^31 (mfma_supported);
//  * FP4 E2M1 (two values per byte, scale 2^1, so codes 0..4 are the
//    integers 0..4): bipolar, unsigned of up to 2 bits, signed of up to 3
//    bits;
//  * FP6 E2M3 (6 bits per value, scale 2^3, so codes 0..16 are the
//    integers 0..16): bipolar, unsigned of up to 4 bits, signed of up to 5
//    bits.
// FP4 and FP6 products accumulate integer products in fp32, exact whenever
// K * max|val(lhs)| * max|val(rhs)| < 2^24 (f4_supported, f6_supported),
// since every partial sum is then an integer below 2^24.  They take half
// and three quarters of the bytes of int8, and one instruction covers
// twice the values.
// Either way the result equals the reference's int32 result bit for bit.
//
// The tile itself (gbs_mfma_tile.h) comes in several shapes; a product runs
// the shape and K split its (M, N, K) class maps to in the tuning table of
// its format (written by gemmbitserial_amd/tune.py and loaded through
// gbs_tune_load, else a size heuristic).  Split products zero the result
// first and add their partial sums with integer atomics, which is exact in
// any order.
#include "gbs_mfma_tile.h"
#include "gbs_skinny.h"

namespace gbs {
namespace {

template <int P, int Q, int WP, int KSB, int DEPTH, int FMT, int KG>
__global__ __launch_bounds__(256 * KG,
                             (P * Q >= 65536 || (P * Q >= 32768 && DEPTH == 0)) ? 1 : 2)
void mfma_kernel(MfmaOperands a, MfmaPlan pl) {
  __shared__ __attribute__((aligned(16))) int8_t lds[mfma_lds_bytes<P, Q, KSB, FMT, DEPTH, KG>()];
  const uint32_t T = gridDim.x, b = blockIdx.x;
  // consecutive logical tiles run on one XCD (blocks are dealt round-robin)
  const uint64_t t = (T % 8 == 0) ? (uint64_t)(b % 8) * (T / 8) + b / 8 : b;
  uint32_t pt, qt, sp;
  plan_decode(pl, t, pt, qt, sp);
  const uint32_t kbeg = sp * pl.kchunk;
  const uint32_t kend = min(a.kbytes, kbeg + pl.kchunk);
  if constexpr (direct_loads(DEPTH, KG))
    mfma_tile_dl<P, Q, WP, KSB, DEPTH, FMT, KG>(a, pt, qt, kbeg, kend, pl.splits > 1, lds);
  else if constexpr (DEPTH == 0)
    mfma_tile_pp<P, Q, WP, FMT>(a, pt, qt, kbeg, kend, pl.splits > 1, lds);
  else
    mfma_tile<P, Q, WP, KSB, DEPTH, FMT>(a, pt, qt, kbeg, kend, pl.splits > 1, lds);
}

// planes -> int8 values: one thread per 64-value word of a copy row (64
// bytes, four 16-byte stores).  Four values at a time, bit b of each code
// is spread into bit b of its byte by one multiply ((x & 0xF) * 0x00204081
// puts bit i of x at bit 8i, with no two partial products meeting); then
// bipolar codes 0/1 become 0xFF/0x01 and signed codes are sign-extended
// bytewise.  Zero outside the real matrix, also for bipolar operands.
__device__ __forceinline__ void expand_i8_word(const ExpandJob &j, uint64_t i) {
  const uint64_t *planes = j.planes, nrows = j.nrows, ncols = j.ncols, wpr = j.wpr, W = j.W;
  const uint32_t nbits = j.nbits;
  const bool issigned = j.issigned;
  const uint64_t ld = j.ld;
  int8_t *out = (int8_t *)j.out;
  const uint64_t wpl = ld / 64;
  const uint64_t r = i / wpl, w = i - r * wpl;
  uint32_t q[16];
#pragma unroll
  for (int t = 0; t < 16; t++) q[t] = 0;
  if (r < nrows && w * 64 < ncols) {
    uint64_t p[8];
#pragma unroll
    for (uint32_t b = 0; b < 8; b++) p[b] = b < nbits ? planes[b * W + r * wpr + w] : 0;
#pragma unroll
    for (int t = 0; t < 16; t++) {
      uint32_t v = 0;
#pragma unroll
      for (uint32_t b = 0; b < 8; b++)
        if (b < nbits) v |= ((((uint32_t)(p[b] >> (4 * t)) & 0xFu) * 0x00204081u) & 0x01010101u) << b;
      q[t] = v;
    }
    if (nbits == 1 && issigned) {
#pragma unroll
      for (int t = 0; t < 16; t++) q[t] |= (q[t] ^ 0x01010101u) * 0xFFu;  // 0 -> -1
    } else if (issigned && nbits < 8) {
      const uint32_t ext = 0xFFu & (0xFFu << nbits);
#pragma unroll
      for (int t = 0; t < 16; t++) q[t] |= ((q[t] >> (nbits - 1)) & 0x01010101u) * ext;
    }
    const uint64_t valid = ncols - w * 64;
    if (valid < 64) {
#pragma unroll
      for (int t = 0; t < 16; t++) {
        const uint64_t k = 4 * (uint64_t)t;
        const uint32_t keep = k >= valid ? 0u : valid - k >= 4 ? 0xFFFFFFFFu
                                                               : (1u << (8 * (valid - k))) - 1u;
        q[t] &= keep;
      }
    }
  }
  uint4 *o = reinterpret_cast<uint4 *>(out + r * ld + w * 64);
#pragma unroll
  for (int v = 0; v < 4; v++) o[v] = make_uint4(q[4 * v], q[4 * v + 1], q[4 * v + 2], q[4 * v + 3]);
}

// planes -> FP4 codes: one thread per 64-value word of a copy row (32
// bytes; value k of the word in nibble k, low nibble first).  Code of an
// integer v (|v| <= 4) under scale 2^1: sign bit 8 | |v|.  Zero outside the
// real matrix, also for bipolar operands.
__device__ __forceinline__ void expand_f4_word(const ExpandJob &j, uint64_t i) {
  const uint64_t *planes = j.planes, nrows = j.nrows, ncols = j.ncols, wpr = j.wpr, W = j.W;
  const uint32_t nbits = j.nbits;
  const bool issigned = j.issigned;
  const uint64_t ld = j.ld;
  uint8_t *out = (uint8_t *)j.out;
  const uint64_t wpl = ld / 32;
  const uint64_t r = i / wpl, w = i - r * wpl;
  uint32_t q[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  if (r < nrows && w * 64 < ncols) {
    uint64_t p[3] = {planes[r * wpr + w], 0, 0};
    if (nbits > 1) p[1] = planes[W + r * wpr + w];
    if (nbits > 2) p[2] = planes[2 * W + r * wpr + w];
    // eight values (one dword of nibbles) at a time: bit b of each code
    // spread into bit b of its nibble by three shift-or-mask steps (bits
    // 0-3 / 4-7 to the two halves, then pairs, then single bits)
    const uint32_t sgn = nbits - 1, mag = (1u << nbits) - 1;
#pragma unroll
    for (int t = 0; t < 8; t++) {
      uint32_t v = 0;
#pragma unroll
      for (uint32_t b = 0; b < 3; b++) {
        if (b >= nbits) break;
        uint32_t x = (uint32_t)(p[b] >> (8 * t)) & 0xFFu;
        x = (x | (x << 12)) & 0x000F000Fu;
        x = (x | (x << 6)) & 0x03030303u;
        x = (x | (x << 3)) & 0x11111111u;
        v |= x << b;
      }
      if (nbits == 1 && issigned) {
        v = 0x11111111u | ((v ^ 0x11111111u) << 3);  // 1 -> +1 (0x1), 0 -> -1 (0x9)
      } else if (issigned) {
        // negative codes: sign bit 8 | (2^nbits - code), nibblewise
        const uint32_t s = (v >> sgn) & 0x11111111u, m = s * mag;
        v = ((v ^ m) + s) | (s << 3);
      }
      q[t] = v;
    }
    const uint64_t valid = ncols - w * 64;
    if (valid < 64) {
#pragma unroll
      for (int t = 0; t < 8; t++) {
        const uint64_t k = 8 * (uint64_t)t;
        q[t] &= k >= valid ? 0u : valid - k >= 8 ? 0xFFFFFFFFu : (1u << (4 * (valid - k))) - 1u;
      }
    }
  }
  uint4 *o = reinterpret_cast<uint4 *>(out + r * ld + w * 32);
  o[0] = make_uint4(q[0], q[1], q[2], q[3]);
  o[1] = make_uint4(q[4], q[5], q[6], q[7]);
}

// planes -> FP6 codes: one thread per 64-value word of a copy row, i.e.
// two lane groups (32 values, 24 bytes each) of one 96-byte k-step.  Value
// j of a group sits in bits 6j..6j+5 of the group's 24 bytes; bytes 0..15
// go to offset 16g of the k-step, bytes 16..23 to offset 64 + 8g.  Code of
// an integer v (|v| <= 16) under scale 2^3: sign bit 32 | |v|.  Zero
// outside the real matrix, also for bipolar operands.
__device__ __forceinline__ void expand_f6_word(const ExpandJob &j, uint64_t i) {
  const uint64_t *planes = j.planes, nrows = j.nrows, ncols = j.ncols, wpr = j.wpr, W = j.W;
  const uint32_t nbits = j.nbits;
  const bool issigned = j.issigned;
  const uint64_t ld = j.ld;
  uint8_t *out = (uint8_t *)j.out;
  const uint64_t wpl = ld / 48;
  const uint64_t r = i / wpl, w = i - r * wpl;
  uint32_t d[2][6] = {{0, 0, 0, 0, 0, 0}, {0, 0, 0, 0, 0, 0}};
  if (r < nrows && w * 64 < ncols) {
    uint64_t p[5] = {0, 0, 0, 0, 0};
#pragma unroll
    for (uint32_t b = 0; b < 5; b++)
      if (b < nbits) p[b] = planes[b * W + r * wpr + w];
    const uint64_t valid = ncols - w * 64;
    const uint32_t mag = (1u << nbits) - 1;
#pragma unroll
    for (int gi = 0; gi < 2; gi++) {
      uint64_t acc[3] = {0, 0, 0};  // the group's 24 bytes
#pragma unroll
      for (int c = 0; c < 8; c++) {
        // four values as four 6-bit fields: bit b of each code spread into
        // bit b of its field (bits 0-3 of x to bits 0, 6, 12, 18)
        uint32_t f = 0;
#pragma unroll
        for (uint32_t b = 0; b < 5; b++) {
          if (b >= nbits) break;
          uint32_t x = (uint32_t)(p[b] >> (32 * gi + 4 * c)) & 0xFu;
          x = (x | (x << 10)) & 0x3003u;
          x = (x | (x << 5)) & 0x41041u;
          f |= x << b;
        }
        if (nbits == 1 && issigned) {
          f = 0x041041u | ((f ^ 0x041041u) << 5);  // 1 -> +1 (0x01), 0 -> -1 (0x21)
        } else if (issigned) {
          // negative codes: sign bit 32 | (2^nbits - code), fieldwise
          const uint32_t sg = (f >> (nbits - 1)) & 0x041041u, m = sg * mag;
          f = ((f ^ m) + sg) | (sg << 5);
        }
        const uint64_t first = 32 * gi + 4 * c;
        if (first + 4 > valid) f &= first >= valid ? 0u : (1u << (6 * (valid - first))) - 1u;
        const int pos = 24 * c, wd = pos >> 6, off = pos & 63;
        acc[wd] |= (uint64_t)f << off;
        if (off > 40) acc[wd + 1] |= (uint64_t)f >> (64 - off);
      }
#pragma unroll
      for (int k = 0; k < 3; k++) {
        d[gi][2 * k] = (uint32_t)acc[k];
        d[gi][2 * k + 1] = (uint32_t)(acc[k] >> 32);
      }
    }
  }
  uint8_t *base = out + r * ld + (w >> 1) * 96;
#pragma unroll
  for (int gi = 0; gi < 2; gi++) {
    const int g = 2 * (int)(w & 1) + gi;
    *reinterpret_cast<uint4 *>(base + 16 * g) = make_uint4(d[gi][0], d[gi][1], d[gi][2], d[gi][3]);
    *reinterpret_cast<uint2 *>(base + 64 + 8 * g) = make_uint2(d[gi][4], d[gi][5]);
  }
}

// Threads (64-value words of a copy row) of a job: every word of its first
// rows_do rows (a rebuild leaves the padding rows' zeros in place).
__host__ __device__ inline uint64_t expand_units(const ExpandJob &j) {
  return j.rows_do * (j.ld / (j.fmt == kFmtF4 ? 32 : j.fmt == kFmtF6 ? 48 : 64));
}

template <int FMT>
__global__ __launch_bounds__(256) void expand_kernel(ExpandJob j) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= expand_units(j)) return;
  if constexpr (FMT == kFmtF4) expand_f4_word(j, i);
  else if constexpr (FMT == kFmtF6) expand_f6_word(j, i);
  else expand_i8_word(j, i);
}

// Many copies in one launch: job k owns blocks from jobs[k].block_begin
// (each job starts on a block of its own, so a block's job is uniform).
__global__ __launch_bounds__(256) void expand_jobs_kernel(const ExpandJob *jobs, uint32_t n) {
  uint32_t lo = 0, hi = n - 1;  // largest k with block_begin <= blockIdx.x
  while (lo < hi) {
    const uint32_t mid = (lo + hi + 1) / 2;
    if (jobs[mid].block_begin <= blockIdx.x) lo = mid;
    else hi = mid - 1;
  }
  const ExpandJob &j = jobs[lo];
  const uint64_t i = (uint64_t)(blockIdx.x - j.block_begin) * blockDim.x + threadIdx.x;
  if (i >= expand_units(j)) return;
  if (j.fmt == kFmtF4) expand_f4_word(j, i);
  else if (j.fmt == kFmtF6) expand_f6_word(j, i);
  else expand_i8_word(j, i);
}

uint64_t max_abs(const Operand &o) {
  if (o.bipolar) return 1;
  if (o.issigned) return 1ull << (o.nbits - 1);
  return (1ull << o.nbits) - 1;
}

template <int P, int Q, int WP, int KSB, int DEPTH, int FMT, int KG = 1>
int run_plan(const MfmaOperands &a, const MfmaPlan &pl, hipStream_t s) {
  if (pl.splits > 1) {
    int rc = launch_fill_i32(a.res, (uint64_t)a.M * a.N, 0, s);
    if (rc) return rc;
  }
  const uint64_t grid = pl.tiles();
  if (grid == 0) return GBS_OK;
  if (grid > 0x7fffffffull) return fail(GBS_ERANGE, "mfma grid too large");
  hipLaunchKernelGGL((mfma_kernel<P, Q, WP, KSB, DEPTH, FMT, KG>), dim3((uint32_t)grid),
                     dim3(256 * KG), 0, s, a, pl);
  GBS_HIP(hipGetLastError());
  return GBS_OK;
}

template <int FMT>
int run_plan_fmt(const MfmaOperands &a, const MfmaPlan &pl, hipStream_t s) {
  switch (pl.shape) {
    case 0: return run_plan<128, 128, 2, 128, 2, FMT>(a, pl, s);
    case 1: return run_plan<64, 128, 2, 128, 2, FMT>(a, pl, s);
    case 2: return run_plan<32, 256, 1, 64, 2, FMT>(a, pl, s);
    case 3: return run_plan<16, 256, 1, 128, 2, FMT>(a, pl, s);
    case 4: return run_plan<64, 64, 2, 128, 2, FMT>(a, pl, s);
    case 5: return run_plan<128, 64, 4, 128, 2, FMT>(a, pl, s);
    case 6: return run_plan<128, 128, 2, 64, 1, FMT>(a, pl, s);
    case 7: return run_plan<128, 256, 2, 64, 2, FMT>(a, pl, s);
    case 11: return run_plan<256, 256, 2, 64, 1, FMT>(a, pl, s);
  }
  if constexpr (FMT != kFmtF6) {  // FP6 plans never name these (fmt_shape)
    switch (pl.shape) {
      case 8: return run_plan<128, 128, 2, 64, 0, FMT>(a, pl, s);
      case 9: return run_plan<128, 256, 2, 64, 0, FMT>(a, pl, s);
      case 10: return run_plan<64, 128, 2, 64, 0, FMT>(a, pl, s);
      case 12: return run_plan<256, 256, 2, 128, 1, FMT>(a, pl, s);
      case 14: return run_plan<64, 64, 2, 256, 3, FMT>(a, pl, s);
    }
  }
  switch (pl.shape) {  // direct-load shapes every format has
    case 13: return run_plan<64, 64, 2, 128, 4, FMT>(a, pl, s);
    case 15: return run_plan<128, 128, 2, 128, 3, FMT>(a, pl, s);
    case 16: return run_plan<128, 64, 4, 128, 3, FMT>(a, pl, s);
    case 17: return run_plan<64, 128, 2, 128, 3, FMT>(a, pl, s);
    case 18: return run_plan<64, 64, 2, 128, 3, FMT>(a, pl, s);
    case 19: return run_plan<64, 64, 2, 128, 2, FMT, 2>(a, pl, s);
    case 20: return run_plan<64, 128, 2, 128, 2, FMT, 2>(a, pl, s);
    case 21: return run_plan<64, 64, 2, 128, 3, FMT, 2>(a, pl, s);
  }

  return fail(GBS_EINVAL, "mfma: tile shape %d not compiled for format %d", pl.shape, FMT);
}

int run_plan_id(const MfmaOperands &a, const MfmaPlan &pl, hipStream_t s) {
  if (pl.fmt == kFmtF4) return run_plan_fmt<kFmtF4>(a, pl, s);
  if (pl.fmt == kFmtF6) return run_plan_fmt<kFmtF6>(a, pl, s);
  return run_plan_fmt<kFmtI8>(a, pl, s);
}

int g_shape[kMfmaFormats][kClasses], g_splits[kMfmaFormats][kClasses];
bool g_table_set[kMfmaFormats] = {false, false, false};
float g_time_us[kMfmaFormats][kClasses];  // measured best time per class, 0 = unknown

int default_shape(uint64_t M, uint64_t N) {
  (void)M;
  (void)N;
  return 4;  // 64 x 64: the best single shape over the tuner's classes
}

}  // namespace

bool mfma_supported(const Operand &l, const Operand &r, uint64_t K) {
  // unsigned 8-bit codes (128..255) do not fit int8
  if (!l.issigned && l.nbits > 7) return false;
  if (!r.issigned && r.nbits > 7) return false;
  const unsigned __int128 bound = (unsigned __int128)K * max_abs(l) * max_abs(r);
  return bound < ((unsigned __int128)1 << 31);
}

// Values an FP4 code holds as an integer under scale 2^1: -4..4 (codes
// 0..4 with a sign bit); 3 is the largest unsigned 2-bit code, -4 the
// smallest signed 3-bit one.
static bool f4_holds(const Operand &o) {
  if (o.bipolar) return true;
  return o.issigned ? o.nbits <= 3 : o.nbits <= 2;
}

bool f4_supported(const Operand &l, const Operand &r, uint64_t K) {
  if (!f4_holds(l) || !f4_holds(r)) return false;
  const unsigned __int128 bound = (unsigned __int128)K * max_abs(l) * max_abs(r);
  return bound < ((unsigned __int128)1 << 24);
}

// Values an FP6 E2M3 code holds as an integer under scale 2^3: -16..16;
// 15 is the largest unsigned 4-bit code, -16 the smallest signed 5-bit one.
static bool f6_holds(const Operand &o) {
  if (o.bipolar) return true;
  return o.issigned ? o.nbits <= 5 : o.nbits <= 4;
}

bool f6_supported(const Operand &l, const Operand &r, uint64_t K) {
  if (!f6_holds(l) || !f6_holds(r)) return false;
  const unsigned __int128 bound = (unsigned __int128)K * max_abs(l) * max_abs(r);
  return bound < ((unsigned __int128)1 << 24);
}

// Format AUTO uses for one product: FP4 where it holds the product (always
// the fastest measured); else FP6 where it holds the product and the
// tuning table measured it at least 5% faster than int8 for the product's
// shape class (it ranges from 0.7x to 1.2x of int8 over the classes);
// else int8.
int best_format(const Operand &l, const Operand &r, uint64_t K) {
  if (f4_supported(l, r, K)) return kFmtF4;
  if (f6_supported(l, r, K)) {
    const int c = shape_class(l.nrows, r.nrows, K);
    const float t6 = g_time_us[kFmtF6][c], t8 = g_time_us[kFmtI8][c];
    if (t6 > 0 && t8 > 0 && t6 < 0.95f * t8) return kFmtF6;
  }
  if (mfma_supported(l, r, K)) return kFmtI8;
  return -1;
}

gbs_algo select_algo(const Operand &l, const Operand &r, uint64_t K) {
  const int fmt = best_format(l, r, K);
  if (fmt < 0) return GBS_ALGO_POPCOUNT;
  const double M = (double)l.nrows, N = (double)r.nrows, Kd = (double)K;
  const double pl = l.bipolar && r.bipolar ? 1 : (double)l.nbits;
  const double pr = l.bipolar && r.bipolar ? 1 : (double)r.nbits;
  // modelled times at the efficiencies these kernels reach (fractions of
  // the 39.3 T lane-op/s VALU, 5 POPS int8 MFMA and 8 TB/s HBM peaks)
  const double valu = M * N * ((K + 31) / 32) * 2 * pl * pr / (0.7 * 39.3e12);
  const double pbytes = (M * Kd * l.nbits + N * Kd * r.nbits) / 8 + 4 * M * N;
  const double t_popc = skinny_supported(l, r) && K < (1ull << 31) ? skinny_time(l, r, K)
                        : (valu > pbytes / (0.6 * 8e12) ? valu : pbytes / (0.6 * 8e12));
  // matrix cores: FP4 / FP6 instructions cover twice the values of int8
  // ones at the same issue rate
  const double vpi = fmt == kFmtI8 ? 1 : 2;                              // values per int8 step
  const double bpv = fmt == kFmtI8 ? 1 : fmt == kFmtF4 ? 0.5 : 0.75;    // copy bytes per value
  const double mops = 2 * M * N * Kd / vpi / (0.35 * 5e15);
  const double ibytes = (M + N) * Kd * bpv + 4 * M * N;
  // a tile row/column only partly filled still costs the whole tile
  int shape_id, splits;
  mfma_config_for(l.nrows, r.nrows, K, fmt, &shape_id, &splits);
  const MfmaTileShape sh = kMfmaShape[shape_id];
  const double waste = (N < sh.P ? sh.P / (N > 1 ? N : 1) : 1) *
                       (M < sh.Q ? sh.Q / (M > 1 ? M : 1) : 1);
  const double t_mfma_c = mops * waste;
  const double t_mfma = (t_mfma_c > ibytes / (0.6 * 8e12) ? t_mfma_c : ibytes / (0.6 * 8e12));
  if (t_mfma >= t_popc) return GBS_ALGO_POPCOUNT;
  if (fmt == kFmtF4 && bits_preferred(l, r, K)) return GBS_ALGO_MFMA_BITS;
  return fmt == kFmtF4 ? GBS_ALGO_MFMA_FP4 : fmt == kFmtF6 ? GBS_ALGO_MFMA_FP6 : GBS_ALGO_MFMA;
}


// AUTO for two matrices in their present state: select_algo's choice,
// except that an int8- or FP6-copy engine whose operands' copies would
// first have to be rebuilt (planes imported or written since the copies
// were made) gives way to the planes engine for large products that engine
// holds exactly.  The rebuild is a pass over the planes that writes 2-8x
// their bytes -- measured on the conv-as-GEMM product (256 x 2304 2-bit
// signed weights, 50176 x 2304 3-bit unsigned activations): 48 us to
// rebuild the activations' int8 copy against 52 us for the product on the
// copies, while the planes engine takes 53 us from the planes alone.
// FP4 copies stay: they cost a quarter of the int8 bytes to rebuild and run
// much faster than the planes engine (the 4096 x 8192 x 512 1-bit product:
// 15.5 against 39.5 us); so do small products, which the planes engine,
// having no K split, runs on few blocks (256 x 2048 x 256: 15 against 5 us).
gbs_algo select_for(const gbs_matrix *l, const gbs_matrix *r) {
  const Operand lo = view(l), ro = view(r);
  const uint64_t K = l->ncols;
  const gbs_algo eng = select_algo(lo, ro, K);
  const int fmt = eng == GBS_ALGO_MFMA_FP4   ? kFmtF4
                  : eng == GBS_ALGO_MFMA_FP6 ? kFmtF6
                  : eng == GBS_ALGO_MFMA     ? kFmtI8
                                             : -1;
  if (fmt < 0 || fmt == kFmtF4 || (copy_current(l, fmt) && copy_current(r, fmt))) return eng;
  const double ops = 2.0 * (double)lo.nrows * (double)ro.nrows * (double)K;
  if (ops >= 8e9 && planes_supported(lo, ro, K)) return GBS_ALGO_MFMA_PLANES;
  return eng;
}

// Geometry of m's copy in format fmt: rows padded to kValsRowAlign; a row
// holds the padded K in int8 bytes, FP4 nibbles or FP6 fields.
static void copy_geometry(const gbs_matrix *m, int fmt, uint64_t &rows_a, uint64_t &ld_a) {
  const uint64_t rows = ((m->nrows + kValsRowAlign - 1) / kValsRowAlign) * kValsRowAlign;
  uint64_t ld, ld0;
  if (fmt == kFmtF4) {
    ld = ((m->ncols + kPackedColAlign - 1) / kPackedColAlign) * kPackedColAlign / 2;
    ld0 = kPackedColAlign / 2;
  } else if (fmt == kFmtF6) {
    ld = ((m->ncols + kPackedColAlign - 1) / kPackedColAlign) * kPackedColAlign / 4 * 3;
    ld0 = kPackedColAlign / 4 * 3;
  } else {
    ld = ((m->ncols + kValsColAlign - 1) / kValsColAlign) * kValsColAlign;
    ld0 = kValsColAlign;
  }
  rows_a = rows ? rows : kValsRowAlign;
  ld_a = ld ? ld : ld0;
}

bool copy_current(const gbs_matrix *m, int fmt) {
  if (fmt == kFmtF4) return m->d_f4 && m->f4_version == m->planes_version;
  if (fmt == kFmtF6) return m->d_f6 && m->f6_version == m->planes_version;
  return m->d_vals && m->vals_version == m->planes_version;
}

void mark_copy_current(gbs_matrix *m, int fmt) {
  (fmt == kFmtF4 ? m->f4_version : fmt == kFmtF6 ? m->f6_version : m->vals_version) =
      m->planes_version;
}

int copy_job(gbs_matrix *m, int fmt, bool may_allocate, bool whole, ExpandJob *j) {
  static const char *const names[] = {"int8", "FP4", "FP6"};
  int8_t **buf = fmt == kFmtF4 ? &m->d_f4 : fmt == kFmtF6 ? &m->d_f6 : &m->d_vals;
  uint64_t &rows = fmt == kFmtF4 ? m->f4_rows : fmt == kFmtF6 ? m->f6_rows : m->vals_rows;
  uint64_t &ld = fmt == kFmtF4 ? m->f4_ld : fmt == kFmtF6 ? m->f6_ld : m->vals_ld;
  if (!*buf) {
    if (!may_allocate)
      return fail(GBS_ERANGE, "mfma: operand has no %s copy yet; call gbs_workspace_reserve",
                  names[fmt]);
    uint64_t rows_a, ld_a;
    copy_geometry(m, fmt, rows_a, ld_a);
    GBS_HIP(hipMalloc((void **)buf, rows_a * ld_a));
    rows = rows_a;
    ld = ld_a;
    whole = true;
  }
  *j = ExpandJob{m->d_planes, m->nrows, m->ncols, m->wpr(), m->wpp(), (uint32_t)m->nbits,
                 m->issigned, rows, ld, whole ? rows : m->nrows, *buf, fmt, 0};
  return GBS_OK;
}

int ensure_copy(gbs_matrix *m, int fmt, hipStream_t s, bool may_allocate) {
  if (copy_current(m, fmt)) return GBS_OK;
  const uint64_t v = fmt == kFmtF4 ? m->f4_version : fmt == kFmtF6 ? m->f6_version : m->vals_version;
  ExpandJob j;
  const int rc = copy_job(m, fmt, may_allocate, v == 0, &j);
  if (rc) return rc;
  const uint32_t blocks = (uint32_t)((expand_units(j) + 255) / 256);
  if (!blocks) {
    mark_copy_current(m, fmt);
    return GBS_OK;
  }
  if (fmt == kFmtF4) hipLaunchKernelGGL(expand_kernel<kFmtF4>, dim3(blocks), dim3(256), 0, s, j);
  else if (fmt == kFmtF6) hipLaunchKernelGGL(expand_kernel<kFmtF6>, dim3(blocks), dim3(256), 0, s, j);
  else hipLaunchKernelGGL(expand_kernel<kFmtI8>, dim3(blocks), dim3(256), 0, s, j);
  GBS_HIP(hipGetLastError());
  mark_copy_current(m, fmt);
  return GBS_OK;
}

int ensure_vals(gbs_matrix *m, hipStream_t s, bool may_allocate) {
  return ensure_copy(m, kFmtI8, s, may_allocate);
}
int ensure_f4(gbs_matrix *m, hipStream_t s, bool may_allocate) {
  return ensure_copy(m, kFmtF4, s, may_allocate);
}
int ensure_f6(gbs_matrix *m, hipStream_t s, bool may_allocate) {
  return ensure_copy(m, kFmtF6, s, may_allocate);
}

uint32_t expand_job_blocks(const ExpandJob &j) {
  return (uint32_t)((expand_units(j) + 255) / 256);
}

int launch_expand_jobs(const ExpandJob *d_jobs, uint32_t n, uint32_t blocks, hipStream_t s) {
  if (!n || !blocks) return GBS_OK;
  hipLaunchKernelGGL(expand_jobs_kernel, dim3(blocks), dim3(256), 0, s, d_jobs, n);
  GBS_HIP(hipGetLastError());
  return GBS_OK;
}

MfmaOperands mfma_operands(const gbs_matrix *l, const gbs_matrix *r, int fmt, int32_t *d_res) {
  MfmaOperands a{};
  const bool f4 = fmt == kFmtF4, f6 = fmt == kFmtF6;
  a.R = f4 ? r->d_f4 : f6 ? r->d_f6 : r->d_vals;
  a.L = f4 ? l->d_f4 : f6 ? l->d_f6 : l->d_vals;
  a.r_ld = f4 ? r->f4_ld : f6 ? r->f6_ld : r->vals_ld;
  a.l_ld = f4 ? l->f4_ld : f6 ? l->f6_ld : l->vals_ld;
  a.res = d_res;
  a.M = (uint32_t)l->nrows;
  a.N = (uint32_t)r->nrows;
  // nominal bytes: 64 per k-step of 64 (int8) or 128 (FP4, FP6) values
  const uint64_t bytes = fmt == kFmtI8 ? l->ncols : (l->ncols + 127) / 128 * 64;
  a.kbytes = (uint32_t)((bytes + kMfmaKS - 1) / kMfmaKS * kMfmaKS);
  return a;
}

int launch_mfma_gemm(const gbs_matrix *l, const gbs_matrix *r, int32_t *d_res,
                     hipStream_t s, int fmt, int forced_shape, bool allow_split,
                     int forced_splits) {
  const uint64_t M = l->nrows, N = r->nrows;
  if (M == 0 || N == 0) return GBS_OK;
  if (M > 0x7fffffffull || N > 0x7fffffffull)
    return fail(GBS_ERANGE, "gemm dimensions exceed 2^31");
  const MfmaOperands a = mfma_operands(l, r, fmt, d_res);
  if (!a.R || !a.L || a.r_ld != a.l_ld || !copy_current(l, fmt) || !copy_current(r, fmt))
    return fail(GBS_EINVAL, "mfma: operand value copies are not current");
  if (a.kbytes == 0) return launch_fill_i32(d_res, M * N, 0, s);
  const MfmaPlan pl = mfma_plan(a.M, a.N, a.kbytes, fmt, forced_shape,
                                forced_splits > 0 ? forced_splits : (allow_split ? 0 : 1));
  return run_plan_id(a, pl, s);
}

void mfma_config_for(uint64_t M, uint64_t N, uint64_t K, int fmt, int *shape, int *splits) {
  *shape = default_shape(M, N);
  *splits = 0;
  if (fmt < 0 || fmt >= kMfmaFormats) fmt = kFmtI8;
  if (g_table_set[fmt]) {
    const int c = shape_class(M, N, K);
    if (g_shape[fmt][c] >= 0 && g_shape[fmt][c] < kMfmaShapes) {
      *shape = g_shape[fmt][c];
      *splits = g_splits[fmt][c];
    }
  }
}

MfmaPlan mfma_plan(uint32_t M, uint32_t N, uint32_t kbytes, int fmt, int forced_shape,
                   int forced_splits) {
  int shape, splits;
  mfma_config_for(M, N, (uint64_t)kbytes * (fmt == kFmtI8 ? 1 : 2), fmt, &shape, &splits);
  if (forced_shape >= 0) {
    shape = forced_shape;
    splits = 0;
  }
  shape = fmt_shape(shape, fmt);
  if (forced_splits > 0) splits = forced_splits;
  const MfmaTileShape sh = kMfmaShape[shape];
  MfmaPlan pl{};
  pl.shape = shape;
  pl.fmt = fmt;
  pl.ptiles = (N + sh.P - 1) / sh.P;
  pl.qtiles = (M + sh.Q - 1) / sh.Q;
  // XCD groups share the larger operand: lhs larger -> rhs tiles fastest
  pl.order = (uint64_t)M > (uint64_t)N ? 1 : 0;
  const uint32_t ksteps = kbytes / kMfmaKS;  // K units
  const uint64_t tiles = (uint64_t)pl.ptiles * pl.qtiles;
  if (splits <= 0) {
    // fill two blocks per CU, keeping at least 1 KiB (nominal) of K per split
    uint64_t want = tiles >= 512 ? 1 : (512 + tiles - 1) / tiles;
    uint64_t maxs = ksteps / 4 ? ksteps / 4 : 1;
    splits = (int)(want < maxs ? want : maxs);
  }
  if (splits < 1) splits = 1;
  if ((uint32_t)splits > (ksteps ? ksteps : 1)) splits = ksteps ? ksteps : 1;
  const uint32_t per = ksteps ? (ksteps + splits - 1) / splits : 1;
  pl.splits = ksteps ? (ksteps + per - 1) / per : 1;
  pl.kchunk = per * kMfmaKS;
  return pl;
}

int set_mfma_times(const float *times_us, int n) {
  for (int f = 0; f < kMfmaFormats; f++)
    for (int i = 0; i < kClasses; i++)
      g_time_us[f][i] = times_us && f * kClasses + i < n ? times_us[f * kClasses + i] : 0.f;
  return GBS_OK;
}

int set_mfma_table(const int *shapes, const int *splits, int n) {
  set_mfma_times(nullptr, 0);
  if (!shapes) {
    for (int f = 0; f < kMfmaFormats; f++) g_table_set[f] = false;
    return GBS_OK;
  }
  if (n <= 0 || n % kClasses != 0 || n > kMfmaFormats * kClasses)
    return fail(GBS_EINVAL,
                "tile table needs %d entries per format (int8, then FP4, then FP6), got %d",
                kClasses, n);
  for (int f = 0; f < kMfmaFormats; f++) {
    if (f * kClasses >= n) {
      g_table_set[f] = false;
      continue;
    }
    bool any = false;
    for (int i = 0; i < kClasses; i++) {
      g_shape[f][i] = shapes[f * kClasses + i];
      g_splits[f][i] = splits ? splits[f * kClasses + i] : 0;
      any |= g_shape[f][i] >= 0;
    }
    g_table_set[f] = any;
  }
  return GBS_OK;
}

}  // namespace gbs
