// gbs_skinny.h -- bit-serial products with few rhs rows (N <= 8) straight
// from the bit-planes: the popcount engine's kernel for GEMV-like products.
//
// A product with a handful of rhs rows is a stream over the lhs planes:
// every lhs word is used N x PR times and never again, so the time is
// HBM bytes as long as the popcount work per word stays under what the
// VALU issues.  The planes are the smallest form of the operands (PL bits
// per lhs value, against 4 to 8 for the matrix cores' value copies), which
// makes this the engine for skinny few-bit products.
//
// Layout of the work: G lanes share one lhs row (G a power of two picked so
// every lane walks a few 128-bit slices of K, and widened while items
// would be too long to keep enough of them in flight) and each lane holds
// two rows, 2 * 256 / G rows per 256-thread pass.  The rhs words of a K
// piece (32 KiB of LDS over the rhs rows and planes: 32768 columns for one
// plane, 16384 for two, 8192 for three or four) are staged once per block
// in LDS, where all rows of the pass read them; the lhs words come
// straight from HBM as pairs of 8-byte words (one 16-byte load each),
// G * 16 consecutive bytes per row, two slices in flight per lane.  Every
// (lhs plane p, rhs plane q) pair adds W[p][q] * popc(a_p & b_q) to one
// int32 accumulator per output (4 AND, 4 chained v_bcnt, one
// v_mad_i32_i24 per 128-bit slice), with
//
//   W[p][q] = wl[p] * wr[q],  wl / wr = +-2^plane (the top plane of a
//   signed operand weighs -2^(bits-1)),
//
// so the planes fold as they are counted (the reference's per-pair passes,
// arch-generic.hpp:169-193, become one pass).  A bipolar operand against a
// regular one is 2b - 1 per value: weight 2 on its plane plus a virtual
// all-ones plane (ones on the real columns) of weight -1, which yields the
// -rowsum term of arch-generic.hpp:39-76 without a separate pass.
// Bipolar x bipolar uses K - 2 popc(a ^ b) (arch-generic.hpp:239-262).
// The G partial sums of an output meet through lane shuffles; every output
// is written once (an item covers all of K).  Integer sums are exact in
// any order: results are bit-identical to the reference's (int32,
// wrapping).
#pragma once
#include "gbs_internal.h"

namespace gbs {

constexpr uint32_t kSkinnyMaxN = 8;          // rhs rows
constexpr uint32_t kSkinnyMaxPlanes = 4;      // per side, virtual ones plane included
constexpr size_t kSkinnyLds = 32768;          // staged rhs piece, bytes
// 16-byte slices of K per (plane, row) in a piece with RQ rhs planes held
// in memory: the LDS split over RQ planes x kSkinnyMaxN rows
template <int RQ>
constexpr uint32_t skinny_piece_slices() {
  return RQ <= 1 ? 256 : RQ == 2 ? 128 : RQ == 3 ? 85 : 64;
}
static_assert(skinny_piece_slices<3>() * 3 * kSkinnyMaxN * 16 <= kSkinnyLds, "rhs piece");

struct SkinnyProb {
  const uint64_t *A, *B;      // plane 0 of lhs / rhs
  uint64_t a_wpr, a_wpp, b_wpr, b_wpp;
  int32_t *res;               // N x M, row n at res + n * M
  uint32_t M, N, K;           // real sizes (K in columns)
  uint32_t kwords;            // 64-bit words of K, rounded up to even
  uint32_t kw_real;           // 64-bit words of K: a row holds at least these
  uint32_t pl, pr;            // planes per side, virtual ones plane included
  uint32_t ones;              // 0, 1 = lhs plane pl-1 is the ones plane, 2 = rhs plane pr-1
  uint32_t xor_mode;          // bipolar x bipolar: K - 2 popc(a ^ b)
  uint32_t g_log2;            // lanes per row = 1 << g_log2
  uint32_t rows;              // rows per item (a multiple of 512 >> g_log2)
  uint32_t items;             // ceil(M / rows)
  int32_t W[kSkinnyMaxPlanes * kSkinnyMaxPlanes];  // W[p * 4 + q]
};

// Host side: can this product take the skinny engine, and its plan.
bool skinny_supported(const Operand &l, const Operand &r);
SkinnyProb skinny_plan(const Operand &l, const Operand &r, uint64_t K, int32_t *res,
                       double item_bytes);
// lhs plane bytes the skinny kernel streams for a product
double skinny_bytes(const Operand &l, const Operand &r, uint64_t K);
// Modelled time (seconds) of a supported product on the skinny kernel.
double skinny_time(const Operand &l, const Operand &r, uint64_t K);

#if defined(__HIPCC__)
// acc + popcount(x) in one v_bcnt_u32_b32 (its second operand is an
// addend).  Left to itself the compiler counts words from zero and adds
// the counts with v_add3_u32, one VALU instruction more per two words.
__device__ __forceinline__ uint32_t bcnt_add(uint32_t x, uint32_t acc) {
  uint32_t r;
  asm("v_bcnt_u32_b32 %0, %1, %2" : "=v"(r) : "v"(x), "v"(acc));
  return r;
}

// acc + t * w for |t|, |w| < 2^23 in one v_mad_i32_i24.
__device__ __forceinline__ int32_t mad_i24(int32_t t, int32_t w, int32_t acc) {
  int32_t r;
  asm("v_mad_i32_i24 %0, %1, %2, %3" : "=v"(r) : "v"(t), "v"(w), "v"(acc));
  return r;
}

__device__ __forceinline__ uint64_t skinny_valid_mask(uint32_t w, uint32_t K) {
  const uint32_t lo = w * 64;
  if (lo + 64 <= K) return ~0ull;
  if (lo >= K) return 0ull;
  return (1ull << (K - lo)) - 1;
}

// popcount of a & b (x: a ^ b) over 128 bits: four chained counts
__device__ __forceinline__ uint32_t popc128(uint4 a, uint4 b, bool x) {
  if (x)
    return bcnt_add(a.w ^ b.w, bcnt_add(a.z ^ b.z, bcnt_add(a.y ^ b.y, bcnt_add(a.x ^ b.x, 0u))));
  return bcnt_add(a.w & b.w, bcnt_add(a.z & b.z, bcnt_add(a.y & b.y, bcnt_add(a.x & b.x, 0u))));
}

// Words w and w + 1 of a row (8-byte aligned; the compiler makes one
// 16-byte load of the pair).  With an odd number of real words the last
// slice's second word lies past the row and reads as zero instead.
template <bool ODD>
__device__ __forceinline__ uint4 load_pair(const uint64_t *q, uint32_t w, uint32_t kw_real) {
  const uint64_t x = q[0];
  const uint64_t y = (!ODD || w + 1 < kw_real) ? q[1] : 0ull;
  return make_uint4((uint32_t)x, (uint32_t)(x >> 32), (uint32_t)y, (uint32_t)(y >> 32));
}

__device__ __forceinline__ uint4 mask128(uint32_t w, uint32_t K) {
  const uint64_t m0 = skinny_valid_mask(w, K), m1 = skinny_valid_mask(w + 1, K);
  return make_uint4((uint32_t)m0, (uint32_t)(m0 >> 32), (uint32_t)m1, (uint32_t)(m1 >> 32));
}

// Stage the rhs words [kb, kb + 2 kq) of every real plane and row in LDS as
// [q][n][slice], skinny_piece_slices<RQ>() per (q, n), so that every read
// of the product loop is one ds_read_b128 with an immediate offset.
template <int RQ>
__device__ __forceinline__ void skinny_stage(const SkinnyProb &p, uint32_t kb, uint32_t kq,
                                             uint4 *lds) {
  __syncthreads();
  const uint32_t N = p.N;
  for (uint32_t i = threadIdx.x; i < RQ * N * kq; i += 256) {
    const uint32_t s = i % kq, qn = i / kq;
    const uint32_t n = qn % N, q = qn / N;
    const uint32_t w = kb + 2 * s;
    const uint64_t *src = p.B + q * p.b_wpp + (uint64_t)n * p.b_wpr + w;
    lds[(q * kSkinnyMaxN + n) * skinny_piece_slices<RQ>() + s] =
        (p.kw_real & 1) ? load_pair<true>(src, w, p.kw_real) : load_pair<false>(src, w, p.kw_real);
  }
  __syncthreads();
}

// U slices of K (s, s + G, ...) for the lane's two rows: all their lhs loads
// are issued before any is used, then every (plane pair, rhs row) adds its
// weighted count.  Slices at or past kq are skipped (their loads too).
template <int PL, int PR, int ONES, bool XOR, int U, bool ODD>
__device__ __forceinline__ void skinny_slices(const SkinnyProb &p, const int32_t (&wt)[PL][PR],
                                              uint32_t kb, uint32_t s, uint32_t G, uint32_t kq,
                                              uint32_t r0, uint32_t r1, bool v0, bool v1,
                                              const uint4 *lds, int32_t (&acc)[2][kSkinnyMaxN]) {
  constexpr bool rhs_ones = ONES == 2, lhs_ones = ONES == 1;
  constexpr int RQ = rhs_ones ? PR - 1 : PR;
  uint4 a[U][2][PL];
#pragma unroll
  for (int u = 0; u < U; u++) {
    const uint32_t su = s + u * G;
    const uint32_t w = kb + 2 * su;
    const bool in = su < kq;
#pragma unroll
    for (int pp = 0; pp < PL; pp++) {
      if (lhs_ones && pp == PL - 1) {
        a[u][0][pp] = a[u][1][pp] = mask128(w, p.K);
      } else {
        const uint64_t *src = p.A + pp * p.a_wpp + w;
        a[u][0][pp] = in && v0 ? load_pair<ODD>(src + (uint64_t)r0 * p.a_wpr, w, p.kw_real)
                               : make_uint4(0, 0, 0, 0);
        a[u][1][pp] = in && v1 ? load_pair<ODD>(src + (uint64_t)r1 * p.a_wpr, w, p.kw_real)
                               : make_uint4(0, 0, 0, 0);
      }
    }
  }
#pragma unroll
  for (int u = 0; u < U; u++) {
    const uint32_t su = s + u * G;
    if (U > 1 && su >= kq) break;
#pragma unroll
    for (int q = 0; q < PR; q++) {
      const bool qones = rhs_ones && q == PR - 1;
      uint4 m;
      if (qones) m = mask128(kb + 2 * su, p.K);
#pragma unroll
      for (int n = 0; n < (int)kSkinnyMaxN; n++) {
        if ((uint32_t)n < p.N) {
          const uint4 b =
              qones ? m : lds[(q * kSkinnyMaxN + n) * skinny_piece_slices<RQ>() + su];
#pragma unroll
          for (int pp = 0; pp < PL; pp++) {
            acc[0][n] = mad_i24((int32_t)popc128(a[u][0][pp], b, XOR), wt[pp][q], acc[0][n]);
            acc[1][n] = mad_i24((int32_t)popc128(a[u][1][pp], b, XOR), wt[pp][q], acc[1][n]);
          }
        }
      }
    }
  }
}

// One item of a skinny product: rows [item * rows, +rows) x all N x all K.
// 256 threads; lds holds kSkinnyLds bytes, 16-byte aligned.  A product of
// one K piece stages its rhs once per item; deeper ones once per piece and
// pass.
template <int PL, int PR, int ONES, bool XOR>
__device__ __forceinline__ void skinny_item_t(const SkinnyProb &p, uint32_t item, uint4 *lds) {
  constexpr int U = PL <= 2 ? 2 : 1;  // slices in flight per lane (registers allow)
  constexpr int RQ = ONES == 2 ? PR - 1 : PR;  // rhs planes held in memory
  int32_t wt[PL][PR];
#pragma unroll
  for (int i = 0; i < PL; i++)
#pragma unroll
    for (int j = 0; j < PR; j++) wt[i][j] = p.W[i * kSkinnyMaxPlanes + j];
  const uint32_t G = 1u << p.g_log2;
  const uint32_t kl = threadIdx.x & (G - 1);
  const uint32_t lg = threadIdx.x >> p.g_log2;       // lane group in the block
  const uint32_t groups = 256u >> p.g_log2;          // rows per half pass
  const uint32_t N = p.N;
  constexpr uint32_t PW = 2 * skinny_piece_slices<RQ>();  // words per piece
  const bool one_piece = p.kwords <= PW;
  if (one_piece) skinny_stage<RQ>(p, 0, p.kwords / 2, lds);
  const uint32_t row_end = min(p.M, (item + 1) * p.rows);
  for (uint32_t base = item * p.rows; base < row_end; base += 2 * groups) {
    const uint32_t r0 = base + lg, r1 = r0 + groups;
    const bool v0 = r0 < row_end, v1 = r1 < row_end;
    int32_t acc[2][kSkinnyMaxN];
#pragma unroll
    for (int n = 0; n < (int)kSkinnyMaxN; n++) acc[0][n] = acc[1][n] = 0;
    for (uint32_t kb = 0; kb < p.kwords; kb += PW) {
      const uint32_t kq = min(PW, p.kwords - kb) / 2;  // 16-byte slices
      if (!one_piece) skinny_stage<RQ>(p, kb, kq, lds);
      if (p.kw_real & 1) {
        for (uint32_t s = kl; s < kq; s += U * G)
          skinny_slices<PL, PR, ONES, XOR, U, true>(p, wt, kb, s, G, kq, r0, r1, v0, v1, lds, acc);
      } else {
        for (uint32_t s = kl; s < kq; s += U * G)
          skinny_slices<PL, PR, ONES, XOR, U, false>(p, wt, kb, s, G, kq, r0, r1, v0, v1, lds, acc);
      }
    }
    // the G partial sums of each output
    for (uint32_t off = G >> 1; off > 0; off >>= 1) {
#pragma unroll
      for (int n = 0; n < (int)kSkinnyMaxN; n++) {
        acc[0][n] += __shfl_xor(acc[0][n], (int)off, 64);
        acc[1][n] += __shfl_xor(acc[1][n], (int)off, 64);
      }
    }
    if (kl == 0) {
      const int32_t c = XOR ? (int32_t)p.K : 0;
#pragma unroll
      for (int n = 0; n < (int)kSkinnyMaxN; n++) {
        if ((uint32_t)n < N) {
          if (v0) p.res[(uint64_t)n * p.M + r0] = acc[0][n] + c;
          if (v1) p.res[(uint64_t)n * p.M + r1] = acc[1][n] + c;
        }
      }
    }
  }
}

template <int PL>
__device__ __forceinline__ void skinny_item_l(const SkinnyProb &p, uint32_t item, uint4 *lds) {
  switch (p.pr) {
    case 1: return skinny_item_t<PL, 1, 0, false>(p, item, lds);
    case 2: return skinny_item_t<PL, 2, 0, false>(p, item, lds);
    case 3: return skinny_item_t<PL, 3, 0, false>(p, item, lds);
    default: return skinny_item_t<PL, 4, 0, false>(p, item, lds);
  }
}

// bipolar x regular: the bipolar side has its plane and the ones plane
__device__ __forceinline__ void skinny_item_ones(const SkinnyProb &p, uint32_t item,
                                                 uint4 *lds) {
  if (p.ones == 1) {
    switch (p.pr) {
      case 1: return skinny_item_t<2, 1, 1, false>(p, item, lds);
      case 2: return skinny_item_t<2, 2, 1, false>(p, item, lds);
      case 3: return skinny_item_t<2, 3, 1, false>(p, item, lds);
      default: return skinny_item_t<2, 4, 1, false>(p, item, lds);
    }
  }
  switch (p.pl) {
    case 1: return skinny_item_t<1, 2, 2, false>(p, item, lds);
    case 2: return skinny_item_t<2, 2, 2, false>(p, item, lds);
    case 3: return skinny_item_t<3, 2, 2, false>(p, item, lds);
    default: return skinny_item_t<4, 2, 2, false>(p, item, lds);
  }
}

__device__ __forceinline__ void skinny_item(const SkinnyProb &p, uint32_t item, uint4 *lds) {
  if (p.xor_mode) return skinny_item_t<1, 1, 0, true>(p, item, lds);
  if (p.ones) return skinny_item_ones(p, item, lds);
  switch (p.pl) {
    case 1: return skinny_item_l<1>(p, item, lds);
    case 2: return skinny_item_l<2>(p, item, lds);
    case 3: return skinny_item_l<3>(p, item, lds);
    default: return skinny_item_l<4>(p, item, lds);
  }
}
#endif

}  // namespace gbs
