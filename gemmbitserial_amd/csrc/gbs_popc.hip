// gbs_popc.hip -- fused bit-serial GEMM on gfx950 VALU (AND/XOR + popcount).
//
// The reference runs one binary GEMM per (lhs plane, rhs plane) pair and
// sweeps the int32 result once per pair (arch-generic.hpp:169-193).  Here a
// lhs row m with PL planes is PL binary rows and a rhs row n with PR planes is
// PR binary rows; one kernel pass computes all PL x PR popcount sums of a
// tile in registers and folds them with their +-2^(i+j) weights once, in the
// epilogue:
//
//   res[n*M + m] = sum_{i<PL, j<PR} s_ij 2^(i+j) popc(A_i[m] & B_j[n])
//
// with s_ij = -1 when exactly one of (i, j) is the top plane of a signed
// (non-bipolar) operand.  Bipolar x bipolar uses popc(a ^ b) and
// K - 2*popc (arch-generic.hpp:239-262); bipolar x regular doubles the sum
// and subtracts the regular operand's row sum (arch-generic.hpp:39-76).
// Integer sums are exact in any order, so results are bit-identical to the
// reference (int32, wrapping modulo 2^32).
//
// Tiling: 256 threads as 16 x 16; thread (tx, ty) owns outputs
// m = m0 + tx + 16 i (i < TM), n = n0 + ty + 16 j (j < TN), i.e. TM*PL lhs
// binary rows x TN*PR rhs binary rows of 32-bit popcount counters.  Plane
// tiles of KW 64-bit words per row are staged in LDS k-major
// ([k][plane*BM + row], k-stride padded by 8 words so the transposing
// writes spread over the banks) and double-buffered through registers; per
// staged word every counter costs one v_and_b32/v_xor_b32 and one
// accumulating v_bcnt_u32_b32 per 32-bit half.  Blocks are numbered so that
// the m-tiles that share one rhs tile run on the same XCD (blocks are dealt
// round-robin over the 8 XCDs), keeping rhs re-reads in that XCD's L2.
// Small grids split K over blocks and add with integer atomics.
//
// N <= 8 with at most 4 planes a side goes to the skinny kernel
// (gbs_skinny.h: lhs planes streamed from HBM, rhs pieces in LDS, plane
// pairs folded as they are counted); other N <= 4 products to gemv_kernel:
// lanes split K, the few rhs rows are read once per word by every lane,
// lhs rows stream from HBM (arch-generic.hpp:264-329).
#include <algorithm>

#include "gbs_internal.h"
#include "gbs_skinny.h"

namespace gbs {
namespace {

constexpr uint32_t kNoPlane = 0xffffffffu;

struct PopcArgs {
  const uint64_t *A, *B;
  uint64_t a_wpr, a_wpp, b_wpr, b_wpp;
  int32_t *res;
  const int32_t *rowsum;
  uint32_t M, N, kwords, kchunk, K;
  uint32_t lneg, rneg;      // plane (within this launch) weighted negative
  uint32_t lshift, rshift;  // weight exponent of plane 0 of each side
  uint32_t mtiles, ntiles;
  int bpreg;                // 0, 1 = lhs bipolar (rowsum[n]), 2 = rhs bipolar (rowsum[m])
  int mode;                 // 0 store, 1 add, 2 atomic add
  int add_base;             // this launch adds K (XOR) / -rowsum (bipolar x regular)
};

// Fold one output's PL x PR counters into its int32 value (mod 2^32).
template <int PL, int PR, bool XOR>
__device__ inline uint32_t fold(const uint32_t (&c)[PL][PR], const PopcArgs &a,
                                bool base, uint32_t m, uint32_t n) {
  if (XOR) return (base ? a.K : 0u) - 2u * c[0][0];
  uint32_t acc = 0;
#pragma unroll
  for (int p = 0; p < PL; p++)
#pragma unroll
    for (int q = 0; q < PR; q++) {
      const uint32_t t = c[p][q] << (a.lshift + a.rshift + p + q);
      acc += (((uint32_t)p == a.lneg) != ((uint32_t)q == a.rneg)) ? 0u - t : t;
    }
  if (a.bpreg) {
    acc *= 2u;
    if (base) acc -= (uint32_t)a.rowsum[a.bpreg == 1 ? n : m];
  }
  return acc;
}

__device__ inline void emit(const PopcArgs &a, uint64_t idx, uint32_t v) {
  if (a.mode == 0) a.res[idx] = (int32_t)v;
  else if (a.mode == 1) a.res[idx] = (int32_t)((uint32_t)a.res[idx] + v);
  else atomicAdd(reinterpret_cast<unsigned int *>(a.res) + idx, v);
}

template <int PL, int PR, bool XOR, int TM, int TN, int KW>
__global__ __launch_bounds__(256) void popc_kernel(PopcArgs a) {
  constexpr int TX = 16, TY = 16;
  constexpr int BM = TX * TM, BN = TY * TN;
  constexpr int AR = PL * BM, BR = PR * BN;
  constexpr int AS = AR + 8, BS = BR + 8;  // padded k-strides (words)
  constexpr int ACH = AR * KW, BCH = BR * KW;
  constexpr int ALD = (ACH + 255) / 256, BLD = (BCH + 255) / 256;
  __shared__ uint64_t sA[2][KW * AS];
  __shared__ uint64_t sB[2][KW * BS];

  // XCD-aware tile order: logical tile L runs on XCD (L / (T/8)).
  const uint32_t T = gridDim.x, b = blockIdx.x;
  uint32_t L = (T % 8 == 0) ? (b % 8) * (T / 8) + b / 8 : b;
  const uint32_t mt = L % a.mtiles;
  L /= a.mtiles;
  const uint32_t nt = L % a.ntiles;
  const uint32_t sp = L / a.ntiles;
  const uint32_t m0 = mt * BM, n0 = nt * BN;
  const uint32_t kbeg = sp * a.kchunk;
  const uint32_t kend = min(a.kwords, kbeg + a.kchunk);

  const int tid = threadIdx.x;
  const int tx = tid % TX, ty = tid / TX;

  uint64_t ra[ALD], rb[BLD];
  auto gload = [&](uint32_t k0) {
#pragma unroll
    for (int i = 0; i < ALD; i++) {
      const int c = tid + 256 * i;
      uint64_t v = 0;
      if (c < ACH) {
        const int kk = c % KW, row = (c / KW) % BM, p = c / (KW * BM);
        const uint32_t m = m0 + row, k = k0 + kk;
        if (m < a.M && k < kend) v = a.A[p * a.a_wpp + (uint64_t)m * a.a_wpr + k];
      }
      ra[i] = v;
    }
#pragma unroll
    for (int i = 0; i < BLD; i++) {
      const int c = tid + 256 * i;
      uint64_t v = 0;
      if (c < BCH) {
        const int kk = c % KW, row = (c / KW) % BN, q = c / (KW * BN);
        const uint32_t n = n0 + row, k = k0 + kk;
        if (n < a.N && k < kend) v = a.B[q * a.b_wpp + (uint64_t)n * a.b_wpr + k];
      }
      rb[i] = v;
    }
  };
  auto swrite = [&](int buf) {
#pragma unroll
    for (int i = 0; i < ALD; i++) {
      const int c = tid + 256 * i;
      if (c < ACH) {
        const int kk = c % KW, r = c / KW;  // r = p*BM + row
        sA[buf][kk * AS + r] = ra[i];
      }
    }
#pragma unroll
    for (int i = 0; i < BLD; i++) {
      const int c = tid + 256 * i;
      if (c < BCH) {
        const int kk = c % KW, r = c / KW;
        sB[buf][kk * BS + r] = rb[i];
      }
    }
  };

  uint32_t cnt[TM][TN][PL][PR];
#pragma unroll
  for (int i = 0; i < TM; i++)
#pragma unroll
    for (int j = 0; j < TN; j++)
#pragma unroll
      for (int p = 0; p < PL; p++)
#pragma unroll
        for (int q = 0; q < PR; q++) cnt[i][j][p][q] = 0;

  if (kbeg < kend) {
    gload(kbeg);
    swrite(0);
  }
  __syncthreads();
  int buf = 0;
  for (uint32_t k0 = kbeg; k0 < kend; k0 += KW) {
    const bool more = k0 + KW < kend;
    if (more) gload(k0 + KW);
    const uint64_t *pa = &sA[buf][0];
    const uint64_t *pb = &sB[buf][0];
#pragma unroll 2
    for (int kk = 0; kk < KW; kk++) {
      uint64_t av[TM][PL], bv[TN][PR];
#pragma unroll
      for (int i = 0; i < TM; i++)
#pragma unroll
        for (int p = 0; p < PL; p++) av[i][p] = pa[kk * AS + p * BM + tx + TX * i];
#pragma unroll
      for (int j = 0; j < TN; j++)
#pragma unroll
        for (int q = 0; q < PR; q++) bv[j][q] = pb[kk * BS + q * BN + ty + TY * j];
#pragma unroll
      for (int i = 0; i < TM; i++)
#pragma unroll
        for (int j = 0; j < TN; j++)
#pragma unroll
          for (int p = 0; p < PL; p++)
#pragma unroll
            for (int q = 0; q < PR; q++) {
              const uint64_t x = XOR ? (av[i][p] ^ bv[j][q]) : (av[i][p] & bv[j][q]);
              cnt[i][j][p][q] =
                  bcnt_add((uint32_t)(x >> 32), bcnt_add((uint32_t)x, cnt[i][j][p][q]));
            }
    }
    if (more) swrite(buf ^ 1);
    __syncthreads();
    buf ^= 1;
  }

  const bool base = a.add_base && sp == 0;
#pragma unroll
  for (int i = 0; i < TM; i++) {
    const uint32_t m = m0 + tx + TX * i;
#pragma unroll
    for (int j = 0; j < TN; j++) {
      const uint32_t n = n0 + ty + TY * j;
      if (m < a.M && n < a.N) emit(a, (uint64_t)n * a.M + m, fold<PL, PR, XOR>(cnt[i][j], a, base, m, n));
    }
  }
}

// N <= NV: G lanes per lhs row split K; 64/G rows per wave in flight.
template <int PL, int PR, bool XOR, int NV, int G>
__global__ __launch_bounds__(256) void gemv_kernel(PopcArgs a) {
  const int lane = threadIdx.x & 63;
  const int kl = lane % G, ro = lane / G;
  constexpr int RPW = 64 / G;
  const uint32_t wave = blockIdx.x * 4 + threadIdx.x / 64;
  const uint32_t nwaves = gridDim.x * 4;
  for (uint32_t mb = wave * RPW; mb < a.M; mb += nwaves * RPW) {
    const uint32_t m = mb + ro;
    uint32_t c[NV][PL][PR];
#pragma unroll
    for (int n = 0; n < NV; n++)
#pragma unroll
      for (int p = 0; p < PL; p++)
#pragma unroll
        for (int q = 0; q < PR; q++) c[n][p][q] = 0;
    if (m < a.M) {
      for (uint32_t k = kl; k < a.kwords; k += G) {
        uint64_t av[PL];
#pragma unroll
        for (int p = 0; p < PL; p++) av[p] = a.A[p * a.a_wpp + (uint64_t)m * a.a_wpr + k];
#pragma unroll
        for (int n = 0; n < NV; n++) {
          if ((uint32_t)n < a.N) {
#pragma unroll
            for (int q = 0; q < PR; q++) {
              const uint64_t bv = a.B[q * a.b_wpp + (uint64_t)n * a.b_wpr + k];
#pragma unroll
              for (int p = 0; p < PL; p++) {
                const uint64_t x = XOR ? (av[p] ^ bv) : (av[p] & bv);
                c[n][p][q] = bcnt_add((uint32_t)(x >> 32), bcnt_add((uint32_t)x, c[n][p][q]));
              }
            }
          }
        }
      }
    }
    // reduce the G partial counters of each row (all lanes participate)
#pragma unroll
    for (int n = 0; n < NV; n++)
#pragma unroll
      for (int p = 0; p < PL; p++)
#pragma unroll
        for (int q = 0; q < PR; q++) {
#pragma unroll
          for (int off = G / 2; off > 0; off >>= 1)
            c[n][p][q] += (uint32_t)__shfl_xor((int)c[n][p][q], off, 64);
        }
    if (kl == 0 && m < a.M) {
#pragma unroll
      for (int n = 0; n < NV; n++)
        if ((uint32_t)n < a.N) emit(a, (uint64_t)n * a.M + m, fold<PL, PR, XOR>(c[n], a, a.add_base != 0, m, n));
    }
  }
}

__global__ __launch_bounds__(256) void skinny_kernel(SkinnyProb p) {
  __shared__ uint4 lds[kSkinnyLds / 16];
  skinny_item(p, blockIdx.x, lds);
}

// ---------------------------------------------------------------- launch
template <int PL, int PR, bool XOR, int TM, int TN, int KW>
int run_tiled(PopcArgs a, int splits_cap, hipStream_t s, bool *used_atomic) {
  constexpr int BM = 16 * TM, BN = 16 * TN;
  a.mtiles = (a.M + BM - 1) / BM;
  a.ntiles = (a.N + BN - 1) / BN;
  const uint64_t tiles = (uint64_t)a.mtiles * a.ntiles;
  uint32_t splits = 1;
  // fill the 256 CUs: split K when there are fewer than two tiles per CU
  if (splits_cap > 1 && tiles < 512 && a.kwords >= 8u * KW) {
    uint64_t want = (512 + tiles - 1) / tiles;
    uint64_t maxs = a.kwords / (4 * KW);
    splits = (uint32_t)(want < maxs ? want : maxs);
    if (splits > (uint32_t)splits_cap) splits = (uint32_t)splits_cap;
    if (splits < 1) splits = 1;
  }
  uint32_t kchunk = (a.kwords + splits - 1) / splits;
  kchunk = (kchunk + KW - 1) / KW * KW;
  if (kchunk == 0) kchunk = KW;
  splits = (a.kwords + kchunk - 1) / kchunk;
  if (splits == 0) splits = 1;
  a.kchunk = kchunk;
  *used_atomic = splits > 1;
  if (splits > 1) a.mode = 2;  // the caller zeroed res when it allowed splitting
  const uint64_t grid = tiles * splits;
  if (grid == 0) return GBS_OK;
  if (grid > 0x7fffffffull) return fail(GBS_ERANGE, "gemm grid too large");
  hipLaunchKernelGGL((popc_kernel<PL, PR, XOR, TM, TN, KW>), dim3((uint32_t)grid),
                     dim3(256), 0, s, a);
  GBS_HIP(hipGetLastError());
  return GBS_OK;
}

template <int PL, int PR, bool XOR>
int run_gemv(PopcArgs a, hipStream_t s) {
  int sms = 256;
  uint64_t rows_per_block;
  if (a.kwords >= 48) {
    rows_per_block = 4;  // G = 64
  } else {
    rows_per_block = 16;  // G = 16
  }
  uint64_t blocks = (a.M + rows_per_block - 1) / rows_per_block;
  if (blocks > (uint64_t)sms * 16) blocks = (uint64_t)sms * 16;
  if (blocks == 0) return GBS_OK;
#define GBS_GEMV(NV)                                                                      \
  do {                                                                                    \
    if (a.kwords >= 48)                                                                   \
      hipLaunchKernelGGL((gemv_kernel<PL, PR, XOR, NV, 64>), dim3((uint32_t)blocks),     \
                         dim3(256), 0, s, a);                                             \
    else                                                                                  \
      hipLaunchKernelGGL((gemv_kernel<PL, PR, XOR, NV, 16>), dim3((uint32_t)blocks),     \
                         dim3(256), 0, s, a);                                             \
  } while (0)
  if (a.N == 1) GBS_GEMV(1);
  else if (a.N == 2) GBS_GEMV(2);
  else GBS_GEMV(4);
#undef GBS_GEMV
  GBS_HIP(hipGetLastError());
  return GBS_OK;
}

constexpr int tdim(int P) { return P == 1 ? 8 : (P == 2 ? 4 : 2); }

// Half the register tile each way when the full one leaves fewer than 512
// blocks and the half one does not: a product that would otherwise split K
// four ways (a handful of K tiles per block, a zeroed result and atomic
// adds) runs unsplit.
bool half_tiles(uint32_t M, uint32_t N, int pl, int pr) {
  const uint64_t BM = 16 * tdim(pl), BN = 16 * tdim(pr);
  const uint64_t full = ((M + BM - 1) / BM) * ((N + BN - 1) / BN);
  const uint64_t half = ((M + BM / 2 - 1) / (BM / 2)) * ((N + BN / 2 - 1) / (BN / 2));
  return full < 512 && half >= 512;
}

template <int PL, int PR, bool XOR>
int run_pair(PopcArgs a, bool allow_split, hipStream_t s, bool *used_atomic) {
  *used_atomic = false;
  if (a.N <= 4 && a.mode != 2) return run_gemv<PL, PR, XOR>(a, s);
  if (half_tiles(a.M, a.N, PL, PR))
    return run_tiled<PL, PR, XOR, tdim(PL) / 2, tdim(PR) / 2, 8>(a, allow_split ? 64 : 1, s,
                                                                 used_atomic);
  return run_tiled<PL, PR, XOR, tdim(PL), tdim(PR), 8>(a, allow_split ? 64 : 1, s,
                                                       used_atomic);
}

int run_group(PopcArgs a, int pl, int pr, bool XOR, bool allow_split, hipStream_t s,
              bool *used_atomic) {
  if (XOR) return run_pair<1, 1, true>(a, allow_split, s, used_atomic);
#define GBS_P(L, R) \
  if (pl == L && pr == R) return run_pair<L, R, false>(a, allow_split, s, used_atomic);
  GBS_P(1, 1) GBS_P(1, 2) GBS_P(1, 3) GBS_P(1, 4)
  GBS_P(2, 1) GBS_P(2, 2) GBS_P(2, 3) GBS_P(2, 4)
  GBS_P(3, 1) GBS_P(3, 2) GBS_P(3, 3) GBS_P(3, 4)
  GBS_P(4, 1) GBS_P(4, 2) GBS_P(4, 3) GBS_P(4, 4)
#undef GBS_P
  return fail(GBS_ENOTSUP, "plane group %dx%d", pl, pr);
}

// Would the tiled launch for this shape split K (and so need atomics)?
bool wants_split(uint32_t M, uint32_t N, uint32_t kwords, int pl, int pr) {
  if (N <= 4) return false;
  const uint32_t hf = half_tiles(M, N, pl, pr) ? 2 : 1;
  const uint32_t BM = 16 * tdim(pl) / hf, BN = 16 * tdim(pr) / hf;
  const uint64_t tiles = (uint64_t)((M + BM - 1) / BM) * ((N + BN - 1) / BN);
  return tiles < 512 && kwords >= 64 && (uint64_t)kwords / 32 >= 2;
}

}  // namespace

// Skinny engine: N <= 8 rhs rows and at most 4 planes a side (a bipolar
// side counts 2: its plane and the ones plane).
bool skinny_supported(const Operand &l, const Operand &r) {
  if (r.nrows == 0 || r.nrows > kSkinnyMaxN || l.nrows == 0) return false;
  if (l.nrows > 0x7fffffffull) return false;
  return (l.bipolar || l.nbits <= 4) && (r.bipolar || r.nbits <= 4);
}

static void plane_weights(const Operand &o, bool other_bipolar, int32_t *w, uint32_t *n,
                          bool *ones) {
  *ones = false;
  if (o.bipolar && other_bipolar) {  // handled by the XOR form
    w[0] = 1;
    *n = 1;
    return;
  }
  if (o.bipolar) {  // 2b - 1: the plane weighs 2, the ones plane -1
    w[0] = 2;
    w[1] = -1;
    *n = 2;
    *ones = true;
    return;
  }
  for (uint32_t i = 0; i < o.nbits; i++) w[i] = 1 << i;
  if (o.issigned) w[o.nbits - 1] = -w[o.nbits - 1];
  *n = o.nbits;
}

double skinny_bytes(const Operand &l, const Operand &r, uint64_t K) {
  (void)r;
  const uint64_t kwords = ((K + 63) / 64 + 1) & ~1ull;
  return (double)l.nrows * (double)kwords * 8 * (double)l.nbits;
}

// Plan of a skinny product.  Lanes per row G: the largest power of two up
// to 16 that leaves every lane at least 4 slices of 128 bits, doubled (up
// to 64, at least one slice a lane) while one pass of 512 / G rows streams
// more than item_bytes of lhs words.  Rows per item: as many passes as
// keep an item near item_bytes.
SkinnyProb skinny_plan(const Operand &l, const Operand &r, uint64_t K, int32_t *res,
                       double item_bytes) {
  SkinnyProb p{};
  p.A = l.planes;
  p.B = r.planes;
  p.a_wpr = l.wpr; p.a_wpp = l.wpp;
  p.b_wpr = r.wpr; p.b_wpp = r.wpp;
  p.res = res;
  p.M = (uint32_t)l.nrows;
  p.N = (uint32_t)r.nrows;
  p.K = (uint32_t)K;
  p.kw_real = (uint32_t)((K + 63) / 64);
  p.kwords = (p.kw_real + 1) & ~1u;
  p.xor_mode = l.bipolar && r.bipolar;
  int32_t wl[4] = {0, 0, 0, 0}, wr[4] = {0, 0, 0, 0};
  bool lo = false, ro = false;
  plane_weights(l, r.bipolar, wl, &p.pl, &lo);
  plane_weights(r, l.bipolar, wr, &p.pr, &ro);
  p.ones = lo ? 1 : ro ? 2 : 0;
  for (int i = 0; i < 4; i++)
    for (int j = 0; j < 4; j++) p.W[i * 4 + j] = p.xor_mode ? -2 : wl[i] * wr[j];
  const uint32_t kq = p.kwords / 2;
  const double lplanes = (double)l.nbits;
  auto pass_bytes = [&](uint32_t gl) {
    return (double)std::min<uint32_t>(p.M, 512u >> gl) * lplanes * p.kwords * 8;
  };
  uint32_t g = 0;
  while (g < 4 && (2u << g) * 4 <= kq) g++;
  while (g < 6 && pass_bytes(g) > item_bytes && (1u << (g + 1)) <= kq) g++;
  p.g_log2 = g;
  const uint32_t pass = 512u >> g;
  const uint32_t reps = (uint32_t)std::max(1.0, std::min(item_bytes / pass_bytes(g), 1e6));
  const uint64_t rows = std::min<uint64_t>((uint64_t)pass * reps, 0x40000000ull);
  p.rows = (uint32_t)rows;
  p.items = (uint32_t)((p.M + rows - 1) / rows);
  return p;
}

// Modelled time on the skinny kernel: the larger of its VALU work (4 AND,
// 4 chained counts and one scaled add per 128-bit slice, row, rhs row and
// plane pair, at 60 % of the 39.3 T lane-op/s peak) and its lhs plane bytes
// (at 70 % of the 8 TB/s HBM peak).
double skinny_time(const Operand &l, const Operand &r, uint64_t K) {
  const double M = (double)l.nrows, N = (double)r.nrows;
  const double kq = (double)((((K + 63) / 64 + 1) & ~1ull) / 2);
  const bool x = l.bipolar && r.bipolar;
  const double pl = x ? 1 : l.bipolar ? 2 : (double)l.nbits;
  const double pr = x ? 1 : r.bipolar ? 2 : (double)r.nbits;
  const double valu = M * kq * N * pl * pr * 9 / (0.6 * 39.3e12);
  const double bytes = M * kq * 16 * (double)l.nbits / (0.7 * 8e12);
  return valu > bytes ? valu : bytes;
}

bool popc_needs_rowsum(const Operand &l, const Operand &r, uint64_t K) {
  return l.bipolar != r.bipolar && !(skinny_supported(l, r) && K < (1ull << 31));
}

int launch_popc_gemm(const Operand &l, const Operand &r, uint64_t kwords,
                     uint64_t K, int32_t *d_res, const int32_t *d_rowsum,
                     hipStream_t s) {
  const uint64_t M = l.nrows, N = r.nrows;
  if (M == 0 || N == 0) return GBS_OK;
  if (M > 0x7fffffffull || N > 0x7fffffffull || kwords > 0x3fffffffull)
    return fail(GBS_ERANGE, "gemm dimensions exceed 2^31");
  const bool XOR = l.bipolar && r.bipolar;
  PopcArgs a{};
  a.res = d_res;
  a.rowsum = d_rowsum;
  a.M = (uint32_t)M;
  a.N = (uint32_t)N;
  a.kwords = (uint32_t)kwords;
  a.K = (uint32_t)K;
  a.a_wpr = l.wpr; a.a_wpp = l.wpp;
  a.b_wpr = r.wpr; a.b_wpp = r.wpp;
  a.bpreg = (l.bipolar != r.bipolar) ? (l.bipolar ? 1 : 2) : 0;
  if (kwords == 0) {
    // empty depth: every dot product is 0 (bipolar x regular: -0 rowsum)
    return launch_fill_i32(d_res, M * N, 0, s);
  }
  if (skinny_supported(l, r) && K < (1ull << 31)) {  // the plan holds K in 32 bits
    // items of about 1/1024 of the product (4 blocks per CU), 16 KiB or more
    const SkinnyProb p = skinny_plan(l, r, K, d_res,
                                     std::max(16384.0, skinny_bytes(l, r, K) / 1024));
    hipLaunchKernelGGL(skinny_kernel, dim3(p.items), dim3(256), 0, s, p);
    GBS_HIP(hipGetLastError());
    return GBS_OK;
  }
  const int gl = (int)((l.nbits + 3) / 4), gr = (int)((r.nbits + 3) / 4);
  const int ngroups = XOR ? 1 : gl * gr;
  // K splitting needs a zeroed result and atomics; decide it up front
  bool split = false;
  if (!XOR) {
    const int pl0 = (int)(l.nbits < 4 ? l.nbits : 4), pr0 = (int)(r.nbits < 4 ? r.nbits : 4);
    split = wants_split(a.M, a.N, a.kwords, pl0, pr0);
  } else {
    split = wants_split(a.M, a.N, a.kwords, 1, 1);
  }
  if (split) {
    int rc = launch_fill_i32(d_res, M * N, 0, s);
    if (rc) return rc;
  }
  bool first = true;
  for (int g = 0; g < (XOR ? 1 : gl); g++) {
    for (int h = 0; h < (XOR ? 1 : gr); h++) {
      const int pl = XOR ? 1 : (int)((l.nbits - 4 * g) < 4 ? (l.nbits - 4 * g) : 4);
      const int pr = XOR ? 1 : (int)((r.nbits - 4 * h) < 4 ? (r.nbits - 4 * h) : 4);
      PopcArgs b = a;
      b.A = l.planes + (uint64_t)4 * g * l.wpp;
      b.B = r.planes + (uint64_t)4 * h * r.wpp;
      b.lshift = 4 * g;
      b.rshift = 4 * h;
      b.lneg = (l.issigned && !l.bipolar && (l.nbits - 1) / 4 == (uint64_t)g)
                   ? (uint32_t)(l.nbits - 1 - 4 * g) : kNoPlane;
      b.rneg = (r.issigned && !r.bipolar && (r.nbits - 1) / 4 == (uint64_t)h)
                   ? (uint32_t)(r.nbits - 1 - 4 * h) : kNoPlane;
      b.add_base = first ? 1 : 0;
      b.mode = split ? 2 : (first ? 0 : 1);
      bool used_atomic = false;
      int rc = run_group(b, pl, pr, XOR, split, s, &used_atomic);
      if (rc) return rc;
      first = false;
    }
  }
  (void)ngroups;
  return GBS_OK;
}

}  // namespace gbs
