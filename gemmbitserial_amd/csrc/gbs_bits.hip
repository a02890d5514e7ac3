// gbs_bits.hip -- matrix-core GEMM of two 1-bit operands read straight from
// their bit-planes: every bit becomes an FP4 (E2M1) code in registers, just
// before v_mfma_scale_f32_16x16x128_f8f6f4 takes it, so no value copy is
// made and each operand row streams 1/4 of the bytes its FP4 copy would.
//
// Same contract as the other engines (reference: gemmBitSerial_generic,
// arch-generic.hpp:331-345; the 1-bit cases are gemmBipolar_generic_naive,
// :239-262, and the one-plane-pair pass of usingBinary, :169-193):
// res[n * M + m] = sum_k val(lhs, m, k) * val(rhs, n, k), with val the
// bipolar value +-1 of the bit (signed 1-bit operands are bipolar,
// gemmbitserial.hpp:74-76) or the unsigned bit.
//
// Code of a bit under the FP4 engine's block scale 2^1: unsigned 1 ->
// 0b0001 (0.5 * 2), 0 -> 0; bipolar 1 -> 0b0001 (+1), 0 -> 0b1001 (-1).
// Where both ports would run under 2^1 at every instruction (bipolar x
// bipolar in bits_kernel, every product in bits_dl_kernel) the instruction
// runs without scales (products of a quarter) and the epilogue multiplies
// the exact fp32 sums by 4: the unscaled encoding issues faster (fp4_mfma).
// K order inside an instruction: both ports take the same (lane, register,
// nibble) -> k assignment, so any bijection works.  Lane l of a fragment
// holds row l % 16 and lane group g = l / 16; per 1024-bit super-step a
// lane loads 32 bytes of its row, the 32-bit words 8g .. 8g + 7 of the
// step, in two halves of four words.  bits_dl_kernel: instruction c of a
// half takes one word, bit 4j + d of it becoming nibble j of register d.
// bits_kernel: instruction D (0..3) of a half takes bit position D of all
// four words, bit 4j + D of word r becoming nibble j of register r, with
// per-instruction block scales for the position's weight (step_reg).
// Either way the instructions of a super-step cover every bit of it once.
//
// Padding: bits past K are zero in a plane row, and words past the row's
// data are read as zero.  A zero bit of an unsigned operand is 0, so it
// adds nothing; under bipolar x bipolar every such position adds (-1)(-1),
// and the epilogue subtracts their number (1024 * steps - K).
//
// Tile: 64 rhs rows (MFMA A port, so output lanes run along m, the
// contiguous dimension of res) x 64 lhs rows, four waves.  bits_kernel:
// each wave owns the whole tile and every fourth super-step of K, loading
// the next one's words into registers as the current one's are spent; the
// four partial sums meet in LDS.  bits_dl_kernel (shallow products): the
// waves split the tile and stream every super-step through an LDS ring.
#include "gbs_internal.h"
#include "gbs_mfma_tile.h"

namespace gbs {
namespace {

struct BitsProb {
  const uint64_t *R, *L;  // plane 0 of rhs / lhs
  uint64_t r_wpr, l_wpr;  // row pitches, 64-bit words
  int32_t *res;           // N x M
  uint32_t M, N;          // lhs rows, rhs rows
  uint32_t kw;            // 64-bit words of a row holding K: ceil(K / 64)
  uint32_t nss;           // super-steps of 16 words: ceil(kw / 16)
  uint32_t ptiles;        // 64-row tiles of rhs
  int32_t corr;           // subtracted from every output
};

constexpr int kBitsTile = 64;
constexpr int kBitsWaves = 4;
constexpr int kBitsPartStride = 68;  // int32 per LDS partial row: 4 rows of a lane group apart by 16 banks

// Registers d = 0..3 of one lane's FP4 fragment from a 32-bit word of bits
// (bit 4j + d -> nibble j of register d); the upper four of the eight
// operand registers are unused by FP4 operands.
template <bool BIP>
__device__ __forceinline__ v8i bits_frag(uint32_t x) {
  v8i f;
  if constexpr (BIP) {
    // sign bit 3 of a nibble set where the bit is 0: one shift and one
    // three-input bitwise op per register
    f[0] = (int)((~(x << 3) & 0x88888888u) | 0x11111111u);
    f[1] = (int)((~(x << 2) & 0x88888888u) | 0x11111111u);
    f[2] = (int)((~(x << 1) & 0x88888888u) | 0x11111111u);
    f[3] = (int)((~x & 0x88888888u) | 0x11111111u);
  } else {
    f[0] = (int)(x & 0x11111111u);
    f[1] = (int)((x >> 1) & 0x11111111u);
    f[2] = (int)((x >> 2) & 0x11111111u);
    f[3] = (int)((x >> 3) & 0x11111111u);
  }
  f[4] = f[5] = f[6] = f[7] = 0;
  return f;
}

// The register-form kernel's fragments take another bijection: k-step D
// (0..3) of a half super-step takes bit position D of all four of the
// lane's 32-bit words, register r holding bit 4j + D of word r in nibble j.
// Bits already at nibble positions 0, 1 and 2 of an unsigned operand are
// FP4 codes as they stand (0.5, 1 and 2: one AND each); position 3 (the
// sign bit) is moved to position 2 (value 2).  Each instruction's block
// scales cancel the step's weight: the unsigned side runs under 2^1, 2^0,
// 2^-1, 2^-1 for D = 0..3, a bipolar side (magnitude 0.5, sign where the
// bit is 0) under 2^1 throughout, so every product of two set bits is
// exactly 1 (unsigned x unsigned: 0.5 * 0.5 * 4, 1 * 1 * 1, 2 * 2 / 4) and
// of two bipolar values +-1.  The unsigned side costs 4 + 4 + 4 + 8 VALU
// operations per fragment over the four steps instead of 4 x 7.
template <bool BIP, int D>
__device__ __forceinline__ uint32_t step_reg(uint32_t x) {
  if constexpr (BIP) {
    if constexpr (D == 3) return (~x & 0x88888888u) | 0x11111111u;
    else return (~(x << (3 - D)) & 0x88888888u) | 0x11111111u;
  } else {
    if constexpr (D == 3) return (x >> 1) & 0x44444444u;
    else return x & (0x11111111u << D);
  }
}

template <bool BIP, int D>
__device__ __forceinline__ v8i step_frag(const uint4 &w) {
  v8i f;
  f[0] = (int)step_reg<BIP, D>(w.x);
  f[1] = (int)step_reg<BIP, D>(w.y);
  f[2] = (int)step_reg<BIP, D>(w.z);
  f[3] = (int)step_reg<BIP, D>(w.w);
  f[4] = f[5] = f[6] = f[7] = 0;
  return f;
}

// E8M0 block scale of a side in step D (see step_reg)
template <bool BIP, int D>
constexpr int step_scale() {
  return BIP ? 128 : (D == 0 ? 128 : D == 1 ? 127 : 126);
}

// The FP4 matrix instruction under E8M0 block scales SA, SB; with both at
// 2^0 the instruction without scales is emitted instead (the builtin with
// zero scale operands compiles to v_mfma_f32_16x16x128_f8f6f4), whose
// products are those of scale 2^0 -- measured equal -- and which issues
// about 8 % faster with two or four waves per SIMD (registers only: 14.2
// against 15.5 ns per instruction per SIMD at two waves).
template <int SA, int SB>
__device__ __forceinline__ v4f fp4_mfma(const v8i &a, const v8i &b, const v4f &c) {
  if constexpr (SA == 127 && SB == 127)
    return __builtin_amdgcn_mfma_scale_f32_16x16x128_f8f6f4(a, b, c, 4, 4, 0, 0, 0, 0);
  else
    return __builtin_amdgcn_mfma_scale_f32_16x16x128_f8f6f4(a, b, c, 4, 4, 0, SA, 0, SB);
}

// Bipolar x bipolar: both sides would run under 2^1 at every step (products
// +-1); they run unscaled instead (products +-1/4, sums exact multiples of
// 1/4 below 2^22) and the epilogue multiplies by 4.  Unsigned steps keep
// their scales except where both are 2^0.
template <bool BR, bool BL, int D>
constexpr int run_scale_r() { return BR && BL ? 127 : step_scale<BR, D>(); }
template <bool BR, bool BL, int D>
constexpr int run_scale_l() { return BR && BL ? 127 : step_scale<BL, D>(); }

// Two 64-bit words of a row at q: one 16-byte load when the row pitch is
// even (rows 16-byte aligned), else two 8-byte loads.
template <bool EVEN>
__device__ __forceinline__ uint4 bits_pair(const uint64_t *q) {
  if constexpr (EVEN) return *reinterpret_cast<const uint4 *>(q);
  const uint64_t x = q[0], y = q[1];
  return make_uint4((uint32_t)x, (uint32_t)(x >> 32), (uint32_t)y, (uint32_t)(y >> 32));
}

// The same for the last, partly filled super-step: words from kw on read
// as zero and are not loaded (w: the row's word index of q[0]).
__device__ __forceinline__ uint4 bits_pair_tail(const uint64_t *q, uint32_t w, uint32_t kw) {
  const uint64_t x = w < kw ? q[0] : 0ull;
  const uint64_t y = w + 1 < kw ? q[1] : 0ull;
  return make_uint4((uint32_t)x, (uint32_t)(x >> 32), (uint32_t)y, (uint32_t)(y >> 32));
}

__device__ __forceinline__ uint32_t word_of(const uint4 &v, int c) {
  return c == 0 ? v.x : c == 1 ? v.y : c == 2 ? v.z : v.w;
}

template <bool BR, bool BL, bool EVEN>
__global__ __launch_bounds__(256, 2) void bits_kernel(BitsProb p) {
  __shared__ int32_t part[kBitsWaves][kBitsTile * kBitsPartStride];
  const uint32_t T = gridDim.x, b = blockIdx.x;
  // consecutive logical tiles (the rhs tiles of one lhs tile) run on one
  // XCD (blocks are dealt round-robin over the 8 XCDs)
  const uint32_t t = (T % 8 == 0) ? (b % 8) * (T / 8) + b / 8 : b;
  const uint32_t pt = t % p.ptiles, qt = t / p.ptiles;
  const uint32_t n0 = pt * kBitsTile, m0 = qt * kBitsTile;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int g = lane >> 4, fr = lane & 15;
  const uint32_t wv = __builtin_amdgcn_readfirstlane(wave);  // uniform: scalar loop control

  // rows past the operand read its last row (in bounds; never stored)
  const uint64_t *ra[4], *la[4];
#pragma unroll
  for (int i = 0; i < 4; i++) {
    ra[i] = p.R + (uint64_t)min(n0 + 16 * i + fr, p.N - 1) * p.r_wpr + 4 * g;
    la[i] = p.L + (uint64_t)min(m0 + 16 * i + fr, p.M - 1) * p.l_wpr + 4 * g;
  }

  v4f acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; i++)
#pragma unroll
    for (int j = 0; j < 4; j++) acc[i][j] = v4f{0.f, 0.f, 0.f, 0.f};

  uint4 xa[4][2], xb[4][2];  // [fragment][half of the lane's 32 bytes]
  auto load_half = [&](uint32_t s, int h) {
    const uint32_t off = 16 * s + 2 * h;  // word offset past the lane's base
#pragma unroll
    for (int i = 0; i < 4; i++) xa[i][h] = bits_pair<EVEN>(ra[i] + off);
#pragma unroll
    for (int i = 0; i < 4; i++) xb[i][h] = bits_pair<EVEN>(la[i] + off);
  };
  auto load_tail = [&](uint32_t s, int h) {
    const uint32_t off = 16 * s + 2 * h, w = off + 4 * g;
#pragma unroll
    for (int i = 0; i < 4; i++) xa[i][h] = bits_pair_tail(ra[i] + off, w, p.kw);
#pragma unroll
    for (int i = 0; i < 4; i++) xb[i][h] = bits_pair_tail(la[i] + off, w, p.kw);
  };
  auto step_d = [&](int h, auto Dc) {
    constexpr int D = decltype(Dc)::value;
    v8i fa[4], fb[4];
#pragma unroll
    for (int i = 0; i < 4; i++) fa[i] = step_frag<BR, D>(xa[i][h]);
#pragma unroll
    for (int j = 0; j < 4; j++) fb[j] = step_frag<BL, D>(xb[j][h]);
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
      for (int j = 0; j < 4; j++)
        acc[i][j] = fp4_mfma<run_scale_r<BR, BL, D>(), run_scale_l<BR, BL, D>()>(fa[i], fb[j],
                                                                               acc[i][j]);
  };
  // the four k-steps of half h
  auto half = [&](int h) {
    step_d(h, std::integral_constant<int, 0>());
    step_d(h, std::integral_constant<int, 1>());
    step_d(h, std::integral_constant<int, 2>());
    step_d(h, std::integral_constant<int, 3>());
  };

  // full super-steps (all 16 words hold data): wave w takes w, w + 4, ...,
  // loading the next one's halves as the current one's are spent
  const uint32_t nfull = p.kw / 16;
  uint32_t s = wv;
  if (s < nfull) {
    load_half(s, 0);
    load_half(s, 1);
  }
  for (; s < nfull; s += kBitsWaves) {
    const bool more = s + kBitsWaves < nfull;
    half(0);
    if (more) load_half(s + kBitsWaves, 0);
    half(1);
    if (more) load_half(s + kBitsWaves, 1);
  }
  // the partly filled last super-step, if any, by the wave whose turn it is
  if (nfull < p.nss && wv == nfull % kBitsWaves) {
    load_tail(nfull, 0);
    load_tail(nfull, 1);
    half(0);
    half(1);
  }

  // the four waves' partial sums (exact integers in fp32: |sum| <= 1024 *
  // steps < 2^24; quarters of them for bipolar x bipolar, scaled back
  // here) meet in LDS; thread tid then owns column m = tid % 64 of
  // rows tid / 64 + 4k, so every store instruction writes 64 consecutive
  // int32 of one res row
  int32_t *mine = part[wave];
#pragma unroll
  for (int i = 0; i < 4; i++)
#pragma unroll
    for (int j = 0; j < 4; j++)
#pragma unroll
      for (int r = 0; r < 4; r++)
        mine[(16 * i + 4 * g + r) * kBitsPartStride + 16 * j + fr] =
            (int32_t)(acc[i][j][r] * (BR && BL ? 4.0f : 1.0f));
  __syncthreads();
  const int m = tid & 63;
  if (m0 + m < p.M) {
#pragma unroll 4
    for (int k = 0; k < 16; k++) {
      const int n = (tid >> 6) + 4 * k;
      if (n0 + n >= p.N) break;
      const int o = n * kBitsPartStride + m;
      p.res[(uint64_t)(n0 + n) * p.M + m0 + m] =
          part[0][o] + part[1][o] + part[2][o] + part[3][o] - p.corr;
    }
  }
}

// The same product through a ring of LDS stages filled by direct loads
// (global_load_lds_dwordx4: lane i of a wave writes the 16 bytes it loads
// at the wave's LDS base + 16 i), so the loads of a super-step are issued
// NBUF - 1 slots ahead without holding registers.  A stage is one
// super-step of the tile's 64 rhs and 64 lhs rows, 128 bytes each (16 KB);
// 16-byte chunk ch of row r sits in slot ch ^ (r % 8), which keeps the
// fragment reads (eight rows of one lane group per 128 bytes) free of bank
// conflicts.  The four waves own 32 x 32 quarters of the tile (2 x 2
// fragments) and every super-step: no K split, no reduction.  Only full
// super-steps go through the ring; a partly filled last one is loaded into
// registers with bounds, as in bits_kernel.  Used for depths of up to two
// super-steps, where the K split of bits_kernel leaves waves idle.
constexpr int kBitsStage = 2 * kBitsTile * 128;  // bytes of one super-step
constexpr int kBitsRing = 4;                     // stages in LDS

template <bool BR, bool BL>
__global__ __launch_bounds__(256, 2) void bits_dl_kernel(BitsProb p) {
  constexpr int NBUF = kBitsRing, D = NBUF - 1, SLOT = kBitsStage;
  constexpr int PER = 4;  // direct loads per wave and stage: 1024 chunks over 256 lanes
  __shared__ __attribute__((aligned(16))) int8_t lds[NBUF * SLOT];
  const uint32_t T = gridDim.x, b = blockIdx.x;
  const uint32_t t = (T % 8 == 0) ? (b % 8) * (T / 8) + b / 8 : b;
  const uint32_t pt = t % p.ptiles, qt = t / p.ptiles;
  const uint32_t n0 = pt * kBitsTile, m0 = qt * kBitsTile;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wp = wave & 1, wq = wave >> 1;
  const int g = lane >> 4, fr = lane & 15;

  // the lane's four direct loads per stage: chunk slot c = (4 i + wave) * 64
  // + lane of the stage; side, row and logical chunk from it (rows past the
  // operand read its last row)
  const uint64_t *src[PER];
#pragma unroll
  for (int i = 0; i < PER; i++) {
    const int c = (4 * i + wave) * 64 + lane;
    const int side = c >> 9, row = (c >> 3) & 63;
    const int ch = (c & 7) ^ (row & 7);
    const uint64_t *base = side ? p.L + (uint64_t)min(m0 + row, p.M - 1) * p.l_wpr
                                : p.R + (uint64_t)min(n0 + row, p.N - 1) * p.r_wpr;
    src[i] = base + 2 * ch;
  }
  const uint32_t nst = p.kw / 16;  // full super-steps
  auto issue = [&](uint32_t st, int buf) {
#pragma unroll
    for (int i = 0; i < PER; i++) {
      const int c0 = (4 * i + wave) * 64;  // the wave's first chunk slot
      __builtin_amdgcn_global_load_lds(
          (const void *)(src[i] + 16 * st),
          (__attribute__((address_space(3))) void *)(lds + buf * SLOT + c0 * 16), 16, 0, 0);
    }
  };

  v4f acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; i++)
#pragma unroll
    for (int j = 0; j < 2; j++) acc[i][j] = v4f{0.f, 0.f, 0.f, 0.f};
  auto mma4 = [&](const uint4 (&xa)[2], const uint4 (&xb)[2]) {
#pragma unroll
    for (int c = 0; c < 4; c++) {
      v8i fa[2], fb[2];
#pragma unroll
      for (int i = 0; i < 2; i++) fa[i] = bits_frag<BR>(word_of(xa[i], c));
#pragma unroll
      for (int j = 0; j < 2; j++) fb[j] = bits_frag<BL>(word_of(xb[j], c));
#pragma unroll
      for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 2; j++)
          // every code 0, +-0.5 or 0.5 in both ports: unscaled (products
          // of a quarter), the epilogue multiplies by 4
          acc[i][j] = fp4_mfma<127, 127>(fa[i], fb[j], acc[i][j]);
    }
  };
  // the lane's 16 bytes (words 8g + 4h .. + 3 of the super-step) of row r
  auto rd = [&](const int8_t *side, int r, int h) {
    return *reinterpret_cast<const uint4 *>(side + r * 128 + 16 * ((2 * g + h) ^ (r & 7)));
  };
  auto compute = [&](int buf) {
    const int8_t *sR = lds + buf * SLOT, *sL = sR + kBitsTile * 128;
#pragma unroll
    for (int h = 0; h < 2; h++) {
      uint4 xa[2], xb[2];
#pragma unroll
      for (int i = 0; i < 2; i++) xa[i] = rd(sR, 32 * wp + 16 * i + fr, h);
#pragma unroll
      for (int j = 0; j < 2; j++) xb[j] = rd(sL, 32 * wq + 16 * j + fr, h);
      mma4(xa, xb);
    }
  };

  static_assert(D == 3, "the waits below count three stages in flight");
#pragma unroll
  for (int d = 0; d < D; d++)
    if ((uint32_t)d < nst) issue(d, d);
  for (uint32_t i = 0; i < nst; i++) {
    // stage i has landed once at most (stages issued after it) x PER loads
    // of this wave are outstanding
    const uint32_t after = nst - 1 - i < (uint32_t)(D - 1) ? nst - 1 - i : (uint32_t)(D - 1);
    if (after >= 2) wait_vmcnt<2 * PER>();
    else if (after == 1) wait_vmcnt<PER>();
    else wait_vmcnt<0>();
    // stage i visible to every wave; every wave is done with stage i - 1,
    // whose buffer the next loads reuse
    asm volatile("" ::: "memory");
    __builtin_amdgcn_s_barrier();
    asm volatile("" ::: "memory");
    if (i + D < nst) issue(i + D, (int)((i + D) % NBUF));
    compute((int)(i % NBUF));
  }
  // the partly filled last super-step, through registers
  if (nst < p.nss) {
    const uint32_t w0 = 16 * nst + 4 * g;
#pragma unroll
    for (int h = 0; h < 2; h++) {
      uint4 xa[2], xb[2];
#pragma unroll
      for (int i = 0; i < 2; i++) {
        const uint32_t n = min(n0 + 32 * wp + 16 * i + fr, p.N - 1);
        const uint32_t m = min(m0 + 32 * wq + 16 * i + fr, p.M - 1);
        xa[i] = bits_pair_tail(p.R + (uint64_t)n * p.r_wpr + w0 + 2 * h, w0 + 2 * h, p.kw);
        xb[i] = bits_pair_tail(p.L + (uint64_t)m * p.l_wpr + w0 + 2 * h, w0 + 2 * h, p.kw);
      }
      mma4(xa, xb);
    }
  }
#pragma unroll
  for (int i = 0; i < 2; i++)
#pragma unroll
    for (int j = 0; j < 2; j++)
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const uint32_t n = n0 + 32 * wp + 16 * i + 4 * g + r;
        const uint32_t m = m0 + 32 * wq + 16 * j + fr;
        if (n < p.N && m < p.M)
          p.res[(uint64_t)n * p.M + m] = (int32_t)(acc[i][j][r] * 4.0f) - p.corr;
      }
}

template <bool BR, bool BL>
void run_bits(const BitsProb &p, bool even, uint32_t grid, hipStream_t s) {
  // bits_kernel splits K over its four waves, which leaves some idle up to
  // three super-steps: up to two the ring kernel (all four waves on every
  // super-step, loads three stages ahead) measured faster, from four the
  // split (three: the split, not measured against the ring)
  if (p.nss <= 2)
    hipLaunchKernelGGL((bits_dl_kernel<BR, BL>), dim3(grid), dim3(256), 0, s, p);
  else if (even)
    hipLaunchKernelGGL((bits_kernel<BR, BL, true>), dim3(grid), dim3(256), 0, s, p);
  else
    hipLaunchKernelGGL((bits_kernel<BR, BL, false>), dim3(grid), dim3(256), 0, s, p);
}

}  // namespace

bool bits_supported(const Operand &l, const Operand &r, uint64_t K) {
  // fp32 partial sums stay exact below 2^24 (padding positions included)
  return l.nbits == 1 && r.nbits == 1 && K + 1024 < (1ull << 24);
}

bool bits_preferred(const Operand &l, const Operand &r, uint64_t K) {
  // measured against the FP4-copy kernel on 1-bit products: 9-18 % faster
  // at depths up to 2048, 2-10 % at 4096 and 8192, while the tiles fit the
  // chip at once or twice (at most four 64 x 64 tiles per CU); slower at
  // 32768 (bipolar) and where many rounds of tiles follow each other
  const uint64_t tiles = ((r.nrows + 63) / 64) * ((l.nrows + 63) / 64);
  return bits_supported(l, r, K) && K <= 8192 && tiles <= 1024;
}

int launch_bits_gemm(const Operand &l, const Operand &r, uint64_t K, int32_t *d_res,
                     hipStream_t s) {
  const uint64_t M = l.nrows, N = r.nrows;
  if (M == 0 || N == 0) return GBS_OK;
  if (!bits_supported(l, r, K))
    return fail(GBS_ENOTSUP, "gemm: the 1-bit matrix-core engine takes 1-bit operands of "
                             "depth below 2^24 - 1024 only");
  if (M > 0x7fffffffull || N > 0x7fffffffull)
    return fail(GBS_ERANGE, "gemm dimensions exceed 2^31");
  if (K == 0) return launch_fill_i32(d_res, M * N, 0, s);
  BitsProb p{};
  p.R = r.planes;
  p.L = l.planes;
  p.r_wpr = r.wpr;
  p.l_wpr = l.wpr;
  p.res = d_res;
  p.M = (uint32_t)M;
  p.N = (uint32_t)N;
  p.kw = (uint32_t)((K + 63) / 64);
  p.nss = (p.kw + 15) / 16;
  if (l.wpr < p.kw || r.wpr < p.kw) return fail(GBS_EINVAL, "gemm: plane rows shorter than K");
  p.ptiles = (uint32_t)((N + 63) / 64);
  const uint64_t tiles = (uint64_t)p.ptiles * ((M + 63) / 64);
  if (tiles > 0x7fffffffull) return fail(GBS_ERANGE, "gemm: too many tiles");
  p.corr = (l.bipolar && r.bipolar) ? (int32_t)(1024ull * p.nss - K) : 0;
  const bool even = (l.wpr % 2 == 0) && (r.wpr % 2 == 0);
  if (r.bipolar) {
    if (l.bipolar) run_bits<true, true>(p, even, (uint32_t)tiles, s);
    else run_bits<true, false>(p, even, (uint32_t)tiles, s);
  } else {
    if (l.bipolar) run_bits<false, true>(p, even, (uint32_t)tiles, s);
    else run_bits<false, false>(p, even, (uint32_t)tiles, s);
  }
  GBS_HIP(hipGetLastError());
  return GBS_OK;
}

}  // namespace gbs
