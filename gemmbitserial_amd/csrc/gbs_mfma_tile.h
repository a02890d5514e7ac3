// gbs_mfma_tile.h -- one output tile of the exact matrix-core GEMM, shared
// by the single-product kernel (gbs_mfma.hip) and the batched persistent
// kernel (gbs_batch.hip).
//
// Tile = P rhs rows (n) x Q lhs rows (m) of res; 256 threads = 4 waves laid
// out WP x (4/WP) over it; each wave owns (P/WP) x (Q/(4/WP)) outputs as
// 16 x 16 fragments.  The rhs operand goes to the MFMA A port so that a
// fragment's output lanes run along m, the contiguous dimension of res.
// Three operand formats share the tile (the same on both ports):
//   FMT 0  int8 values, v_mfma_i32_16x16x64_i8: a k-step is 64 values,
//          64 bytes of a row, 16 bytes per lane;
//   FMT 1  FP4 (E2M1) codes, two per byte, v_mfma_scale_f32_16x16x128_f8f6f4
//          with block scales 2^1: a k-step is 128 values, 64 bytes, 16 per lane;
//   FMT 2  FP6 (E2M3) codes, 6 bits each, the same instruction with block
//          scales 2^3: a k-step is 128 values, 96 bytes, 24 per lane.
// K positions are counted in nominal bytes, 64 per k-step, for every
// format (KSB, kbeg, kend, kbytes); an FP6 row holds 96 real bytes per 64
// nominal ones.  Both operand fragments of an instruction take the same
// k-to-(lane, byte) assignment on both ports, so the sum over k does not
// depend on the order the instruction adds in.  K is staged KSB nominal
// bytes at a time in LDS, double-buffered, with 16-byte global loads held
// in registers one or two stages ahead of the stage being multiplied.
//
// FP6 copies store each 96-byte k-step as the four lane groups' first 16
// bytes (64 bytes) followed by their last 8 bytes (32 bytes), so a lane's
// 24 bytes are one aligned 16-byte and one aligned 8-byte LDS read.
#pragma once

#include <type_traits>

#include "gbs_internal.h"

namespace gbs {

typedef int v2i __attribute__((ext_vector_type(2)));
typedef int v4i __attribute__((ext_vector_type(4)));
typedef int v8i __attribute__((ext_vector_type(8)));
typedef float v4f __attribute__((ext_vector_type(4)));

// Operand formats of the matrix-core engine.
constexpr int kFmtI8 = 0, kFmtF4 = 1, kFmtF6 = 2, kMfmaFormats = 3;
// E8M0 block scales: 2^1 makes FP4 code c (|c| <= 4: 0, 0.5, 1, 1.5, 2)
// the integer c; 2^3 makes FP6 E2M3 code c (|c| <= 16: subnormals c/8,
// then (1 + m/8) * 2^(e-1)) the integer c.  Products and their fp32 sums
// are then exact integers while |sum| < 2^24.
constexpr int kF4Scale = 128, kF6Scale = 130;

template <int FMT>
constexpr int step_bytes() { return FMT == kFmtF6 ? 96 : 64; }  // row bytes per k-step

// one lane's share of an operand fragment
template <int FMT>
struct MfmaFrag { v4i h; };
template <>
struct MfmaFrag<kFmtF6> { v8i x; };  // lanes 0..5 used, loaded in place (no register copies)

template <int FMT>
using MfmaAcc = typename std::conditional<FMT == kFmtI8, v4i, v4f>::type;

// LDS offset of the last 8 bytes of an FP6 fragment, hidden from the
// optimiser when APART: otherwise it pairs these reads across fragments
// into ds_read2 instructions whose halves must then be moved into each
// fragment's 6-register operand (measured: the standalone FP6 tiles'
// LDS-and-MFMA time drops by a fifth without the pairing; inside the
// persistent kernel, whose registers are nearly all taken, the pairing
// measures 3 % faster and is kept there).
template <bool APART>
__device__ __forceinline__ int f6_tail_offset(int off) {
  if constexpr (APART) asm volatile("" : "+v"(off));
  return off;
}

template <int FMT>
__device__ __forceinline__ MfmaAcc<FMT> mfma_op(const MfmaFrag<FMT> &a, const MfmaFrag<FMT> &b,
                                                const MfmaAcc<FMT> &c) {
  if constexpr (FMT == kFmtI8) {
    return __builtin_amdgcn_mfma_i32_16x16x64_i8(a.h, b.h, c, 0, 0, 0);
  } else if constexpr (FMT == kFmtF4) {
    // FP4 operands occupy the low four registers of the 8-register slots
    const v8i a8 = {a.h[0], a.h[1], a.h[2], a.h[3], 0, 0, 0, 0};
    const v8i b8 = {b.h[0], b.h[1], b.h[2], b.h[3], 0, 0, 0, 0};
    return __builtin_amdgcn_mfma_scale_f32_16x16x128_f8f6f4(a8, b8, c, 4, 4, 0, kF4Scale, 0,
                                                            kF4Scale);
  } else {
    // FP6 operands occupy the low six registers
    return __builtin_amdgcn_mfma_scale_f32_16x16x128_f8f6f4(a.x, b.x, c, 2, 2, 0, kF6Scale, 0,
                                                            kF6Scale);
  }
}

constexpr int kMfmaKS = 256;  // padding unit of K (nominal bytes): every stage size divides it

// Operands of one product: value copies (one format on both sides)
// [rows >= tile-aligned][ld bytes].
struct MfmaOperands {
  const int8_t *R, *L;   // rhs / lhs values
  uint64_t r_ld, l_ld;   // row strides (real bytes)
  int32_t *res;          // N x M
  uint32_t M, N;         // lhs rows, rhs rows
  uint32_t kbytes;       // nominal bytes of K in the copies, a multiple of kMfmaKS (zero padded)
};

// Fused requantization of a product's results into the input of the next
// layer (the serving loop): code = min(maxv, max(0, v >> shift)) written as
// int8 into vals (row stride ld) and as bit-planes into planes
// ([nbits][rows][wpr] words, wpp words per plane), so the int32 sums never
// reach memory; with vals null the int32 results go to out (N x M,
// row-major) instead.  Whole-K tiles convert in their epilogue; a product
// split along K adds its pieces into MfmaOperands::res and is converted
// after the launch (launch_views).
struct MfmaRequant {
  int8_t *vals;
  uint64_t *planes;
  int32_t *out;
  uint64_t ld, wpr, wpp;
  uint32_t shift, nbits, maxv;
};

__device__ __forceinline__ uint32_t requant_code(int32_t v, const MfmaRequant &q) {
  const int32_t a = v >> q.shift;  // arithmetic shift
  return a < 0 ? 0u : (a > (int32_t)q.maxv ? q.maxv : (uint32_t)a);
}

__device__ __forceinline__ int32_t acc_value(int32_t x) { return x; }
__device__ __forceinline__ int32_t acc_value(float x) { return (int32_t)x; }  // exact integer

// Results of one wave's FP x FQ fragments (lane holds rhs rows
// nb + 16 i + 4 (lane / 16) + r, r < 4, of lhs column mb + 16 j + lane % 16):
// int32 stores into res (atomic adds for a K piece), or with RQ and a
// whole-K tile the requantized codes -- int8 bytes, and per plane one
// 16-bit piece of a plane word per 16 columns, gathered with one ballot.
template <int FP, int FQ, bool RQ, class AccT>
__device__ __forceinline__ void store_fragments(const MfmaOperands &a, const AccT (&acc)[FP][FQ],
                                                uint32_t nb, uint32_t mb, bool atomic,
                                                const MfmaRequant &q) {
  const int lane = threadIdx.x & 63, fr = lane & 15, g = lane >> 4;
  if constexpr (RQ) {
    if (!atomic && q.vals) {
#pragma unroll
      for (int i = 0; i < FP; i++)
#pragma unroll
        for (int j = 0; j < FQ; j++)
#pragma unroll
          for (int r = 0; r < 4; r++) {
            const uint32_t n = nb + i * 16 + 4 * g + r;
            const uint32_t m0 = mb + j * 16, m = m0 + fr;
            const bool ok = n < a.N && m < a.M;
            const uint32_t code = ok ? requant_code(acc_value(acc[i][j][r]), q) : 0u;
            if (ok) q.vals[(uint64_t)n * q.ld + m] = (int8_t)code;
            for (uint32_t b = 0; b < q.nbits; b++) {
              const uint64_t w = __ballot((code >> b) & 1u);
              if (fr == 0 && n < a.N && m0 < a.M)
                reinterpret_cast<uint16_t *>(q.planes + b * q.wpp + (uint64_t)n * q.wpr)[m0 / 16] =
                    (uint16_t)(w >> (16 * g));
            }
          }
      return;
    }
  }
#pragma unroll
  for (int i = 0; i < FP; i++)
#pragma unroll
    for (int j = 0; j < FQ; j++)
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const uint32_t n = nb + i * 16 + 4 * g + r;
        const uint32_t m = mb + j * 16 + fr;
        if (n < a.N && m < a.M) {
          const uint64_t idx = (uint64_t)n * a.M + m;
          const int32_t v = acc_value(acc[i][j][r]);
          if (atomic) atomicAdd(a.res + idx, v);
          else a.res[idx] = v;
        }
      }
}

// Tile shapes the engine is compiled for: P x Q outputs, WP waves along P,
// KSB bytes of K per LDS stage, KG groups of 4 waves per block (each group
// multiplies every KG-th stage; partial sums are added in LDS at the end),
// DEPTH: with KG = 1, 1 or 2 stages of global loads held in registers
// ahead (mfma_tile), 0 the software-pipelined tile (mfma_tile_pp), 3 or
// more a ring of DEPTH LDS slots filled by direct loads (mfma_tile_dl);
// with KG > 1 always direct loads into a ring of DEPTH slots of KG stages.
struct MfmaTileShape { int P, Q, WP, KSB, DEPTH, KG; };
constexpr int kMfmaShapes = 22;
constexpr MfmaTileShape kMfmaShape[kMfmaShapes] = {
    {128, 128, 2, 128, 2, 1}, {64, 128, 2, 128, 2, 1}, {32, 256, 1, 64, 2, 1},
    {16, 256, 1, 128, 2, 1},  {64, 64, 2, 128, 2, 1},  {128, 64, 4, 128, 2, 1},
    {128, 128, 2, 64, 1, 1},  {128, 256, 2, 64, 2, 1}, {128, 128, 2, 64, 0, 1},
    {128, 256, 2, 64, 0, 1},  {64, 128, 2, 64, 0, 1},  {256, 256, 2, 64, 1, 1},
    {256, 256, 2, 128, 1, 1}, {64, 64, 2, 128, 4, 1},  {64, 64, 2, 256, 3, 1},
    {128, 128, 2, 128, 3, 1}, {128, 64, 4, 128, 3, 1}, {64, 128, 2, 128, 3, 1},
    {64, 64, 2, 128, 3, 1},   {64, 64, 2, 128, 2, 2},  {64, 128, 2, 128, 2, 2},
    {64, 64, 2, 128, 3, 2}};
__host__ __device__ constexpr bool direct_loads(int DEPTH, int KG) { return DEPTH >= 3 || KG > 1; }

// LDS row layout of a stage.  128- and 256-byte stages use rows of exactly
// KSB bytes with the 16-byte chunks of row r XOR-permuted by (r / (256/KSB))
// mod (KSB/16): any 16 consecutive rows at one logical chunk, and any 256
// consecutive bytes written, then fall in 64 distinct banks.  64-byte
// stages pad each row by 16 bytes instead.
template <int KSB>
constexpr int lds_row_bytes() { return KSB >= 128 ? KSB : KSB + 16; }
template <int KSB>
__device__ __forceinline__ int lds_chunk(int row, int ch) {
  if constexpr (KSB >= 128) return ch ^ ((row / (256 / KSB)) % (KSB / 16));
  else return ch;
}


// LDS row stride of a stage in format FMT: FP6 rows (96 bytes per k-step)
// are padded by 16 bytes, which keeps 16 rows at one offset in distinct
// banks for every stage length used here.
template <int KSB, int FMT>
constexpr int lds_row_stride() {
  return FMT == kFmtF6 ? KSB / 64 * 96 + 16 : lds_row_bytes<KSB>();
}
template <int P, int Q, int KSB, int FMT = kFmtI8, int DEPTH = 2, int KG = 1>
constexpr size_t mfma_lds_bytes() {
  return direct_loads(DEPTH, KG) ? (size_t)DEPTH * KG * (P + Q) * (KSB / 64 * step_bytes<FMT>())
                                 : (size_t)2 * (P + Q) * lds_row_stride<KSB, FMT>();
}
constexpr size_t kLdsPerCU = 160 * 1024;
// persistent kernel: the largest carried shape (64 x 64, three FP6 stages)
constexpr size_t kMfmaLdsMax = (size_t)3 * 128 * 192;

template <int P, int Q, int WP, int KSB, int DEPTH, int FMT = kFmtI8, bool TAILS_APART = true,
          bool RQ = false>
__device__ __forceinline__ void mfma_tile(const MfmaOperands &a, uint32_t pt, uint32_t qt,
                                          uint32_t kbeg, uint32_t kend, bool atomic,
                                          int8_t *lds, const MfmaRequant &q = MfmaRequant{}) {
  constexpr int WQ = 4 / WP;
  constexpr int PW = P / WP, QW = Q / WQ;       // wave tile
  constexpr int FP = PW / 16, FQ = QW / 16;     // fragments per wave
  constexpr int S = KSB / 64;                   // k-steps per stage
  constexpr int SB = S * step_bytes<FMT>();     // real bytes per row and stage
  constexpr int LDR = lds_row_stride<KSB, FMT>();  // LDS row stride (bytes)
  constexpr int CPR = SB / 16;                  // 16-byte chunks per row and stage
  constexpr int RCH = P * CPR, LCH = Q * CPR;   // chunks per stage
  constexpr int RLD = (RCH + 255) / 256, LLD = (LCH + 255) / 256;
  static_assert(FP >= 1 && FQ >= 1 && PW % 16 == 0 && QW % 16 == 0, "bad tile");
  static_assert(KSB % 64 == 0 && kMfmaKS % KSB == 0, "bad stage size");
  static_assert(DEPTH == 1 || DEPTH == 2, "prefetch depth");
  static_assert(mfma_lds_bytes<P, Q, KSB, FMT>() <= kLdsPerCU, "LDS");
  int8_t *sR0 = lds, *sL0 = lds + P * LDR;
  constexpr int stage = (P + Q) * LDR;
  // LDS slot of chunk ch of a row: swizzled for int8 / FP4, in place for FP6
  auto slot = [](int row, int ch) { return FMT == kFmtF6 ? ch : lds_chunk<KSB>(row, ch); };

  const uint32_t p0 = pt * P, q0 = qt * Q;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wp = wave % WP, wq = wave / WP;

  v4i gr[DEPTH][RLD], gl[DEPTH][LLD];
  auto gload = [&](v4i(&r)[RLD], v4i(&l)[LLD], uint32_t k0) {
    const uint64_t kb = (uint64_t)k0 / 64 * step_bytes<FMT>();  // real byte offset
#pragma unroll
    for (int i = 0; i < RLD; i++) {
      const int c = tid + 256 * i, row = c / CPR, ch = c % CPR;
      if (RCH % 256 == 0 || c < RCH)
        r[i] = *reinterpret_cast<const v4i *>(a.R + (uint64_t)(p0 + row) * a.r_ld + kb + ch * 16);
    }
#pragma unroll
    for (int i = 0; i < LLD; i++) {
      const int c = tid + 256 * i, row = c / CPR, ch = c % CPR;
      if (LCH % 256 == 0 || c < LCH)
        l[i] = *reinterpret_cast<const v4i *>(a.L + (uint64_t)(q0 + row) * a.l_ld + kb + ch * 16);
    }
  };
  auto swrite = [&](const v4i(&r)[RLD], const v4i(&l)[LLD], int buf) {
    int8_t *sR = sR0 + buf * stage, *sL = sL0 + buf * stage;
#pragma unroll
    for (int i = 0; i < RLD; i++) {
      const int c = tid + 256 * i, row = c / CPR, ch = c % CPR;
      if (RCH % 256 == 0 || c < RCH)
        *reinterpret_cast<v4i *>(&sR[row * LDR + slot(row, ch) * 16]) = r[i];
    }
#pragma unroll
    for (int i = 0; i < LLD; i++) {
      const int c = tid + 256 * i, row = c / CPR, ch = c % CPR;
      if (LCH % 256 == 0 || c < LCH)
        *reinterpret_cast<v4i *>(&sL[row * LDR + slot(row, ch) * 16]) = l[i];
    }
  };
  // lane's fragment of LDS row `row`, k-step kk of the stage
  auto frag = [&](const int8_t *base, int row, int kk) {
    MfmaFrag<FMT> f;
    const int g = lane >> 4;
    if constexpr (FMT == kFmtF6) {
      const int8_t *p = base + row * LDR + kk * 96;
      f.x.s0123 = *reinterpret_cast<const v4i *>(p + 16 * g);
      f.x.s45 = *reinterpret_cast<const v2i *>(p + f6_tail_offset<TAILS_APART>(64 + 8 * g));
      f.x.s67 = v2i{0, 0};
    } else {
      f.h = *reinterpret_cast<const v4i *>(base + row * LDR +
                                           lds_chunk<KSB>(row, kk * 4 + g) * 16);
    }
    return f;
  };

  MfmaAcc<FMT> acc[FP][FQ];
#pragma unroll
  for (int i = 0; i < FP; i++)
#pragma unroll
    for (int j = 0; j < FQ; j++) acc[i][j] = MfmaAcc<FMT>{0, 0, 0, 0};

  const int fr = lane & 15;
  auto compute = [&](int buf) {
    const int8_t *sR = sR0 + buf * stage, *sL = sL0 + buf * stage;
#pragma unroll
    for (int kk = 0; kk < S; kk++) {
      MfmaFrag<FMT> af[FP], bf[FQ];
#pragma unroll
      for (int i = 0; i < FP; i++) af[i] = frag(sR, wp * PW + i * 16 + fr, kk);
#pragma unroll
      for (int j = 0; j < FQ; j++) bf[j] = frag(sL, wq * QW + j * 16 + fr, kk);
#pragma unroll
      for (int i = 0; i < FP; i++)
#pragma unroll
        for (int j = 0; j < FQ; j++)
          acc[i][j] = mfma_op<FMT>(af[i], bf[j], acc[i][j]);
    }
  };

  const uint32_t nst = kend > kbeg ? (kend - kbeg) / KSB : 0;  // stages
  if constexpr (DEPTH == 1) {
    if (nst) {
      gload(gr[0], gl[0], kbeg);
      swrite(gr[0], gl[0], 0);
    }
    __syncthreads();
    for (uint32_t i = 0; i < nst; i++) {
      const bool more = i + 1 < nst;
      if (more) gload(gr[0], gl[0], kbeg + (i + 1) * KSB);
      compute(i & 1);
      if (more) swrite(gr[0], gl[0], (i + 1) & 1);
      __syncthreads();
    }
  } else {
    // stage i sits in LDS buffer i%2 and, before that, in register set i%2;
    // loads run two stages ahead.  The loop is unrolled by two so that the
    // register sets are indexed statically.
    if (nst) gload(gr[0], gl[0], kbeg);
    if (nst > 1) gload(gr[DEPTH - 1], gl[DEPTH - 1], kbeg + KSB);
    if (nst) swrite(gr[0], gl[0], 0);
    __syncthreads();
    auto step = [&](uint32_t i, auto cur) {
      constexpr int c = decltype(cur)::value, n = c ^ 1;
      if (i + 2 < nst) gload(gr[c], gl[c], kbeg + (i + 2) * KSB);
      compute(c);
      if (i + 1 < nst) swrite(gr[n], gl[n], n);
      __syncthreads();
    };
    for (uint32_t i = 0; i < nst; i += 2) {
      step(i, std::integral_constant<int, 0>());
      if (i + 1 < nst) step(i + 1, std::integral_constant<int, DEPTH - 1>());
    }
  }

  store_fragments<FP, FQ, RQ>(a, acc, p0 + wp * PW, q0 + wq * QW, atomic, q);
}

// Block-wide barrier that waits for this wave's LDS operations only (not
// for global loads still in flight), and keeps the compiler from moving
// memory operations across it.
__device__ __forceinline__ void lds_barrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// Wait until at most n vector-memory operations of this wave are
// outstanding (n a compile-time constant below 64); LDS and export
// counters are left alone.
template <int n>
__device__ __forceinline__ void wait_vmcnt() {
  static_assert(n >= 0 && n < 64, "vmcnt range");
  __builtin_amdgcn_s_waitcnt((n & 15) | ((n >> 4) << 14) | (7 << 4) | (15 << 8));
}

// Tile whose K stages go from HBM straight into LDS with
// global_load_lds_dwordx4 (lane i of a wave writes the 16 bytes it loads
// at the wave's LDS base + 16 i), through a ring of NBUF stage buffers:
// the loads of stage i + NBUF - 1 are issued while stage i is multiplied,
// and nothing passes through registers.  Rows are unpadded; each lane
// picks the global chunk that its LDS slot holds under a per-row chunk
// permutation that keeps fragment reads free of bank conflicts:
//   int8 / FP4 (rows of KSB bytes): slot s of row r holds logical chunk
//     s ^ f(r), the swizzle of mfma_tile;
//   FP6 (rows of 12 chunks per 128 nominal bytes): slot s of row r holds
//     logical chunk (s - (r / 4) mod 4) mod CPR -- rows 4 apart share a
//     bank offset (192-byte stride) and the rotation moves them onto
//     distinct 16-byte bank groups.
// Stages of 128 or 256 nominal bytes (FP6: 128), and tile sides whose
// chunk counts are multiples of 256 (every wave issues the same number of
// loads, which the vmcnt waits count on).  With KG wave groups a ring slot
// holds KG consecutive stages and group g multiplies the slot's stage g;
// the groups' partial sums meet in LDS after the last stage.
template <int KSB, int FMT>
__device__ __forceinline__ int dl_slot(int row, int chunk) {  // LDS slot of a logical chunk
  constexpr int CPR = KSB / 64 * step_bytes<FMT>() / 16;
  if constexpr (FMT == kFmtF6) return (chunk + ((row >> 2) & 3)) % CPR;
  else return lds_chunk<KSB>(row, chunk);
}
template <int KSB, int FMT>
__device__ __forceinline__ int dl_chunk(int row, int slot) {  // logical chunk of a slot
  constexpr int CPR = KSB / 64 * step_bytes<FMT>() / 16;
  if constexpr (FMT == kFmtF6) return (slot + CPR - ((row >> 2) & 3)) % CPR;
  else return lds_chunk<KSB>(row, slot);
}
template <int P, int Q, int WP, int KSB, int NBUF, int FMT = kFmtI8, int KG = 1,
          bool TAILS_APART = true, bool RQ = false>
__device__ __forceinline__ void mfma_tile_dl(const MfmaOperands &a, uint32_t pt, uint32_t qt,
                                             uint32_t kbeg, uint32_t kend, bool atomic,
                                             int8_t *lds, const MfmaRequant &q = MfmaRequant{}) {
  constexpr int NW = 4 * KG;                       // waves per block
  constexpr int WQ = 4 / WP;
  constexpr int PW = P / WP, QW = Q / WQ;
  constexpr int FP = PW / 16, FQ = QW / 16;
  constexpr int S = KSB / 64;
  constexpr int SB = S * step_bytes<FMT>();        // real bytes per row and stage
  constexpr int CPR = SB / 16;
  constexpr int RCH = P * CPR, LCH = Q * CPR;
  constexpr int RLD = RCH / 256, LLD = LCH / 256;  // load instructions per wave and stage
  constexpr int PER = RLD + LLD;
  constexpr int D = NBUF - 1;                      // ring slots in flight ahead
  static_assert(FMT == kFmtF6 ? KSB == 128 : (KSB == 128 || KSB == 256), "stage size");
  static_assert(RCH % 256 == 0 && LCH % 256 == 0, "uneven loads per wave");
  static_assert(NBUF >= 2 && PER * (NBUF - 1) < 64, "ring depth");
  static_assert(FP >= 1 && FQ >= 1 && PW % 16 == 0 && QW % 16 == 0, "bad tile");
  static_assert(KG == 1 || KG == 2, "wave groups");
  constexpr int stage = (P + Q) * SB;
  constexpr int slot = KG * stage;                 // one ring slot: KG stages
  static_assert((size_t)NBUF * slot <= kLdsPerCU, "LDS");

  const uint32_t p0 = pt * P, q0 = qt * Q;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int grp = wave >> 2, wl = wave & 3;        // wave group, wave within the group
  const int wp = wl % WP, wq = wl / WP;
  const int fr = lane & 15;
  const uint32_t nst = kend > kbeg ? (kend - kbeg) / KSB : 0;  // stages
  const uint32_t nsl = (nst + KG - 1) / KG;                    // ring slots of work

  // loads of ring slot `sl` (stages sl*KG ..) into ring buffer `buf`.  Every
  // wave issues RLD + LLD loads; one for a stage past the end re-reads this
  // product's first stage (in bounds) into a part of LDS nobody reads.
  auto issue = [&](uint32_t sl, int buf) {
    int8_t *base = lds + buf * slot;
#pragma unroll
    for (int i = 0; i < RLD; i++) {
      const int c0 = (i * NW + wave) * 64;           // first chunk of this wave's load
      const int g = c0 / RCH, cs = c0 % RCH + lane;  // stage within the slot, chunk slot
      const uint32_t st = sl * KG + g;
      const uint64_t kb = (uint64_t)(kbeg + (st < nst ? st : 0) * KSB) / 64 * step_bytes<FMT>();
      const int row = cs / CPR, lc = dl_chunk<KSB, FMT>(row, cs % CPR);
      __builtin_amdgcn_global_load_lds(
          (const void *)(a.R + (uint64_t)(p0 + row) * a.r_ld + kb + lc * 16),
          (__attribute__((address_space(3))) void *)(base + g * stage + (c0 % RCH) * 16), 16,
          0, 0);
    }
#pragma unroll
    for (int i = 0; i < LLD; i++) {
      const int c0 = (i * NW + wave) * 64;
      const int g = c0 / LCH, cs = c0 % LCH + lane;
      const uint32_t st = sl * KG + g;
      const uint64_t kb = (uint64_t)(kbeg + (st < nst ? st : 0) * KSB) / 64 * step_bytes<FMT>();
      const int row = cs / CPR, lc = dl_chunk<KSB, FMT>(row, cs % CPR);
      __builtin_amdgcn_global_load_lds(
          (const void *)(a.L + (uint64_t)(q0 + row) * a.l_ld + kb + lc * 16),
          (__attribute__((address_space(3))) void *)(base + g * stage + P * SB +
                                                      (c0 % LCH) * 16),
          16, 0, 0);
    }
  };
  // lane's fragment of LDS row `row`, k-step kk of the stage
  auto frag = [&](const int8_t *base, int row, int kk) {
    MfmaFrag<FMT> f;
    const int g = lane >> 4;
    const int8_t *p = base + row * SB;
    if constexpr (FMT == kFmtF6) {
      f.x.s0123 = *reinterpret_cast<const v4i *>(p + dl_slot<KSB, FMT>(row, kk * 6 + g) * 16);
      f.x.s45 = *reinterpret_cast<const v2i *>(
          p + f6_tail_offset<TAILS_APART>(dl_slot<KSB, FMT>(row, kk * 6 + 4 + (g >> 1)) * 16 +
                                           (g & 1) * 8));
      f.x.s67 = v2i{0, 0};
    } else {
      f.h = *reinterpret_cast<const v4i *>(p + dl_slot<KSB, FMT>(row, kk * 4 + g) * 16);
    }
    return f;
  };

  MfmaAcc<FMT> acc[FP][FQ];
#pragma unroll
  for (int i = 0; i < FP; i++)
#pragma unroll
    for (int j = 0; j < FQ; j++) acc[i][j] = MfmaAcc<FMT>{0, 0, 0, 0};

  auto compute = [&](int buf) {
    const int8_t *sR = lds + buf * slot + grp * stage, *sL = sR + P * SB;
#pragma unroll
    for (int kk = 0; kk < S; kk++) {
      MfmaFrag<FMT> af[FP], bf[FQ];
#pragma unroll
      for (int i = 0; i < FP; i++) af[i] = frag(sR, wp * PW + i * 16 + fr, kk);
#pragma unroll
      for (int j = 0; j < FQ; j++) bf[j] = frag(sL, wq * QW + j * 16 + fr, kk);
#pragma unroll
      for (int i = 0; i < FP; i++)
#pragma unroll
        for (int j = 0; j < FQ; j++) acc[i][j] = mfma_op<FMT>(af[i], bf[j], acc[i][j]);
    }
  };

#pragma unroll
  for (int d = 0; d < D; d++)
    if ((uint32_t)d < nsl) issue(d, d);
  for (uint32_t i = 0; i < nsl; i++) {
    // slot i has landed once at most (slots issued after it) x PER loads of
    // this wave are outstanding
    const uint32_t after = nsl - 1 - i < (uint32_t)(D - 1) ? nsl - 1 - i : (uint32_t)(D - 1);
    if constexpr (D >= 3) {
      if (after >= 2) wait_vmcnt<2 * PER>();
      else if (after == 1) wait_vmcnt<PER>();
      else wait_vmcnt<0>();
    } else if constexpr (D == 2) {
      if (after >= 1) wait_vmcnt<PER>();
      else wait_vmcnt<0>();
    } else {
      wait_vmcnt<0>();
    }
    // stage i visible to every wave; every wave is done with stage i - 1,
    // whose ring slot the next loads reuse (its LDS reads fed MFMAs that
    // have issued, so they are complete).  The vmcnt wait above is the
    // only one needed: no fence, which would wait for every load in flight.
    asm volatile("" ::: "memory");
    __builtin_amdgcn_s_barrier();
    asm volatile("" ::: "memory");
    if (i + D < nsl) issue(i + D, (int)((i + D) % NBUF));
    if (KG == 1 || i * KG + grp < nst) compute((int)(i % NBUF));
  }
  // the LDS is reused by the next tile of a persistent block, or by the
  // groups' reduction
  lds_barrier();
  if constexpr (KG > 1) {
    // groups 1.. park their sums in LDS (lane-contiguous: no bank
    // conflicts), group 0 adds them and writes the tile
    typedef typename std::conditional<FMT == kFmtI8, int32_t, float>::type E;
    constexpr int PART = FP * FQ * 4 * 256;  // elements per group
    static_assert((size_t)(KG - 1) * PART * sizeof(E) <= (size_t)NBUF * slot, "reduction");
    E *red = reinterpret_cast<E *>(lds);
    const int at = wl * 64 + lane;
    if (grp > 0) {
      E *mine = red + (grp - 1) * PART;
#pragma unroll
      for (int i = 0; i < FP; i++)
#pragma unroll
        for (int j = 0; j < FQ; j++)
#pragma unroll
          for (int r = 0; r < 4; r++) mine[((i * FQ + j) * 4 + r) * 256 + at] = acc[i][j][r];
    }
    lds_barrier();
    if (grp > 0) return;
#pragma unroll
    for (int g = 0; g < KG - 1; g++)
#pragma unroll
      for (int i = 0; i < FP; i++)
#pragma unroll
        for (int j = 0; j < FQ; j++)
#pragma unroll
          for (int r = 0; r < 4; r++)
            acc[i][j][r] += red[g * PART + ((i * FQ + j) * 4 + r) * 256 + at];
  }

  store_fragments<FP, FQ, RQ>(a, acc, p0 + wp * PW, q0 + wq * QW, atomic, q);
}

// Software-pipelined tile: K in 64-byte stages.  In the iteration that
// multiplies stage i (fragments already in registers), the block writes
// stage i+1 from registers into the other LDS buffer, issues the global
// loads of stage i+2, passes one LDS-only barrier, and reads stage i+1's
// fragments -- those LDS reads and the global loads overlap the MFMAs of
// stage i.
template <int P, int Q, int WP, int FMT = kFmtI8, bool RQ = false>
__device__ __forceinline__ void mfma_tile_pp(const MfmaOperands &a, uint32_t pt, uint32_t qt,
                                             uint32_t kbeg, uint32_t kend, bool atomic,
                                             int8_t *lds, const MfmaRequant &q = MfmaRequant{}) {
  static_assert(FMT != kFmtF6, "the pipelined tile takes 16-byte fragments only");
  constexpr int KSB = 64, LDR = KSB + 16;
  constexpr int WQ = 4 / WP;
  constexpr int PW = P / WP, QW = Q / WQ;
  constexpr int FP = PW / 16, FQ = QW / 16;
  constexpr int CPR = KSB / 16;
  constexpr int RCH = P * CPR, LCH = Q * CPR;
  constexpr int RLD = (RCH + 255) / 256, LLD = (LCH + 255) / 256;
  static_assert(FP >= 1 && FQ >= 1 && PW % 16 == 0 && QW % 16 == 0, "bad tile");
  static_assert((size_t)2 * (P + Q) * LDR <= kLdsPerCU, "LDS");
  int8_t *sR0 = lds, *sL0 = lds + P * LDR;
  constexpr int stage = (P + Q) * LDR;
  const uint32_t p0 = pt * P, q0 = qt * Q;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wp = wave % WP, wq = wave / WP;
  const int fr = lane & 15, fk = (lane >> 4) * 16;

  v4i gr[RLD], gl[LLD];
  auto gload = [&](uint32_t k0) {
#pragma unroll
    for (int i = 0; i < RLD; i++) {
      const int c = tid + 256 * i, row = c / CPR, ch = c % CPR;
      if (RCH % 256 == 0 || c < RCH)
        gr[i] = *reinterpret_cast<const v4i *>(a.R + (uint64_t)(p0 + row) * a.r_ld + k0 + ch * 16);
    }
#pragma unroll
    for (int i = 0; i < LLD; i++) {
      const int c = tid + 256 * i, row = c / CPR, ch = c % CPR;
      if (LCH % 256 == 0 || c < LCH)
        gl[i] = *reinterpret_cast<const v4i *>(a.L + (uint64_t)(q0 + row) * a.l_ld + k0 + ch * 16);
    }
  };
  auto swrite = [&](int buf) {
    int8_t *sR = sR0 + buf * stage, *sL = sL0 + buf * stage;
#pragma unroll
    for (int i = 0; i < RLD; i++) {
      const int c = tid + 256 * i, row = c / CPR, ch = c % CPR;
      if (RCH % 256 == 0 || c < RCH) *reinterpret_cast<v4i *>(&sR[row * LDR + ch * 16]) = gr[i];
    }
#pragma unroll
    for (int i = 0; i < LLD; i++) {
      const int c = tid + 256 * i, row = c / CPR, ch = c % CPR;
      if (LCH % 256 == 0 || c < LCH) *reinterpret_cast<v4i *>(&sL[row * LDR + ch * 16]) = gl[i];
    }
  };
  MfmaFrag<FMT> fa[2][FP], fb[2][FQ];
  auto lread = [&](int buf, MfmaFrag<FMT>(&xa)[FP], MfmaFrag<FMT>(&xb)[FQ]) {
    const int8_t *sR = sR0 + buf * stage, *sL = sL0 + buf * stage;
#pragma unroll
    for (int i = 0; i < FP; i++)
      xa[i].h = *reinterpret_cast<const v4i *>(&sR[(wp * PW + i * 16 + fr) * LDR + fk]);
#pragma unroll
    for (int j = 0; j < FQ; j++)
      xb[j].h = *reinterpret_cast<const v4i *>(&sL[(wq * QW + j * 16 + fr) * LDR + fk]);
  };
  MfmaAcc<FMT> acc[FP][FQ];
#pragma unroll
  for (int i = 0; i < FP; i++)
#pragma unroll
    for (int j = 0; j < FQ; j++) acc[i][j] = MfmaAcc<FMT>{0, 0, 0, 0};
  auto mma = [&](const MfmaFrag<FMT>(&xa)[FP], const MfmaFrag<FMT>(&xb)[FQ]) {
#pragma unroll
    for (int i = 0; i < FP; i++)
#pragma unroll
      for (int j = 0; j < FQ; j++)
        acc[i][j] = mfma_op<FMT>(xa[i], xb[j], acc[i][j]);
  };

  const uint32_t nst = kend > kbeg ? (kend - kbeg) / KSB : 0;
  if (nst) {
    gload(kbeg);
    swrite(0);
    if (nst > 1) gload(kbeg + KSB);
  }
  lds_barrier();
  if (nst) lread(0, fa[0], fb[0]);
  auto step = [&](uint32_t i, auto cur) {
    constexpr int c = decltype(cur)::value, n = c ^ 1;
    if (i + 1 < nst) swrite(n);
    if (i + 2 < nst) gload(kbeg + (i + 2) * KSB);
    lds_barrier();
    if (i + 1 < nst) lread(n, fa[n], fb[n]);
    mma(fa[c], fb[c]);
  };
  for (uint32_t i = 0; i < nst; i += 2) {
    step(i, std::integral_constant<int, 0>());
    if (i + 1 < nst) step(i + 1, std::integral_constant<int, 1>());
  }
  // the LDS is reused by the next tile of a persistent block
  lds_barrier();

  store_fragments<FP, FQ, RQ>(a, acc, p0 + wp * PW, q0 + wq * QW, atomic, q);
}

// Shape run for shape s in format fmt: FP6 has no pipelined tile (it takes
// the staged tile of the same size), no 256 x 256 tile with 128-byte
// stages (its LDS would not fit) and no 256-byte direct-load stages.
__host__ __device__ constexpr int fmt_shape(int s, int fmt) {
  return fmt != kFmtF6 ? s
         : s == 8      ? 6
         : s == 9      ? 7
         : s == 10     ? 1
         : s == 12     ? 11
         : s == 14     ? 18
                       : s;
}

// Shapes the persistent kernel carries: all but the 128 x 256 and 256 x 256
// ones, whose registers or LDS would make the shared kernel spill or drop to
// one block per CU; they map to the 128 x 128 shape of the same pipeline.
// Its 64 x 64 tile is the direct-load one (four stages; three for FP6);
// the other direct-load shapes map to their register-staged counterparts.
// With FP4 copies the 128 x 128 tile with 128-byte stages maps to the
// 64-byte-stage one (the FP4 code path holds more registers); with FP6
// copies (wider LDS rows, more registers) only 32 x 256 and 64 x 64 are
// carried.
__host__ __device__ constexpr int batch_shape(int s, int fmt = 0) {
  return s == 4 || s == 13 || s == 14 || s == 18 || s == 19 || s == 21
             ? (fmt == 2 ? 18 : 13)  // 64 x 64: direct loads, one wave group
         : s == 20          ? batch_shape(1, fmt)
         : s == 15          ? batch_shape(0, fmt)
         : s == 16          ? batch_shape(5, fmt)
         : s == 17          ? batch_shape(1, fmt)
         : fmt == 2         ? (s == 2 || s == 3 ? 2 : 18)
         : fmt == 1 && (s == 0 || s == 7 || s == 11 || s == 12) ? 6
         : s == 7 || s == 11 || s == 12                          ? 0
         : s == 9                                                  ? 8
                                                                   : s;
}

constexpr bool batch_shapes_fit_lds() {
  for (int f = 0; f < kMfmaFormats; f++)
    for (int s = 0; s < kMfmaShapes; s++) {
      const MfmaTileShape t = kMfmaShape[batch_shape(s, f)];
      const int row = f == kFmtF6 ? t.KSB / 64 * 96 + 16 : t.KSB >= 128 ? t.KSB : t.KSB + 16;
      const size_t bytes = direct_loads(t.DEPTH, t.KG)
                               ? (size_t)t.DEPTH * t.KG * (t.P + t.Q) *
                                     (t.KSB / 64 * (f == kFmtF6 ? 96 : 64))
                               : (size_t)2 * (t.P + t.Q) * row;
      if (t.KG != 1) return false;  // the persistent kernel runs 256-thread blocks
      if (bytes > kMfmaLdsMax) return false;
    }
  return true;
}
static_assert(batch_shapes_fit_lds(), "a shape of the persistent kernel overflows its LDS");

// Dispatch one tile by shape id (the persistent kernel's per-product
// switch); RQ: the serving loop's launch, whose whole-K tiles may be
// requantized in the epilogue (MfmaRequant).
template <int FMT, bool RQ = false>
__device__ __forceinline__ void mfma_tile_by_shape(int shape, const MfmaOperands &a,
                                                   uint32_t pt, uint32_t qt, uint32_t kbeg,
                                                   uint32_t kend, bool atomic, int8_t *lds,
                                                   const MfmaRequant &q = MfmaRequant{}) {
  if constexpr (FMT == kFmtF6) {
    if (shape == 2)
      mfma_tile<32, 256, 1, 64, 2, FMT, false, RQ>(a, pt, qt, kbeg, kend, atomic, lds, q);
    else mfma_tile_dl<64, 64, 2, 128, 3, FMT, 1, false, RQ>(a, pt, qt, kbeg, kend, atomic, lds, q);
  } else {
    switch (shape) {
      case 0:
        if constexpr (FMT == kFmtI8)
          mfma_tile<128, 128, 2, 128, 2, FMT, true, RQ>(a, pt, qt, kbeg, kend, atomic, lds, q);
        break;
      case 1: mfma_tile<64, 128, 2, 128, 2, FMT, true, RQ>(a, pt, qt, kbeg, kend, atomic, lds, q); break;
      case 2: mfma_tile<32, 256, 1, 64, 2, FMT, true, RQ>(a, pt, qt, kbeg, kend, atomic, lds, q); break;
      case 3: mfma_tile<16, 256, 1, 128, 2, FMT, true, RQ>(a, pt, qt, kbeg, kend, atomic, lds, q); break;
      case 13: mfma_tile_dl<64, 64, 2, 128, 4, FMT, 1, true, RQ>(a, pt, qt, kbeg, kend, atomic, lds, q); break;
      case 5: mfma_tile<128, 64, 4, 128, 2, FMT, true, RQ>(a, pt, qt, kbeg, kend, atomic, lds, q); break;
      case 6: mfma_tile<128, 128, 2, 64, 1, FMT, true, RQ>(a, pt, qt, kbeg, kend, atomic, lds, q); break;
      case 8: mfma_tile_pp<128, 128, 2, FMT, RQ>(a, pt, qt, kbeg, kend, atomic, lds, q); break;
      default: mfma_tile_pp<64, 128, 2, FMT, RQ>(a, pt, qt, kbeg, kend, atomic, lds, q); break;
    }
  }
}

// Launch geometry of one product: tile shape, K split and tile order.  The
// tiles of a product are numbered split-major; inside a split, the tiles
// that share the larger operand are consecutive (order 0: lhs tiles
// fastest, so consecutive tiles share an rhs tile; order 1: rhs tiles
// fastest), and consecutive tiles land on one XCD, whose L2 then holds that
// operand's rows.
struct MfmaPlan {
  int shape, fmt;
  uint32_t splits, kchunk, ptiles, qtiles, order;
  uint64_t tiles() const { return (uint64_t)ptiles * qtiles * splits; }
};

__host__ __device__ inline void plan_decode(const MfmaPlan &pl, uint64_t t, uint32_t &pt,
                                            uint32_t &qt, uint32_t &sp) {
  const uint64_t per = (uint64_t)pl.ptiles * pl.qtiles;
  sp = (uint32_t)(t / per);
  const uint32_t r = (uint32_t)(t - (uint64_t)sp * per);
  const uint32_t fast = pl.order ? pl.ptiles : pl.qtiles;
  const uint32_t lo = r % fast, hi = r / fast;
  qt = pl.order ? hi : lo;
  pt = pl.order ? lo : hi;
}

// Plan for a product whose copies hold kbytes of K in format fmt (uses the
// tuned table; see mfma_config_for).
MfmaPlan mfma_plan(uint32_t M, uint32_t N, uint32_t kbytes, int fmt, int forced_shape = -1,
                   int forced_splits = 0);

// Operand views of a product's value copies in format fmt.
MfmaOperands mfma_operands(const gbs_matrix *l, const gbs_matrix *r, int fmt, int32_t *d_res);

// Shape class of a product, for the tile autotuner's table: buckets of
// M, N and K.  kClassM x kClassN x kClassK classes.
constexpr int kClassM = 5, kClassN = 6, kClassK = 3;
constexpr int kClasses = kClassM * kClassN * kClassK;
inline int shape_class(uint64_t M, uint64_t N, uint64_t K) {
  const int cm = M < 256 ? 0 : M < 1024 ? 1 : M < 4096 ? 2 : M < 16384 ? 3 : 4;
  const int cn = N <= 16 ? 0 : N <= 32 ? 1 : N <= 64 ? 2 : N <= 256 ? 3 : N <= 2048 ? 4 : 5;
  const int ck = K < 1024 ? 0 : K < 4096 ? 1 : 2;
  return (cm * kClassN + cn) * kClassK + ck;
}

// Batch class of a product inside a persistent launch, for the batch tile
// table: rhs rows <= 16 / <= 32 / <= 64 / <= 128 / <= 256 / more, times
// depth (values) < 1024 / < 4096 / more.
constexpr int kBatchClassN = 6, kBatchClassK = 3, kBatchClasses = kBatchClassN * kBatchClassK;
int batch_class(uint64_t N, uint64_t K);

// Tile shape and K split chosen for a product of depth K (values) in
// format fmt (table set by gbs_tune_load; else a size heuristic).
// splits 0 = decide from the tile count.
void mfma_config_for(uint64_t M, uint64_t N, uint64_t K, int fmt, int *shape, int *splits);

// Persistent launch over at most 16 products given as int8 operand views
// (gbs_batch.hip), each with its requantization (rq[i].vals set) or int32
// destination (rq[i].out): views[i].res are int32 accumulators that
// products split along K add into, zero before the call and left zero (a
// second, small launch converts those products).  *counter is zero at the
// call; the launch zeroes *next, the counter of the following call.
int launch_views(const MfmaOperands *views, const MfmaRequant *rq, int n, unsigned int *counter,
                 unsigned int *next, hipStream_t s);

}  // namespace gbs
