// gbs_planes.hip -- matrix-core GEMM of multi-bit operands read straight
// from their bit-planes: no value copies, no expansion pass.
//
// Same contract as the other engines (reference: gemmBitSerial_generic,
// arch-generic.hpp:331-345, which runs one binary GEMM per plane pair,
// gemmBitSerial_generic_usingBinary :169-193):
// res[n * M + m] = sum_k val(lhs, m, k) * val(rhs, n, k).
//
// The two operands play different roles:
//   plane side  the operand streamed through the chip (more rows, or the
//               one whose values do not fit an FP4 code).  Each wave loads
//               its rows' plane words into registers and turns them, two
//               planes at a time, into FP4 (E2M1) codes of 2-bit chunks of
//               the value just before the matrix instruction takes them:
//               v = sum_c 4^c chunk_c, chunk_c in [0, 3] (the top chunk of
//               a signed value in [-2, 1], or -b for an odd top plane), so
//               a p-plane operand costs ceil(p / 2) instructions instead of
//               p (one binary product per plane) or a copy of its values.
//   code side   the operand whose whole values fit an FP4 code (bipolar,
//               <= 2 bits unsigned, <= 3 bits signed: |v| <= 4) -- or, when
//               neither operand does, the one with fewer rows taken as two
//               2-bit chunks like the plane side (3-4 bits unsigned, 4 bits
//               signed; CP x 2 instructions per k-step).  The block turns
//               its plane words into FP4 codes once per half super-step,
//               into LDS, and every wave reads them from there.
// Both use the block scale of v_mfma_scale_f32_16x16x128_f8f6f4 for the
// chunk weight: FP4 code c (c = 0..4 -> 0, 0.5, 1, 1.5, 2) under scale
// 2^p is c * 2^(p - 1), so chunk c of the plane side (planes 2c, 2c + 1)
// runs under scale 2^(2c) and a code-side chunk cc under 2^(2 cc): every
// product is a quarter of its integer, and the epilogue multiplies the
// sums by 4.  Chunk 0 x chunk 0 (every instruction of a product whose two
// sides are single chunks) thus runs under 2^0 on both sides, i.e. the
// instruction without scales, which issues about 8 % faster.  Every fp32
// partial sum is a quarter of an integer of magnitude below K * maxP *
// maxV, checked against 2^24 on the host, so the sums are exact whatever
// order the hardware adds in.
//
// K order: bit 4j + d of 32-bit word w of a row becomes nibble j of
// register d of the fragment that covers word w, on both sides (any
// bijection works when both ports use the same one).  Lane l of a
// fragment holds row l % 16 and lane group g = l / 16; per 1024-bit
// super-step a lane covers words 8g .. 8g + 7 of its row, and k-step c
// (0..7) of the super-step takes word 8g + c.  Words past K are zero in a
// plane row and those past the row's data are read as zero; a zero bit
// codes 0 except on a bipolar operand (-1), and the engine never pairs two
// bipolar operands (the 1-bit engine does), so padding adds nothing.
//
// Tile: 4 waves; the block owns BP = 64 * WPG plane-side rows (each wave
// 16 * WPG of them) x BV = 16 * VG code-side rows (shared by the waves).
// The rhs operand goes to the MFMA A port so that output lanes run along m,
// the contiguous dimension of res.
//
// The same tile body runs one product per launch (planes_kernel, the
// plane side's encoding fixed at compile time) and many products in one
// persistent launch (planes_batch_kernel: lhs = plane side, its encoding
// from masks at run time, items = code-side tile x plane-side tile x K
// piece, pieces added into zeroed results with integer atomics).
#include <stdlib.h>
#include <string.h>

#include <type_traits>

#include "gbs_internal.h"
#include "gbs_mfma_tile.h"

namespace gbs {
namespace {

// code-side encodings: FP4 codes of whole values
enum : int { kCodeU1 = 0, kCodeBip = 1, kCodeU2 = 2, kCodeS2 = 3, kCodeS3 = 4 };
// code-side encodings of two chunks (values that do not fit one FP4 code):
// chunk 0 = planes 0, 1 (unsigned [0, 3]), chunk 1 = plane 2 (U3: [0, 1]),
// planes 2, 3 (U4: [0, 3]) or planes 2, 3 signed (S4: [-2, 1])
enum : int { kCodeU3 = 5, kCodeU4 = 6, kCodeS4 = 7 };

struct PlanesProb {
  const uint64_t *P, *V;   // plane 0 of the plane side / code side
  uint64_t p_wpr, v_wpr;   // row pitches, 64-bit words
  uint64_t p_wpp, v_wpp;   // words per plane
  uint32_t p_bytes, v_bytes;  // bytes of all planes (< 4 GiB)
  int32_t *res;            // N x M
  uint32_t M, N;           // lhs rows, rhs rows
  uint32_t NP, NV;         // plane-side rows, code-side rows
  uint32_t kw;             // 64-bit words of a row holding K: ceil(K / 64)
  uint32_t nss;            // super-steps of 16 words: ceil(kw / 16)
  uint32_t vtiles;         // code-side tiles of the grid
  int v_enc;               // kCode*
};

constexpr uint32_t kNib = 0x11111111u;


// bits 4j + d of x -> bit 0 of nibble j
template <int D>
__device__ __forceinline__ uint32_t nib(uint32_t x) {
  return (x >> D) & kNib;
}

// dword c of a 16-byte load (c is a constant after unrolling)
__device__ __forceinline__ uint32_t dword_of(const uint4 &v, int c) {
  return c == 0 ? v.x : c == 1 ? v.y : c == 2 ? v.z : v.w;
}

// Register D of an FP4 fragment of the plane side's chunk covering planes
// lo (and hi): the chunk's codes, nibble j for bit 4j + D.
template <int D, int KIND>
__device__ __forceinline__ uint32_t chunk_reg(uint32_t lo, uint32_t hi) {
  // KIND: 0 unsigned 1 plane, 1 signed 1 plane (-b), 2 unsigned 2 planes,
  // 3 signed 2 planes ([-2, 1]), 4 bipolar (+-1)
  if constexpr (KIND == 0) {
    return nib<D>(lo);
  } else if constexpr (KIND == 1) {
    const uint32_t t = nib<D>(lo);
    return t | (t << 3);  // 1001 = -0.5
  } else if constexpr (KIND == 2) {
    return nib<D>(lo) | (nib<D>(hi) << 1);
  } else if constexpr (KIND == 3) {
    // (hi, lo): 00 -> 0000, 01 -> 0001, 10 -> 1010 (-2), 11 -> 1001 (-1)
    const uint32_t m0 = nib<D>(lo), s = nib<D>(hi);
    return m0 | ((s & ~m0) << 1) | (s << 3);
  } else {
    // sign bit 3 of a nibble set where the bit is 0
    return (~(lo << (3 - D)) & 0x88888888u) | kNib;
  }
}

// Register D of the top chunk of a two-chunk code-side value, unsigned
// (sg = 0: the chunk_reg KIND 2 codes) or signed (sg = ~0: KIND 3) from one
// formula, so the kernel carries no branch or select between the two.
template <int D>
__device__ __forceinline__ uint32_t top_chunk_reg(uint32_t lo, uint32_t hi, uint32_t sg) {
  const uint32_t m0 = nib<D>(lo), h = nib<D>(hi), s = h & sg;
  return m0 | (((h & ~sg) | (s & ~m0)) << 1) | (s << 3);
}

template <int KIND>
__device__ __forceinline__ v8i chunk_frag(uint32_t lo, uint32_t hi) {
  v8i f;
  f[0] = (int)chunk_reg<0, KIND>(lo, hi);
  f[1] = (int)chunk_reg<1, KIND>(lo, hi);
  f[2] = (int)chunk_reg<2, KIND>(lo, hi);
  f[3] = (int)chunk_reg<3, KIND>(lo, hi);
  f[4] = f[5] = f[6] = f[7] = 0;
  return f;
}

// Register D of the code side's FP4 fragment: one formula for every code
// encoding, from the words (b0, b1, s) the encoding's masks make of its
// planes (see CodeMasks).  The value is sign-magnitude: m0 = b0,
// m1 = b1 ^ (s & b0), m2 = s & ~(b0 | b1), sign s -- for a signed 3-bit
// value (b0, b1, s = b2) the magnitude of a negative one is 4 - (b0 + 2 b1);
// a signed 2-bit value is the case b1 = s (magnitude 2 - b0); unsigned
// values have s = 0; a bipolar bit x is the case b0 = 1, b1 = s = ~x
// (codes 0001 and 1001); all-zero inputs give code 0.
template <int D>
__device__ __forceinline__ uint32_t code_reg(uint32_t b0, uint32_t b1, uint32_t s) {
  const uint32_t x0 = nib<D>(b0), x1 = nib<D>(b1), sg = nib<D>(s);
  const uint32_t m1 = x1 ^ (sg & x0), m2 = sg & ~(x0 | x1);
  return x0 | (m1 << 1) | (m2 << 2) | (sg << 3);
}

// Word masks that turn a code-side operand's plane words (p0, p1, p2, the
// planes past the operand's repeating its last one) into code_reg's inputs:
// b0 = p0 | or0, b1 = (p1 & and1) ^ xor1, s = (p2 & and2) ^ xor2.
struct CodeMasks { uint32_t or0, and1, xor1, and2, xor2; };
__host__ __device__ inline CodeMasks code_masks(int enc) {
  switch (enc) {
    case kCodeU1: return {0u, 0u, 0u, 0u, 0u};
    case kCodeU2: return {0u, ~0u, 0u, 0u, 0u};
    case kCodeS2: return {0u, ~0u, 0u, ~0u, 0u};    // p2 repeats p1: s = b1
    case kCodeS3: return {0u, ~0u, 0u, ~0u, 0u};
    default: return {~0u, ~0u, ~0u, ~0u, ~0u};      // bipolar: p1 = p2 = p0
  }
}

// Plane-side encoding known only at run time (the batched launch, whose
// products differ): the top chunk's codes come from three words made of its
// low plane word lo and high plane word hi (zero when the chunk has one
// plane) under these masks -- m0 = lo | or0, s = ((hi & sgm) | (lo & sgl))
// ^ xs, t = hi & (tu | ~lo) -- and code = m0 | t << 1 | s << 3 per nibble:
//   one unsigned plane       all zero                  code b
//   one signed plane (-b)    sgl                       1001 (-0.5) for b
//   two unsigned planes      tu                        lo | hi << 1
//   two signed ([-2, 1])     sgm                       chunk_reg KIND 3
//   bipolar (+-1)            or0, sgl, xs              0001 / 1001
// Lower chunks are always two unsigned planes.
struct PlaneMasks { uint32_t or0, sgm, sgl, xs, tu; };

__host__ __device__ inline PlaneMasks plane_masks(uint32_t pp, bool issigned, bool bipolar) {
  if (bipolar) return {~0u, 0u, ~0u, ~0u, 0u};
  if (pp % 2) return {0u, 0u, issigned ? ~0u : 0u, 0u, 0u};
  return {0u, issigned ? ~0u : 0u, 0u, 0u, issigned ? 0u : ~0u};
}

template <int D>
__device__ __forceinline__ uint32_t gen_reg(uint32_t m0, uint32_t t, uint32_t s) {
  return nib<D>(m0) | (nib<D>(t) << 1) | (nib<D>(s) << 3);
}

// One output tile of a product: plane-side rows from pt * BPR, code-side
// rows from vt * BV, full super-steps jb .. je - 1 and, with `with_tail`,
// the partly filled last one (jb = 0, je = kw / 16 and the tail for the
// whole of K); `add` adds the tile's sums into res with integer atomics
// (K pieces of one product) instead of storing them.
//
// PSIGNED with PP == 1: bipolar.  NW waves per block.  CC: chunks of the
// code side (1: whole values as FP4 codes; 2: two 2-bit chunks, each an
// FP4 code under its own block scale, CP x 2 instructions per k-step).
// GEN: the plane side's encoding comes from `pm` and `pp_rt` (PP is then
// 2 CP plane slots; slots at or past pp_rt read zeros, through a buffer
// resource of no records, which reads no memory).
template <bool PLANE_RHS, int PP, bool PSIGNED, int NW, int WPG, int VG, int CC, bool GEN>
__device__ __forceinline__ void planes_tile(const PlanesProb &p, const PlaneMasks &pm,
                                            uint32_t pp_rt, uint32_t pt, uint32_t vt,
                                            uint32_t jb, uint32_t je, bool with_tail, bool add,
                                            uint8_t *lds) {
  constexpr int NT = NW * 64;                  // threads
  constexpr int CP = (PP + 1) / 2;             // chunks of the plane side
  constexpr int BPR = NW * WPG * 16;           // plane-side rows of the block
  constexpr int BV = 16 * VG;                  // code-side rows of the block
  constexpr int TASKS = (BV * 4 + NT - 1) / NT;  // code-side (row, group) tasks per thread
  constexpr bool ALL_TASKS = TASKS * NT == BV * 4;
  constexpr int BUF = 4 * BV * 64;             // bytes of one LDS buffer: 4 k-steps x BV rows
  constexpr bool BIP = !GEN && PSIGNED && PP == 1;
  static_assert(!GEN || PP % 2 == 0, "run-time plane sides take whole chunks of slots");

  const uint32_t v0 = vt * BV, p0 = pt * BPR;
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);  // uniform
  const int g = lane >> 4, fr = lane & 15;

  // Operands are read through buffer resources (the host checks that each
  // matrix is below 2 GiB): a load is a 32-bit per-lane byte offset plus a
  // uniform one, so the loads in flight hold no 64-bit addresses, and a
  // load past the matrix returns zeros.
  const __amdgpu_buffer_rsrc_t rp =
      __builtin_amdgcn_make_buffer_rsrc((void *)p.P, (short)0, (int)p.p_bytes, 0x00020000);
  const __amdgpu_buffer_rsrc_t rv =
      __builtin_amdgcn_make_buffer_rsrc((void *)p.V, (short)0, (int)p.v_bytes, 0x00020000);
  // plane slots of a run-time plane side: absent planes through a resource
  // of no records (every load returns zero)
  __amdgpu_buffer_rsrc_t rps[PP];
#pragma unroll
  for (int q = 0; q < PP; q++)
    rps[q] = !GEN || (uint32_t)q < pp_rt
                 ? rp
                 : __builtin_amdgcn_make_buffer_rsrc((void *)p.P, (short)0, 0, 0x00020000);
  // plane-side rows of this lane (rows past the operand read its last row,
  // in bounds; their outputs are never stored): byte offset of the row
  uint32_t prow[WPG];
#pragma unroll
  for (int rg = 0; rg < WPG; rg++) {
    const uint32_t r = min(p0 + (uint32_t)((wave * WPG + rg) * 16 + fr), p.NP - 1);
    prow[rg] = r * (uint32_t)p.p_wpr * 8u;
  }
  // code-side tasks: task u = tid + NT i -> (row group u / 64, lane group
  // (u / 16) % 4, row u % 16)
  uint32_t vrow[TASKS];   // byte offset of the task's row
  int vunit[TASKS];       // LDS unit of the task's k-step 0: (v * 4 + g) * 16 + fr
  int vgrp[TASKS];        // the task's lane group
  bool vtask[TASKS];
#pragma unroll
  for (int i = 0; i < TASKS; i++) {
    const int u = tid + NT * i;
    vtask[i] = ALL_TASKS || wave * 64 + NT * i < BV * 4;  // whole waves: uniform
    const int vv = u / 64, vg = (u / 16) & 3, vf = u & 15;
    vunit[i] = (vv * 4 + vg) * 16 + vf;
    vgrp[i] = vg;
    const uint32_t r = min(v0 + (uint32_t)(16 * vv + vf), p.NV - 1);
    vrow[i] = r * (uint32_t)p.v_wpr * 8u;
  }
  const int npv = p.v_enc == kCodeU4 || p.v_enc == kCodeS4 ? 4
                  : p.v_enc == kCodeS3 || p.v_enc == kCodeU3 ? 3
                  : (p.v_enc == kCodeU2 || p.v_enc == kCodeS2) ? 2 : 1;
  const CodeMasks cm = code_masks(p.v_enc);
  // two-chunk code sides: plane 3 is absent for U3, the sign plane for S4
  const uint32_t hi_mask = p.v_enc == kCodeU3 ? 0u : ~0u;
  const uint32_t sg_mask = p.v_enc == kCodeS4 ? ~0u : 0u;

  v4f acc[WPG][VG];
#pragma unroll
  for (int i = 0; i < WPG; i++)
#pragma unroll
    for (int j = 0; j < VG; j++) acc[i][j] = v4f{0.f, 0.f, 0.f, 0.f};

  uint4 xp[WPG][PP][2];     // plane words: [row group][plane][half]
  constexpr int NPV = CC == 2 ? 4 : 3;  // code-side planes loaded
  uint4 xv[TASKS][NPV];     // code-side plane words of the next half

  // K is walked in stages of four k-steps (half super-steps).  In a full
  // super-step lane group g covers dwords 8g .. 8g + 7 and k-step 4H + c
  // takes dword 8g + 4H + c.  The partly filled last super-step (tail_dw of
  // its 32 dwords hold data) is walked as two compact halves instead: in
  // half H lane group g covers dwords 16H + 4g .. + 3 and k-step c takes
  // dword 16H + 4g + c, so only the k-steps with data are multiplied
  // (min(4, tail_dw - 16H) of them).  Both forms load 16 contiguous bytes
  // per lane and plane; past a row's data they read the next row's words
  // (or, past the matrix, the buffer's zeros), and the code side's codes
  // for those positions are zeroed -- FP4 codes have no infinities, so
  // 0 * x = 0 whatever the plane side holds there.
  const uint32_t nfull = p.kw / 16;
  const uint32_t tail_dw = 2 * p.kw - 32 * nfull;

  auto ld128 = [&](const __amdgpu_buffer_rsrc_t &rs, uint32_t voff, uint32_t soff)
      __attribute__((always_inline)) -> uint4 {
    const auto v = __builtin_amdgcn_raw_buffer_load_b128(rs, voff, soff, 0);
    return make_uint4(v[0], v[1], v[2], v[3]);
  };
  // byte offset of stage (j, H) within a row: uniform part, and the lane
  // group's part (32 g in full super-steps, 16 g in the compact tail)
  auto stage_soff = [&](uint32_t j, int H) __attribute__((always_inline)) -> uint32_t {
    return 128u * j + (j < nfull ? 16u : 64u) * (uint32_t)H;
  };
  auto lane_off = [&](uint32_t j, int grp) __attribute__((always_inline)) -> uint32_t {
    return (uint32_t)grp << (j < nfull ? 5 : 4);
  };
  auto load_plane = [&](uint32_t j, auto Hc) __attribute__((always_inline)) {
    constexpr int H = decltype(Hc)::value;
    const uint32_t lo = lane_off(j, g), so = stage_soff(j, H);
#pragma unroll
    for (int rg = 0; rg < WPG; rg++)
#pragma unroll
      for (int q = 0; q < PP; q++)
        xp[rg][q][H] = ld128(rps[q], prow[rg] + lo, (uint32_t)q * (uint32_t)p.p_wpp * 8u + so);
  };
  // (code-side planes past the operand's repeat its last plane, so every
  // load is unconditional and no load sits in a branch)
  auto load_code = [&](uint32_t j, auto Hc) __attribute__((always_inline)) {
    constexpr int H = decltype(Hc)::value;
    const uint32_t so = stage_soff(j, H);
#pragma unroll
    for (int u = 0; u < TASKS; u++)
#pragma unroll
      for (int q = 0; q < NPV; q++)
        xv[u][q] = ld128(rv, vrow[u] + lane_off(j, vgrp[u]),
                         (uint32_t)min(q, npv - 1) * (uint32_t)p.v_wpp * 8u + so);
  };
  // codes of the loaded half, stage (j, B), into LDS buffer B: unit
  // (k-step c, row group v, lane group g, row fr) at ((c * VG + v) * 4 + g)
  // * 16 + fr, so a wave's fragment read is 1 KB of consecutive 16-byte
  // units; in the compact tail, positions past the data get code 0
  // (k-step c only: the caller spreads the four over the matrix work)
  auto store_codes_c = [&](auto Bc, uint32_t j, int c) __attribute__((always_inline)) {
    constexpr int B = decltype(Bc)::value;
    const bool tail = j == nfull;
#pragma unroll
    for (int u = 0; u < TASKS; u++) {
      if (!vtask[u]) continue;  // uniform: whole waves
      const uint32_t valid =
          (!tail || 16u * B + 4u * (uint32_t)vgrp[u] + c < tail_dw) ? ~0u : 0u;
      if constexpr (CC == 1) {
        const uint32_t b0 = (dword_of(xv[u][0], c) | cm.or0) & valid;
        const uint32_t b1 = ((dword_of(xv[u][1], c) & cm.and1) ^ cm.xor1) & valid;
        const uint32_t sg = ((dword_of(xv[u][2], c) & cm.and2) ^ cm.xor2) & valid;
        const uint4 codes = make_uint4(code_reg<0>(b0, b1, sg), code_reg<1>(b0, b1, sg),
                                       code_reg<2>(b0, b1, sg), code_reg<3>(b0, b1, sg));
        *reinterpret_cast<uint4 *>(lds + B * BUF + (c * VG * 64 + vunit[u]) * 16) = codes;
      } else {
        // chunk 0: planes 0, 1; chunk 1: plane 2 (U3), planes 2, 3 (U4), or
        // planes 2, 3 signed (S4); the same chunk codes as the plane side
        const uint32_t w0 = dword_of(xv[u][0], c) & valid, w1 = dword_of(xv[u][1], c) & valid;
        const uint32_t w2 = dword_of(xv[u][2], c) & valid;
        const uint32_t w3 = dword_of(xv[u][3], c) & valid & hi_mask;
        const uint4 c0 = make_uint4(chunk_reg<0, 2>(w0, w1), chunk_reg<1, 2>(w0, w1),
                                    chunk_reg<2, 2>(w0, w1), chunk_reg<3, 2>(w0, w1));
        const uint4 c1 = make_uint4(top_chunk_reg<0>(w2, w3, sg_mask), top_chunk_reg<1>(w2, w3, sg_mask),
                                    top_chunk_reg<2>(w2, w3, sg_mask), top_chunk_reg<3>(w2, w3, sg_mask));
        const int at = (c * VG * 64 + vunit[u]) * 16;
        *reinterpret_cast<uint4 *>(lds + (B * 2 + 0) * BUF + at) = c0;
        *reinterpret_cast<uint4 *>(lds + (B * 2 + 1) * BUF + at) = c1;
      }
    }
  };

  auto store_codes = [&](auto Bc, uint32_t j) __attribute__((always_inline)) {
#pragma unroll
    for (int c = 0; c < 4; c++) store_codes_c(Bc, j, c);
  };

  // one k-step c of the half in plane registers [H] x LDS buffer H
  auto kstep = [&](auto Hc, int c) __attribute__((always_inline)) {
    constexpr int H = decltype(Hc)::value;
    v8i pf[WPG][CP];
#pragma unroll
    for (int rg = 0; rg < WPG; rg++) {
#pragma unroll
      for (int ch = 0; ch < CP; ch++) {
        const int lo = 2 * ch, hi = 2 * ch + 1;
        const uint32_t xl = dword_of(xp[rg][lo][H], c);
        const uint32_t xh = hi < PP ? dword_of(xp[rg][hi < PP ? hi : lo][H], c) : 0u;
        const bool top = ch == CP - 1;
        if (GEN && top) {
          const uint32_t m0 = xl | pm.or0, s = ((xh & pm.sgm) | (xl & pm.sgl)) ^ pm.xs;
          const uint32_t t = xh & (pm.tu | ~xl);
          pf[rg][ch] = v8i{(int)gen_reg<0>(m0, t, s), (int)gen_reg<1>(m0, t, s),
                           (int)gen_reg<2>(m0, t, s), (int)gen_reg<3>(m0, t, s), 0, 0, 0, 0};
        } else if (BIP) pf[rg][ch] = chunk_frag<4>(xl, 0u);
        else if (hi < PP)
          pf[rg][ch] = (top && PSIGNED) ? chunk_frag<3>(xl, xh) : chunk_frag<2>(xl, xh);
        else
          pf[rg][ch] = (top && PSIGNED) ? chunk_frag<1>(xl, 0u) : chunk_frag<0>(xl, 0u);
      }
    }
#pragma unroll
    for (int cc = 0; cc < CC; cc++) {
#pragma unroll
      for (int v = 0; v < VG; v++) {
        const uint4 cw = *reinterpret_cast<const uint4 *>(
            lds + (H * CC + cc) * BUF + (((c * VG + v) * 4 + g) * 16 + fr) * 16);
        const v8i cf = {(int)cw.x, (int)cw.y, (int)cw.z, (int)cw.w, 0, 0, 0, 0};
        const int cs = 127 + 2 * cc;  // code chunk cc: 2^(2 cc)
#pragma unroll
        for (int rg = 0; rg < WPG; rg++)
#pragma unroll
          for (int ch = 0; ch < CP; ch++) {
            const int ps = 127 + 2 * ch;  // 2^(2 ch)
            // both scales 2^0 (chunk 0 x chunk 0): the instruction without
            // scales (the builtin with zero scale operands), which issues
            // faster and multiplies the same
            if (ps == 127 && cs == 127) {
              if constexpr (PLANE_RHS)
                acc[rg][v] = __builtin_amdgcn_mfma_scale_f32_16x16x128_f8f6f4(
                    pf[rg][ch], cf, acc[rg][v], 4, 4, 0, 0, 0, 0);
              else
                acc[rg][v] = __builtin_amdgcn_mfma_scale_f32_16x16x128_f8f6f4(
                    cf, pf[rg][ch], acc[rg][v], 4, 4, 0, 0, 0, 0);
            } else if constexpr (PLANE_RHS) {
              acc[rg][v] = __builtin_amdgcn_mfma_scale_f32_16x16x128_f8f6f4(
                  pf[rg][ch], cf, acc[rg][v], 4, 4, 0, ps, 0, cs);
            } else {
              acc[rg][v] = __builtin_amdgcn_mfma_scale_f32_16x16x128_f8f6f4(
                  cf, pf[rg][ch], acc[rg][v], 4, 4, 0, cs, 0, ps);
            }
          }
      }
    }
  };

  using H0 = std::integral_constant<int, 0>;
  using H1 = std::integral_constant<int, 1>;
  // Stage (j, H) multiplies plane registers [H] with LDS buffer H.
  // Prologue: plane words of both halves of iteration 0, codes of its half
  // 0 in buffer 0, code words of its half 1 in registers.
  load_plane(jb, H0{});
  load_plane(jb, H1{});
  load_code(jb, H0{});
  store_codes(H0{}, jb);
  load_code(jb, H1{});
  __syncthreads();
  // Full super-steps.  Within a half, the codes of the next half go to the
  // other buffer (its last readers passed the barrier ending the previous
  // half) in the same basic block as this half's matrix instructions, so
  // the scheduler can interleave their VALU work with them; then the plane
  // registers just spent take the next iteration's words and the code
  // registers the words of the half after next.  The last iteration's
  // prefetch reads the tail (same form) or, without one, words no one uses
  // (in-bounds rows, or the buffer's zeros past the matrix).
  for (uint32_t j = jb; j < je; j++) {
#pragma unroll
    for (int c = 0; c < 4; c++) {
      kstep(H0{}, c);
      store_codes_c(H1{}, j, c);
      __builtin_amdgcn_sched_barrier(0);
    }
    load_plane(j + 1, H0{});
    load_code(j + 1, H0{});
    __syncthreads();
#pragma unroll
    for (int c = 0; c < 4; c++) {
      kstep(H1{}, c);
      store_codes_c(H0{}, j + 1, c);
      __builtin_amdgcn_sched_barrier(0);
    }
    load_plane(j + 1, H1{});
    load_code(j + 1, H1{});
    __syncthreads();
  }
  // the partly filled last super-step: only its k-steps with data
  if (with_tail && tail_dw) {
    const uint32_t nk0 = min(tail_dw, 4u);
#pragma unroll
    for (int c = 0; c < 4; c++)
      if ((uint32_t)c < nk0) kstep(H0{}, c);
    if (tail_dw > 16) {
      store_codes(H1{}, nfull);
      __syncthreads();
      const uint32_t nk1 = min(tail_dw - 16, 4u);
#pragma unroll
      for (int c = 0; c < 4; c++)
        if ((uint32_t)c < nk1) kstep(H1{}, c);
    }
  }

#pragma unroll
  for (int rg = 0; rg < WPG; rg++)
#pragma unroll
    for (int v = 0; v < VG; v++)
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const uint32_t prow_i = p0 + (uint32_t)((wave * WPG + rg) * 16);
        const uint32_t n = PLANE_RHS ? prow_i + 4 * g + r : v0 + 16 * v + 4 * g + r;
        const uint32_t m = PLANE_RHS ? v0 + 16 * v + fr : prow_i + fr;
        if (n < p.N && m < p.M) {
          int32_t *o = p.res + (uint64_t)n * p.M + m;
          const int32_t val = (int32_t)(acc[rg][v][r] * 4.0f);  // quarters (see kstep)
          if (add) atomicAdd(o, val);
          else *o = val;
        }
      }
}

template <bool PLANE_RHS, int PP, bool PSIGNED, int NW, int WPG, int VG, int CC = 1>
__global__ __launch_bounds__(NW * 64, 8 / NW) void planes_kernel(PlanesProb p) {
  __shared__ __attribute__((aligned(16))) uint8_t lds[2 * CC * 4 * 16 * VG * 64];  // [half][chunk]
  const uint32_t T = gridDim.x, b = blockIdx.x;
  // consecutive logical tiles (the code-side tiles of one plane-side tile)
  // run on one XCD (blocks are dealt round-robin over the 8 XCDs)
  const uint32_t t = (T % 8 == 0) ? (b % 8) * (T / 8) + b / 8 : b;
  planes_tile<PLANE_RHS, PP, PSIGNED, NW, WPG, VG, CC, false>(
      p, PlaneMasks{}, PP, t / p.vtiles, t % p.vtiles, 0, p.kw / 16, true, false, lds);
}

// Tile configurations: waves per block, plane-side row groups per wave,
// code-side row groups per block.  Two-chunk code sides take tile 1 only
// (its LDS holds both chunks; the 8-wave forms spill registers).  An
// 8-wave block of 256 code-side rows spills at every plane count and is
// not carried.
struct PlanesTile { int NW, WPG, VG; };
constexpr int kPlanesTiles = 4;
constexpr PlanesTile kPlanesTile[kPlanesTiles] = {{4, 2, 8}, {4, 2, 4}, {4, 2, 2}, {8, 2, 8}};

template <bool PR, int PP, bool PS>
void run_pp2(const PlanesProb &p, uint32_t grid, hipStream_t s) {
  hipLaunchKernelGGL((planes_kernel<PR, PP, PS, 4, 2, 4, 2>), dim3(grid), dim3(256), 0, s, p);
}

template <bool PR, int PP, bool PS>
void run_pp(int tile, const PlanesProb &p, uint32_t grid, hipStream_t s) {
  switch (tile) {
    case 0:
      hipLaunchKernelGGL((planes_kernel<PR, PP, PS, 4, 2, 8>), dim3(grid), dim3(256), 0, s, p);
      break;
    case 1:
      hipLaunchKernelGGL((planes_kernel<PR, PP, PS, 4, 2, 4>), dim3(grid), dim3(256), 0, s, p);
      break;
    case 2:
      hipLaunchKernelGGL((planes_kernel<PR, PP, PS, 4, 2, 2>), dim3(grid), dim3(256), 0, s, p);
      break;
    default:
      hipLaunchKernelGGL((planes_kernel<PR, PP, PS, 8, 2, 8>), dim3(grid), dim3(512), 0, s, p);
      break;
  }
}

template <bool PR>
void run_planes(int pp, bool ps, int cc, int tile, const PlanesProb &p, uint32_t grid,
                hipStream_t s) {
  if (cc == 2) {
    switch (pp * 2 + (ps ? 1 : 0)) {
      case 2: run_pp2<PR, 1, false>(p, grid, s); break;
      case 3: run_pp2<PR, 1, true>(p, grid, s); break;
      case 4: run_pp2<PR, 2, false>(p, grid, s); break;
      case 5: run_pp2<PR, 2, true>(p, grid, s); break;
      case 6: run_pp2<PR, 3, false>(p, grid, s); break;
      case 7: run_pp2<PR, 3, true>(p, grid, s); break;
      case 8: run_pp2<PR, 4, false>(p, grid, s); break;
      default: run_pp2<PR, 4, true>(p, grid, s); break;
    }
    return;
  }
  switch (pp * 2 + (ps ? 1 : 0)) {
    case 2: run_pp<PR, 1, false>(tile, p, grid, s); break;
    case 3: run_pp<PR, 1, true>(tile, p, grid, s); break;
    case 4: run_pp<PR, 2, false>(tile, p, grid, s); break;
    case 5: run_pp<PR, 2, true>(tile, p, grid, s); break;
    case 6: run_pp<PR, 3, false>(tile, p, grid, s); break;
    case 7: run_pp<PR, 3, true>(tile, p, grid, s); break;
    case 8: run_pp<PR, 4, false>(tile, p, grid, s); break;
    default: run_pp<PR, 4, true>(tile, p, grid, s); break;
  }
}

// FP4 code encoding of a code-side operand, -1 when its values do not fit
int code_enc(const Operand &o) {
  if (o.bipolar) return kCodeBip;
  if (!o.issigned) return o.nbits == 1 ? kCodeU1 : o.nbits == 2 ? kCodeU2 : -1;
  return o.nbits == 2 ? kCodeS2 : o.nbits == 3 ? kCodeS3 : -1;
}

// Two-chunk encoding of a code-side operand that does not fit one FP4
// code, -1 when it has more than 4 planes (or is bipolar / fits one code)
int code_enc2(const Operand &o) {
  if (o.bipolar) return -1;
  if (!o.issigned) return o.nbits == 3 ? kCodeU3 : o.nbits == 4 ? kCodeU4 : -1;
  return o.nbits == 4 ? kCodeS4 : -1;
}

// Largest magnitude a code-side value's codes add up to (for two chunks,
// the sum of |chunk maxima| x chunk weight: every partial sum is bounded by
// K times it times the plane side's)
double code_max(int enc) {
  return enc == kCodeU2 ? 3 : enc == kCodeS2 ? 2 : enc == kCodeS3 ? 4
         : enc == kCodeU3 ? 7 : enc == kCodeU4 ? 15 : enc == kCodeS4 ? 11 : 1;
}

// Largest sum of |chunk maxima| * chunk weight of a plane-side operand:
// bounds every partial sum together with K and the code side's maximum.
double plane_max(const Operand &o) {
  if (o.bipolar) return 1;
  const int pp = (int)o.nbits;
  if (!o.issigned || pp % 2 == 1) return (double)((1u << pp) - 1);
  const int cp = (pp + 1) / 2;
  return 3.0 * (double)(1u << (2 * (cp - 1))) - 1.0;
}

struct PlanesPlan {
  bool ok;
  bool plane_rhs;      // rhs is the plane side
  int enc, pp, cc, tile;
  double bound;        // K * maxP * maxV
};

PlanesPlan planes_plan(const Operand &l, const Operand &r, uint64_t K) {
  PlanesPlan pl{};
  pl.ok = false;
  if (l.bipolar && r.bipolar) return pl;  // the 1-bit engine's product
  int el = code_enc(l), er = code_enc(r);
  // code side: the operand that fits FP4 codes; of two that do, the one
  // with fewer rows (the plane side is the streamed one); else an operand
  // of 3-4 bits as two chunks, again the one with fewer rows
  pl.cc = 1;
  if (el < 0 && er < 0) {
    el = code_enc2(l);
    er = code_enc2(r);
    pl.cc = 2;
  }
  bool plane_rhs;
  if (el >= 0 && er >= 0) plane_rhs = r.nrows >= l.nrows;
  else if (el >= 0) plane_rhs = true;
  else if (er >= 0) plane_rhs = false;
  else return pl;
  const Operand &ps = plane_rhs ? r : l, &cs = plane_rhs ? l : r;
  if (ps.nbits > 4) return pl;
  pl.plane_rhs = plane_rhs;
  pl.enc = plane_rhs ? el : er;
  pl.pp = (int)ps.nbits;
  pl.bound = (double)K * plane_max(ps) * code_max(pl.enc);
  // code-side rows per block: 128 when there are that many, fewer for
  // narrow code sides (each plane-side fragment then feeds fewer
  // instructions, so more plane rows per wave); 8-wave blocks once the
  // plane side fills several hundred of them.  GBS_PLANES_TILE (an index
  // into kPlanesTile) forces a tile, for tuning.
  const uint64_t nv = cs.nrows;
  pl.tile = nv > 64 ? (ps.nrows >= 32768 ? 3 : 0) : nv > 32 ? 1 : 2;
  if (const char *e = getenv("GBS_PLANES_TILE")) pl.tile = atoi(e) % kPlanesTiles;
  if (pl.cc == 2) pl.tile = 1;
  pl.ok = pl.bound < 16777216.0;
  return pl;
}

// One product of the batched launch: lhs = plane side (its encoding in
// masks), rhs = code side; work item i of the product (i from tile_begin)
// is code-side tile i % vtiles, plane-side tile (i / vtiles) % ptiles, K
// piece i / (vtiles ptiles) of jper full super-steps (the last piece also
// takes the partly filled super-step).  variant: (CP - 1) + 2 (CC - 1)
// + 4 (VG == 4).
struct PlanesBatchProb {
  PlanesProb p;
  PlaneMasks pm;
  uint32_t pp, ptiles, splits, jper, variant;
  uint64_t tile_begin;
};
constexpr int kPlanesBatchLds = 2 * 2 * 4 * 64 * 64;  // CC 2, VG 4: 64 KiB

template <int CP, int CC, int VG>
__device__ __forceinline__ void planes_batch_item(const PlanesBatchProb &d, uint32_t pt,
                                                  uint32_t vt, uint32_t jb, uint32_t je,
                                                  bool last, bool add, uint8_t *lds) {
  planes_tile<false, 2 * CP, false, 4, 2, VG, CC, true>(d.p, d.pm, d.pp, pt, vt, jb, je, last,
                                                        add, lds);
}

// Persistent launch over the products' work items: a block takes the next
// item from one queue counter, finds its product in the per-item table and
// runs that product's tile form (a uniform branch over the eight forms);
// every block leaves once the counter passes the last item.
__global__ __launch_bounds__(256, 2) void planes_batch_kernel(const PlanesBatchProb *probs,
                                                              uint32_t total,
                                                              unsigned int *counter,
                                                              const uint32_t *item_prod) {
  __shared__ __attribute__((aligned(16))) uint8_t lds[kPlanesBatchLds];
  __shared__ uint32_t s_item;
  for (;;) {
    if (threadIdx.x == 0) s_item = atomicAdd(counter, 1u);
    __syncthreads();
    const uint32_t item = __builtin_amdgcn_readfirstlane(s_item);
    __syncthreads();
    if (item >= total) break;
    const PlanesBatchProb &d = probs[item_prod[item]];
    const uint32_t local = item - (uint32_t)d.tile_begin;
    const uint32_t vt = local % d.p.vtiles, rest = local / d.p.vtiles;
    const uint32_t pt = rest % d.ptiles, sp = rest / d.ptiles;
    const uint32_t nfull = d.p.kw / 16, jb = sp * d.jper, je = min(nfull, jb + d.jper);
    const bool last = sp + 1 == d.splits, add = d.splits > 1;
    switch (d.variant) {
      case 0: planes_batch_item<1, 1, 2>(d, pt, vt, jb, je, last, add, lds); break;
      case 1: planes_batch_item<2, 1, 2>(d, pt, vt, jb, je, last, add, lds); break;
      case 2: planes_batch_item<1, 2, 2>(d, pt, vt, jb, je, last, add, lds); break;
      case 3: planes_batch_item<2, 2, 2>(d, pt, vt, jb, je, last, add, lds); break;
      case 4: planes_batch_item<1, 1, 4>(d, pt, vt, jb, je, last, add, lds); break;
      case 5: planes_batch_item<2, 1, 4>(d, pt, vt, jb, je, last, add, lds); break;
      case 6: planes_batch_item<1, 2, 4>(d, pt, vt, jb, je, last, add, lds); break;
      default: planes_batch_item<2, 2, 4>(d, pt, vt, jb, je, last, add, lds); break;
    }
    __syncthreads();
  }
}

}  // namespace

bool planes_supported(const Operand &l, const Operand &r, uint64_t K) {
  return planes_plan(l, r, K).ok && K > 0 && l.nrows > 0 && r.nrows > 0;
}

int launch_planes_gemm(const Operand &l, const Operand &r, uint64_t K, int32_t *d_res,
                       hipStream_t s) {
  const uint64_t M = l.nrows, N = r.nrows;
  if (M == 0 || N == 0) return GBS_OK;
  if (K == 0) return launch_fill_i32(d_res, M * N, 0, s);
  const PlanesPlan pl = planes_plan(l, r, K);
  if (!pl.ok)
    return fail(GBS_ENOTSUP,
                "gemm: the planes matrix-core engine cannot hold this product exactly "
                "(bitwidths %u x %u, depth %llu)", l.nbits, r.nbits, (unsigned long long)K);
  if (M > 0x7fffffffull || N > 0x7fffffffull)
    return fail(GBS_ERANGE, "gemm dimensions exceed 2^31");
  const Operand &ps = pl.plane_rhs ? r : l, &cs = pl.plane_rhs ? l : r;
  PlanesProb p{};
  p.P = ps.planes;
  p.V = cs.planes;
  p.p_wpr = ps.wpr;
  p.v_wpr = cs.wpr;
  p.p_wpp = ps.wpp;
  p.v_wpp = cs.wpp;
  const uint64_t pb = (uint64_t)ps.nbits * ps.wpp * 8, vb = (uint64_t)cs.nbits * cs.wpp * 8;
  if (pb >= (1ull << 31) || vb >= (1ull << 31))
    return fail(GBS_ENOTSUP, "gemm: the planes engine takes operands below 2 GiB of planes");
  p.p_bytes = (uint32_t)pb;
  p.v_bytes = (uint32_t)vb;
  p.res = d_res;
  p.M = (uint32_t)M;
  p.N = (uint32_t)N;
  p.NP = (uint32_t)ps.nrows;
  p.NV = (uint32_t)cs.nrows;
  p.kw = (uint32_t)((K + 63) / 64);
  p.nss = (p.kw + 15) / 16;
  if (ps.wpr < p.kw || cs.wpr < p.kw) return fail(GBS_EINVAL, "gemm: plane rows shorter than K");
  p.v_enc = pl.enc;
  const PlanesTile tl = kPlanesTile[pl.tile];
  const uint64_t bp = 16ull * tl.NW * tl.WPG, bv = 16ull * tl.VG;
  p.vtiles = (uint32_t)((p.NV + bv - 1) / bv);
  const uint64_t tiles = (uint64_t)p.vtiles * ((p.NP + bp - 1) / bp);
  if (tiles > 0x7fffffffull) return fail(GBS_ERANGE, "gemm: too many tiles");
  const bool psigned = ps.issigned;  // with one plane: bipolar
  if (pl.plane_rhs) run_planes<true>(pl.pp, psigned, pl.cc, pl.tile, p, (uint32_t)tiles, s);
  else run_planes<false>(pl.pp, psigned, pl.cc, pl.tile, p, (uint32_t)tiles, s);
  GBS_HIP(hipGetLastError());
  return GBS_OK;
}

PlanesBatchPlan planes_batch_plan(const Operand &l, const Operand &r, uint64_t K,
                                  uint32_t min_pieces) {
  PlanesBatchPlan pl{};
  pl.ok = false;
  if (K == 0 || l.nrows == 0 || r.nrows == 0 || l.nbits > 4) return pl;
  if (l.bipolar && r.bipolar) return pl;  // padding bits would pair -1 with -1
  if (l.nrows > 0x7fffffffull || r.nrows > 0x7fffffffull) return pl;
  if ((uint64_t)l.nbits * l.wpp * 8 >= (1ull << 31) || (uint64_t)r.nbits * r.wpp * 8 >= (1ull << 31))
    return pl;
  int enc = code_enc(r), cc = 1;
  if (enc < 0) {
    enc = code_enc2(r);
    cc = 2;
  }
  if (enc < 0) return pl;
  if ((double)K * plane_max(l) * code_max(enc) >= 16777216.0) return pl;
  // 32 code-side rows per tile for narrow rhs operands, else 64
  const uint32_t vg = r.nrows <= 32 ? 2 : 4;
  const uint32_t cp = (l.nbits + 1) / 2;
  pl.variant = (cp - 1) + 2 * (uint32_t)(cc - 1) + (vg == 4 ? 4u : 0u);
  pl.vtiles = (uint32_t)((r.nrows + 16 * vg - 1) / (16 * vg));
  pl.ptiles = (uint32_t)((l.nrows + 127) / 128);
  // K pieces of at most 8 super-steps (8192 values, as the copy formats'
  // 4 KiB pieces of FP4 codes), more when asked (at least 2 super-steps)
  const uint32_t nfull = (uint32_t)((K + 63) / 64 / 16);
  uint32_t want = nfull ? (nfull + 7) / 8 : 1;
  const uint32_t most = nfull / 2 ? nfull / 2 : 1;
  if (want < min_pieces) want = min_pieces < most ? min_pieces : most;
  pl.jper = nfull ? (nfull + want - 1) / want : 0;
  pl.splits = nfull ? (nfull + pl.jper - 1) / pl.jper : 1;
  const double kvals = nfull ? (double)pl.jper * 1024 : (double)K;
  const uint32_t code_planes = cc == 2 ? 4 : 3;  // loads per row, repeats included
  pl.item_bytes = (128.0 * l.nbits + 16.0 * vg * code_planes) * kvals / 8;
  pl.ok = true;
  return pl;
}

size_t planes_batch_desc_size() { return sizeof(PlanesBatchProb); }

void planes_batch_desc(const Operand &l, const Operand &r, uint64_t K, int32_t *d_res,
                       const PlanesBatchPlan &pl, uint64_t tile_begin, void *out) {
  PlanesBatchProb d{};
  PlanesProb &p = d.p;
  p.P = l.planes;
  p.V = r.planes;
  p.p_wpr = l.wpr;
  p.v_wpr = r.wpr;
  p.p_wpp = l.wpp;
  p.v_wpp = r.wpp;
  p.p_bytes = (uint32_t)((uint64_t)l.nbits * l.wpp * 8);
  p.v_bytes = (uint32_t)((uint64_t)r.nbits * r.wpp * 8);
  p.res = d_res;
  p.M = p.NP = (uint32_t)l.nrows;
  p.N = p.NV = (uint32_t)r.nrows;
  p.kw = (uint32_t)((K + 63) / 64);
  p.nss = (p.kw + 15) / 16;
  p.vtiles = pl.vtiles;
  const int enc1 = code_enc(r);
  p.v_enc = enc1 >= 0 ? enc1 : code_enc2(r);
  d.pm = plane_masks(l.nbits, l.issigned, l.bipolar);
  d.pp = l.nbits;
  d.ptiles = pl.ptiles;
  d.splits = pl.splits;
  d.jper = pl.jper;
  d.variant = pl.variant;
  d.tile_begin = tile_begin;
  memcpy(out, &d, sizeof d);
}

int launch_planes_batch(const void *d_desc, uint32_t total, unsigned int *counter,
                        const uint32_t *item_prod, uint32_t grid, hipStream_t s) {
  hipLaunchKernelGGL(planes_batch_kernel, dim3(grid), dim3(256), 0, s,
                     (const PlanesBatchProb *)d_desc, total, counter, item_prod);
  GBS_HIP(hipGetLastError());
  return GBS_OK;
}

}  // namespace gbs
