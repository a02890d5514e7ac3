// gbs_mfma.hip -- exact int8 MFMA engine for the bit-serial GEMM on gfx950.
//
// Few-bit values fit int8 exactly: bipolar +-1, signed two's complement of
// up to 8 bits, unsigned codes of up to 7 bits.  A matrix's planes are
// expanded once per change into an int8 copy ("unpacked values",
// [align(rows,256)][align(K,256)], zero padded), and the product is a plain
// int8 GEMM on the matrix cores:
//
//   res[n*M + m] = sum_k val(rhs[n,k]) * val(lhs[m,k])      (int32)
//
// It is exact -- no wrap, no rounding -- whenever
// K * max|val(lhs)| * max|val(rhs)| < 2^31 (checked by mfma_supported), in
// which case it equals the reference's int32 result bit for bit.
//
// Kernel: 256 threads = 2 x 2 waves over a P (rhs rows) x Q (lhs rows)
// tile; each wave owns (P/2) x (Q/2) outputs as 16 x 16 fragments of
// v_mfma_i32_16x16x64_i8.  K is staged 128 bytes at a time in LDS
// (row stride 144 B, so the 16 rows a fragment load touches fall in
// distinct banks), double-buffered through registers with 16-byte loads.
// The rhs operand goes to the MFMA A port so that a fragment's output
// lanes run along m, the contiguous dimension of res.  Both operand
// fragments take 16 contiguous bytes of a row at k-offset 16*(lane/16):
// the same k-to-(lane, byte) assignment on both ports, so the dot product
// over k is the same whatever order the instruction sums in.  Small grids
// split K and add with integer atomics.
#include "gbs_internal.h"

namespace gbs {
namespace {

typedef int v4i __attribute__((ext_vector_type(4)));

constexpr int KS = 128;       // K bytes per LDS stage
constexpr int LDR = KS + 16;  // LDS row stride (bytes)

struct MfmaArgs {
  const int8_t *R, *L;  // rhs / lhs int8 values
  uint64_t r_ld, l_ld;  // row strides (bytes)
  int32_t *res;
  uint32_t M, N;        // lhs rows, rhs rows
  uint32_t kbytes;      // padded depth (multiple of KS)
  uint32_t kchunk;      // depth per split (multiple of KS)
  uint32_t ptiles, qtiles;
  int atomic;
};

template <int P, int Q>
__global__ __launch_bounds__(256) void mfma_kernel(MfmaArgs a) {
  constexpr int FP = P / 32, FQ = Q / 32;         // fragments per wave
  constexpr int CPR = KS / 16;                    // 16-byte chunks per row
  constexpr int RLD = P * CPR / 256, LLD = Q * CPR / 256;
  static_assert(RLD * 256 == P * CPR && LLD * 256 == Q * CPR, "tile/thread mismatch");
  __shared__ __attribute__((aligned(16))) int8_t sR[2][P * LDR];
  __shared__ __attribute__((aligned(16))) int8_t sL[2][Q * LDR];

  const uint32_t T = gridDim.x, b = blockIdx.x;
  uint32_t t = (T % 8 == 0) ? (b % 8) * (T / 8) + b / 8 : b;  // XCD-aware order
  const uint32_t qt = t % a.qtiles;
  t /= a.qtiles;
  const uint32_t pt = t % a.ptiles;
  const uint32_t sp = t / a.ptiles;
  const uint32_t p0 = pt * P, q0 = qt * Q;
  const uint32_t kbeg = sp * a.kchunk;
  const uint32_t kend = min(a.kbytes, kbeg + a.kchunk);

  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wp = wave & 1, wq = wave >> 1;

  v4i gr[RLD], gl[LLD];
  auto gload = [&](uint32_t k0) {
#pragma unroll
    for (int i = 0; i < RLD; i++) {
      const int c = tid + 256 * i, row = c / CPR, ch = c % CPR;
      gr[i] = *reinterpret_cast<const v4i *>(a.R + (uint64_t)(p0 + row) * a.r_ld + k0 + ch * 16);
    }
#pragma unroll
    for (int i = 0; i < LLD; i++) {
      const int c = tid + 256 * i, row = c / CPR, ch = c % CPR;
      gl[i] = *reinterpret_cast<const v4i *>(a.L + (uint64_t)(q0 + row) * a.l_ld + k0 + ch * 16);
    }
  };
  auto swrite = [&](int buf) {
#pragma unroll
    for (int i = 0; i < RLD; i++) {
      const int c = tid + 256 * i, row = c / CPR, ch = c % CPR;
      *reinterpret_cast<v4i *>(&sR[buf][row * LDR + ch * 16]) = gr[i];
    }
#pragma unroll
    for (int i = 0; i < LLD; i++) {
      const int c = tid + 256 * i, row = c / CPR, ch = c % CPR;
      *reinterpret_cast<v4i *>(&sL[buf][row * LDR + ch * 16]) = gl[i];
    }
  };

  v4i acc[FP][FQ];
#pragma unroll
  for (int i = 0; i < FP; i++)
#pragma unroll
    for (int j = 0; j < FQ; j++) acc[i][j] = v4i{0, 0, 0, 0};

  if (kbeg < kend) {
    gload(kbeg);
    swrite(0);
  }
  __syncthreads();
  int buf = 0;
  const int fr = lane & 15, fk = (lane >> 4) * 16;
  for (uint32_t k0 = kbeg; k0 < kend; k0 += KS) {
    const bool more = k0 + KS < kend;
    if (more) gload(k0 + KS);
#pragma unroll
    for (int kk = 0; kk < KS / 64; kk++) {
      v4i af[FP], bf[FQ];
#pragma unroll
      for (int i = 0; i < FP; i++)
        af[i] = *reinterpret_cast<const v4i *>(
            &sR[buf][(wp * (P / 2) + i * 16 + fr) * LDR + kk * 64 + fk]);
#pragma unroll
      for (int j = 0; j < FQ; j++)
        bf[j] = *reinterpret_cast<const v4i *>(
            &sL[buf][(wq * (Q / 2) + j * 16 + fr) * LDR + kk * 64 + fk]);
#pragma unroll
      for (int i = 0; i < FP; i++)
#pragma unroll
        for (int j = 0; j < FQ; j++)
          acc[i][j] = __builtin_amdgcn_mfma_i32_16x16x64_i8(af[i], bf[j], acc[i][j], 0, 0, 0);
    }
    if (more) swrite(buf ^ 1);
    __syncthreads();
    buf ^= 1;
  }

  // fragment output: lane holds rows 4*(lane/16)+r (rhs, n) of column lane%16 (lhs, m)
#pragma unroll
  for (int i = 0; i < FP; i++)
#pragma unroll
    for (int j = 0; j < FQ; j++)
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const uint32_t n = p0 + wp * (P / 2) + i * 16 + 4 * (lane >> 4) + r;
        const uint32_t m = q0 + wq * (Q / 2) + j * 16 + fr;
        if (n < a.N && m < a.M) {
          const uint64_t idx = (uint64_t)n * a.M + m;
          if (a.atomic) atomicAdd(a.res + idx, acc[i][j][r]);
          else a.res[idx] = acc[i][j][r];
        }
      }
}

// planes -> int8 values, 4 bytes per thread.  Zero outside the real matrix.
__global__ __launch_bounds__(256) void expand_kernel(
    const uint64_t *planes, uint64_t nrows, uint64_t ncols, uint64_t wpr, uint64_t W,
    uint32_t nbits, bool issigned, uint64_t rows_a, uint64_t ld, int8_t *out) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint64_t per_row = ld / 4;
  if (i >= rows_a * per_row) return;
  const uint64_t r = i / per_row, k0 = (i - r * per_row) * 4;
  uint32_t packed = 0;
  if (r < nrows) {
#pragma unroll
    for (int j = 0; j < 4; j++) {
      const uint64_t k = k0 + j;
      if (k >= ncols) break;
      uint32_t code = 0;
      for (uint32_t bb = 0; bb < nbits; bb++)
        code |= (uint32_t)((planes[(uint64_t)bb * W + r * wpr + (k >> 6)] >> (k & 63)) & 1) << bb;
      int32_t v;
      if (nbits == 1 && issigned) v = code ? 1 : -1;
      else if (issigned && (code >> (nbits - 1))) v = (int32_t)code - (int32_t)(1u << nbits);
      else v = (int32_t)code;
      packed |= ((uint32_t)v & 0xffu) << (8 * j);
    }
  }
  reinterpret_cast<uint32_t *>(out)[i] = packed;
}

uint64_t max_abs(const Operand &o) {
  if (o.bipolar) return 1;
  if (o.issigned) return 1ull << (o.nbits - 1);
  return (1ull << o.nbits) - 1;
}

template <int P, int Q>
int run_mfma(MfmaArgs a, hipStream_t s) {
  a.ptiles = (a.N + P - 1) / P;
  a.qtiles = (a.M + Q - 1) / Q;
  const uint64_t tiles = (uint64_t)a.ptiles * a.qtiles;
  uint32_t splits = 1;
  const uint32_t ksteps = a.kbytes / KS;
  if (tiles < 512 && ksteps >= 16) {
    uint64_t want = (512 + tiles - 1) / tiles;
    uint64_t maxs = ksteps / 8;
    splits = (uint32_t)(want < maxs ? want : maxs);
    if (splits < 1) splits = 1;
  }
  uint32_t per = (ksteps + splits - 1) / splits;
  splits = (ksteps + per - 1) / per;
  a.kchunk = per * KS;
  a.atomic = splits > 1;
  if (a.atomic) {
    int rc = launch_fill_i32(a.res, (uint64_t)a.M * a.N, 0, s);
    if (rc) return rc;
  }
  const uint64_t grid = tiles * splits;
  if (grid == 0) return GBS_OK;
  if (grid > 0x7fffffffull) return fail(GBS_ERANGE, "mfma grid too large");
  hipLaunchKernelGGL((mfma_kernel<P, Q>), dim3((uint32_t)grid), dim3(256), 0, s, a);
  GBS_HIP(hipGetLastError());
  return GBS_OK;
}

}  // namespace

bool mfma_supported(const Operand &l, const Operand &r, uint64_t K) {
  // unsigned 8-bit codes (128..255) do not fit int8
  if (!l.issigned && l.nbits > 7) return false;
  if (!r.issigned && r.nbits > 7) return false;
  const unsigned __int128 bound = (unsigned __int128)K * max_abs(l) * max_abs(r);
  return bound < ((unsigned __int128)1 << 31);
}

gbs_algo select_algo(const Operand &l, const Operand &r, uint64_t K) {
  if (!mfma_supported(l, r, K)) return GBS_ALGO_POPCOUNT;
  const double M = (double)l.nrows, N = (double)r.nrows, Kd = (double)K;
  const double pl = l.bipolar && r.bipolar ? 1 : (double)l.nbits;
  const double pr = l.bipolar && r.bipolar ? 1 : (double)r.nbits;
  // modelled times at the efficiencies these kernels reach (fractions of
  // the 39.3 T lane-op/s VALU, 5 POPS int8 MFMA and 8 TB/s HBM peaks)
  const double valu = M * N * ((K + 31) / 32) * 2 * pl * pr / (0.55 * 39.3e12);
  const double pbytes = (M * Kd * l.nbits + N * Kd * r.nbits) / 8 + 4 * M * N;
  const double t_popc = (valu > pbytes / (0.6 * 8e12) ? valu : pbytes / (0.6 * 8e12));
  const double mops = 2 * M * N * Kd / (0.35 * 5e15);
  const double ibytes = (M + N) * Kd + 4 * M * N;
  // the MFMA tile is 128 x 128: skinny products waste most of it
  const double waste = (N < 128 ? 128 / (N > 1 ? N : 1) : 1) * (M < 128 ? 128 / (M > 1 ? M : 1) : 1);
  const double t_mfma_c = mops * waste;
  const double t_mfma = (t_mfma_c > ibytes / (0.6 * 8e12) ? t_mfma_c : ibytes / (0.6 * 8e12));
  return t_mfma < t_popc ? GBS_ALGO_MFMA : GBS_ALGO_POPCOUNT;
}

int ensure_vals(gbs_matrix *m, hipStream_t s, bool may_allocate) {
  if (m->vals_version == m->planes_version && m->d_vals) return GBS_OK;
  const uint64_t rows = ((m->nrows + kValsRowAlign - 1) / kValsRowAlign) * kValsRowAlign;
  const uint64_t ld = ((m->ncols + kValsColAlign - 1) / kValsColAlign) * kValsColAlign;
  const uint64_t rows_a = rows ? rows : kValsRowAlign, ld_a = ld ? ld : kValsColAlign;
  if (!m->d_vals) {
    if (!may_allocate)
      return fail(GBS_ERANGE, "mfma: operand has no int8 copy yet; call gbs_workspace_reserve");
    GBS_HIP(hipMalloc((void **)&m->d_vals, rows_a * ld_a));
    m->vals_rows = rows_a;
    m->vals_ld = ld_a;
  }
  const uint64_t n = m->vals_rows * m->vals_ld / 4;
  hipLaunchKernelGGL(expand_kernel, dim3((uint32_t)((n + 255) / 256)), dim3(256), 0, s,
                     m->d_planes, m->nrows, m->ncols, m->wpr(), m->wpp(),
                     (uint32_t)m->nbits, m->issigned, m->vals_rows, m->vals_ld, m->d_vals);
  GBS_HIP(hipGetLastError());
  m->vals_version = m->planes_version;
  return GBS_OK;
}

int launch_mfma_gemm(const gbs_matrix *l, const gbs_matrix *r, int32_t *d_res,
                     hipStream_t s) {
  const uint64_t M = l->nrows, N = r->nrows;
  if (M == 0 || N == 0) return GBS_OK;
  if (M > 0x7fffffffull || N > 0x7fffffffull)
    return fail(GBS_ERANGE, "gemm dimensions exceed 2^31");
  if (l->vals_ld != r->vals_ld || !l->d_vals || !r->d_vals)
    return fail(GBS_EINVAL, "mfma: operand int8 copies are not current");
  MfmaArgs a{};
  a.R = r->d_vals;
  a.L = l->d_vals;
  a.r_ld = r->vals_ld;
  a.l_ld = l->vals_ld;
  a.res = d_res;
  a.M = (uint32_t)M;
  a.N = (uint32_t)N;
  a.kbytes = (uint32_t)((l->ncols + KS - 1) / KS * KS);
  if (a.kbytes == 0) return launch_fill_i32(d_res, M * N, 0, s);
  return run_mfma<128, 128>(a, s);
}

}  // namespace gbs
