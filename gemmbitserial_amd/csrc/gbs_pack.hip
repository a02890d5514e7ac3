// gbs_pack.hip -- regular <-> bit-serial conversion kernels for gfx950.
//
//   pack (row-major)  one-byte elements in 16-byte aligned rows: one
//                     thread per plane word, four 16-byte loads, bits
//                     gathered with u8 dot products, one 8-byte store per
//                     plane.  Other element types or alignments: one thread
//                     per 16 consecutive columns of a row (a quarter of a
//                     plane word), quantize, and per plane one 16-bit store
//                     -- the four threads of a word write its four
//                     quarters, so a wave stores 128 contiguous bytes per
//                     plane.  Threads past the real matrix write the zero
//                     padding.  (The threshold import keeps a ballot kernel:
//                     one wave per 64 plane words, one __ballot per plane.)
//   pack (col-major)  one wave per 64 rows x 1 word; lane = row, so loads of
//                     src[c*nrows + r] are coalesced; each lane shifts its
//                     64 bits into its own words.
//   pack (im2col)     the im2col matrix of an NCHW tensor packed without
//                     building it: lane = output pixel, so for each
//                     (c, ky, kx) the 64 lanes read 64 neighbouring input
//                     elements of one image row (0 outside the input).
//   export / sum_rows the inverse and the row-wise sums.
//
// Semantics follow gemmbitserial.hpp:169-186 (quantize), 212-275
// (importRegular and its raw uint8_t* overload), 281-308
// (importRegularAndQuantize), 312-333 (exportRegular) and
// arch-generic.hpp:6-37 (sumRows).  For integral element types the code's
// low nbits equal the element modulo 2^nbits (the signed branch of quantize
// adds 2^nbits before truncating to uint8), bipolar takes "element > 0",
// and the raw overload takes the element's low bits as they are.
#include <type_traits>

#include "gbs_internal.h"

namespace gbs {
namespace {

constexpr int kWave = 64;

// --- element -> code ---------------------------------------------------
template <typename T, bool RAW>
struct CodeOf {
  __device__ static inline uint32_t get(T e, bool bipolar) {
    if (RAW) return (uint32_t)(uint8_t)e;
    if (bipolar) return e > (T)0 ? 1u : 0u;
    return (uint32_t)(uint8_t)e;  // low 8 bits, two's complement
  }
};

// Source functors: code of real element (r, c).
template <typename T, bool RAW>
struct RowMajorSrc {
  const T *p;
  uint64_t nrows, ncols;
  bool bipolar;
  __device__ inline uint32_t code(uint64_t r, uint64_t c) const {
    return CodeOf<T, RAW>::get(p[r * ncols + c], bipolar);
  }
};

template <typename T, bool RAW>
struct ColMajorSrc {
  const T *p;
  uint64_t nrows, ncols;
  bool bipolar;
  __device__ inline uint32_t code(uint64_t r, uint64_t c) const {
    return CodeOf<T, RAW>::get(p[c * nrows + r], bipolar);
  }
};

// importRegularAndQuantize<T>: index of the first threshold >= e (ascending
// per row), compared in T; nthres when e exceeds all of them; e itself when
// there are no thresholds.  The value is then converted to uint8_t
// (gemmbitserial.hpp:289-299).
template <typename T>
__device__ inline uint32_t to_u8(T e) {
  if constexpr (std::is_floating_point<T>::value) {
    // truncation toward zero; the low 8 bits of the int32 (0 beyond +-2^31, NaN)
    return (e > (T)-2147483648.0 && e < (T)2147483648.0) ? (uint32_t)(uint8_t)(int32_t)e : 0u;
  } else {
    return (uint32_t)(uint8_t)e;
  }
}

template <typename T>
struct ThresholdSrc {
  const T *p, *thr;
  uint64_t nrows, ncols;
  int nthres;
  bool col_major;
  __device__ inline uint32_t code(uint64_t r, uint64_t c) const {
    const T e = col_major ? p[c * nrows + r] : p[r * ncols + c];
    if (nthres == 0) return to_u8(e);
    uint32_t q = (uint32_t)nthres;
    for (int t = 0; t < nthres; t++) {
      if (e <= thr[(uint64_t)t * nrows + r]) { q = (uint32_t)t; break; }
    }
    return q & 0xffu;
  }
};

// --- row-major pack: ballot per plane word -----------------------------
// Each wave owns flattened words f0 .. f0+63 of every plane (f = r*wpr + w).
template <typename Src>
__global__ __launch_bounds__(256) void pack_rows_kernel(
    Src src, uint64_t *planes, uint64_t nrows, uint64_t ncols, uint64_t wpr,
    uint64_t W, uint32_t nbits) {
  const uint32_t lane = threadIdx.x & (kWave - 1);
  const uint64_t wave = (uint64_t)blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave;
  const uint64_t f0 = wave * kWave;
  if (f0 >= W) return;
  uint64_t r = f0 / wpr, w = f0 - r * wpr;
  const uint32_t tmax = (uint32_t)((W - f0) < kWave ? (W - f0) : kWave);
  uint64_t out[8];
#pragma unroll
  for (int b = 0; b < 8; b++) out[b] = 0;
  for (uint32_t t = 0; t < tmax; t++) {
    const uint64_t c = w * kWave + lane;
    uint32_t code = 0;
    if (r < nrows && c < ncols) code = src.code(r, c);
#pragma unroll
    for (int b = 0; b < 8; b++) {
      if ((uint32_t)b < nbits) {
        uint64_t m = __ballot((code >> b) & 1u);
        out[b] = (lane == t) ? m : out[b];
      }
    }
    if (++w == wpr) { w = 0; ++r; }
  }
  if (lane < tmax) {
#pragma unroll
    for (int b = 0; b < 8; b++)
      if ((uint32_t)b < nbits) planes[(uint64_t)b * W + f0 + lane] = out[b];
  }
}

// --- row-major pack, 16 columns per thread ----------------------------------
// VEC: rows start 16-byte aligned and hold a multiple of 16 bytes, so the 16
// elements of a full quarter are 16 * sizeof(T) / 16 aligned vector loads.
template <typename T, bool RAW, bool VEC>
__global__ __launch_bounds__(256) void pack_rows16_kernel(
    const T *src, uint64_t nrows, uint64_t ncols, uint64_t nrows_a, uint64_t wpr,
    uint64_t W, uint32_t nbits, bool bipolar, uint16_t *planes16) {
  const uint64_t qpr = wpr * 4;  // quarters per row
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nrows_a * qpr) return;
  const uint64_t r = i / qpr, q = i - r * qpr, c0 = q * 16;
  uint32_t code[16];
#pragma unroll
  for (int j = 0; j < 16; j++) code[j] = 0;
  if (r < nrows && c0 < ncols) {
    const T *row = src + r * ncols;
    if (VEC && c0 + 16 <= ncols) {
      constexpr int NV = 16 * (int)sizeof(T) / 16;
      alignas(16) T e[16];
#pragma unroll
      for (int v = 0; v < NV; v++)
        *reinterpret_cast<uint4 *>(&e[v * 16 / sizeof(T)]) =
            reinterpret_cast<const uint4 *>(row + c0)[v];
#pragma unroll
      for (int j = 0; j < 16; j++) code[j] = CodeOf<T, RAW>::get(e[j], bipolar);
    } else {
#pragma unroll
      for (int j = 0; j < 16; j++)
        if (c0 + j < ncols) code[j] = CodeOf<T, RAW>::get(row[c0 + j], bipolar);
    }
  }
#pragma unroll
  for (int b = 0; b < 8; b++) {
    if ((uint32_t)b >= nbits) break;
    uint32_t bits = 0;
#pragma unroll
    for (int j = 0; j < 16; j++) bits |= ((code[j] >> b) & 1u) << j;
    planes16[((uint64_t)b * W + r * wpr) * 4 + q] = (uint16_t)bits;
  }
}

// --- column-major pack: lane = row, each lane builds its own words ------
template <typename Src>
__global__ __launch_bounds__(256) void pack_cols_kernel(
    Src src, uint64_t *planes, uint64_t nrows, uint64_t ncols, uint64_t nrows_a,
    uint64_t wpr, uint64_t W, uint32_t nbits) {
  const uint32_t lane = threadIdx.x & (kWave - 1);
  const uint64_t wave = (uint64_t)blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave;
  const uint64_t row_groups = (nrows_a + kWave - 1) / kWave;
  if (wave >= row_groups * wpr) return;
  const uint64_t w = wave / row_groups;
  const uint64_t r = (wave - w * row_groups) * kWave + lane;
  uint64_t out[8];
#pragma unroll
  for (int b = 0; b < 8; b++) out[b] = 0;
  if (r < nrows) {
    const uint64_t c0 = w * kWave;
    const uint32_t jmax = c0 >= ncols ? 0u
                          : (uint32_t)((ncols - c0) < kWave ? (ncols - c0) : kWave);
    // whole words unrolled by 16 so that the loads of a group overlap
    uint32_t j = 0;
    for (; j + 16 <= jmax; j += 16) {
      uint32_t code[16];
#pragma unroll
      for (int u = 0; u < 16; u++) code[u] = src.code(r, c0 + j + u);
#pragma unroll
      for (int u = 0; u < 16; u++)
#pragma unroll
        for (int b = 0; b < 8; b++)
          if ((uint32_t)b < nbits) out[b] |= (uint64_t)((code[u] >> b) & 1u) << (j + u);
    }
    for (; j < jmax; j++) {
      uint32_t code = src.code(r, c0 + j);
#pragma unroll
      for (int b = 0; b < 8; b++)
        if ((uint32_t)b < nbits) out[b] |= (uint64_t)((code >> b) & 1u) << j;
    }
  }
  if (r < nrows_a) {
#pragma unroll
    for (int b = 0; b < 8; b++)
      if ((uint32_t)b < nbits) planes[(uint64_t)b * W + r * wpr + w] = out[b];
  }
}

// --- row-major pack of one-byte elements, one plane word per thread ------
// Rows start 16-byte aligned and hold a multiple of 16 bytes.  A thread
// loads its word's 64 bytes with four 16-byte loads (the 36 threads of a
// 2304-column row read it contiguously); the codes are the bytes themselves
// (bipolar: 1 where the element is > 0), and plane b's bits are bit b of
// each byte, gathered four bytes at a time by a u8 dot product with
// weights 1, 2, 4, 8 (16, 32, 64, 128 for the next four bytes).  One 8-byte
// store per plane.  A row's last, partial word and the padding rows take
// the element loop.
template <typename T, bool RAW>
__global__ __launch_bounds__(256) void pack_bytes64_kernel(
    const T *src, uint64_t nrows, uint64_t ncols, uint64_t nrows_a, uint64_t wpr,
    uint64_t W, uint32_t nbits, bool bipolar, uint64_t *planes) {
  static_assert(sizeof(T) == 1, "one-byte elements");
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nrows_a * wpr) return;
  const uint64_t r = i / wpr, w = i - r * wpr, c0 = w * 64;
  uint64_t out[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  if (r < nrows && c0 + 64 <= ncols) {
    const uint4 *p = reinterpret_cast<const uint4 *>(src + r * ncols + c0);
    uint32_t x[16];
#pragma unroll
    for (int v = 0; v < 4; v++) {
      const uint4 q = p[v];
      x[4 * v] = q.x, x[4 * v + 1] = q.y, x[4 * v + 2] = q.z, x[4 * v + 3] = q.w;
    }
    if (!RAW && bipolar) {
#pragma unroll
      for (int d = 0; d < 16; d++) {
        // bit 7 of each byte: byte != 0 (and, for signed bytes, >= 0)
        uint32_t nz = ((x[d] & 0x7f7f7f7fu) + 0x7f7f7f7fu) | x[d];
        if (T(-1) < T(0)) nz &= ~x[d];
        x[d] = (nz >> 7) & 0x01010101u;
      }
    }
#pragma unroll
    for (int b = 0; b < 8; b++) {
      if ((uint32_t)b >= nbits) break;
      const uint32_t m = 0x01010101u << b;
      uint32_t half[2];
#pragma unroll
      for (int h = 0; h < 2; h++) {
        uint32_t acc = 0;
#pragma unroll
        for (int k = 0; k < 4; k++) {
          uint32_t part = __builtin_amdgcn_udot4(x[8 * h + 2 * k] & m, 0x08040201u, 0u, false);
          part = __builtin_amdgcn_udot4(x[8 * h + 2 * k + 1] & m, 0x80402010u, part, false);
          acc |= (part >> b) << (8 * k);
        }
        half[h] = acc;
      }
      out[b] = (uint64_t)half[0] | ((uint64_t)half[1] << 32);
    }
  } else if (r < nrows && c0 < ncols) {
    const T *row = src + r * ncols;
    const uint32_t jmax = (uint32_t)(ncols - c0 < 64 ? ncols - c0 : 64);
    for (uint32_t j = 0; j < jmax; j++) {
      const uint32_t code = CodeOf<T, RAW>::get(row[c0 + j], bipolar);
#pragma unroll
      for (int b = 0; b < 8; b++)
        if ((uint32_t)b < nbits) out[b] |= (uint64_t)((code >> b) & 1u) << j;
    }
  }
#pragma unroll
  for (int b = 0; b < 8; b++)
    if ((uint32_t)b < nbits) planes[(uint64_t)b * W + r * wpr + w] = out[b];
}

template <typename T, bool RAW>
int pack_typed(const gbs_matrix *m, const void *d_src, bool col_major,
               hipStream_t s) {
  const uint64_t W = m->wpp();
  const uint32_t nb = (uint32_t)m->nbits;
  if (W == 0) return GBS_OK;
  if (!col_major) {
    const uint64_t n = m->nrows_a * m->wpr() * 4;
    const bool vec = ((uintptr_t)d_src % 16 == 0) && (m->ncols * sizeof(T)) % 16 == 0;
    const dim3 grid((uint32_t)((n + 255) / 256));
    if constexpr (sizeof(T) == 1) {
      if (vec) {
        const uint64_t words = m->nrows_a * m->wpr();
        hipLaunchKernelGGL((pack_bytes64_kernel<T, RAW>), dim3((uint32_t)((words + 255) / 256)),
                           dim3(256), 0, s, (const T *)d_src, m->nrows, m->ncols, m->nrows_a,
                           m->wpr(), W, nb, m->bipolar(), m->d_planes);
        GBS_HIP(hipGetLastError());
        return GBS_OK;
      }
    }
    if (vec)
      hipLaunchKernelGGL((pack_rows16_kernel<T, RAW, true>), grid, dim3(256), 0, s,
                         (const T *)d_src, m->nrows, m->ncols, m->nrows_a, m->wpr(), W, nb,
                         m->bipolar(), (uint16_t *)m->d_planes);
    else
      hipLaunchKernelGGL((pack_rows16_kernel<T, RAW, false>), grid, dim3(256), 0, s,
                         (const T *)d_src, m->nrows, m->ncols, m->nrows_a, m->wpr(), W, nb,
                         m->bipolar(), (uint16_t *)m->d_planes);
  } else {
    ColMajorSrc<T, RAW> src{(const T *)d_src, m->nrows, m->ncols, m->bipolar()};
    const uint64_t waves = (m->nrows_a + kWave - 1) / kWave * m->wpr();
    const uint64_t blocks = (waves + 3) / 4;
    hipLaunchKernelGGL((pack_cols_kernel<ColMajorSrc<T, RAW>>), dim3((uint32_t)blocks),
                       dim3(256), 0, s, src, m->d_planes, m->nrows, m->ncols,
                       m->nrows_a, m->wpr(), W, nb);
  }
  GBS_HIP(hipGetLastError());
  return GBS_OK;
}

// im2col of an NCHW tensor: row r = output pixel (n, oy, ox), column
// c = (ch, ky, kx); element input[n][ch][oy*sh - ph + ky*dh][ox*sw - pw +
// kx*dw], 0 outside the input.  One wave per 64 rows x one plane word;
// lane = row, so for each column the wave reads 64 neighbouring pixels of
// one input row.  Per lane the pixel's offsets are computed once; (ch, ky,
// kx) advance as wave-uniform counters.  An image is below 2^31 elements
// and the operand's rows and columns below 2^31 (checked on the host).
struct Im2colGeom {
  uint32_t C, H, W, KH, KW, OW, OHW, sh, sw, ph, pw, dh, dw;
};

template <typename T, bool RAW>
__global__ __launch_bounds__(256) void pack_im2col_kernel(
    const T *src, Im2colGeom g, uint64_t nrows, uint64_t ncols, uint64_t nrows_a,
    uint64_t wpr, uint64_t W, uint32_t nbits, bool bipolar, uint64_t *planes) {
  const uint32_t lane = threadIdx.x & (kWave - 1);
  const uint64_t wave = (uint64_t)blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave;
  const uint64_t row_groups = (nrows_a + kWave - 1) / kWave;
  if (wave >= row_groups * wpr) return;
  const uint64_t w = wave / row_groups;
  const uint64_t r = (wave - w * row_groups) * kWave + lane;
  uint64_t out[8];
#pragma unroll
  for (int b = 0; b < 8; b++) out[b] = 0;
  const uint64_t c0 = w * kWave;
  if (r < nrows && c0 < ncols) {
    const uint32_t n = (uint32_t)(r / g.OHW), rem = (uint32_t)(r - (uint64_t)n * g.OHW);
    const uint32_t oy = rem / g.OW, ox = rem - oy * g.OW;
    const int iy0 = (int)(oy * g.sh) - (int)g.ph, ix0 = (int)(ox * g.sw) - (int)g.pw;
    const T *img = src + (uint64_t)n * g.C * g.H * g.W;
    const uint32_t jmax = (uint32_t)((ncols - c0) < kWave ? (ncols - c0) : kWave);
    // wave-uniform position of column c0
    uint32_t ch = (uint32_t)(c0 / (g.KH * g.KW));
    const uint32_t kr = (uint32_t)(c0 - (uint64_t)ch * g.KH * g.KW);
    uint32_t ky = kr / g.KW, kx = kr - ky * g.KW;
    // G columns at a time: G loads issued before any is used (an element
    // outside the input loads element 0 of the image and is replaced by
    // 0), then per plane the G bits are gathered from the codes, four
    // bytes per u8 dot product
    constexpr int G = 16;  // columns whose loads are in flight together (32: no faster)
    for (uint32_t j0 = 0; j0 < jmax; j0 += G) {
      T e[G];
      bool ok[G];
#pragma unroll
      for (int u = 0; u < G; u++) {
        const int iy = iy0 + (int)(ky * g.dh), ix = ix0 + (int)(kx * g.dw);
        ok[u] = j0 + u < jmax && (uint32_t)iy < g.H && (uint32_t)ix < g.W;
        e[u] = img[ok[u] ? (ch * g.H + (uint32_t)iy) * g.W + (uint32_t)ix : 0u];
        // next column, without branches (they would split the loads'
        // basic block and serialise them)
        const bool wx = kx + 1 == g.KW, wy = wx && ky + 1 == g.KH;
        kx = wx ? 0u : kx + 1;
        ky = wy ? 0u : (wx ? ky + 1 : ky);
        ch += wy ? 1u : 0u;
      }
      uint32_t x[G / 4];
#pragma unroll
      for (int d = 0; d < G / 4; d++) {
        uint32_t c[4];
#pragma unroll
        for (int k = 0; k < 4; k++) c[k] = ok[4 * d + k] ? CodeOf<T, RAW>::get(e[4 * d + k], bipolar) : 0u;
        x[d] = c[0] | (c[1] << 8) | (c[2] << 16) | (c[3] << 24);
      }
#pragma unroll
      for (int b = 0; b < 8; b++) {
        if ((uint32_t)b >= nbits) break;
        const uint32_t m = 0x01010101u << b;
        uint32_t bits = 0;
#pragma unroll
        for (int k = 0; k < G / 8; k++) {
          uint32_t part = __builtin_amdgcn_udot4(x[2 * k] & m, 0x08040201u, 0u, false);
          part = __builtin_amdgcn_udot4(x[2 * k + 1] & m, 0x80402010u, part, false);
          bits |= (part >> b) << (8 * k);
        }
        out[b] |= (uint64_t)bits << j0;
      }
    }
  }
  if (r < nrows_a) {
#pragma unroll
    for (int b = 0; b < 8; b++)
      if ((uint32_t)b < nbits) planes[(uint64_t)b * W + r * wpr + w] = out[b];
  }
}

template <typename T, bool RAW>
int pack_im2col_typed(const gbs_matrix *m, const void *d_src, const gbs_conv2d &c,
                      uint32_t oh, uint32_t ow, hipStream_t s) {
  const uint64_t W = m->wpp();
  if (W == 0) return GBS_OK;
  const Im2colGeom g{c.channels, c.height, c.width, c.kh, c.kw, ow, oh * ow, c.stride_h,
                     c.stride_w, c.pad_h, c.pad_w, c.dil_h, c.dil_w};
  const uint64_t waves = (m->nrows_a + kWave - 1) / kWave * m->wpr();
  hipLaunchKernelGGL((pack_im2col_kernel<T, RAW>), dim3((uint32_t)((waves + 3) / 4)),
                     dim3(256), 0, s, (const T *)d_src, g, m->nrows, m->ncols, m->nrows_a,
                     m->wpr(), W, (uint32_t)m->nbits, m->bipolar(), m->d_planes);
  GBS_HIP(hipGetLastError());
  return GBS_OK;
}

// --- export ---------------------------------------------------------------
__global__ __launch_bounds__(256) void export_kernel(
    const uint64_t *planes, uint64_t nrows, uint64_t ncols, uint64_t wpr,
    uint64_t W, uint32_t nbits, bool issigned, int32_t *out) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nrows * ncols) return;
  const uint64_t r = i / ncols, c = i - r * ncols;
  const uint64_t off = r * wpr + (c >> 6);
  const uint32_t sh = (uint32_t)(c & 63);
  int32_t v;
  if (nbits == 1 && issigned) {
    v = ((planes[off] >> sh) & 1) ? 1 : -1;
  } else {
    uint32_t acc = 0;
    for (uint32_t b = 0; b < nbits; b++) {
      uint32_t bit = (uint32_t)((planes[(uint64_t)b * W + off] >> sh) & 1);
      uint32_t t = bit << b;
      acc += (issigned && b == nbits - 1) ? 0u - t : t;
    }
    v = (int32_t)acc;
  }
  out[i] = v;
}

// --- row sums: one wave per row --------------------------------------------
__global__ __launch_bounds__(256) void sum_rows_kernel(
    const uint64_t *planes, uint64_t nrows, uint64_t wpr, uint64_t W,
    uint64_t kwords, uint32_t nbits, bool issigned, uint32_t ncols,
    int32_t *out) {
  const uint32_t lane = threadIdx.x & (kWave - 1);
  const uint64_t r = (uint64_t)blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave;
  if (r >= nrows) return;
  uint32_t acc = 0;
  const bool bip = nbits == 1 && issigned;
  for (uint32_t b = 0; b < nbits; b++) {
    uint32_t pc = 0;
    const uint64_t *row = planes + (uint64_t)b * W + r * wpr;
    for (uint64_t k = lane; k < kwords; k += kWave) pc += (uint32_t)__popcll(row[k]);
    if (bip) {
      acc = pc;
    } else {
      uint32_t t = pc << b;
      acc += (issigned && b == nbits - 1) ? 0u - t : t;
    }
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) acc += (uint32_t)__shfl_xor((int)acc, off, kWave);
  if (lane == 0) out[r] = bip ? (int32_t)(2u * acc - ncols) : (int32_t)acc;
}

__global__ void fill_i32_kernel(int32_t *d, uint64_t n, int32_t v) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  // 16-byte stores over the aligned body, scalar stores for the ends
  const uint64_t head = ((16 - ((uintptr_t)d & 15)) & 15) / 4;
  const uint64_t h = head < n ? head : n;
  const uint64_t nv = (n - h) / 4;
  int4 *dv = reinterpret_cast<int4 *>(d + h);
  const int4 vv = make_int4(v, v, v, v);
  for (uint64_t j = i; j < nv; j += stride) dv[j] = vv;
  for (uint64_t j = i; j < h; j += stride) d[j] = v;
  for (uint64_t j = h + 4 * nv + i; j < n; j += stride) d[j] = v;
}

}  // namespace

size_t dtype_size(gbs_dtype dt) {
  switch (dt) {
    case GBS_I8: case GBS_U8: case GBS_U8_RAW: return 1;
    case GBS_I16: case GBS_U16: return 2;
    case GBS_I32: case GBS_U32: return 4;
    case GBS_I64: case GBS_U64: return 8;
    default: return 0;
  }
}

size_t dtype_size_any(gbs_dtype dt) {
  switch (dt) {
    case GBS_F32: return 4;
    case GBS_F64: return 8;
    default: return dtype_size(dt);
  }
}

int launch_pack(const gbs_matrix *m, const void *d_src, gbs_dtype dt,
                bool col_major, hipStream_t s) {
  switch (dt) {
    case GBS_I8: return pack_typed<int8_t, false>(m, d_src, col_major, s);
    case GBS_U8: return pack_typed<uint8_t, false>(m, d_src, col_major, s);
    case GBS_U8_RAW: return pack_typed<uint8_t, true>(m, d_src, col_major, s);
    case GBS_I16: return pack_typed<int16_t, false>(m, d_src, col_major, s);
    case GBS_U16: return pack_typed<uint16_t, false>(m, d_src, col_major, s);
    case GBS_I32: return pack_typed<int32_t, false>(m, d_src, col_major, s);
    case GBS_U32: return pack_typed<uint32_t, false>(m, d_src, col_major, s);
    case GBS_I64: return pack_typed<int64_t, false>(m, d_src, col_major, s);
    case GBS_U64: return pack_typed<uint64_t, false>(m, d_src, col_major, s);
  }
  return fail(GBS_EINVAL, "gbs_import: unknown dtype %d", (int)dt);
}

int launch_pack_im2col(const gbs_matrix *m, const void *d_src, gbs_dtype dt,
                       const gbs_conv2d &g, uint32_t oh, uint32_t ow, hipStream_t s) {
  switch (dt) {
    case GBS_I8: return pack_im2col_typed<int8_t, false>(m, d_src, g, oh, ow, s);
    case GBS_U8: return pack_im2col_typed<uint8_t, false>(m, d_src, g, oh, ow, s);
    case GBS_U8_RAW: return pack_im2col_typed<uint8_t, true>(m, d_src, g, oh, ow, s);
    case GBS_I16: return pack_im2col_typed<int16_t, false>(m, d_src, g, oh, ow, s);
    case GBS_U16: return pack_im2col_typed<uint16_t, false>(m, d_src, g, oh, ow, s);
    case GBS_I32: return pack_im2col_typed<int32_t, false>(m, d_src, g, oh, ow, s);
    case GBS_U32: return pack_im2col_typed<uint32_t, false>(m, d_src, g, oh, ow, s);
    case GBS_I64: return pack_im2col_typed<int64_t, false>(m, d_src, g, oh, ow, s);
    case GBS_U64: return pack_im2col_typed<uint64_t, false>(m, d_src, g, oh, ow, s);
  }
  return fail(GBS_EINVAL, "gbs_import_im2col: unknown dtype %d", (int)dt);
}

template <typename T>
int pack_quantize_typed(const gbs_matrix *m, const void *d_src, const void *d_thr,
                        int nthres, bool col_major, hipStream_t s) {
  const uint64_t W = m->wpp();
  if (W == 0) return GBS_OK;
  ThresholdSrc<T> src{(const T *)d_src, (const T *)d_thr, m->nrows, m->ncols, nthres,
                      col_major};
  if (!col_major) {
    const uint64_t blocks = ((W + kWave - 1) / kWave + 3) / 4;
    hipLaunchKernelGGL((pack_rows_kernel<ThresholdSrc<T>>), dim3((uint32_t)blocks), dim3(256),
                       0, s, src, m->d_planes, m->nrows, m->ncols, m->wpr(), W,
                       (uint32_t)m->nbits);
  } else {
    const uint64_t waves = (m->nrows_a + kWave - 1) / kWave * m->wpr();
    hipLaunchKernelGGL((pack_cols_kernel<ThresholdSrc<T>>), dim3((uint32_t)((waves + 3) / 4)),
                       dim3(256), 0, s, src, m->d_planes, m->nrows, m->ncols,
                       m->nrows_a, m->wpr(), W, (uint32_t)m->nbits);
  }
  GBS_HIP(hipGetLastError());
  return GBS_OK;
}

int launch_pack_quantize(const gbs_matrix *m, const void *d_src, const void *d_thr,
                         gbs_dtype dt, int nthres, bool col_major, hipStream_t s) {
  switch (dt) {
    case GBS_I8: return pack_quantize_typed<int8_t>(m, d_src, d_thr, nthres, col_major, s);
    case GBS_U8:
    case GBS_U8_RAW: return pack_quantize_typed<uint8_t>(m, d_src, d_thr, nthres, col_major, s);
    case GBS_I16: return pack_quantize_typed<int16_t>(m, d_src, d_thr, nthres, col_major, s);
    case GBS_U16: return pack_quantize_typed<uint16_t>(m, d_src, d_thr, nthres, col_major, s);
    case GBS_I32: return pack_quantize_typed<int32_t>(m, d_src, d_thr, nthres, col_major, s);
    case GBS_U32: return pack_quantize_typed<uint32_t>(m, d_src, d_thr, nthres, col_major, s);
    case GBS_I64: return pack_quantize_typed<int64_t>(m, d_src, d_thr, nthres, col_major, s);
    case GBS_U64: return pack_quantize_typed<uint64_t>(m, d_src, d_thr, nthres, col_major, s);
    case GBS_F32: return pack_quantize_typed<float>(m, d_src, d_thr, nthres, col_major, s);
    case GBS_F64: return pack_quantize_typed<double>(m, d_src, d_thr, nthres, col_major, s);
  }
  return fail(GBS_EINVAL, "gbs_import_quantize_typed: unknown dtype %d", (int)dt);
}

int launch_export(const gbs_matrix *m, int32_t *d_dst, hipStream_t s) {
  const uint64_t n = m->nrows * m->ncols;
  if (n == 0) return GBS_OK;
  hipLaunchKernelGGL(export_kernel, dim3((uint32_t)((n + 255) / 256)), dim3(256), 0, s,
                     m->d_planes, m->nrows, m->ncols, m->wpr(), m->wpp(),
                     (uint32_t)m->nbits, m->issigned, d_dst);
  GBS_HIP(hipGetLastError());
  return GBS_OK;
}

int launch_sum_rows(const Operand &o, uint64_t kwords, uint64_t ncols,
                    int32_t *d_dst, hipStream_t s) {
  if (o.nrows == 0) return GBS_OK;
  hipLaunchKernelGGL(sum_rows_kernel, dim3((uint32_t)((o.nrows + 3) / 4)), dim3(256), 0,
                     s, o.planes, o.nrows, o.wpr, o.wpp, kwords, o.nbits,
                     o.issigned, (uint32_t)ncols, d_dst);
  GBS_HIP(hipGetLastError());
  return GBS_OK;
}

int launch_fill_i32(int32_t *d, uint64_t n, int32_t v, hipStream_t s) {
  if (n == 0) return GBS_OK;
  uint64_t blocks = (n / 4 + 255) / 256 + 1;
  if (blocks > 2048) blocks = 2048;
  hipLaunchKernelGGL(fill_i32_kernel, dim3((uint32_t)blocks), dim3(256), 0, s, d, n, v);
  GBS_HIP(hipGetLastError());
  return GBS_OK;
}

}  // namespace gbs
