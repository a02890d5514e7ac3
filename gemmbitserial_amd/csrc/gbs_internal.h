// gbs_internal.h -- shared declarations of libgbs_hip.so (not installed).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "gbs_hip.h"

// A device-resident bit-serial matrix.  Plane words live in HBM in the
// reference layout [nbits][nrows_a][ncols_a/64] (gemmbitserial.hpp:44-50),
// so host mirrors are plain copies.
struct gbs_matrix {
  uint64_t nbits, nrows, ncols, nrows_a, ncols_a;
  bool issigned;
  int device;
  uint64_t *d_planes;
  // int8 copy of the values for the MFMA engine, made from the planes on
  // first use after each change: [vals_rows][vals_ld], zero padded.
  int8_t *d_vals = nullptr;
  uint64_t vals_rows = 0, vals_ld = 0;
  // FP4 copy (two values per byte) for operands of at most 2 bits
  // (3 signed): [f4_rows][f4_ld], zero padded, same lifecycle.
  int8_t *d_f4 = nullptr;
  uint64_t f4_rows = 0, f4_ld = 0;
  // FP6 copy (6 bits per value) for operands of at most 4 bits (5 signed):
  // [f6_rows][f6_ld], zero padded, same lifecycle.
  int8_t *d_f6 = nullptr;
  uint64_t f6_rows = 0, f6_ld = 0;
  uint64_t planes_version = 1, vals_version = 0, f4_version = 0, f6_version = 0;
  uint64_t wpr() const { return ncols_a / 64; }
  uint64_t wpp() const { return nrows_a * wpr(); }
  uint64_t words() const { return nbits * wpp(); }
  bool bipolar() const { return nbits == 1 && issigned; }
  // words holding real columns; the rest of each row is zero padding
  uint64_t kwords() const { return (ncols + 63) / 64; }
};

struct gbs_workspace {
  int device = 0;
  int32_t *d_res = nullptr;      // result buffer for host-result calls
  size_t res_cap = 0;            // bytes
  int32_t *d_rowsum = nullptr;   // bipolar x regular correction
  size_t rowsum_cap = 0;         // bytes
  void *d_stage = nullptr;       // staging for host sources (import)
  size_t stage_cap = 0;          // bytes
};

namespace gbs {

// error plumbing: record a message for gbs_last_error, return code
int fail(int code, const char *fmt, ...);
int hip_fail(hipError_t e, const char *what);
#define GBS_HIP(call)                                   \
  do {                                                  \
    hipError_t e_ = (call);                             \
    if (e_ != hipSuccess) return ::gbs::hip_fail(e_, #call); \
  } while (0)

hipStream_t thread_stream();
int ensure_device();
// make *p hold at least need bytes; an outgrown buffer is freed after
// waiting for stream s, the one the buffer serves
int grow(void **p, size_t *cap, size_t need, hipStream_t s);

// operand view handed to the GEMM launchers
struct Operand {
  const uint64_t *planes;
  uint64_t nrows, wpr, wpp;
  uint32_t nbits;
  bool issigned, bipolar;
};
inline Operand view(const gbs_matrix *m) {
  return Operand{m->d_planes, m->nrows, m->wpr(), m->wpp(), (uint32_t)m->nbits,
                 m->issigned, m->bipolar()};
}

// kernels (gbs_pack.hip)
int launch_pack(const gbs_matrix *m, const void *d_src, gbs_dtype dt,
                bool col_major, hipStream_t s);
int launch_pack_im2col(const gbs_matrix *m, const void *d_src, gbs_dtype dt,
                       const gbs_conv2d &g, uint32_t oh, uint32_t ow, hipStream_t s);
int launch_pack_quantize(const gbs_matrix *m, const void *d_src, const void *d_thr,
                         gbs_dtype dt, int nthres, bool col_major, hipStream_t s);
int launch_export(const gbs_matrix *m, int32_t *d_dst, hipStream_t s);
int launch_sum_rows(const Operand &o, uint64_t kwords, uint64_t ncols,
                    int32_t *d_dst, hipStream_t s);
int launch_fill_i32(int32_t *d, uint64_t n, int32_t v, hipStream_t s);

// GEMM engines.  d_rowsum: row sums of the regular operand when exactly one
// side is bipolar (nullptr otherwise).
// does the popcount engine want the regular operand's row sums for this
// bipolar x regular product (the skinny kernel folds them in itself)
bool popc_needs_rowsum(const Operand &l, const Operand &r, uint64_t K);
int launch_popc_gemm(const Operand &l, const Operand &r, uint64_t kwords,
                     uint64_t K, int32_t *d_res, const int32_t *d_rowsum,
                     hipStream_t s);
// Matrix-core engine of two 1-bit operands on their planes (gbs_bits.hip):
// no value copies; supported for nbits == 1 on both sides, K < 2^24 - 1024.
bool bits_supported(const Operand &l, const Operand &r, uint64_t K);
bool bits_preferred(const Operand &l, const Operand &r, uint64_t K);  // AUTO takes it
int launch_bits_gemm(const Operand &l, const Operand &r, uint64_t K, int32_t *d_res,
                     hipStream_t s);
// Matrix-core engine of multi-bit operands on their planes (gbs_planes.hip):
// no value copies; one side of <= 4 planes in 2-bit chunks, the other
// whole-valued FP4 codes (bipolar, <= 2 bits unsigned, <= 3 bits signed),
// every partial sum below 2^24.
bool planes_supported(const Operand &l, const Operand &r, uint64_t K);
int launch_planes_gemm(const Operand &l, const Operand &r, uint64_t K, int32_t *d_res,
                       hipStream_t s);
// The same engine as one persistent launch over many products (the batch's
// planes group): lhs = plane side (<= 4 planes), rhs = code side (one FP4
// code, or two chunks of <= 4 planes), not both bipolar.  A plan gives the
// product's work items (code-side tiles x plane-side tiles x K pieces of
// jper super-steps of 1024 values); planes_batch_desc writes its device
// descriptor (planes_batch_desc_size() bytes) with the index of its first
// item; launch_planes_batch runs `total` items over a grid of persistent
// blocks (item_prod: the descriptor of every item; results of products of
// more than one K piece must be zero before it, the pieces add into them).
struct PlanesBatchPlan {
  bool ok;
  uint32_t vtiles, ptiles, splits, jper, variant;
  double item_bytes;  // plane bytes one item loads
  uint64_t tiles() const { return (uint64_t)vtiles * ptiles * splits; }
};
PlanesBatchPlan planes_batch_plan(const Operand &l, const Operand &r, uint64_t K,
                                  uint32_t min_pieces = 1);
size_t planes_batch_desc_size();
void planes_batch_desc(const Operand &l, const Operand &r, uint64_t K, int32_t *d_res,
                       const PlanesBatchPlan &pl, uint64_t tile_begin, void *out);
int launch_planes_batch(const void *d_desc, uint32_t total, unsigned int *counter,
                        const uint32_t *item_prod, uint32_t grid, hipStream_t s);
// Matrix-core engine (gbs_mfma.hip).  ensure_vals / ensure_f4 / ensure_f6
// make the int8 / FP4 / FP6 copy current (allocating it on first use when
// may_allocate); ensure_copy picks by format (0 int8, 1 FP4, 2 FP6).
int ensure_vals(gbs_matrix *m, hipStream_t s, bool may_allocate);
int ensure_f4(gbs_matrix *m, hipStream_t s, bool may_allocate);
int ensure_f6(gbs_matrix *m, hipStream_t s, bool may_allocate);
int ensure_copy(gbs_matrix *m, int fmt, hipStream_t s, bool may_allocate);
// Copies rebuilt many at a time (the batch's refresh): copy_current tells
// whether m's copy in fmt matches its planes; copy_job describes the
// rebuild (allocating the copy on first use when may_allocate; `whole`
// also writes the padding rows, which a copy built before already holds
// as zeros); jobs laid
// out with block_begin = the prefix of expand_job_blocks run in one
// launch_expand_jobs over d_jobs (device memory), after which the caller
// calls mark_copy_current on each.
struct ExpandJob {
  const uint64_t *planes;
  uint64_t nrows, ncols, wpr, W;  // the matrix: rows, K, row pitch, words per plane
  uint32_t nbits;
  bool issigned;
  uint64_t rows_a, ld;            // the copy: rows (padded), bytes per row
  uint64_t rows_do;               // rows written: rows_a on the first build (the
                                  // padding rows' zeros), the matrix's rows after
  void *out;
  int fmt;
  uint32_t block_begin;           // first block of the job in a many-job launch
};
bool copy_current(const gbs_matrix *m, int fmt);
int copy_job(gbs_matrix *m, int fmt, bool may_allocate, bool whole, ExpandJob *j);
void mark_copy_current(gbs_matrix *m, int fmt);
uint32_t expand_job_blocks(const ExpandJob &j);
int launch_expand_jobs(const ExpandJob *d_jobs, uint32_t n, uint32_t blocks, hipStream_t s);
int launch_mfma_gemm(const gbs_matrix *l, const gbs_matrix *r, int32_t *d_res,
                     hipStream_t s, int fmt, int forced_shape = -1, bool allow_split = true,
                     int forced_splits = 0);
// tuning table of the MFMA engine: tile shape and K split per shape class
// (shapes == nullptr restores the built-in heuristic; splits 0 = automatic)
int set_mfma_table(const int *shapes, const int *splits, int n);
// measured best time (microseconds) per format and class, as the table
// lists them (int8 classes, then FP4, then FP6; 0 = unknown); used to pick
// between FP6 and int8 copies.  set_mfma_table clears them.
int set_mfma_times(const float *times_us, int n);
bool mfma_supported(const Operand &l, const Operand &r, uint64_t K);  // int8 copies exact
bool f4_supported(const Operand &l, const Operand &r, uint64_t K);    // FP4 copies exact
bool f6_supported(const Operand &l, const Operand &r, uint64_t K);    // FP6 copies exact
// copy format AUTO uses (1 FP4, 2 FP6, 0 int8), -1 if none
int best_format(const Operand &l, const Operand &r, uint64_t K);
gbs_algo select_algo(const Operand &l, const Operand &r, uint64_t K);
// AUTO for two matrices: select_algo, with the state of their value copies
gbs_algo select_for(const gbs_matrix *l, const gbs_matrix *r);
constexpr uint64_t kValsRowAlign = 256;  // rows of the int8 copy
constexpr uint64_t kValsColAlign = 256;  // values per row of the int8 copy (bytes)
constexpr uint64_t kPackedColAlign = 512;  // values per row of the FP4 / FP6 copies

size_t dtype_size(gbs_dtype dt);      // integral element types (0 otherwise)
size_t dtype_size_any(gbs_dtype dt);  // also the floating types

// the GEMM enqueue shared by gbs_gemm, gbs_gemm_async and the batch engine
int gemm_enqueue_public(gbs_workspace *ws, const gbs_matrix *l, const gbs_matrix *r,
                        int32_t *d_res, gbs_algo algo, hipStream_t s, bool may_allocate);

}  // namespace gbs
