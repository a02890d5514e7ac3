/*
 * gbs_hip.h -- C ABI of libgbs_hip.so, the MI355X (gfx950) bit-serial GEMM.
 *
 * Plain C: pointers, sizes and status codes; no C++ or torch types.  The C++
 * drop-in header include/gemmbitserial.hpp is written on top of it, and any
 * FFI (ctypes, cgo, JNI) can bind it directly.
 *
 * Each entry point names the reference interface it replaces (paths are
 * relative to the gemmbitserial checkout).  The reference is a header-only
 * C++ library with no FFI of its own; the functions below are what its
 * BitSerialMatrix / GEMMContext methods call in the drop-in header.
 *
 * Conventions
 *   - Every function returns GBS_OK (0) or a positive gbs_status; the message
 *     of the last failure on the calling thread is gbs_last_error().
 *   - Plane layout, on the host and on the device: uint64 words
 *     [nbits][nrows_a][ncols_a/64]; column c of a row is bit (c & 63) of word
 *     (c >> 6); every bit outside the real nrows x ncols is zero
 *     (gemmbitserial.hpp:44-50, 89-129).
 *   - rhs is given as rhsRows x depth ("transposed"); the result is
 *     rhsRows x lhsRows row-major int32, res[n*M + m] (README.md "Note that the
 *     right-hand-side matrix must be provided in transposed form").
 *   - Synchronous entry points return with their results in place.  The
 *     *_async entry points only enqueue on the given HIP stream (NULL = the
 *     calling thread's library stream).
 *   - Thread safety: calls on distinct objects from distinct threads are
 *     safe; each thread owns its own library stream.  Calls that share a
 *     gbs_workspace must not overlap, on the host or on the device (one
 *     stream at a time per workspace).
 */
#ifndef GBS_HIP_H
#define GBS_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum gbs_status {
  GBS_OK = 0,
  GBS_EINVAL = 1,   /* bad argument: shape mismatch, null pointer, bitwidth */
  GBS_ENOMEM = 2,   /* device or pinned allocation failed */
  GBS_EDEVICE = 3,  /* HIP runtime or launch error */
  GBS_ERANGE = 4,   /* request exceeds a stated limit */
  GBS_ENOTSUP = 5   /* combination this build does not implement */
} gbs_status;

/* Element type of a regular (unpacked) matrix handed to gbs_import.
 * GBS_U8 follows the reference's template importRegular<T> with
 * T = (const) uint8_t; GBS_U8_RAW follows its non-template
 * importRegular(uint8_t*) overload, which takes the raw low bits without the
 * bipolar "> 0" test (gemmbitserial.hpp:212-245 vs 249-275).  The floating
 * types are accepted by the threshold import only (its comparisons are
 * meaningful for any ordered T; the plain import's quantize is defined for
 * integral T only, gemmbitserial.hpp:169-186). */
typedef enum gbs_dtype {
  GBS_I8 = 0, GBS_U8 = 1, GBS_U8_RAW = 2, GBS_I16 = 3, GBS_U16 = 4,
  GBS_I32 = 5, GBS_U32 = 6, GBS_I64 = 7, GBS_U64 = 8,
  GBS_F32 = 9, GBS_F64 = 10
} gbs_dtype;

/* GEMM engine.  AUTO picks per shape from the measured crossover; every
 * engine returns identical int32 results. */
typedef enum gbs_algo {
  GBS_ALGO_AUTO = 0,
  GBS_ALGO_POPCOUNT = 1, /* fused AND/XOR + popcount over bit-planes; products
                            of <= 8 rhs rows and <= 4 planes a side stream
                            the lhs planes once (the smallest operand form) */
  GBS_ALGO_MFMA = 2,     /* planes expanded to int8, v_mfma_i32_*_i8 */
  GBS_ALGO_MFMA_FP4 = 3, /* planes expanded to FP4 codes (operands of <= 2
                            bits, <= 3 signed), v_mfma_scale_f32_*_f8f6f4 */
  GBS_ALGO_MFMA_FP6 = 4, /* planes expanded to FP6 codes (operands of <= 4
                            bits, <= 5 signed), v_mfma_scale_f32_*_f8f6f4 */
  GBS_ALGO_MFMA_BITS = 5, /* 1-bit x 1-bit only: the bit-planes themselves,
                            each bit turned into an FP4 code in registers
                            (no value copies), v_mfma_scale_f32_*_f8f6f4 */
  GBS_ALGO_MFMA_PLANES = 6 /* the bit-planes themselves (no value copies):
                            one operand of <= 4 planes taken two planes at a
                            time as FP4 codes of 2-bit chunks, the other as
                            FP4 codes of its whole values (bipolar, <= 2 bits
                            unsigned, <= 3 bits signed) or, of 3-4 bits, of
                            two 2-bit chunks; v_mfma_scale_f32_*_f8f6f4 */
} gbs_algo;

typedef struct gbs_matrix gbs_matrix;       /* device-resident bit-serial matrix */
typedef struct gbs_workspace gbs_workspace; /* device result buffer + scratch */

/* Library / device ---------------------------------------------------- */

/* Select and initialise the HIP device used by this thread's calls
 * (default 0).  No reference counterpart: the reference runs on the host. */
int gbs_init(int device);

/* Message of the last failing call on this thread ("" if none). */
const char *gbs_last_error(void);

/* Library version, major*10000 + minor*100 + patch. */
int gbs_version(void);

/* Pinned host memory for plane mirrors and results (hipHostMalloc).
 * Replaces the `new uint64_t[]` / `new int32_t[]` of
 * gemmbitserial.hpp:33 and :450. */
int gbs_host_alloc(size_t bytes, void **out);
int gbs_host_free(void *p);

/* Bit-serial matrices -------------------------------------------------- */

/* Allocate a device matrix of nbits planes, nrows x ncols real elements,
 * padded to nrows_a x ncols_a (ncols_a % 64 == 0, nrows_a >= nrows,
 * ncols_a >= ncols).  Planes start zero.  1 <= nbits <= 8.
 * Replaces BitSerialMatrix::alloc (gemmbitserial.hpp:24-35). */
int gbs_matrix_create(uint64_t nbits, uint64_t nrows, uint64_t ncols,
                      int issigned, uint64_t nrows_a, uint64_t ncols_a,
                      gbs_matrix **out);

/* Free a matrix.  Replaces BitSerialMatrix::dealloc (gemmbitserial.hpp:38-40).
 * Waits for the calling thread's stream only: work the caller enqueued with
 * this matrix on other streams (gbs_gemm_async, batches) must be complete. */
int gbs_matrix_destroy(gbs_matrix *m);

/* Device address of the plane words (layout above). */
int gbs_matrix_device_planes(const gbs_matrix *m, uint64_t **d_planes);

/* Tell the library that the plane words were written directly on the
 * device (through gbs_matrix_device_planes), so derived copies -- the int8
 * values of the MFMA engine -- are rebuilt before the next GEMM.  No
 * reference counterpart (the reference has no derived copies). */
int gbs_matrix_mark_modified(gbs_matrix *m);

/* Pack a regular matrix into the planes on the GPU (one thread per plane
 * word or quarter word for row-major sources, plane bits gathered by u8
 * dot products; one lane per row for column-major ones).  src holds
 * nrows x ncols elements of type dt, row-major, or
 * column-major (src[c*nrows + r]) when read_col_major; src_on_device says
 * whether src is a device pointer.  When host_planes is non-NULL the packed
 * words (all nbits*nrows_a*ncols_a/64 of them) are copied there too.
 * Replaces BitSerialMatrix::importRegular<T> / importRegular(uint8_t*) /
 * importRegular_naive (gemmbitserial.hpp:135-164, 192-275). */
int gbs_import(gbs_matrix *m, const void *src, gbs_dtype dt, int read_col_major,
               int src_on_device, uint64_t *host_planes);

/* Threshold-quantizing import for unsigned matrices.  src and thresholds
 * are both of element type dt (any gbs_dtype; GBS_U8_RAW reads as GBS_U8);
 * thresholds holds thr[t*nrows + r] for t < nthres, ascending per row.
 * Element e of row r becomes the index of the first threshold t with
 * e <= thr[t*nrows + r], compared in dt, or nthres when e exceeds all of
 * them; with nthres == 0 it stays e itself.  The low nbits of that value
 * converted to uint8_t are packed: integers modulo 256; floating values
 * truncated toward zero first (C++ leaves a float outside (-1, 256)
 * undefined there; here it is the low 8 bits of the truncated int32, 0
 * beyond +-2^31 and for NaN, which is what x86-64 code computes).
 * Replaces BitSerialMatrix::importRegularAndQuantize<T>
 * (gemmbitserial.hpp:281-308). */
int gbs_import_quantize_typed(gbs_matrix *m, const void *src,
                              const void *thresholds, gbs_dtype dt, int nthres,
                              int read_col_major, int src_on_device,
                              uint64_t *host_planes);

/* gbs_import_quantize_typed with int32 elements and thresholds. */
int gbs_import_quantize(gbs_matrix *m, const int32_t *src,
                        const int32_t *thresholds, int nthres,
                        int read_col_major, int src_on_device,
                        uint64_t *host_planes);

/* Geometry of a 2-D convolution whose activation operand is packed by
 * gbs_import_im2col: NCHW input of batch x channels x height x width,
 * kernel kh x kw, strides, zero padding and dilations.  The operand has one
 * row per output pixel (n, oy, ox) in that order and one column per
 * (c, ky, kx) in that order -- the depth order of an [O][C][kh][kw] weight
 * tensor flattened row-major, which is then the lhs. */
typedef struct gbs_conv2d {
  uint32_t batch, channels, height, width;
  uint32_t kh, kw, stride_h, stride_w, pad_h, pad_w, dil_h, dil_w;
} gbs_conv2d;

/* Rows (batch*out_h*out_w) and columns (channels*kh*kw) of the operand;
 * GBS_EINVAL for an empty or inconsistent geometry.  Host only. */
int gbs_conv2d_shape(const gbs_conv2d *g, uint64_t *rows, uint64_t *cols,
                     uint32_t *out_h, uint32_t *out_w);

/* Pack the im2col matrix of an NCHW tensor straight into m's bit-planes,
 * without building it: element (row, col) is input[n][c][oy*stride_h -
 * pad_h + ky*dil_h][ox*stride_w - pad_w + kx*dil_w], or 0 outside the
 * input, quantized as gbs_import quantizes (so padding codes as 0, and as
 * -1 for a bipolar matrix, exactly as importing a zero-padded im2col matrix
 * would).  m must be rows x cols of gbs_conv2d_shape.  No reference
 * counterpart: it generalises importRegular (gemmbitserial.hpp:212-275) to
 * a convolution's activations, which the reference's callers would first
 * expand to a K x N matrix on the host (the CaffeNet list,
 * benchmark/benchmark.cpp:173-185, holds such GEMMs). */
int gbs_import_im2col(gbs_matrix *m, const void *src, gbs_dtype dt,
                      const gbs_conv2d *g, int src_on_device, uint64_t *host_planes);

/* Packed-matrix files: a 64-byte header (magic "GBSPLN01", then uint64
 * nbits, nrows, ncols, nrows_a, ncols_a, issigned, word count, little
 * endian) followed by the plane words in the reference layout
 * [nbits][nrows_a][ncols_a/64] -- byte for byte the BitSerialMatrix::data
 * array of gemmbitserial.hpp:44-50 for a matrix of that geometry.  No
 * reference counterpart (the reference keeps planes in memory only). */
int gbs_matrix_save(const gbs_matrix *m, const char *path);
/* Validates the header and size before touching the device. */
int gbs_matrix_load(const char *path, gbs_matrix **out);

/* Copy plane words between a host buffer (nbits*nrows_a*ncols_a/64 words)
 * and the device.  Used by the drop-in header to keep BitSerialMatrix::data
 * (gemmbitserial.hpp:50) and the device copy coherent after host-side
 * set/unset/word/clearAll (gemmbitserial.hpp:95-114). */
int gbs_planes_from_host(gbs_matrix *m, const uint64_t *host_planes);
int gbs_planes_to_host(const gbs_matrix *m, uint64_t *host_planes);

/* Unpack to nrows x ncols int32 values, row-major: bipolar -> +-1, else the
 * two's-complement or unsigned value.
 * Replaces BitSerialMatrix::exportRegular (gemmbitserial.hpp:312-333). */
int gbs_export(const gbs_matrix *m, int32_t *dst, int dst_on_device);

/* Row sums of the values (nrows int32).
 * Replaces sumRows_generic (arch-generic.hpp:6-37). */
int gbs_sum_rows(const gbs_matrix *m, int32_t *dst, int dst_on_device);

/* GEMM ------------------------------------------------------------------- */

/* A workspace holds the device result buffer and the row-sum scratch one
 * GEMM needs; it grows on demand.  NULL workspaces mean the calling
 * thread's own.  A workspace serves one stream at a time: the popcount
 * engine's bipolar x regular products write row sums into it on the
 * call's stream, so calls that share one (NULL included) are enqueued on
 * the same stream or separated by a synchronisation.  No reference
 * counterpart (the reference GEMMContext owns a host result,
 * gemmbitserial.hpp:379-411). */
int gbs_workspace_create(gbs_workspace **out);
int gbs_workspace_destroy(gbs_workspace *ws);

/* res[n*M + m] = sum_k val(lhs[m,k]) * val(rhs[n,k]) in int32 (wrapping),
 * M = lhs->nrows, N = rhs->nrows; lhs->ncols must equal rhs->ncols.
 * Values: bipolar (1-bit signed) +-1; signed two's complement with the top
 * plane weighted -2^(n-1); unsigned otherwise.  res is a host pointer
 * (copied back before return) or, with res_on_device, a device pointer.
 * Replaces gemmBitSerial_generic and everything it dispatches to
 * (arch-generic.hpp:39-76, 106-193, 239-345). */
int gbs_gemm(gbs_workspace *ws, const gbs_matrix *lhs, const gbs_matrix *rhs,
             int32_t *res, int res_on_device, gbs_algo algo);

/* Same product, enqueued on `stream` into the device buffer d_res
 * (N*M int32) without synchronising.  Nothing is allocated when ws already
 * has room (see gbs_workspace_reserve).  For timing loops and graph
 * capture. */
int gbs_gemm_async(gbs_workspace *ws, const gbs_matrix *lhs,
                   const gbs_matrix *rhs, int32_t *d_res, gbs_algo algo,
                   void *stream);

/* Rows of the regular operand the allocation-free entry point's row-sum
 * scratch holds (1 MiB of int32 per device). */
#define GBS_NOALLOC_MAX_ROWSUM_ROWS 262144

/* The product of gbs_gemm_async, enqueued without allocating anything: no
 * hipMalloc, no workspace.  No reference counterpart; it is for callers
 * that cannot allocate (code under HIP graph capture, tight serving loops,
 * harnesses that forbid allocation).  The caller passes the operands and
 * the device result buffer d_res (N*M int32).  Row sums of the regular
 * operand of a bipolar x regular product on the popcount engine go to a
 * fixed per-device scratch array (at most GBS_NOALLOC_MAX_ROWSUM_ROWS rows,
 * else GBS_ERANGE) that every call on that device shares: calls that
 * overlap on one GPU (other streams, other threads) race on it, so
 * serialise them.  A matrix-core engine asked for explicitly needs the
 * operands' value copies to exist already (gbs_workspace_reserve, or any
 * earlier allocating call on them), else GBS_ERANGE; GBS_ALGO_AUTO picks
 * among the engines that need nothing allocated (a matrix-core engine
 * whose copies both operands hold, else the 1-bit planes engine or the
 * popcount engine), all of which give the same int32 results. */
int gbs_gemm_async_noalloc(const gbs_matrix *lhs, const gbs_matrix *rhs,
                           int32_t *d_res, gbs_algo algo, void *stream);

/* Grow ws for a product with these two operands, so that gbs_gemm_async
 * allocates nothing. */
int gbs_workspace_reserve(gbs_workspace *ws, const gbs_matrix *lhs,
                          const gbs_matrix *rhs);

/* The engine GBS_ALGO_AUTO resolves to for these operands in their present
 * state: a matrix-core engine on value copies where the copies are current
 * (or the product is small), the planes engine for large products whose
 * copies would first have to be rebuilt from freshly imported planes. */
int gbs_gemm_select(const gbs_matrix *lhs, const gbs_matrix *rhs,
                    gbs_algo *chosen);

/* Streams and timing (HIP events on the stream the kernels run on). ---- */

/* The calling thread's library stream (a hipStream_t). */
int gbs_stream(void **stream);
int gbs_stream_sync(void *stream);
/* Elapsed milliseconds between two points recorded on a stream. */
int gbs_event_create(void **ev);
int gbs_event_destroy(void *ev);
int gbs_event_record(void *ev, void *stream);
int gbs_event_elapsed_ms(void *start, void *stop, float *ms);

/* Batched GEMMs ---------------------------------------------------------- */
/* Many independent products res_i = lhs_i * rhs_i^T (uneven shapes, mixed
 * bitwidths and signedness) run by one persistent kernel that takes output
 * tiles from a device-side work queue.  No reference counterpart: the
 * reference's callers loop over gemmBitSerial (benchmark/benchmark.cpp:
 * 215-220, the CaffeNet layer list). */
typedef struct gbs_batch gbs_batch;
typedef struct gbs_problem {
  const gbs_matrix *lhs, *rhs;
  int32_t *d_res;               /* device, rhs->nrows x lhs->nrows int32 */
} gbs_problem;

/* Plan a batch (validates shapes, builds the tile list and the device
 * descriptors).  Products the planes group takes (gbs_batch_set_planes;
 * it holds an lhs of <= 4 planes with an rhs of <= 4 planes or bipolar,
 * not both bipolar, every partial sum below 2^24) are read straight from
 * their bit-planes; the others get their operands' value copies made current in the narrowest
 * format both operands fit: FP4, FP6 or int8.  The matrices and result
 * buffers must outlive the batch.  Operands may be re-imported between
 * runs (same shapes): each run first rebuilds the copies of operands
 * changed since, on the run's stream, outside the graph of
 * gbs_batch_run_graph. */
int gbs_batch_create(const gbs_problem *problems, uint64_t count, gbs_batch **out);
int gbs_batch_destroy(gbs_batch *b);
/* Enqueue the whole batch: one persistent launch for the planes group and
 * one per copy format present (plus separate popcount launches for
 * products the matrix-core engine cannot hold exactly). */
int gbs_batch_run(gbs_batch *b, void *stream);
/* Same, captured into a hipGraph on the first call and replayed after. */
int gbs_batch_run_graph(gbs_batch *b, void *stream);
/* Baseline: one gbs_gemm_async launch sequence per product. */
int gbs_batch_run_per_call(gbs_batch *b, void *stream);
int gbs_batch_info(const gbs_batch *b, uint64_t *tiles, uint64_t *persistent,
                   uint64_t *separate);
/* The plan a batch made for problem i: group (0 int8, 1 FP4, 2 FP6 copies,
 * 3 planes; -1 for a product run by the popcount engine), tile shape (for
 * the planes group its tile form: plane-side chunks - 1 + 2 x (code-side
 * chunks - 1) + 4 when the tile holds 64 rhs rows, else 32), K splits and
 * K piece (units of 256 nominal bytes; planes group: super-steps of 1024
 * values). */
int gbs_batch_plan_of(const gbs_batch *b, uint64_t problem, int *fmt, int *shape,
                      uint32_t *splits, uint32_t *piece_units);
/* Which products batches planned from now on put into the planes group:
 * 0 (the default) none, every product from value copies (faster on the
 * measured mixed workloads, also with operands re-imported before every
 * run); 1 those it holds with at most 16 rhs rows whose operands do not
 * both fit FP4 codes (the class where it is faster than FP6 copies); 2
 * every product it holds (no value copies are made for them: less HBM).
 * No reference counterpart. */
int gbs_batch_set_planes(int mode);
int gbs_batch_get_planes(int *mode);

/* Batch tile table: the persistent launch's tile shape and largest K piece
 * per batch class (rhs-row bucket x depth bucket) and copy format, measured
 * on whole persistent launches of a mixed workload by
 * gemmbitserial_amd/tune.py --batch and loaded at gbs_init from
 * batch_tune_table.txt beside the library ($GBS_BATCH_TUNE_FILE
 * overrides).  A class without an entry keeps the built-in rule (least
 * padding for the rhs row count, pieces of at most 16 units).  Used by
 * gbs_batch_create and the serving loop's step launches; plans already
 * made keep theirs.  Process-wide: set it while no other thread plans.
 * No reference counterpart (the reference sizes CPU cache blocks,
 * gemmbitserial.hpp:343-377). */
int gbs_batch_tune_classes(int *n_classes);
int gbs_batch_tune_class_of(uint64_t N, uint64_t K, int *cls);
/* The tile shapes the persistent kernel carries for a format. */
int gbs_batch_tune_shapes(int fmt, int *shapes, int cap, int *n);
/* shape -1: the built-in rule for that class; piece_units in [1, 4096]. */
int gbs_batch_tune_set(int fmt, int cls, int shape, int piece_units);
int gbs_batch_tune_get(int fmt, int cls, int *shape, int *piece_units);
/* Text table, one "class format shape piece_units [microseconds]" line per
 * entry, '#' comments; replaces the whole table (classes not listed revert
 * to the built-in rule). */
int gbs_batch_tune_load(const char *path);

/* Tile autotuning of the MFMA engine -------------------------------------- */
/* Products fall into shape classes (buckets of M, N and K); each class maps
 * to one of the compiled tile shapes.  No reference counterpart (the
 * reference sizes CPU cache blocks with computeBlockSize,
 * gemmbitserial.hpp:343-377). */
int gbs_tune_classes(int *n_classes, int *n_shapes);
int gbs_tune_class_of(uint64_t M, uint64_t N, uint64_t K, int *cls);
/* shapes[cls], splits[cls] for every class, n = n_classes, 2 * n_classes
 * or 3 * n_classes (int8 classes, then FP4 classes, then FP6 classes).  shapes NULL: built-in
 * heuristic; a shape of -1: heuristic for that class; splits NULL or 0: K
 * split chosen from the tile count. */
int gbs_tune_set(const int *shapes, const int *splits, int n);
/* Text table, one "class shape [splits [format [microseconds]]]" line per
 * class and format (0 int8, 1 FP4, 2 FP6), '#' comments.  The measured
 * times decide between FP6 and int8 copies for GBS_ALGO_AUTO.  The table is
 * process-wide: set or load it while no other thread is planning a GEMM. */
int gbs_tune_load(const char *path);
/* One matrix-core product (algo GBS_ALGO_MFMA, _MFMA_FP4 or _MFMA_FP6)
 * with a forced tile shape and K split (the tuner's probe).  Operands must
 * already have current copies in that format (gbs_workspace_reserve). */
int gbs_gemm_async_config(gbs_workspace *ws, const gbs_matrix *lhs,
                          const gbs_matrix *rhs, int32_t *d_res, gbs_algo algo,
                          int shape, int splits, void *stream);

/* Serving -------------------------------------------------------------- */
/* A continuous-batching request loop over a stack of layers whose weights
 * stay resident in HBM in bit-serial form (plus their int8 copy).  Each
 * request is 1..max_rows rows of act_bits-bit unsigned activations; between
 * layers int32 results are requantized to min(2^act_bits - 1,
 * max(0, acc >> shift_l)) and packed into bit-planes (and int8 codes) on the
 * GPU by the GEMM launch itself, in its epilogue (generalising
 * importRegularAndQuantize, gemmbitserial.hpp:281-308, to a shift-and-clamp
 * requantization on the device).  Every step admits waiting requests as row
 * capacity and request slots free up, and runs the GEMMs and
 * requantizations of all requests in flight (each at its own layer) in one
 * persistent launch.  No reference counterpart: the reference ships only
 * the single-GEMM call (its MNIST test, test/test.cpp:157-172, runs one
 * layer).  A server is driven by one host thread. */
typedef struct gbs_server gbs_server;
typedef struct gbs_server_config {
  uint32_t max_requests; /* requests in flight at once; one step admits at
                            most max(1, max_requests / n_layers) of them */
  uint32_t max_rows;     /* rows per request */
  uint64_t mem_budget;   /* bytes for activation pools and staging; 0 = a
                            quarter of the free HBM at creation */
} gbs_server_config;
typedef struct gbs_server_counters {
  uint64_t steps, done_requests, done_rows;
  double ops;             /* 2*M*K*N summed over every GEMM row computed */
  uint32_t row_capacity;  /* rows one cohort may hold (from mem_budget) */
  uint64_t bytes_reserved;
  uint32_t waiting, in_flight;
  uint64_t layer_steps;   /* layer GEMMs run, summed over steps */
  double hbm_bytes;       /* algorithmic bytes of those GEMMs: weight planes,
                             input and output activation planes (int32 rows
                             out of the last layer) */
} gbs_server_counters;

/* Nanoseconds of the steady clock the server stamps requests with. */
int gbs_clock_ns(int64_t *now);
/* weights[l]: h_{l+1} x h_l, each must be exact on the int8 engine and
 * outlive the server; they may be re-imported between steps (same shapes),
 * each step multiplies by the current values.  shifts[l] for
 * l < n_layers - 1. */
int gbs_server_create(const gbs_matrix *const *weights, const uint32_t *shifts,
                      uint32_t n_layers, uint32_t act_bits,
                      const gbs_server_config *cfg, gbs_server **out);
int gbs_server_destroy(gbs_server *s);
/* Queue a request: x is rows x h_0 codes (uint8, low act_bits bits used).
 * arrival_ns (gbs_clock_ns time base) 0 = now.  Copies x. */
int gbs_server_submit(gbs_server *s, uint64_t id, const uint8_t *x, uint32_t rows,
                      int64_t arrival_ns);
/* Enqueue one step (admit waiting requests, compute one layer for every
 * request in flight) and complete the previous step: wait for it and make
 * the requests that left its last layer ready (*finished_now of them).  The
 * host thus plans a step while the GPU runs the one before; a request is
 * ready one call after the step that computed it. */
int gbs_server_step(gbs_server *s, uint32_t *finished_now);
/* Collect a finished request: *ready = 0 while it is not done.  out holds
 * rows x h_L int32 (the last layer's accumulators). */
int gbs_server_take(gbs_server *s, uint64_t id, int32_t *out, uint64_t cap_elems,
                    int *ready, uint32_t *rows, int64_t *arrival_ns, int64_t *done_ns);
int gbs_server_stats(const gbs_server *s, gbs_server_counters *st);
/* Ids of up to cap finished requests not yet taken (ascending), *n of them. */
int gbs_server_finished(const gbs_server *s, uint64_t *ids, uint32_t cap, uint32_t *n);
/* The bit-serial input of the cohort that sits at `layer` for the next
 * step: an act_bits-plane matrix of row capacity x h_layer (planes in the
 * layout above; the first *rows rows are the cohort's, its requests in
 * admission order), valid until the next step.  Waits for the server's
 * stream.  For layer > 0 these are the requantized outputs of layer - 1,
 * written by the step launch's epilogue. */
int gbs_server_layer_input(gbs_server *s, uint32_t layer, const gbs_matrix **out,
                           uint32_t *rows);

#ifdef __cplusplus
}
#endif

#endif /* GBS_HIP_H */
