// gbs_runtime.hip -- the C ABI of libgbs_hip.so (include/gbs_hip.h):
// device selection, per-thread streams, matrices, workspaces and the
// synchronous / asynchronous GEMM entry points.
#include <dlfcn.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <mutex>
#include <string>
#include <vector>

#include "gbs_mfma_tile.h"

namespace gbs {

namespace {
thread_local std::string t_err;
thread_local int t_device = -1;
thread_local hipStream_t t_stream = nullptr;
thread_local int t_stream_device = -1;
thread_local gbs_workspace *t_ws = nullptr;
}  // namespace

// Scratch of the allocation-free GEMM entry point (gbs_gemm_async_noalloc):
// the regular operand's row sums of a bipolar x regular product on the
// popcount engine.  Every device the library runs on holds its own copy
// until the process ends; host code uses the address hipGetSymbolAddress
// returns for the current device, never the array's name.
constexpr size_t kScratchInts = GBS_NOALLOC_MAX_ROWSUM_ROWS;
static __device__ __align__(256) int32_t scratch[kScratchInts];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
int32_t *g_scratch_ptr[kMaxDevices] = {};
}  // namespace

int fail(int code, const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  t_err = buf;
  return code;
}

int hip_fail(hipError_t e, const char *what) {
  return fail(e == hipErrorOutOfMemory ? GBS_ENOMEM : GBS_EDEVICE, "%s: %s", what,
              hipGetErrorString(e));
}

// A file beside this library, or the path $env names.
static std::string table_path(const char *env, const char *name) {
  if (const char *e = getenv(env)) return e;
  Dl_info info;
  if (!dladdr((void *)&table_path, &info) || !info.dli_fname) return std::string();
  std::string path = info.dli_fname;
  const size_t slash = path.rfind('/');
  return (slash == std::string::npos ? std::string(".") : path.substr(0, slash)) + "/" + name;
}

// The tuned tile tables: the standalone one ($GBS_TUNE_FILE, else
// tune_table.txt beside the library) and the persistent launch's
// ($GBS_BATCH_TUNE_FILE, else batch_tune_table.txt); the built-in rules
// where a table does not exist.
static void load_default_table() {
  static std::once_flag once;
  std::call_once(once, [] {
    const std::string saved = t_err;
    const std::string path = table_path("GBS_TUNE_FILE", "tune_table.txt");
    FILE *f = path.empty() ? nullptr : fopen(path.c_str(), "r");
    if (f) {
      fclose(f);
      if (gbs_tune_load(path.c_str()) != GBS_OK) t_err = saved;
    }
    const std::string bpath = table_path("GBS_BATCH_TUNE_FILE", "batch_tune_table.txt");
    f = bpath.empty() ? nullptr : fopen(bpath.c_str(), "r");
    if (f) {
      fclose(f);
      if (gbs_batch_tune_load(bpath.c_str()) != GBS_OK) t_err = saved;
    }
  });
}

int ensure_device() {
  load_default_table();
  if (t_device < 0) {
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n == 0)
      return fail(GBS_EDEVICE, "no HIP device available (%s)", hipGetErrorString(e));
    t_device = 0;
  }
  GBS_HIP(hipSetDevice(t_device));
  return GBS_OK;
}

hipStream_t thread_stream() {
  if (t_stream == nullptr || t_stream_device != t_device) {
    hipStream_t s = nullptr;
    // default priority, non-blocking w.r.t. the null stream
    if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) return nullptr;
    t_stream = s;
    t_stream_device = t_device;
  }
  return t_stream;
}

int grow(void **p, size_t *cap, size_t need, hipStream_t s) {
  if (*cap >= need && *p) return GBS_OK;
  if (*p) {
    // the buffer's last user is the stream it serves (a workspace serves one
    // stream at a time): wait for that stream only
    GBS_HIP(hipStreamSynchronize(s));
    GBS_HIP(hipFree(*p));
    *p = nullptr;
    *cap = 0;
  }
  size_t sz = need < 256 ? 256 : need;
  GBS_HIP(hipMalloc(p, sz));
  *cap = sz;
  return GBS_OK;
}

static gbs_workspace *thread_ws() {
  if (!t_ws) t_ws = new gbs_workspace();
  return t_ws;
}

static int check_matrix(const gbs_matrix *m, const char *fn) {
  if (!m) return fail(GBS_EINVAL, "%s: null matrix", fn);
  return GBS_OK;
}

// Shared body of gbs_gemm / gbs_gemm_async / the batch fallbacks and the
// allocation-free entry point.  Row sums of the regular operand (popcount
// engine, bipolar x regular) go to ws's scratch, grown when may_allocate,
// or, when ws is null, to `rowsum_buf` (rowsum_cap bytes, never grown).
static int gemm_enqueue_into(gbs_workspace *ws, const gbs_matrix *l, const gbs_matrix *r,
                             int32_t *d_res, gbs_algo algo, hipStream_t s,
                             bool may_allocate, int32_t *rowsum_buf, size_t rowsum_cap) {
  if (l->ncols != r->ncols)
    return fail(GBS_EINVAL, "gemm: lhs has %llu columns, rhs %llu (depths must match)",
                (unsigned long long)l->ncols, (unsigned long long)r->ncols);
  Operand lo = view(l), ro = view(r);
  const uint64_t K = l->ncols;
  const uint64_t kwords = (K + 63) / 64;
  gbs_algo eng = algo == GBS_ALGO_AUTO ? select_for(l, r) : algo;
  if (eng == GBS_ALGO_MFMA_BITS) return launch_bits_gemm(lo, ro, K, d_res, s);
  if (eng == GBS_ALGO_MFMA_PLANES) return launch_planes_gemm(lo, ro, K, d_res, s);
  if (eng == GBS_ALGO_MFMA || eng == GBS_ALGO_MFMA_FP4 || eng == GBS_ALGO_MFMA_FP6) {
    const int fmt = eng == GBS_ALGO_MFMA_FP4 ? kFmtF4 : eng == GBS_ALGO_MFMA_FP6 ? kFmtF6 : kFmtI8;
    const bool ok = fmt == kFmtF4   ? f4_supported(lo, ro, K)
                    : fmt == kFmtF6 ? f6_supported(lo, ro, K)
                                    : mfma_supported(lo, ro, K);
    if (!ok)
      return fail(GBS_ENOTSUP,
                  "gemm: the %s matrix-core engine cannot hold this product exactly "
                  "(bitwidths %llu x %llu, depth %llu)",
                  fmt == kFmtF4 ? "FP4" : fmt == kFmtF6 ? "FP6" : "int8",
                  (unsigned long long)l->nbits, (unsigned long long)r->nbits,
                  (unsigned long long)K);
    int rc = ensure_copy(const_cast<gbs_matrix *>(l), fmt, s, may_allocate);
    if (!rc) rc = ensure_copy(const_cast<gbs_matrix *>(r), fmt, s, may_allocate);
    if (rc) return rc;
    return launch_mfma_gemm(l, r, d_res, s, fmt);
  }
  if (eng != GBS_ALGO_POPCOUNT) return fail(GBS_EINVAL, "gemm: unknown algo %d", (int)algo);
  const int32_t *d_rowsum = nullptr;
  if (popc_needs_rowsum(lo, ro, K) && l->nrows && r->nrows && kwords) {
    const Operand &reg = lo.bipolar ? ro : lo;
    const size_t need = reg.nrows * sizeof(int32_t);
    int32_t *buf = rowsum_buf;
    if (ws) {
      if (ws->rowsum_cap < need) {
        if (!may_allocate)
          return fail(GBS_ERANGE, "gemm_async: workspace too small, call gbs_workspace_reserve");
        int rc = grow((void **)&ws->d_rowsum, &ws->rowsum_cap, need, s);
        if (rc) return rc;
      }
      buf = ws->d_rowsum;
    } else if (need > rowsum_cap) {
      return fail(GBS_ERANGE, "gemm: the regular operand's %llu rows exceed the %llu of the "
                              "row-sum scratch", (unsigned long long)reg.nrows,
                  (unsigned long long)(rowsum_cap / sizeof(int32_t)));
    }
    int rc = launch_sum_rows(reg, kwords, K, buf, s);
    if (rc) return rc;
    d_rowsum = buf;
  }
  return launch_popc_gemm(lo, ro, kwords, K, d_res, d_rowsum, s);
}

int gemm_enqueue(gbs_workspace *ws, const gbs_matrix *l, const gbs_matrix *r,
                 int32_t *d_res, gbs_algo algo, hipStream_t s, bool may_allocate) {
  return gemm_enqueue_into(ws, l, r, d_res, algo, s, may_allocate, nullptr, 0);
}

// Device address of the current device's scratch array, looked up once per
// device (with that device current).
static int scratch_of_device(int dev, int32_t **out) {
  if (dev < 0 || dev >= kMaxDevices)
    return fail(GBS_ERANGE, "device %d beyond the %d the library tracks", dev, kMaxDevices);
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void *p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return hip_fail(e, "hipGetSymbolAddress(scratch)");
    g_scratch_ptr[dev] = static_cast<int32_t *>(p);
  }
  *out = g_scratch_ptr[dev];
  return GBS_OK;
}

int gemm_enqueue_public(gbs_workspace *ws, const gbs_matrix *l, const gbs_matrix *r,
                        int32_t *d_res, gbs_algo algo, hipStream_t s, bool may_allocate) {
  return gemm_enqueue(ws, l, r, d_res, algo, s, may_allocate);
}

}  // namespace gbs

using namespace gbs;

extern "C" {

int gbs_init(int device) {
  load_default_table();
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  if (e != hipSuccess || n == 0)
    return fail(GBS_EDEVICE, "gbs_init: no HIP device (%s)", hipGetErrorString(e));
  if (device < 0 || device >= n)
    return fail(GBS_EINVAL, "gbs_init: device %d out of range [0, %d)", device, n);
  t_device = device;
  GBS_HIP(hipSetDevice(device));
  if (!thread_stream()) return fail(GBS_EDEVICE, "gbs_init: stream creation failed");
  int32_t *sp = nullptr;
  return scratch_of_device(device, &sp);
}

const char *gbs_last_error(void) { return t_err.c_str(); }

int gbs_version(void) { return 10100; }

int gbs_host_alloc(size_t bytes, void **out) {
  if (!out) return fail(GBS_EINVAL, "gbs_host_alloc: null out");
  *out = nullptr;
  int rc = ensure_device();
  if (rc) return rc;
  GBS_HIP(hipHostMalloc(out, bytes ? bytes : 8, hipHostMallocDefault));
  return GBS_OK;
}

int gbs_host_free(void *p) {
  if (p) GBS_HIP(hipHostFree(p));
  return GBS_OK;
}

int gbs_matrix_create(uint64_t nbits, uint64_t nrows, uint64_t ncols, int issigned,
                      uint64_t nrows_a, uint64_t ncols_a, gbs_matrix **out) {
  if (!out) return fail(GBS_EINVAL, "gbs_matrix_create: null out");
  *out = nullptr;
  if (nbits < 1 || nbits > 8)
    return fail(GBS_EINVAL, "gbs_matrix_create: nbits %llu outside [1, 8]",
                (unsigned long long)nbits);
  if (ncols_a % 64 || nrows_a < nrows || ncols_a < ncols)
    return fail(GBS_EINVAL,
                "gbs_matrix_create: allocation %llu x %llu cannot hold %llu x %llu "
                "(ncols_a must be a multiple of 64)",
                (unsigned long long)nrows_a, (unsigned long long)ncols_a,
                (unsigned long long)nrows, (unsigned long long)ncols);
  int rc = ensure_device();
  if (rc) return rc;
  gbs_matrix *m = new gbs_matrix();
  m->nbits = nbits;
  m->nrows = nrows;
  m->ncols = ncols;
  m->nrows_a = nrows_a;
  m->ncols_a = ncols_a;
  m->issigned = issigned != 0;
  m->device = t_device;
  m->d_planes = nullptr;
  const size_t bytes = m->words() * 8;
  hipError_t e = hipMalloc((void **)&m->d_planes, bytes ? bytes : 8);
  if (e != hipSuccess) {
    delete m;
    return hip_fail(e, "gbs_matrix_create: hipMalloc");
  }
  e = hipMemsetAsync(m->d_planes, 0, bytes ? bytes : 8, thread_stream());
  if (e == hipSuccess) e = hipStreamSynchronize(thread_stream());
  if (e != hipSuccess) {
    (void)hipFree(m->d_planes);
    delete m;
    return hip_fail(e, "gbs_matrix_create: clear");
  }
  *out = m;
  return GBS_OK;
}

int gbs_matrix_destroy(gbs_matrix *m) {
  if (!m) return GBS_OK;
  int rc = ensure_device();
  if (rc) return rc;
  GBS_HIP(hipStreamSynchronize(thread_stream()));
  GBS_HIP(hipFree(m->d_planes));
  if (m->d_vals) GBS_HIP(hipFree(m->d_vals));
  if (m->d_f4) GBS_HIP(hipFree(m->d_f4));
  if (m->d_f6) GBS_HIP(hipFree(m->d_f6));
  delete m;
  return GBS_OK;
}

int gbs_matrix_device_planes(const gbs_matrix *m, uint64_t **d_planes) {
  int rc = check_matrix(m, "gbs_matrix_device_planes");
  if (rc) return rc;
  if (!d_planes) return fail(GBS_EINVAL, "gbs_matrix_device_planes: null out");
  *d_planes = m->d_planes;
  return GBS_OK;
}

// Stage a host source in the thread workspace; returns the device pointer.
static int stage_host(const void *src, size_t bytes, const void **d_src, hipStream_t s) {
  gbs_workspace *ws = thread_ws();
  int rc = grow(&ws->d_stage, &ws->stage_cap, bytes, s);
  if (rc) return rc;
  GBS_HIP(hipMemcpyAsync(ws->d_stage, src, bytes, hipMemcpyHostToDevice, s));
  *d_src = ws->d_stage;
  return GBS_OK;
}

int gbs_import(gbs_matrix *m, const void *src, gbs_dtype dt, int read_col_major,
               int src_on_device, uint64_t *host_planes) {
  int rc = check_matrix(m, "gbs_import");
  if (rc) return rc;
  const size_t es = dtype_size(dt);
  if (es == 0) return fail(GBS_EINVAL, "gbs_import: unknown dtype %d", (int)dt);
  const size_t bytes = (size_t)(m->nrows * m->ncols) * es;
  if (!src && bytes) return fail(GBS_EINVAL, "gbs_import: null source");
  if ((rc = ensure_device())) return rc;
  hipStream_t s = thread_stream();
  const void *d_src = src;
  if (!src_on_device && bytes) {
    if ((rc = stage_host(src, bytes, &d_src, s))) return rc;
  }
  if ((rc = launch_pack(m, d_src, dt, read_col_major != 0, s))) return rc;
  m->planes_version++;
  if (host_planes && m->words())
    GBS_HIP(hipMemcpyAsync(host_planes, m->d_planes, m->words() * 8,
                           hipMemcpyDeviceToHost, s));
  GBS_HIP(hipStreamSynchronize(s));
  return GBS_OK;
}

int gbs_conv2d_shape(const gbs_conv2d *g, uint64_t *rows, uint64_t *cols, uint32_t *out_h,
                     uint32_t *out_w) {
  if (!g) return fail(GBS_EINVAL, "gbs_conv2d_shape: null geometry");
  if (!g->batch || !g->channels || !g->height || !g->width || !g->kh || !g->kw ||
      !g->stride_h || !g->stride_w || !g->dil_h || !g->dil_w)
    return fail(GBS_EINVAL, "gbs_conv2d_shape: sizes, strides and dilations must be >= 1");
  const uint64_t ekh = (uint64_t)g->dil_h * (g->kh - 1) + 1;
  const uint64_t ekw = (uint64_t)g->dil_w * (g->kw - 1) + 1;
  const uint64_t ph = (uint64_t)g->height + 2ull * g->pad_h, pw = (uint64_t)g->width + 2ull * g->pad_w;
  if (ph < ekh || pw < ekw)
    return fail(GBS_EINVAL, "gbs_conv2d_shape: the kernel does not fit the padded input");
  const uint64_t oh = (ph - ekh) / g->stride_h + 1, ow = (pw - ekw) / g->stride_w + 1;
  const uint64_t r = (uint64_t)g->batch * oh * ow, c = (uint64_t)g->channels * g->kh * g->kw;
  if (r >= (1ull << 31) || c >= (1ull << 31) || ph >= (1ull << 31) || pw >= (1ull << 31) ||
      (uint64_t)g->channels * g->height * g->width >= (1ull << 31))
    return fail(GBS_ERANGE, "gbs_conv2d_shape: operand or image dimensions exceed 2^31");
  if (rows) *rows = r;
  if (cols) *cols = c;
  if (out_h) *out_h = (uint32_t)oh;
  if (out_w) *out_w = (uint32_t)ow;
  return GBS_OK;
}

int gbs_import_im2col(gbs_matrix *m, const void *src, gbs_dtype dt, const gbs_conv2d *g,
                      int src_on_device, uint64_t *host_planes) {
  int rc = check_matrix(m, "gbs_import_im2col");
  if (rc) return rc;
  uint64_t rows = 0, cols = 0;
  uint32_t oh = 0, ow = 0;
  if ((rc = gbs_conv2d_shape(g, &rows, &cols, &oh, &ow))) return rc;
  if (m->nrows != rows || m->ncols != cols)
    return fail(GBS_EINVAL,
                "gbs_import_im2col: matrix is %llu x %llu, the convolution needs %llu x %llu",
                (unsigned long long)m->nrows, (unsigned long long)m->ncols,
                (unsigned long long)rows, (unsigned long long)cols);
  const size_t es = dtype_size(dt);
  if (es == 0) return fail(GBS_EINVAL, "gbs_import_im2col: unknown dtype %d", (int)dt);
  const size_t bytes = (size_t)g->batch * g->channels * g->height * g->width * es;
  if (!src) return fail(GBS_EINVAL, "gbs_import_im2col: null source");
  if ((rc = ensure_device())) return rc;
  hipStream_t s = thread_stream();
  const void *d_src = src;
  if (!src_on_device && (rc = stage_host(src, bytes, &d_src, s))) return rc;
  if ((rc = launch_pack_im2col(m, d_src, dt, *g, oh, ow, s))) return rc;
  m->planes_version++;
  if (host_planes && m->words())
    GBS_HIP(hipMemcpyAsync(host_planes, m->d_planes, m->words() * 8,
                           hipMemcpyDeviceToHost, s));
  GBS_HIP(hipStreamSynchronize(s));
  return GBS_OK;
}

static const char kPlaneMagic[8] = {'G', 'B', 'S', 'P', 'L', 'N', '0', '1'};

int gbs_matrix_save(const gbs_matrix *m, const char *path) {
  int rc = check_matrix(m, "gbs_matrix_save");
  if (rc) return rc;
  if (!path) return fail(GBS_EINVAL, "gbs_matrix_save: null path");
  std::vector<uint64_t> words(m->words());
  if (!words.empty()) {
    if ((rc = ensure_device())) return rc;
    hipStream_t s = thread_stream();
    GBS_HIP(hipMemcpyAsync(words.data(), m->d_planes, words.size() * 8, hipMemcpyDeviceToHost,
                           s));
    GBS_HIP(hipStreamSynchronize(s));
  }
  const uint64_t head[7] = {m->nbits, m->nrows, m->ncols, m->nrows_a,
                            m->ncols_a, m->issigned ? 1ull : 0ull, (uint64_t)words.size()};
  FILE *f = fopen(path, "wb");
  if (!f) return fail(GBS_EINVAL, "gbs_matrix_save: cannot create %s", path);
  bool ok = fwrite(kPlaneMagic, 1, 8, f) == 8 && fwrite(head, 8, 7, f) == 7 &&
            (words.empty() || fwrite(words.data(), 8, words.size(), f) == words.size());
  ok = (fclose(f) == 0) && ok;
  if (!ok) return fail(GBS_EINVAL, "gbs_matrix_save: short write to %s", path);
  return GBS_OK;
}

int gbs_matrix_load(const char *path, gbs_matrix **out) {
  if (!out) return fail(GBS_EINVAL, "gbs_matrix_load: null out");
  *out = nullptr;
  if (!path) return fail(GBS_EINVAL, "gbs_matrix_load: null path");
  FILE *f = fopen(path, "rb");
  if (!f) return fail(GBS_EINVAL, "gbs_matrix_load: cannot open %s", path);
  char magic[8];
  uint64_t head[7];
  if (fread(magic, 1, 8, f) != 8 || memcmp(magic, kPlaneMagic, 8) != 0 ||
      fread(head, 8, 7, f) != 7) {
    fclose(f);
    return fail(GBS_EINVAL, "gbs_matrix_load: %s is not a packed-matrix file", path);
  }
  const uint64_t nbits = head[0], nrows = head[1], ncols = head[2], nrows_a = head[3],
                 ncols_a = head[4], issigned = head[5], nwords = head[6];
  const unsigned __int128 want =
      (unsigned __int128)nbits * nrows_a * (ncols_a / 64);
  long end = -1;
  if (fseek(f, 0, SEEK_END) == 0) end = ftell(f);
  if (nbits < 1 || nbits > 8 || issigned > 1 || ncols_a % 64 || nrows > nrows_a ||
      ncols > ncols_a || want != nwords || nwords > (1ull << 40) || end < 0 ||
      (uint64_t)end != 64 + nwords * 8) {
    fclose(f);
    return fail(GBS_EINVAL, "gbs_matrix_load: %s has an inconsistent header or size", path);
  }
  std::vector<uint64_t> words(nwords);
  bool ok = fseek(f, 64, SEEK_SET) == 0 &&
            (nwords == 0 || fread(words.data(), 8, nwords, f) == nwords);
  fclose(f);
  if (!ok) return fail(GBS_EINVAL, "gbs_matrix_load: short read from %s", path);
  gbs_matrix *m = nullptr;
  int rc = gbs_matrix_create(nbits, nrows, ncols, (int)issigned, nrows_a, ncols_a, &m);
  if (rc) return rc;
  if (nwords && (rc = gbs_planes_from_host(m, words.data()))) {
    gbs_matrix_destroy(m);
    return rc;
  }
  *out = m;
  return GBS_OK;
}

int gbs_import_quantize_typed(gbs_matrix *m, const void *src, const void *thresholds,
                              gbs_dtype dt, int nthres, int read_col_major, int src_on_device,
                              uint64_t *host_planes) {
  int rc = check_matrix(m, "gbs_import_quantize_typed");
  if (rc) return rc;
  if (m->issigned)
    return fail(GBS_EINVAL, "gbs_import_quantize_typed: threshold quantization needs an "
                            "unsigned matrix (gemmbitserial.hpp:283)");
  if (nthres < 0) return fail(GBS_EINVAL, "gbs_import_quantize_typed: nthres < 0");
  const size_t es = dtype_size_any(dt);
  if (es == 0) return fail(GBS_EINVAL, "gbs_import_quantize_typed: unknown dtype %d", (int)dt);
  const size_t bytes = (size_t)(m->nrows * m->ncols) * es;
  const size_t tbytes = (size_t)nthres * m->nrows * es;
  if ((bytes && !src) || (tbytes && !thresholds))
    return fail(GBS_EINVAL, "gbs_import_quantize_typed: null source or thresholds");
  if ((rc = ensure_device())) return rc;
  hipStream_t s = thread_stream();
  const void *d_src = src, *d_thr = thresholds;
  if (!src_on_device && (bytes || tbytes)) {
    // both arrays in the thread's staging buffer, thresholds 16-byte aligned
    gbs_workspace *ws = thread_ws();
    const size_t off = (bytes + 15) & ~(size_t)15;
    if ((rc = grow(&ws->d_stage, &ws->stage_cap, off + tbytes, s))) return rc;
    if (bytes) GBS_HIP(hipMemcpyAsync(ws->d_stage, src, bytes, hipMemcpyHostToDevice, s));
    if (tbytes)
      GBS_HIP(hipMemcpyAsync((char *)ws->d_stage + off, thresholds, tbytes,
                             hipMemcpyHostToDevice, s));
    d_src = ws->d_stage;
    d_thr = (const char *)ws->d_stage + off;
  }
  if ((rc = launch_pack_quantize(m, d_src, d_thr, dt, nthres, read_col_major != 0, s)))
    return rc;
  m->planes_version++;
  if (host_planes && m->words())
    GBS_HIP(hipMemcpyAsync(host_planes, m->d_planes, m->words() * 8,
                           hipMemcpyDeviceToHost, s));
  GBS_HIP(hipStreamSynchronize(s));
  return GBS_OK;
}

int gbs_import_quantize(gbs_matrix *m, const int32_t *src, const int32_t *thresholds,
                        int nthres, int read_col_major, int src_on_device,
                        uint64_t *host_planes) {
  return gbs_import_quantize_typed(m, src, thresholds, GBS_I32, nthres, read_col_major,
                                   src_on_device, host_planes);
}

int gbs_planes_from_host(gbs_matrix *m, const uint64_t *host_planes) {
  int rc = check_matrix(m, "gbs_planes_from_host");
  if (rc) return rc;
  if (!m->words()) return GBS_OK;
  if (!host_planes) return fail(GBS_EINVAL, "gbs_planes_from_host: null source");
  if ((rc = ensure_device())) return rc;
  hipStream_t s = thread_stream();
  GBS_HIP(hipMemcpyAsync(m->d_planes, host_planes, m->words() * 8,
                         hipMemcpyHostToDevice, s));
  GBS_HIP(hipStreamSynchronize(s));
  m->planes_version++;
  return GBS_OK;
}

int gbs_matrix_mark_modified(gbs_matrix *m) {
  int rc = check_matrix(m, "gbs_matrix_mark_modified");
  if (rc) return rc;
  m->planes_version++;
  return GBS_OK;
}

int gbs_planes_to_host(const gbs_matrix *m, uint64_t *host_planes) {
  int rc = check_matrix(m, "gbs_planes_to_host");
  if (rc) return rc;
  if (!m->words()) return GBS_OK;
  if (!host_planes) return fail(GBS_EINVAL, "gbs_planes_to_host: null destination");
  if ((rc = ensure_device())) return rc;
  hipStream_t s = thread_stream();
  GBS_HIP(hipMemcpyAsync(host_planes, m->d_planes, m->words() * 8,
                         hipMemcpyDeviceToHost, s));
  GBS_HIP(hipStreamSynchronize(s));
  return GBS_OK;
}

// Run a kernel that writes n int32 to dst (device or host).
}  // extern "C"
template <typename F>
static int to_i32(size_t n, int32_t *dst, int dst_on_device, F &&launch) {
  if (n == 0) return GBS_OK;
  if (!dst) return fail(GBS_EINVAL, "null destination");
  int rc = ensure_device();
  if (rc) return rc;
  hipStream_t s = thread_stream();
  int32_t *d = dst;
  gbs_workspace *ws = thread_ws();
  if (!dst_on_device) {
    if ((rc = grow((void **)&ws->d_res, &ws->res_cap, n * 4, s))) return rc;
    d = ws->d_res;
  }
  if ((rc = launch(d, s))) return rc;
  if (!dst_on_device) GBS_HIP(hipMemcpyAsync(dst, d, n * 4, hipMemcpyDeviceToHost, s));
  GBS_HIP(hipStreamSynchronize(s));
  return GBS_OK;
}
extern "C" {

int gbs_export(const gbs_matrix *m, int32_t *dst, int dst_on_device) {
  int rc = check_matrix(m, "gbs_export");
  if (rc) return rc;
  return to_i32(m->nrows * m->ncols, dst, dst_on_device,
                [&](int32_t *d, hipStream_t s) { return launch_export(m, d, s); });
}

int gbs_sum_rows(const gbs_matrix *m, int32_t *dst, int dst_on_device) {
  int rc = check_matrix(m, "gbs_sum_rows");
  if (rc) return rc;
  return to_i32(m->nrows, dst, dst_on_device, [&](int32_t *d, hipStream_t s) {
    return launch_sum_rows(view(m), m->kwords(), m->ncols, d, s);
  });
}

int gbs_workspace_create(gbs_workspace **out) {
  if (!out) return fail(GBS_EINVAL, "gbs_workspace_create: null out");
  int rc = ensure_device();
  if (rc) return rc;
  *out = new gbs_workspace();
  (*out)->device = t_device;
  return GBS_OK;
}

int gbs_workspace_destroy(gbs_workspace *ws) {
  if (!ws) return GBS_OK;
  int rc = ensure_device();
  if (rc) return rc;
  GBS_HIP(hipDeviceSynchronize());
  if (ws->d_res) GBS_HIP(hipFree(ws->d_res));
  if (ws->d_rowsum) GBS_HIP(hipFree(ws->d_rowsum));
  if (ws->d_stage) GBS_HIP(hipFree(ws->d_stage));
  delete ws;
  return GBS_OK;
}

int gbs_workspace_reserve(gbs_workspace *ws, const gbs_matrix *lhs, const gbs_matrix *rhs) {
  if (!ws) ws = thread_ws();
  int rc = check_matrix(lhs, "gbs_workspace_reserve");
  if (!rc) rc = check_matrix(rhs, "gbs_workspace_reserve");
  if (rc) return rc;
  if ((rc = ensure_device())) return rc;
  const size_t rows = lhs->nrows > rhs->nrows ? lhs->nrows : rhs->nrows;
  if ((rc = grow((void **)&ws->d_rowsum, &ws->rowsum_cap, rows * 4, thread_stream())))
    return rc;
  if (lhs->ncols == rhs->ncols) {
    hipStream_t s = thread_stream();
    if (mfma_supported(view(lhs), view(rhs), lhs->ncols)) {
      if ((rc = ensure_vals(const_cast<gbs_matrix *>(lhs), s, true))) return rc;
      if ((rc = ensure_vals(const_cast<gbs_matrix *>(rhs), s, true))) return rc;
    }
    if (f4_supported(view(lhs), view(rhs), lhs->ncols)) {
      if ((rc = ensure_f4(const_cast<gbs_matrix *>(lhs), s, true))) return rc;
      if ((rc = ensure_f4(const_cast<gbs_matrix *>(rhs), s, true))) return rc;
    }
    if (f6_supported(view(lhs), view(rhs), lhs->ncols)) {
      if ((rc = ensure_f6(const_cast<gbs_matrix *>(lhs), s, true))) return rc;
      if ((rc = ensure_f6(const_cast<gbs_matrix *>(rhs), s, true))) return rc;
    }
    GBS_HIP(hipStreamSynchronize(s));
  }
  return GBS_OK;
}

int gbs_gemm(gbs_workspace *ws, const gbs_matrix *lhs, const gbs_matrix *rhs,
             int32_t *res, int res_on_device, gbs_algo algo) {
  int rc = check_matrix(lhs, "gbs_gemm");
  if (!rc) rc = check_matrix(rhs, "gbs_gemm");
  if (rc) return rc;
  if (!ws) ws = thread_ws();
  const size_t n = (size_t)(lhs->nrows * rhs->nrows);
  if (n && !res) return fail(GBS_EINVAL, "gbs_gemm: null result pointer");
  if ((rc = ensure_device())) return rc;
  hipStream_t s = thread_stream();
  int32_t *d = res;
  if (!res_on_device && n) {
    if ((rc = grow((void **)&ws->d_res, &ws->res_cap, n * 4, s))) return rc;
    d = ws->d_res;
  }
  if ((rc = gemm_enqueue(ws, lhs, rhs, d, algo, s, true))) return rc;
  if (!res_on_device && n) GBS_HIP(hipMemcpyAsync(res, d, n * 4, hipMemcpyDeviceToHost, s));
  GBS_HIP(hipStreamSynchronize(s));
  return GBS_OK;
}

int gbs_gemm_async(gbs_workspace *ws, const gbs_matrix *lhs, const gbs_matrix *rhs,
                   int32_t *d_res, gbs_algo algo, void *stream) {
  int rc = check_matrix(lhs, "gbs_gemm_async");
  if (!rc) rc = check_matrix(rhs, "gbs_gemm_async");
  if (rc) return rc;
  if (!ws) ws = thread_ws();
  if (lhs->nrows * rhs->nrows && !d_res)
    return fail(GBS_EINVAL, "gbs_gemm_async: null result pointer");
  if ((rc = ensure_device())) return rc;
  hipStream_t s = stream ? (hipStream_t)stream : thread_stream();
  return gemm_enqueue(ws, lhs, rhs, d_res, algo, s, false);
}

int gbs_gemm_async_noalloc(const gbs_matrix *lhs, const gbs_matrix *rhs, int32_t *d_res,
                           gbs_algo algo, void *stream) {
  int rc = check_matrix(lhs, "gbs_gemm_async_noalloc");
  if (!rc) rc = check_matrix(rhs, "gbs_gemm_async_noalloc");
  if (rc) return rc;
  if (lhs->nrows * rhs->nrows && !d_res)
    return fail(GBS_EINVAL, "gbs_gemm_async_noalloc: null result pointer");
  if (lhs->ncols != rhs->ncols)
    return fail(GBS_EINVAL, "gbs_gemm_async_noalloc: depths differ (%llu, %llu)",
                (unsigned long long)lhs->ncols, (unsigned long long)rhs->ncols);
  if ((rc = ensure_device())) return rc;
  int32_t *sp = nullptr;
  if ((rc = scratch_of_device(t_device, &sp))) return rc;
  hipStream_t s = stream ? (hipStream_t)stream : thread_stream();
  gbs_algo eng = algo;
  if (algo == GBS_ALGO_AUTO) {
    // AUTO among the engines that need nothing allocated: a matrix-core
    // engine whose value copies both operands already hold, else the
    // engines that read the planes themselves
    const Operand lo = view(lhs), ro = view(rhs);
    eng = select_for(lhs, rhs);
    const bool have =
        eng == GBS_ALGO_MFMA_FP4   ? (lhs->d_f4 && rhs->d_f4)
        : eng == GBS_ALGO_MFMA_FP6 ? (lhs->d_f6 && rhs->d_f6)
        : eng == GBS_ALGO_MFMA     ? (lhs->d_vals && rhs->d_vals)
                                   : true;
    if (!have) eng = bits_supported(lo, ro, lhs->ncols) ? GBS_ALGO_MFMA_BITS : GBS_ALGO_POPCOUNT;
  }
  return gemm_enqueue_into(nullptr, lhs, rhs, d_res, eng, s, false, sp,
                           kScratchInts * sizeof(int32_t));
}

int gbs_gemm_select(const gbs_matrix *lhs, const gbs_matrix *rhs, gbs_algo *chosen) {
  int rc = check_matrix(lhs, "gbs_gemm_select");
  if (!rc) rc = check_matrix(rhs, "gbs_gemm_select");
  if (rc) return rc;
  if (!chosen) return fail(GBS_EINVAL, "gbs_gemm_select: null out");
  *chosen = select_for(lhs, rhs);
  return GBS_OK;
}

int gbs_stream(void **stream) {
  if (!stream) return fail(GBS_EINVAL, "gbs_stream: null out");
  int rc = ensure_device();
  if (rc) return rc;
  *stream = (void *)thread_stream();
  return *stream ? GBS_OK : fail(GBS_EDEVICE, "gbs_stream: creation failed");
}

int gbs_stream_sync(void *stream) {
  int rc = ensure_device();
  if (rc) return rc;
  GBS_HIP(hipStreamSynchronize(stream ? (hipStream_t)stream : thread_stream()));
  return GBS_OK;
}

int gbs_event_create(void **ev) {
  if (!ev) return fail(GBS_EINVAL, "gbs_event_create: null out");
  int rc = ensure_device();
  if (rc) return rc;
  hipEvent_t e;
  GBS_HIP(hipEventCreate(&e));
  *ev = (void *)e;
  return GBS_OK;
}

int gbs_event_destroy(void *ev) {
  if (ev) GBS_HIP(hipEventDestroy((hipEvent_t)ev));
  return GBS_OK;
}

int gbs_event_record(void *ev, void *stream) {
  int rc = ensure_device();
  if (rc) return rc;
  GBS_HIP(hipEventRecord((hipEvent_t)ev, stream ? (hipStream_t)stream : thread_stream()));
  return GBS_OK;
}

int gbs_event_elapsed_ms(void *start, void *stop, float *ms) {
  if (!ms) return fail(GBS_EINVAL, "gbs_event_elapsed_ms: null out");
  GBS_HIP(hipEventSynchronize((hipEvent_t)stop));
  GBS_HIP(hipEventElapsedTime(ms, (hipEvent_t)start, (hipEvent_t)stop));
  return GBS_OK;
}

int gbs_gemm_async_config(gbs_workspace *ws, const gbs_matrix *lhs, const gbs_matrix *rhs,
                          int32_t *d_res, gbs_algo algo, int shape, int splits, void *stream) {
  int rc = check_matrix(lhs, "gbs_gemm_async_config");
  if (!rc) rc = check_matrix(rhs, "gbs_gemm_async_config");
  if (rc) return rc;
  (void)ws;
  if (lhs->ncols != rhs->ncols) return fail(GBS_EINVAL, "gbs_gemm_async_config: depths differ");
  if (algo != GBS_ALGO_MFMA && algo != GBS_ALGO_MFMA_FP4 && algo != GBS_ALGO_MFMA_FP6)
    return fail(GBS_EINVAL, "gbs_gemm_async_config: algo must be a matrix-core engine");
  const int fmt = algo == GBS_ALGO_MFMA_FP4 ? kFmtF4 : algo == GBS_ALGO_MFMA_FP6 ? kFmtF6 : kFmtI8;
  const Operand lo = view(lhs), ro = view(rhs);
  const bool ok = fmt == kFmtF4   ? f4_supported(lo, ro, lhs->ncols)
                  : fmt == kFmtF6 ? f6_supported(lo, ro, lhs->ncols)
                                  : mfma_supported(lo, ro, lhs->ncols);
  if (!ok) return fail(GBS_ENOTSUP, "gbs_gemm_async_config: not exact in this operand format");
  if (shape < 0 || shape >= kMfmaShapes || splits < 1)
    return fail(GBS_EINVAL, "gbs_gemm_async_config: shape %d / splits %d out of range", shape,
                splits);
  if (lhs->nrows * rhs->nrows && !d_res)
    return fail(GBS_EINVAL, "gbs_gemm_async_config: null result pointer");
  if ((rc = ensure_device())) return rc;
  hipStream_t s = stream ? (hipStream_t)stream : thread_stream();
  if ((rc = ensure_copy(const_cast<gbs_matrix *>(lhs), fmt, s, false))) return rc;
  if ((rc = ensure_copy(const_cast<gbs_matrix *>(rhs), fmt, s, false))) return rc;
  return launch_mfma_gemm(lhs, rhs, d_res, s, fmt, shape, true, splits);
}

int gbs_tune_classes(int *n_classes, int *n_shapes) {
  if (n_classes) *n_classes = kClasses;
  if (n_shapes) *n_shapes = kMfmaShapes;
  return GBS_OK;
}

int gbs_tune_class_of(uint64_t M, uint64_t N, uint64_t K, int *cls) {
  if (!cls) return fail(GBS_EINVAL, "gbs_tune_class_of: null out");
  *cls = shape_class(M, N, K);
  return GBS_OK;
}

int gbs_tune_set(const int *shapes, const int *splits, int n) {
  return set_mfma_table(shapes, splits, n);
}

int gbs_tune_load(const char *path) {
  if (!path) return fail(GBS_EINVAL, "gbs_tune_load: null path");
  FILE *f = fopen(path, "r");
  if (!f) return fail(GBS_EINVAL, "gbs_tune_load: cannot open %s", path);
  int shapes[kMfmaFormats * kClasses], splits[kMfmaFormats * kClasses];
  float times[kMfmaFormats * kClasses];
  for (int i = 0; i < kMfmaFormats * kClasses; i++) {
    shapes[i] = -1;
    splits[i] = 0;
    times[i] = 0.f;
  }
  char line[512];
  int n = 0;
  while (fgets(line, sizeof(line), f)) {
    int c, sh, sp = 0, fmt = 0;
    float us = 0.f;
    if (line[0] == '#') continue;
    const int got = sscanf(line, "%d %d %d %d %f", &c, &sh, &sp, &fmt, &us);
    if (got >= 2 && c >= 0 && c < kClasses && sh >= 0 && sh < kMfmaShapes && fmt >= 0 &&
        fmt < kMfmaFormats) {
      shapes[fmt * kClasses + c] = sh;
      splits[fmt * kClasses + c] = (got >= 3 && sp > 0) ? sp : 0;
      times[fmt * kClasses + c] = (got >= 5 && us > 0.f) ? us : 0.f;
      n++;
    }
  }
  fclose(f);
  if (n == 0) return fail(GBS_EINVAL, "gbs_tune_load: no entries in %s", path);
  int rc = set_mfma_table(shapes, splits, kMfmaFormats * kClasses);
  if (!rc) rc = set_mfma_times(times, kMfmaFormats * kClasses);
  return rc;
}

}  // extern "C"
