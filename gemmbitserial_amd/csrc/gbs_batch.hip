// gbs_batch.hip -- many independent bit-serial GEMMs (uneven M/K/N, mixed
// bitwidths and signedness) in one persistent kernel launch.
//
// Plan (gbs_batch_create): every product that the matrix-core engine holds
// exactly becomes a run of work items in the narrowest copy format both
// operands fit (FP4, else FP6, else int8: inside one launch an item's time
// follows the bytes of K it streams) -- output tiles of a shape picked by
// the rhs row count (16 x 256 for N <= 16, 32 x 256 for N <= 32, else
// 64 x 64: the least padding, and no grid to fill), times K splits (deep
// products are cut into pieces of at most 4 KiB of nominal K, whose
// partial sums are added with integer atomics into results zeroed first;
// measured: 2 KiB pieces lose 6 % on the 512-product workload to pipeline
// fill per item, 8 KiB pieces leave too few items for GEMV-heavy lists).  Products are ordered by
// the cost of one item, largest first, so the long items start early and
// the short ones fill the tail.  A device array holds one descriptor per product and
// the prefix sum of tile counts.  Products the MFMA engine cannot hold
// exactly (unsigned 8-bit codes, very deep K) go to the popcount engine as
// separate launches after the persistent kernel.
//
// Products are grouped by copy format (int8, FP4, FP6); each group has its
// own descriptor range, tile prefix sums and queue counter, and its own
// persistent kernel (one kernel carrying every format's tile code would
// spill registers).
//
// Run: one small kernel zeroes the queue counters (and the results of
// split products, which their pieces add to atomically), then per format group a
// grid of 2 blocks x 256 CUs loops: thread 0 takes the next tile index with
// one atomicAdd, the block finds its product in a per-tile product table
// built at plan time (one load; launches planned per call, like the serving
// loop's, binary-search the prefix sums instead) and computes that tile with
// the shape's code path.  Every
// block leaves when the index passes the group's last tile.  The whole
// sequence can be captured once into a hipGraph and replayed.
#include <algorithm>
#include <vector>

#include "gbs_mfma_tile.h"

struct gbs_batch {
  int device = 0;
  std::vector<gbs_problem> probs;          // as given
  std::vector<uint32_t> mfma_idx;          // persistent products, in run order
  std::vector<uint8_t> mfma_fmt;           // their copy formats
  std::vector<uint32_t> separate_idx;      // popcount products
  void *d_desc = nullptr;                  // BatchProb[], grouped by format
  void *d_zero = nullptr;                  // ZeroSpan[] of split products
  uint32_t n_zero = 0;
  unsigned int *d_counter = nullptr;       // one queue counter per format
  uint32_t *d_item_prod = nullptr;         // product of every tile, per group
  uint64_t total_tiles = 0;
  uint32_t first[3] = {0, 0, 0}, count[3] = {0, 0, 0};  // descriptor range per format
  uint64_t tiles[3] = {0, 0, 0};
  uint64_t item0[3] = {0, 0, 0};            // group's first entry in d_item_prod
  uint32_t grid[3] = {0, 0, 0};
  gbs_workspace *ws = nullptr;             // for separate products
  hipGraphExec_t exec = nullptr;
  hipGraph_t graph = nullptr;
  hipStream_t graph_stream = nullptr;
};

namespace gbs {
namespace {

struct BatchProb {
  MfmaOperands ops;
  MfmaPlan plan;
  uint64_t tile_begin;
};

struct ZeroSpan {
  int32_t *p;
  uint64_t n;
};

__global__ void reset_counter_kernel(unsigned int *c) {
  if (threadIdx.x < 3) c[threadIdx.x] = 0;
}

// Zero the results of split products (their partial sums are added
// atomically); blockIdx.y picks the span.  Block (0, 0) also resets the
// queue counters, which saves the launch of reset_counter_kernel.
__global__ __launch_bounds__(256) void zero_spans_kernel(const ZeroSpan *spans,
                                                         unsigned int *counters) {
  if (blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x < 3) counters[threadIdx.x] = 0;
  const ZeroSpan z = spans[blockIdx.y];
  for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < z.n;
       i += (uint64_t)gridDim.x * 256)
    z.p[i] = 0;
}

template <int FMT>
__global__ __launch_bounds__(256, 2) void batch_kernel(const BatchProb *probs, uint32_t nprobs,
                                                       uint64_t total, unsigned int *counter,
                                                       const uint32_t *item_prod) {
  __shared__ __attribute__((aligned(16))) int8_t lds[kMfmaLdsMax];
  __shared__ uint32_t s_item;
  for (;;) {
    if (threadIdx.x == 0) s_item = atomicAdd(counter, 1u);
    __syncthreads();
    const uint64_t item = s_item;
    __syncthreads();
    if (item >= total) break;
    uint32_t lo = 0;
    if (item_prod) {
      lo = item_prod[item];
    } else {
      // largest p with tile_begin <= item
      uint32_t hi = nprobs - 1;
      while (lo < hi) {
        const uint32_t mid = (lo + hi + 1) / 2;
        if (probs[mid].tile_begin <= item) lo = mid;
        else hi = mid - 1;
      }
    }
    const MfmaOperands ops = probs[lo].ops;
    const MfmaPlan pl = probs[lo].plan;
    uint32_t pt, qt, sp;
    plan_decode(pl, item - probs[lo].tile_begin, pt, qt, sp);
    const uint32_t kbeg = sp * pl.kchunk;
    const uint32_t kend = min(ops.kbytes, kbeg + pl.kchunk);
    mfma_tile_by_shape<FMT>(pl.shape, ops, pt, qt, kbeg, kend, pl.splits > 1, lds);
    __syncthreads();
  }
}

}  // namespace

// Plan of one product inside a persistent launch: the tile shape with the
// least padding for its rhs row count (the tuning table's shapes are
// chosen to fill the chip with one product, which a shared queue does not
// need), and deep products cut into pieces of at most 4 KiB (nominal) of K
// so that no item is much longer than the rest -- or into up to
// min_pieces pieces of at least 1 KiB, for products that would otherwise
// dominate the queue.
static MfmaPlan batch_plan(uint32_t M, uint32_t N, uint32_t kbytes, int fmt,
                           uint32_t min_pieces = 1) {
  const int shape = batch_shape(N <= 16 ? 3 : N <= 32 ? 2 : 4, fmt);
  const uint32_t units = kbytes / kMfmaKS;
  uint32_t want = (units + 15) / 16;
  const uint32_t most = units / 4 ? units / 4 : 1;  // pieces of at least 1 KiB
  if (want < min_pieces) want = min_pieces < most ? min_pieces : most;
  return mfma_plan(M, N, kbytes, fmt, shape, (int)(want ? want : 1));
}

static int step_bytes_of(int fmt) { return fmt == kFmtF6 ? step_bytes<kFmtF6>() : 64; }

// Operand bytes one work item of a plan streams (its tile's rows times its
// K piece, in the copy format's real bytes).
static double item_bytes(const MfmaPlan &pl) {
  const MfmaTileShape sh = kMfmaShape[pl.shape];
  return (double)(sh.P + sh.Q) * pl.kchunk * step_bytes_of(pl.fmt) / 64.0;
}

// Balance: a product whose items each stream more than 1/1024 of the whole
// launch's bytes (two per queue slot) would leave its last items running
// alone -- fully connected GEMVs are the case; it is re-planned with
// proportionally more K pieces (of at least 1 KiB).
static void rebalance(BatchProb *d, int n, double total_bytes) {
  const double target = total_bytes / 1024;
  for (int i = 0; i < n; i++) {
    const MfmaPlan &pl = d[i].plan;
    if (target <= 0 || item_bytes(pl) <= target) continue;
    const double per_tile = item_bytes(pl) * pl.splits;  // one tile, all of K
    d[i].plan = batch_plan(d[i].ops.M, d[i].ops.N, d[i].ops.kbytes, pl.fmt,
                           (uint32_t)std::min(per_tile / target + 1, 1e6));
  }
}

// The serving loop's per-step batch: at most kMaxViews products whose
// descriptors travel in the kernel arguments (no copy ahead of the
// launch); the caller zeroes the queue counter after the launch, with
// work it launches next anyway; split products add their pieces
// atomically into results the caller keeps zero between launches.
constexpr int kMaxViews = 16;
struct ViewBatch {
  BatchProb p[kMaxViews];
  uint32_t n;
  uint64_t total;
};

__global__ __launch_bounds__(256, 2) void views_kernel(ViewBatch V, unsigned int *counters) {
  __shared__ __attribute__((aligned(16))) int8_t lds[kMfmaLdsMax];
  __shared__ uint32_t s_item;
  for (;;) {
    if (threadIdx.x == 0) s_item = atomicAdd(counters, 1u);
    __syncthreads();
    const uint64_t item = s_item;
    __syncthreads();
    if (item >= V.total) break;
    uint32_t lo = 0, hi = V.n - 1;  // largest p with tile_begin <= item
    while (lo < hi) {
      const uint32_t mid = (lo + hi + 1) / 2;
      if (V.p[mid].tile_begin <= item) lo = mid;
      else hi = mid - 1;
    }
    const MfmaOperands ops = V.p[lo].ops;
    const MfmaPlan pl = V.p[lo].plan;
    uint32_t pt, qt, sp;
    plan_decode(pl, item - V.p[lo].tile_begin, pt, qt, sp);
    const uint32_t kbeg = sp * pl.kchunk;
    const uint32_t kend = min(ops.kbytes, kbeg + pl.kchunk);
    mfma_tile_by_shape<kFmtI8>(pl.shape, ops, pt, qt, kbeg, kend, pl.splits > 1, lds);
    __syncthreads();
  }
}

// Plan and launch products given as int8 operand views.  *counter must be
// zero at the launch; the caller zeroes it again after the launch (in
// stream order).  The results of products the plan splits along K must be
// zero before the launch.
int launch_views(const MfmaOperands *views, int n, unsigned int *counter, hipStream_t s) {
  if (n > kMaxViews) return fail(GBS_EINVAL, "launch_views: %d products (at most %d)", n, kMaxViews);
  ViewBatch V{};
  BatchProb *h = V.p;
  int m = 0;
  double total = 0;
  for (int i = 0; i < n; i++) {
    const MfmaOperands &v = views[i];
    if (v.M == 0 || v.N == 0) continue;
    BatchProb d{};
    d.ops = v;
    d.plan = batch_plan(v.M, v.N, v.kbytes, kFmtI8);
    total += item_bytes(d.plan) * (double)d.plan.tiles();
    h[m++] = d;
  }
  // a step of few GEMV-like products would otherwise run on a few CUs:
  // the same re-split as a planned batch, then longest items first
  rebalance(h, m, total);
  std::sort(h, h + m, [](const BatchProb &x, const BatchProb &y) {
    return item_bytes(x.plan) > item_bytes(y.plan);
  });
  uint64_t tiles = 0;
  for (int i = 0; i < m; i++) {
    h[i].tile_begin = tiles;
    tiles += h[i].plan.tiles();
  }
  if (tiles == 0) return GBS_OK;
  V.n = (uint32_t)m;
  V.total = tiles;
  const uint32_t grid = (uint32_t)(tiles < 512 ? tiles : 512);
  hipLaunchKernelGGL(views_kernel, dim3(grid), dim3(256), 0, s, V, counter);
  GBS_HIP(hipGetLastError());
  return GBS_OK;
}


// Operands re-imported since the plan get their copies rebuilt before a
// run, on the run's stream.  The plan stays valid: a copy buffer, once
// allocated, keeps its address and layout for the matrix's lifetime.
static int refresh_copies(gbs_batch *b, hipStream_t s) {
  for (size_t k = 0; k < b->mfma_idx.size(); k++) {
    const gbs_problem &p = b->probs[b->mfma_idx[k]];
    int rc = ensure_copy(const_cast<gbs_matrix *>(p.lhs), b->mfma_fmt[k], s, false);
    if (!rc) rc = ensure_copy(const_cast<gbs_matrix *>(p.rhs), b->mfma_fmt[k], s, false);
    if (rc) return rc;
  }
  return GBS_OK;
}

static int batch_launch(gbs_batch *b, hipStream_t s) {
  if (b->total_tiles) {
    if (b->n_zero)
      hipLaunchKernelGGL(zero_spans_kernel, dim3(64, b->n_zero), dim3(256), 0, s,
                         (const ZeroSpan *)b->d_zero, b->d_counter);
    else
      hipLaunchKernelGGL(reset_counter_kernel, dim3(1), dim3(64), 0, s, b->d_counter);
    GBS_HIP(hipGetLastError());
    const BatchProb *desc = (const BatchProb *)b->d_desc;
    if (b->tiles[kFmtI8])
      hipLaunchKernelGGL((batch_kernel<kFmtI8>), dim3(b->grid[kFmtI8]), dim3(256), 0, s,
                         desc + b->first[kFmtI8], b->count[kFmtI8], b->tiles[kFmtI8],
                         b->d_counter + kFmtI8, b->d_item_prod + b->item0[kFmtI8]);
    if (b->tiles[kFmtF4])
      hipLaunchKernelGGL((batch_kernel<kFmtF4>), dim3(b->grid[kFmtF4]), dim3(256), 0, s,
                         desc + b->first[kFmtF4], b->count[kFmtF4], b->tiles[kFmtF4],
                         b->d_counter + kFmtF4, b->d_item_prod + b->item0[kFmtF4]);
    if (b->tiles[kFmtF6])
      hipLaunchKernelGGL((batch_kernel<kFmtF6>), dim3(b->grid[kFmtF6]), dim3(256), 0, s,
                         desc + b->first[kFmtF6], b->count[kFmtF6], b->tiles[kFmtF6],
                         b->d_counter + kFmtF6, b->d_item_prod + b->item0[kFmtF6]);
    GBS_HIP(hipGetLastError());
  }
  for (uint32_t i : b->separate_idx) {
    const gbs_problem &p = b->probs[i];
    int rc = gemm_enqueue_public(b->ws, p.lhs, p.rhs, p.d_res, GBS_ALGO_POPCOUNT, s, false);
    if (rc) return rc;
  }
  return GBS_OK;
}

}  // namespace gbs

using namespace gbs;

extern "C" {

int gbs_batch_create(const gbs_problem *problems, uint64_t count, gbs_batch **out) {
  if (!out) return fail(GBS_EINVAL, "gbs_batch_create: null out");
  *out = nullptr;
  if (count && !problems) return fail(GBS_EINVAL, "gbs_batch_create: null problems");
  int rc = ensure_device();
  if (rc) return rc;
  hipStream_t s = thread_stream();
  gbs_batch *b = new gbs_batch();
  b->probs.assign(problems, problems + count);
  if ((rc = gbs_workspace_create(&b->ws))) {
    delete b;
    return rc;
  }
  struct Item { uint32_t idx; double item_cost; MfmaPlan plan; MfmaOperands ops; };
  std::vector<Item> items;
  for (uint64_t i = 0; i < count; i++) {
    const gbs_problem &p = problems[i];
    if (!p.lhs || !p.rhs) {
      gbs_batch_destroy(b);
      return fail(GBS_EINVAL, "gbs_batch_create: problem %llu has a null matrix",
                  (unsigned long long)i);
    }
    if (p.lhs->ncols != p.rhs->ncols) {
      gbs_batch_destroy(b);
      return fail(GBS_EINVAL, "gbs_batch_create: problem %llu depths differ (%llu vs %llu)",
                  (unsigned long long)i, (unsigned long long)p.lhs->ncols,
                  (unsigned long long)p.rhs->ncols);
    }
    if (p.lhs->nrows * p.rhs->nrows && !p.d_res) {
      gbs_batch_destroy(b);
      return fail(GBS_EINVAL, "gbs_batch_create: problem %llu has no result buffer",
                  (unsigned long long)i);
    }
    if (p.lhs->nrows == 0 || p.rhs->nrows == 0) continue;
    // room for any engine the per-call baseline may pick
    if ((rc = gbs_workspace_reserve(b->ws, p.lhs, p.rhs))) {
      gbs_batch_destroy(b);
      return rc;
    }
    const Operand lo = view(p.lhs), ro = view(p.rhs);
    const int fmt = f4_supported(lo, ro, p.lhs->ncols)   ? kFmtF4
                    : f6_supported(lo, ro, p.lhs->ncols) ? kFmtF6
                    : mfma_supported(lo, ro, p.lhs->ncols) ? kFmtI8
                                                           : -1;
    if (fmt >= 0 && p.lhs->ncols > 0) {
      rc = ensure_copy(const_cast<gbs_matrix *>(p.lhs), fmt, s, true);
      if (!rc) rc = ensure_copy(const_cast<gbs_matrix *>(p.rhs), fmt, s, true);
      if (rc) {
        gbs_batch_destroy(b);
        return rc;
      }
      const MfmaOperands ops = mfma_operands(p.lhs, p.rhs, fmt, p.d_res);
      const MfmaPlan pl = batch_plan(ops.M, ops.N, ops.kbytes, fmt);
      const MfmaTileShape sh = kMfmaShape[pl.shape];
      items.push_back({(uint32_t)i, (double)sh.P * sh.Q * (double)pl.kchunk, pl, ops});
    } else {
      b->separate_idx.push_back((uint32_t)i);
    }
  }
  // re-split products whose items would dominate the queue (rebalance)
  {
    double total = 0;
    std::vector<BatchProb> tmp(items.size());
    for (size_t i = 0; i < items.size(); i++) {
      tmp[i].ops = items[i].ops;
      tmp[i].plan = items[i].plan;
      total += item_bytes(items[i].plan) * (double)items[i].plan.tiles();
    }
    rebalance(tmp.data(), (int)tmp.size(), total);
    for (size_t i = 0; i < items.size(); i++) {
      const MfmaTileShape sh = kMfmaShape[tmp[i].plan.shape];
      items[i].plan = tmp[i].plan;
      items[i].item_cost = (double)sh.P * sh.Q * (double)tmp[i].plan.kchunk;
    }
  }
  // grouped by format; inside a group the longest items first: they start
  // early and the short ones fill the tail
  std::stable_sort(items.begin(), items.end(), [](const Item &x, const Item &y) {
    return x.plan.fmt != y.plan.fmt ? x.plan.fmt < y.plan.fmt : x.item_cost > y.item_cost;
  });
  std::vector<BatchProb> desc;
  std::vector<ZeroSpan> zero;
  const uint64_t slots = 256ull * 2;
  for (const Item &it : items) {
    const gbs_problem &p = problems[it.idx];
    const int f = it.plan.fmt;
    if (b->count[f] == 0) b->first[f] = (uint32_t)desc.size();
    BatchProb d{};
    d.ops = it.ops;
    d.plan = it.plan;
    d.tile_begin = b->tiles[f];
    b->tiles[f] += d.plan.tiles();
    b->count[f]++;
    if (d.plan.splits > 1) zero.push_back(ZeroSpan{p.d_res, (uint64_t)d.ops.M * d.ops.N});
    desc.push_back(d);
    b->mfma_idx.push_back(it.idx);
    b->mfma_fmt.push_back((uint8_t)f);
  }
  for (int f = 0; f < 3; f++) {
    b->item0[f] = b->total_tiles;
    b->total_tiles += b->tiles[f];
    b->grid[f] = (uint32_t)(b->tiles[f] < slots ? b->tiles[f] : slots);
  }
  if (b->total_tiles >= (1ull << 32)) {
    gbs_batch_destroy(b);
    return fail(GBS_ERANGE, "gbs_batch_create: more than 2^32 tiles");
  }
  // product (index within its format group) of every tile
  std::vector<uint32_t> item_prod(b->total_tiles);
  for (int f = 0; f < 3; f++)
    for (uint32_t k = 0; k < b->count[f]; k++) {
      const BatchProb &d = desc[b->first[f] + k];
      const uint64_t at = b->item0[f] + d.tile_begin;
      std::fill(item_prod.begin() + at, item_prod.begin() + at + d.plan.tiles(), k);
    }
  if (!desc.empty()) {
    hipError_t e = hipMalloc(&b->d_desc, desc.size() * sizeof(BatchProb));
    if (e == hipSuccess) e = hipMalloc((void **)&b->d_counter, 256);
    if (e == hipSuccess) e = hipMalloc((void **)&b->d_item_prod, item_prod.size() * 4);
    if (e == hipSuccess)
      e = hipMemcpyAsync(b->d_item_prod, item_prod.data(), item_prod.size() * 4,
                         hipMemcpyHostToDevice, s);
    if (e == hipSuccess && !zero.empty()) {
      e = hipMalloc(&b->d_zero, zero.size() * sizeof(ZeroSpan));
      if (e == hipSuccess)
        e = hipMemcpyAsync(b->d_zero, zero.data(), zero.size() * sizeof(ZeroSpan),
                           hipMemcpyHostToDevice, s);
      b->n_zero = (uint32_t)zero.size();
    }
    if (e == hipSuccess)
      e = hipMemcpyAsync(b->d_desc, desc.data(), desc.size() * sizeof(BatchProb),
                         hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) {
      gbs_batch_destroy(b);
      return hip_fail(e, "gbs_batch_create");
    }
  } else {
    hipError_t e = hipStreamSynchronize(s);
    if (e != hipSuccess) {
      gbs_batch_destroy(b);
      return hip_fail(e, "gbs_batch_create");
    }
  }
  b->device = 0;
  *out = b;
  return GBS_OK;
}

int gbs_batch_destroy(gbs_batch *b) {
  if (!b) return GBS_OK;
  int rc = ensure_device();
  if (rc) return rc;
  (void)hipDeviceSynchronize();
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  if (b->d_desc) (void)hipFree(b->d_desc);
  if (b->d_counter) (void)hipFree(b->d_counter);
  if (b->d_item_prod) (void)hipFree(b->d_item_prod);
  if (b->d_zero) (void)hipFree(b->d_zero);
  if (b->ws) gbs_workspace_destroy(b->ws);
  delete b;
  return GBS_OK;
}

int gbs_batch_run(gbs_batch *b, void *stream) {
  if (!b) return fail(GBS_EINVAL, "gbs_batch_run: null batch");
  int rc = ensure_device();
  if (rc) return rc;
  hipStream_t s = stream ? (hipStream_t)stream : thread_stream();
  if ((rc = refresh_copies(b, s))) return rc;
  return batch_launch(b, s);
}

int gbs_batch_run_graph(gbs_batch *b, void *stream) {
  if (!b) return fail(GBS_EINVAL, "gbs_batch_run_graph: null batch");
  int rc = ensure_device();
  if (rc) return rc;
  hipStream_t s = stream ? (hipStream_t)stream : thread_stream();
  if ((rc = refresh_copies(b, s))) return rc;  // outside the graph: only when stale
  if (!b->exec) {
    GBS_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    rc = batch_launch(b, s);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(s, &g);
    if (rc) {
      if (g) (void)hipGraphDestroy(g);
      return rc;
    }
    if (e != hipSuccess) return hip_fail(e, "gbs_batch_run_graph: capture");
    b->graph = g;
    GBS_HIP(hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0));
  }
  GBS_HIP(hipGraphLaunch(b->exec, s));
  return GBS_OK;
}

int gbs_batch_run_per_call(gbs_batch *b, void *stream) {
  if (!b) return fail(GBS_EINVAL, "gbs_batch_run_per_call: null batch");
  int rc = ensure_device();
  if (rc) return rc;
  hipStream_t s = stream ? (hipStream_t)stream : thread_stream();
  for (const gbs_problem &p : b->probs) {
    if ((rc = gemm_enqueue_public(b->ws, p.lhs, p.rhs, p.d_res, GBS_ALGO_AUTO, s, false)))
      return rc;
  }
  return GBS_OK;
}

int gbs_batch_info(const gbs_batch *b, uint64_t *tiles, uint64_t *persistent,
                   uint64_t *separate) {
  if (!b) return fail(GBS_EINVAL, "gbs_batch_info: null batch");
  if (tiles) *tiles = b->total_tiles;
  if (persistent) *persistent = b->mfma_idx.size();
  if (separate) *separate = b->separate_idx.size();
  return GBS_OK;
}

}  // extern "C"
