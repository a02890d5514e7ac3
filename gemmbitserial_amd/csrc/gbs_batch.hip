// gbs_batch.hip -- many independent bit-serial GEMMs (uneven M/K/N, mixed
// bitwidths and signedness) in one persistent kernel launch.
//
// Plan (gbs_batch_create): every product that the matrix-core engine holds
// exactly becomes a run of work items in the narrowest copy format both
// operands fit (FP4, else FP6, else int8: inside one launch an item's time
// follows the bytes of K it streams) -- output tiles times K pieces, whose
// partial sums are added with integer atomics into results zeroed first.
// Tile shape and largest K piece come from the batch tile table per batch
// class (rhs-row bucket x depth bucket x format; batch_plan), measured on
// whole persistent launches by the tuner; without an entry the built-in
// rule: the tile with the least padding for the rhs row count (16 x 256
// for N <= 16, 32 x 256 for N <= 32, else 64 x 64) and pieces of at most
// 4 KiB of nominal K (measured: 2 KiB pieces lose 6 % on the 512-product
// workload to pipeline fill per item, 8 KiB pieces leave too few items for
// GEMV-heavy lists).  Products are ordered by the cost of one item, largest
// first, so the long items start early and the short ones fill the tail.
// A device array holds one descriptor per product and the prefix sum of
// tile counts.  Products the MFMA engine cannot hold exactly (unsigned
// 8-bit codes, very deep K) go to the popcount engine as separate launches
// after the persistent kernel.
//
// Products are grouped by copy format (int8, FP4, FP6); each group has its
// own descriptor range, tile prefix sums and queue counter, and its own
// persistent kernel (one kernel carrying every format's tile code would
// spill registers).  On request (gbs_batch_set_planes) products of <= 4-bit
// operands form a fourth group instead, read straight from their
// bit-planes by the planes engine's persistent kernel (gbs_planes.hip): no
// value copies, at more matrix instructions and code expansion per tile.
//
// Copies of operands re-imported since the plan are rebuilt at the start
// of a run, every stale (matrix, format) pair once and all of them in one
// launch over a job table.
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
  std::vector<uint8_t> mfma_fmt;           // their groups: copy format, or planes (3)
  std::vector<gbs::MfmaPlan> mfma_plans;   // their plans (planes: variant, splits, jper)
  std::vector<int32_t> plan_of;            // per problem: index into mfma_plans, or -1
  std::vector<uint32_t> separate_idx;      // popcount products
  void *d_desc = nullptr;                  // BatchProb[], grouped by format
  void *d_pdesc = nullptr;                 // descriptors of the planes group
  void *d_zero = nullptr;                  // ZeroSpan[] of split products
  uint32_t n_zero = 0;
  unsigned int *d_counter = nullptr;       // one queue counter per group
  uint32_t *d_item_prod = nullptr;         // product of every tile, per group
  uint64_t total_tiles = 0;
  // per group (int8, FP4, FP6 copies; planes): descriptor range, tiles,
  // first entry in d_item_prod, grid
  uint32_t first[4] = {0, 0, 0, 0}, count[4] = {0, 0, 0, 0};
  uint64_t tiles[4] = {0, 0, 0, 0};
  uint64_t item0[4] = {0, 0, 0, 0};
  uint32_t grid[4] = {0, 0, 0, 0};
  gbs_workspace *ws = nullptr;             // for separate products
  // copy refresh in one launch: job table (pinned host, device), capacity
  // in jobs, and the event after the last refresh launch (its table may be
  // rewritten once that has passed)
  gbs::ExpandJob *h_jobs = nullptr, *d_jobs = nullptr;
  uint32_t jobs_cap = 0;
  hipEvent_t jobs_done = nullptr;
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

constexpr int kGroupPlanes = kMfmaFormats, kGroups = kMfmaFormats + 1;

__global__ void reset_counter_kernel(unsigned int *c) {
  if (threadIdx.x < kGroups) c[threadIdx.x] = 0;
}

// Zero the results of split products (their partial sums are added
// atomically); blockIdx.y picks the span.  Block (0, 0) also resets the
// queue counters, which saves the launch of reset_counter_kernel.
__global__ __launch_bounds__(256) void zero_spans_kernel(const ZeroSpan *spans,
                                                         unsigned int *counters) {
  if (blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x < kGroups) counters[threadIdx.x] = 0;
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

// Tile table of the persistent launch, separate from the standalone one
// (whose shapes are chosen to fill the chip with one product, which a
// shared queue does not need).  Per batch class (rhs row count bucket x
// depth bucket) and copy format: a tile shape the persistent kernel
// carries and the largest K piece of an item, in K units (kMfmaKS nominal
// bytes).  Filled by gbs_batch_tune_load / _set (the tuner times whole
// persistent launches of a mixed workload); shape -1 keeps the built-in
// rule for that class.
struct BatchTune {
  int shape = -1, piece = 0;
};
BatchTune g_batch_tune[kMfmaFormats][kBatchClasses];
// Which products go to the planes group (gbs_batch_set_planes): 0 none
// (the default), 1 those it holds that have at most kPlanesRhsRows rhs rows
// and do not fit FP4 codes on both sides, 2 every product it holds.
// Measured on the 512-product uneven workload: per class alone in a
// launch, beyond FP4 with <= 16 rhs rows the planes group beats FP6 copies
// (83 against 97 us: half the operand bytes), with 17-32 rows they are
// even, and wider or FP4-format products run 1.3-2x faster from copies,
// whose expansion happens once, before the launch, where the planes group
// expands the plane side once per code-side tile and issues 2-4x the
// matrix instructions.  In the whole batch the group's own launch (and the
// tail it adds) outweighs that class's gain: copies 301-307 us, mode 1
// 322-329, mode 2 365-370; with every rhs re-imported before each run,
// copies (rebuilt in one launch) 344 us, mode 2 377 us.  The group needs
// no value copies at all (787 MB less HBM on that workload).
constexpr uint64_t kPlanesRhsRows = 16;
int g_batch_planes = 0;

int batch_class(uint64_t N, uint64_t K) {
  const int cn = N <= 16 ? 0 : N <= 32 ? 1 : N <= 64 ? 2 : N <= 128 ? 3 : N <= 256 ? 4 : 5;
  const int ck = K < 1024 ? 0 : K < 4096 ? 1 : 2;
  return cn * kBatchClassK + ck;
}

// Plan of one product inside a persistent launch: the tuned tile shape and
// K piece of its batch class, else the built-in rule -- the tile shape with
// the least padding for its rhs row count and pieces of at most 4 KiB
// (nominal) of K so that no item is much longer than the rest; products
// that would otherwise dominate the queue are cut into up to min_pieces
// pieces of at least 1 KiB.
static MfmaPlan batch_plan(uint32_t M, uint32_t N, uint32_t kbytes, int fmt,
                           uint32_t min_pieces = 1) {
  const uint64_t K = (uint64_t)kbytes * (fmt == kFmtI8 ? 1 : 2);  // values
  const BatchTune t = g_batch_tune[fmt][batch_class(N, K)];
  const int shape = batch_shape(t.shape >= 0 ? t.shape : N <= 16 ? 3 : N <= 32 ? 2 : 4, fmt);
  const uint32_t piece = t.shape >= 0 && t.piece > 0 ? (uint32_t)t.piece : 16;
  const uint32_t units = kbytes / kMfmaKS;
  uint32_t want = (units + piece - 1) / piece;
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
// launch).  Whole-K tiles requantize their results into the next layer's
// input, or store them as int32, in their epilogue (MfmaRequant); products
// split along K add their pieces into zeroed accumulators, which
// finish_split_kernel converts after the launch.  The launch zeroes the
// queue counter of the next launch (the two alternate).
constexpr int kMaxViews = 16;
struct ViewBatch {
  BatchProb p[kMaxViews];
  MfmaRequant q[kMaxViews];
  uint32_t n;
  uint64_t total;
};

__global__ __launch_bounds__(256, 2) void views_kernel(ViewBatch V, unsigned int *counter,
                                                       unsigned int *next) {
  __shared__ __attribute__((aligned(16))) int8_t lds[kMfmaLdsMax];
  __shared__ uint32_t s_item;
  if (blockIdx.x == 0 && threadIdx.x == 0) *next = 0;  // its launch follows this one
  for (;;) {
    if (threadIdx.x == 0) s_item = atomicAdd(counter, 1u);
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
    mfma_tile_by_shape<kFmtI8, true>(pl.shape, ops, pt, qt, kbeg, kend, pl.splits > 1, lds,
                                     V.q[lo]);
    __syncthreads();
  }
}

// Products of a step launch that were split along K, converted after it
// (the launch boundary orders every piece's atomic add before these reads).
struct FinishJob {
  int32_t *acc;
  uint32_t M, N, chunks;  // chunks: 64-column units of a row
  MfmaRequant q;
  uint64_t unit_begin;    // prefix of N * chunks over the jobs
};
struct FinishJobs {
  FinishJob j[kMaxViews];
  int n;
};

// One wave per (product, row, 64 columns), lane = column: the sum read and
// zeroed for the next step, then requantized into codes and one plane word
// per plane (a ballot), or stored as int32.
__global__ __launch_bounds__(256) void finish_split_kernel(FinishJobs J, uint64_t units) {
  const uint32_t lane = threadIdx.x & 63;
  const uint64_t u = (uint64_t)blockIdx.x * 4 + threadIdx.x / 64;
  if (u >= units) return;
  int k = 0;
  while (k + 1 < J.n && J.j[k + 1].unit_begin <= u) k++;
  const FinishJob &f = J.j[k];
  const uint64_t local = u - f.unit_begin;
  const uint64_t n = local / f.chunks, mc = (local - n * f.chunks) * 64, m = mc + lane;
  int32_t v = 0;
  if (m < f.M) {
    v = f.acc[n * f.M + m];
    f.acc[n * f.M + m] = 0;
  }
  if (f.q.vals) {
    const uint32_t code = m < f.M ? requant_code(v, f.q) : 0u;
    if (m < f.M) f.q.vals[n * f.q.ld + m] = (int8_t)code;
    for (uint32_t b = 0; b < f.q.nbits; b++) {
      const uint64_t w = __ballot((code >> b) & 1u);
      if (lane == 0) f.q.planes[b * f.q.wpp + n * f.q.wpr + mc / 64] = w;
    }
  } else if (m < f.M) {
    f.q.out[n * f.M + m] = v;
  }
}

// Plan and launch products given as int8 operand views with their
// requantization or output (see launch_views in gbs_mfma_tile.h).
int launch_views(const MfmaOperands *views, const MfmaRequant *rq, int n, unsigned int *counter,
                 unsigned int *next, hipStream_t s) {
  if (n > kMaxViews) return fail(GBS_EINVAL, "launch_views: %d products (at most %d)", n, kMaxViews);
  ViewBatch V{};
  BatchProb *h = V.p;
  MfmaRequant hq[kMaxViews];
  int m = 0;
  double total = 0;
  for (int i = 0; i < n; i++) {
    const MfmaOperands &v = views[i];
    if (v.M == 0 || v.N == 0) continue;
    BatchProb d{};
    d.ops = v;
    d.plan = batch_plan(v.M, v.N, v.kbytes, kFmtI8);
    d.tile_begin = (uint64_t)i;  // the product's index until sorted
    total += item_bytes(d.plan) * (double)d.plan.tiles();
    hq[i] = rq[i];
    h[m++] = d;
  }
  // a step of few GEMV-like products would otherwise run on a few CUs:
  // the same re-split as a planned batch, then longest items first
  rebalance(h, m, total);
  std::sort(h, h + m, [](const BatchProb &x, const BatchProb &y) {
    return item_bytes(x.plan) > item_bytes(y.plan);
  });
  uint64_t tiles = 0;
  FinishJobs F{};
  uint64_t units = 0;
  for (int i = 0; i < m; i++) {
    const MfmaRequant q = hq[h[i].tile_begin];
    if (h[i].plan.splits > 1) {
      FinishJob &f = F.j[F.n++];
      f.acc = h[i].ops.res;
      f.M = h[i].ops.M;
      f.N = h[i].ops.N;
      f.chunks = (f.M + 63) / 64;
      f.q = q;
      f.unit_begin = units;
      units += (uint64_t)f.N * f.chunks;
    } else if (!q.vals) {
      // a whole-K product without requantization stores its int32 rows
      // straight into the destination
      h[i].ops.res = q.out;
    }
    V.q[i] = q;
    h[i].tile_begin = tiles;
    tiles += h[i].plan.tiles();
  }
  if (tiles == 0) return GBS_OK;
  V.n = (uint32_t)m;
  V.total = tiles;
  const uint32_t grid = (uint32_t)(tiles < 512 ? tiles : 512);
  hipLaunchKernelGGL(views_kernel, dim3(grid), dim3(256), 0, s, V, counter, next);
  GBS_HIP(hipGetLastError());
  if (units) {
    hipLaunchKernelGGL(finish_split_kernel, dim3((uint32_t)((units + 3) / 4)), dim3(256), 0, s,
                       F, units);
  }
  GBS_HIP(hipGetLastError());
  return GBS_OK;
}


// Operands re-imported since the plan get their copies rebuilt before a
// run, on the run's stream: every stale (matrix, format) pair once, all in
// one launch when there are more than two (one launch per copy costs
// ~4 us of launch and host work: 512 fresh activations took 1.9 ms that
// way).  The plan stays valid: a copy buffer, once allocated, keeps its
// address and layout for the matrix's lifetime.
static int refresh_copies(gbs_batch *b, hipStream_t s) {
  // each stale pair once: a pair is marked current as it is listed (and
  // put back if the refresh fails before its launch)
  struct Stale { gbs_matrix *m; int f; uint64_t version; };
  std::vector<Stale> stale;
  auto version_of = [](gbs_matrix *m, int f) -> uint64_t & {
    return f == kFmtF4 ? m->f4_version : f == kFmtF6 ? m->f6_version : m->vals_version;
  };
  for (size_t k = 0; k < b->mfma_idx.size(); k++) {
    const int f = b->mfma_fmt[k];
    if (f == kGroupPlanes) continue;  // reads the planes themselves
    const gbs_problem &p = b->probs[b->mfma_idx[k]];
    for (gbs_matrix *m : {const_cast<gbs_matrix *>(p.lhs), const_cast<gbs_matrix *>(p.rhs)})
      if (!copy_current(m, f)) {
        stale.push_back({m, f, version_of(m, f)});
        mark_copy_current(m, f);
      }
  }
  if (stale.empty()) return GBS_OK;
  auto undo = [&](size_t from) {
    for (size_t i = from; i < stale.size(); i++) version_of(stale[i].m, stale[i].f) = stale[i].version;
  };
  if (stale.size() <= 2 || stale.size() > b->jobs_cap) {
    for (size_t i = 0; i < stale.size(); i++) {
      version_of(stale[i].m, stale[i].f) = stale[i].version;
      const int rc = ensure_copy(stale[i].m, stale[i].f, s, false);
      if (rc) {
        undo(i + 1);
        return rc;
      }
    }
    return GBS_OK;
  }
  // the previous refresh launch has read its table
  hipError_t e = hipEventSynchronize(b->jobs_done);
  if (e != hipSuccess) {
    undo(0);
    return hip_fail(e, "gbs_batch_run: copy refresh");
  }
  uint64_t blocks = 0;
  for (size_t i = 0; i < stale.size(); i++) {
    int rc = copy_job(stale[i].m, stale[i].f, false, stale[i].version == 0, &b->h_jobs[i]);
    if (!rc && blocks + expand_job_blocks(b->h_jobs[i]) >= 0x7fffffffull)
      rc = fail(GBS_ERANGE, "gbs_batch_run: copy refresh too large");
    if (rc) {
      undo(0);
      return rc;
    }
    b->h_jobs[i].block_begin = (uint32_t)blocks;
    blocks += expand_job_blocks(b->h_jobs[i]);
  }
  e = hipMemcpyAsync(b->d_jobs, b->h_jobs, stale.size() * sizeof(ExpandJob),
                     hipMemcpyHostToDevice, s);
  if (e != hipSuccess) {
    undo(0);
    return hip_fail(e, "gbs_batch_run: copy refresh");
  }
  const int rc = launch_expand_jobs(b->d_jobs, (uint32_t)stale.size(), (uint32_t)blocks, s);
  if (rc) {
    undo(0);
    return rc;
  }
  GBS_HIP(hipEventRecord(b->jobs_done, s));
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
    if (b->tiles[kGroupPlanes]) {
      const int rc = launch_planes_batch(b->d_pdesc, (uint32_t)b->tiles[kGroupPlanes],
                                         b->d_counter + kGroupPlanes,
                                         b->d_item_prod + b->item0[kGroupPlanes],
                                         b->grid[kGroupPlanes], s);
      if (rc) return rc;
    }
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
  struct Item {
    uint32_t idx;
    int grp;  // copy format, or kGroupPlanes
    double item_cost;
    MfmaPlan plan;
    MfmaOperands ops;
    PlanesBatchPlan pplan;
  };
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
    const uint64_t K = p.lhs->ncols;
    if (g_batch_planes == 2 ||
        (g_batch_planes == 1 && ro.nrows <= kPlanesRhsRows && !f4_supported(lo, ro, K))) {
      const PlanesBatchPlan pp = planes_batch_plan(lo, ro, K);
      if (pp.ok) {
        items.push_back({(uint32_t)i, kGroupPlanes, pp.item_bytes, MfmaPlan{}, MfmaOperands{}, pp});
        continue;
      }
    }
    const int fmt = f4_supported(lo, ro, K)     ? kFmtF4
                    : f6_supported(lo, ro, K)   ? kFmtF6
                    : mfma_supported(lo, ro, K) ? kFmtI8
                                                : -1;
    if (fmt >= 0 && K > 0) {
      rc = ensure_copy(const_cast<gbs_matrix *>(p.lhs), fmt, s, true);
      if (!rc) rc = ensure_copy(const_cast<gbs_matrix *>(p.rhs), fmt, s, true);
      if (rc) {
        gbs_batch_destroy(b);
        return rc;
      }
      const MfmaOperands ops = mfma_operands(p.lhs, p.rhs, fmt, p.d_res);
      const MfmaPlan pl = batch_plan(ops.M, ops.N, ops.kbytes, fmt);
      const MfmaTileShape sh = kMfmaShape[pl.shape];
      items.push_back({(uint32_t)i, fmt, (double)sh.P * sh.Q * (double)pl.kchunk, pl, ops,
                       PlanesBatchPlan{}});
    } else {
      b->separate_idx.push_back((uint32_t)i);
    }
  }
  // re-split products whose items would dominate the queue (rebalance; for
  // the planes group the same rule on its items' plane bytes)
  {
    double total = 0;
    std::vector<BatchProb> tmp;
    for (const Item &it : items)
      total += it.grp == kGroupPlanes ? it.pplan.item_bytes * (double)it.pplan.tiles()
                                      : item_bytes(it.plan) * (double)it.plan.tiles();
    for (Item &it : items) {
      if (it.grp == kGroupPlanes) {
        const double target = total / 1024;
        if (target > 0 && it.pplan.item_bytes > target) {
          const gbs_problem &p = problems[it.idx];
          const double per_tile = it.pplan.item_bytes * it.pplan.splits;
          it.pplan = planes_batch_plan(view(p.lhs), view(p.rhs), p.lhs->ncols,
                                       (uint32_t)std::min(per_tile / target + 1, 1e6));
          it.item_cost = it.pplan.item_bytes;
        }
        continue;
      }
      BatchProb t{};
      t.ops = it.ops;
      t.plan = it.plan;
      rebalance(&t, 1, total);
      const MfmaTileShape sh = kMfmaShape[t.plan.shape];
      it.plan = t.plan;
      it.item_cost = (double)sh.P * sh.Q * (double)t.plan.kchunk;
    }
  }
  // grouped (the planes group first, then the copy formats); inside a group
  // the longest items first: they start early and the short ones fill the
  // tail
  auto order = [](int g) { return g == kGroupPlanes ? -1 : g; };
  std::stable_sort(items.begin(), items.end(), [&](const Item &x, const Item &y) {
    return x.grp != y.grp ? order(x.grp) < order(y.grp) : x.item_cost > y.item_cost;
  });
  std::vector<BatchProb> desc;
  std::vector<uint8_t> pdesc;
  const size_t psz = planes_batch_desc_size();
  std::vector<ZeroSpan> zero;
  const uint64_t slots = 256ull * 2;
  // per group: (descriptor index within the group, its first tile, tiles)
  struct Span { uint32_t k; uint64_t begin, n; };
  std::vector<Span> spans[kGroups];
  for (const Item &it : items) {
    const gbs_problem &p = problems[it.idx];
    const int f = it.grp;
    if (f == kGroupPlanes) {
      const PlanesBatchPlan &pp = it.pplan;
      spans[f].push_back({b->count[f], b->tiles[f], pp.tiles()});
      pdesc.resize(pdesc.size() + psz);
      planes_batch_desc(view(p.lhs), view(p.rhs), p.lhs->ncols, p.d_res, pp, b->tiles[f],
                        pdesc.data() + pdesc.size() - psz);
      b->tiles[f] += pp.tiles();
      b->count[f]++;
      if (pp.splits > 1) zero.push_back(ZeroSpan{p.d_res, p.lhs->nrows * p.rhs->nrows});
      MfmaPlan rec{};
      rec.fmt = kGroupPlanes;
      rec.shape = (int)pp.variant;
      rec.splits = pp.splits;
      rec.kchunk = pp.jper;
      b->mfma_idx.push_back(it.idx);
      b->mfma_fmt.push_back((uint8_t)f);
      b->mfma_plans.push_back(rec);
      continue;
    }
    if (b->count[f] == 0) b->first[f] = (uint32_t)desc.size();
    BatchProb d{};
    d.ops = it.ops;
    d.plan = it.plan;
    d.tile_begin = b->tiles[f];
    spans[f].push_back({b->count[f], b->tiles[f], d.plan.tiles()});
    b->tiles[f] += d.plan.tiles();
    b->count[f]++;
    if (d.plan.splits > 1) zero.push_back(ZeroSpan{p.d_res, (uint64_t)d.ops.M * d.ops.N});
    desc.push_back(d);
    b->mfma_idx.push_back(it.idx);
    b->mfma_fmt.push_back((uint8_t)f);
    b->mfma_plans.push_back(d.plan);
  }
  b->plan_of.assign(count, -1);
  for (size_t k = 0; k < b->mfma_idx.size(); k++) b->plan_of[b->mfma_idx[k]] = (int32_t)k;
  for (int f = 0; f < kGroups; f++) {
    b->item0[f] = b->total_tiles;
    b->total_tiles += b->tiles[f];
    b->grid[f] = (uint32_t)(b->tiles[f] < slots ? b->tiles[f] : slots);
  }
  if (b->total_tiles >= (1ull << 32)) {
    gbs_batch_destroy(b);
    return fail(GBS_ERANGE, "gbs_batch_create: more than 2^32 tiles");
  }
  // product (descriptor index within its group) of every tile
  std::vector<uint32_t> item_prod(b->total_tiles);
  for (int f = 0; f < kGroups; f++)
    for (const Span &sp : spans[f]) {
      const uint64_t at = b->item0[f] + sp.begin;
      std::fill(item_prod.begin() + at, item_prod.begin() + at + sp.n, sp.k);
    }
  if (!pdesc.empty()) {
    hipError_t e = hipMalloc(&b->d_pdesc, pdesc.size());
    if (e == hipSuccess)
      e = hipMemcpyAsync(b->d_pdesc, pdesc.data(), pdesc.size(), hipMemcpyHostToDevice, s);
    if (e != hipSuccess) {
      gbs_batch_destroy(b);
      return hip_fail(e, "gbs_batch_create");
    }
  }
  if (!desc.empty() || !pdesc.empty()) {
    hipError_t e = desc.empty() ? hipSuccess : hipMalloc(&b->d_desc, desc.size() * sizeof(BatchProb));
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
    if (e == hipSuccess && !desc.empty())
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
  // copy refresh table: at most two copies per copy-format product
  {
    uint32_t n_copy = 0;
    for (uint8_t f : b->mfma_fmt) n_copy += f != kGroupPlanes;
    if (n_copy) {
      b->jobs_cap = 2 * n_copy;
      hipError_t e = hipHostMalloc((void **)&b->h_jobs, b->jobs_cap * sizeof(ExpandJob));
      if (e == hipSuccess) e = hipMalloc((void **)&b->d_jobs, b->jobs_cap * sizeof(ExpandJob));
      if (e == hipSuccess) e = hipEventCreateWithFlags(&b->jobs_done, hipEventDisableTiming);
      if (e != hipSuccess) {
        gbs_batch_destroy(b);
        return hip_fail(e, "gbs_batch_create");
      }
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
  if (b->d_pdesc) (void)hipFree(b->d_pdesc);
  if (b->h_jobs) (void)hipHostFree(b->h_jobs);
  if (b->d_jobs) (void)hipFree(b->d_jobs);
  if (b->jobs_done) (void)hipEventDestroy(b->jobs_done);
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

int gbs_batch_plan_of(const gbs_batch *b, uint64_t problem, int *fmt, int *shape,
                      uint32_t *splits, uint32_t *piece_units) {
  if (!b) return fail(GBS_EINVAL, "gbs_batch_plan_of: null batch");
  if (problem >= b->probs.size())
    return fail(GBS_ERANGE, "gbs_batch_plan_of: problem %llu of %zu",
                (unsigned long long)problem, b->probs.size());
  const int k = b->plan_of.empty() ? -1 : b->plan_of[problem];
  const MfmaPlan *pl = k >= 0 ? &b->mfma_plans[k] : nullptr;
  if (fmt) *fmt = pl ? pl->fmt : -1;
  if (shape) *shape = pl ? pl->shape : -1;
  if (splits) *splits = pl ? pl->splits : 0;
  if (piece_units)
    *piece_units = !pl ? 0 : pl->fmt == kGroupPlanes ? pl->kchunk : pl->kchunk / kMfmaKS;
  return GBS_OK;
}

int gbs_batch_set_planes(int mode) {
  if (mode < 0 || mode > 2) return fail(GBS_EINVAL, "gbs_batch_set_planes: mode %d", mode);
  g_batch_planes = mode;
  return GBS_OK;
}

int gbs_batch_get_planes(int *mode) {
  if (!mode) return fail(GBS_EINVAL, "gbs_batch_get_planes: null out");
  *mode = g_batch_planes;
  return GBS_OK;
}

int gbs_batch_tune_classes(int *n_classes) {
  if (!n_classes) return fail(GBS_EINVAL, "gbs_batch_tune_classes: null out");
  *n_classes = kBatchClasses;
  return GBS_OK;
}

int gbs_batch_tune_class_of(uint64_t N, uint64_t K, int *cls) {
  if (!cls) return fail(GBS_EINVAL, "gbs_batch_tune_class_of: null out");
  *cls = batch_class(N, K);
  return GBS_OK;
}

int gbs_batch_tune_shapes(int fmt, int *shapes, int cap, int *n) {
  if (!n || (cap > 0 && !shapes)) return fail(GBS_EINVAL, "gbs_batch_tune_shapes: null out");
  if (fmt < 0 || fmt >= kMfmaFormats)
    return fail(GBS_EINVAL, "gbs_batch_tune_shapes: format %d", fmt);
  int k = 0;
  for (int s = 0; s < kMfmaShapes; s++)
    if (batch_shape(s, fmt) == s) {
      if (k < cap) shapes[k] = s;
      k++;
    }
  *n = k;
  return GBS_OK;
}

int gbs_batch_tune_set(int fmt, int cls, int shape, int piece_units) {
  if (fmt < 0 || fmt >= kMfmaFormats || cls < 0 || cls >= kBatchClasses)
    return fail(GBS_EINVAL, "gbs_batch_tune_set: format %d class %d", fmt, cls);
  if (shape >= 0 && (shape >= kMfmaShapes || batch_shape(shape, fmt) != shape))
    return fail(GBS_EINVAL,
                "gbs_batch_tune_set: shape %d is not carried by the persistent kernel "
                "for format %d",
                shape, fmt);
  if (shape >= 0 && (piece_units < 1 || piece_units > 4096))
    return fail(GBS_EINVAL, "gbs_batch_tune_set: K piece of %d units outside [1, 4096]",
                piece_units);
  g_batch_tune[fmt][cls].shape = shape < 0 ? -1 : shape;
  g_batch_tune[fmt][cls].piece = shape < 0 ? 0 : piece_units;
  return GBS_OK;
}

int gbs_batch_tune_get(int fmt, int cls, int *shape, int *piece_units) {
  if (fmt < 0 || fmt >= kMfmaFormats || cls < 0 || cls >= kBatchClasses)
    return fail(GBS_EINVAL, "gbs_batch_tune_get: format %d class %d", fmt, cls);
  if (shape) *shape = g_batch_tune[fmt][cls].shape;
  if (piece_units) *piece_units = g_batch_tune[fmt][cls].piece;
  return GBS_OK;
}

int gbs_batch_tune_load(const char *path) {
  if (!path) return fail(GBS_EINVAL, "gbs_batch_tune_load: null path");
  FILE *f = fopen(path, "r");
  if (!f) return fail(GBS_EINVAL, "gbs_batch_tune_load: cannot open %s", path);
  BatchTune t[kMfmaFormats][kBatchClasses];
  char line[512];
  int n = 0, bad = 0;
  while (fgets(line, sizeof(line), f)) {
    int c, fmt, sh, pc;
    if (line[0] == '#') continue;
    if (sscanf(line, "%d %d %d %d", &c, &fmt, &sh, &pc) != 4) continue;
    if (c < 0 || c >= kBatchClasses || fmt < 0 || fmt >= kMfmaFormats || sh < 0 ||
        sh >= kMfmaShapes || batch_shape(sh, fmt) != sh || pc < 1 || pc > 4096) {
      bad++;
      continue;
    }
    t[fmt][c].shape = sh;
    t[fmt][c].piece = pc;
    n++;
  }
  fclose(f);
  if (bad) return fail(GBS_EINVAL, "gbs_batch_tune_load: %d invalid entries in %s", bad, path);
  if (n == 0) return fail(GBS_EINVAL, "gbs_batch_tune_load: no entries in %s", path);
  for (int fmt = 0; fmt < kMfmaFormats; fmt++)
    for (int c = 0; c < kBatchClasses; c++) g_batch_tune[fmt][c] = t[fmt][c];
  return GBS_OK;
}

}  // extern "C"
