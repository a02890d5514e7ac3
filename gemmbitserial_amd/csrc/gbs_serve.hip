// gbs_serve.hip -- continuous-batching serving loop for a stack of
// bit-serial layers with weights resident in HBM.
//
// A server holds L weight matrices W_l (h_{l+1} x h_l, bit-planes plus their
// int8 copy) and, per layer, two activation pools (step parity): bit-planes
// and int8 values of up to R rows of width h_l.  Requests (1..max_rows rows
// of act_bits-bit unsigned activations) wait in a FIFO.  Every step t:
//
//   1. admission: waiting requests join a new cohort while its rows fit the
//      pool capacity R (set from the memory budget at creation), the
//      number of requests in flight stays under max_requests and the
//      cohort holds fewer than max_requests / L of them; the cohort's
//      inputs are packed on the GPU into pool 0 (parity t), read from pinned
//      host memory by the packing kernel itself;
//   2. one persistent launch computes, for every cohort in flight (cohort l
//      sits at layer l), acc_l = pool_l * W_l^T -- all layers at once, so
//      requests that arrived at different times share the device -- and
//      requantizes the result in the same launch: min(2^act_bits - 1,
//      max(0, acc >> s_l)) as int8 codes and bit-planes into pool l+1 of
//      parity t+1 (the pool cohort l+1 reads in this launch has parity t,
//      so the two never meet), the last layer's int32 rows straight into
//      pinned host memory.  Whole-K output tiles convert in their epilogue,
//      from registers; products split along K add their pieces into zeroed
//      accumulators, which a second, small launch converts and zeroes
//      again.  Its descriptors ride in the kernel arguments and it zeroes
//      the queue counter of the next launch (two alternate);
//   3. those requests finish.
//
#include <string.h>

#include <chrono>
#include <algorithm>
#include <deque>
#include <map>
#include <vector>

#include "gbs_mfma_tile.h"

namespace {

int64_t now_ns() {
  return std::chrono::duration_cast<std::chrono::nanoseconds>(
             std::chrono::steady_clock::now().time_since_epoch())
      .count();
}

struct Req {
  uint64_t id;
  uint32_t rows;
  uint32_t off;  // first row in its cohort's pool
  int64_t arrival, done;
  std::vector<uint8_t> x;
  // a finished request's rows: in the pinned output half `half` (from
  // `rows_at`) until that half is about to be reused, then in y
  std::vector<int32_t> y;
  const int32_t *rows_at = nullptr;
  int half = -1;
};

constexpr int kMaxLayers = 16;

// An admitted cohort's input codes (rows x h, row-major, pinned host
// memory) into a pool: int8 copy and bit-planes.  One wave per (row, plane
// word), lane = column of that word, one ballot per plane.
__global__ __launch_bounds__(256) void input_pack_kernel(const uint8_t *in, uint64_t *planes,
                                                         int8_t *vals, uint64_t wpr, uint64_t wpp,
                                                         uint64_t ld, uint32_t h, uint64_t units,
                                                         uint32_t nbits, uint32_t maxv) {
  const uint32_t lane = threadIdx.x & 63;
  const uint64_t u = (uint64_t)blockIdx.x * 4 + threadIdx.x / 64;
  if (u >= units) return;
  const uint64_t r = u / wpr, w = u - r * wpr;
  const uint64_t c = w * 64 + lane;
  uint32_t code = 0;
  if (c < h) {
    code = in[r * h + c] & maxv;
    vals[r * ld + c] = (int8_t)code;
  }
  for (uint32_t b = 0; b < nbits; b++) {
    const uint64_t m = __ballot((code >> b) & 1u);
    if (lane == 0) planes[b * wpp + r * wpr + w] = m;
  }
}

}  // namespace

struct gbs_server {
  uint32_t L = 0, act_bits = 0, R = 0, max_requests = 0, max_rows = 0;
  std::vector<const gbs_matrix *> W;
  std::vector<uint32_t> h, shift;
  std::vector<gbs_matrix *> pool;              // pool[2 * l + parity]
  std::vector<int32_t *> acc;                  // split products' sums, zero between steps
  unsigned int *d_counter = nullptr;           // the GEMM launches' queue counters (two)
  uint64_t launches = 0;
  // pinned host staging, two halves each, read and written by the kernels
  // in place (no copies): inputs of an admitted cohort, last-layer rows
  uint8_t *h_in = nullptr;
  int32_t *h_out = nullptr;
  size_t in_half = 0, out_half = 0;  // bytes / elements per half
  hipEvent_t ev[2] = {nullptr, nullptr};        // end of the step using each half
  std::vector<Req *> completing[2];             // requests whose rows land in h_out half
  std::vector<Req *> resident[2];               // finished, untaken, rows still in that half
  bool pending[2] = {false, false};
  uint64_t bytes_reserved = 0;
  std::deque<Req *> waiting;
  std::vector<std::vector<Req *>> cohort;  // cohort[l]: requests at layer l
  std::vector<uint32_t> cohort_rows;
  std::map<uint64_t, Req *> finished;
  uint64_t steps = 0, done_requests = 0, done_rows = 0, layer_steps = 0;
  double ops_done = 0, hbm_bytes = 0;
};

using namespace gbs;

static void server_free(gbs_server *s) {
  if (!s) return;
  (void)hipDeviceSynchronize();
  for (gbs_matrix *m : s->pool) gbs_matrix_destroy(m);
  for (int32_t *a : s->acc) (void)hipFree(a);
  if (s->d_counter) (void)hipFree(s->d_counter);
  if (s->h_in) (void)hipHostFree(s->h_in);
  if (s->h_out) (void)hipHostFree(s->h_out);
  for (int k = 0; k < 2; k++) {
    if (s->ev[k]) (void)hipEventDestroy(s->ev[k]);
    for (Req *r : s->completing[k]) delete r;
  }
  for (Req *r : s->waiting) delete r;
  for (auto &c : s->cohort)
    for (Req *r : c) delete r;
  for (auto &kv : s->finished) delete kv.second;
  delete s;
}

static uint64_t align_up(uint64_t x, uint64_t a) { return (x + a - 1) / a * a; }

extern "C" {

int gbs_clock_ns(int64_t *now) {
  if (!now) return fail(GBS_EINVAL, "gbs_clock_ns: null out");
  *now = now_ns();
  return GBS_OK;
}

int gbs_server_create(const gbs_matrix *const *weights, const uint32_t *shifts,
                      uint32_t n_layers, uint32_t act_bits, const gbs_server_config *cfg,
                      gbs_server **out) {
  if (!out) return fail(GBS_EINVAL, "gbs_server_create: null out");
  *out = nullptr;
  if (!weights || !cfg || n_layers == 0 || n_layers > kMaxLayers)
    return fail(GBS_EINVAL, "gbs_server_create: need 1..%d layers and a config", kMaxLayers);
  if (act_bits < 1 || act_bits > 7)
    return fail(GBS_EINVAL, "gbs_server_create: act_bits %u outside [1, 7]", act_bits);
  if (n_layers > 1 && !shifts) return fail(GBS_EINVAL, "gbs_server_create: null shifts");
  if (cfg->max_rows == 0 || cfg->max_requests == 0)
    return fail(GBS_EINVAL, "gbs_server_create: max_rows and max_requests must be > 0");
  int rc = ensure_device();
  if (rc) return rc;
  hipStream_t st = thread_stream();
  gbs_server *s = new gbs_server();
  s->L = n_layers;
  s->act_bits = act_bits;
  s->max_requests = cfg->max_requests;
  s->max_rows = cfg->max_rows;
  s->h.push_back((uint32_t)weights[0]->ncols);
  for (uint32_t l = 0; l < n_layers; l++) {
    const gbs_matrix *w = weights[l];
    if (!w || w->ncols != s->h[l]) {
      server_free(s);
      return fail(GBS_EINVAL, "gbs_server_create: layer %u width does not chain", l);
    }
    const Operand wo = view(w);
    Operand ao{};
    ao.nbits = act_bits;
    if (!mfma_supported(wo, ao, w->ncols)) {
      server_free(s);
      return fail(GBS_ENOTSUP, "gbs_server_create: layer %u not exact on the int8 engine", l);
    }
    if ((rc = ensure_vals(const_cast<gbs_matrix *>(w), st, true))) {
      server_free(s);
      return rc;
    }
    s->W.push_back(w);
    s->h.push_back((uint32_t)w->nrows);
    s->shift.push_back(l + 1 < n_layers ? shifts[l] : 0);
  }
  // bytes one row of capacity costs across every pool and buffer: the
  // staging, two pools per layer and the accumulators
  uint64_t per_row = s->h[0] + 4ull * s->h[n_layers];
  for (uint32_t l = 0; l < n_layers; l++)
    per_row += 2 * (act_bits * align_up(s->h[l], 64) / 8 + align_up(s->h[l], kValsColAlign)) +
               4ull * s->h[l + 1];
  uint64_t budget = cfg->mem_budget;
  if (budget == 0) {
    size_t fr = 0, tot = 0;
    if (hipMemGetInfo(&fr, &tot) != hipSuccess) fr = 1ull << 30;
    budget = fr / 4;
  }
  uint64_t R = budget / per_row;
  if (R > 65536) R = 65536;
  if (R < cfg->max_rows) {
    server_free(s);
    return fail(GBS_ERANGE,
                "gbs_server_create: memory budget %llu B holds %llu rows, less than one "
                "request of max_rows %u",
                (unsigned long long)budget, (unsigned long long)R, cfg->max_rows);
  }
  s->R = (uint32_t)R;
  s->cohort.resize(n_layers);
  s->cohort_rows.assign(n_layers, 0);
  const uint64_t Ra = align_up(R, kValsRowAlign);
  for (uint32_t l = 0; l < n_layers && !rc; l++) {
    for (int par = 0; par < 2 && !rc; par++) {
      gbs_matrix *p = nullptr;
      rc = gbs_matrix_create(act_bits, R, s->h[l], 0, R, align_up(s->h[l], 64), &p);
      if (rc) break;
      s->pool.push_back(p);
      const uint64_t ld = align_up(s->h[l], kValsColAlign);
      hipError_t e = hipMalloc((void **)&p->d_vals, Ra * ld);
      if (e == hipSuccess) e = hipMemset(p->d_vals, 0, Ra * ld);
      if (e != hipSuccess) {
        rc = hip_fail(e, "gbs_server_create: pool");
        break;
      }
      p->vals_rows = Ra;
      p->vals_ld = ld;
      p->vals_version = p->planes_version;  // kept current by the step launches
    }
    if (rc) break;
    int32_t *a = nullptr;
    hipError_t e = hipMalloc((void **)&a, R * s->h[l + 1] * 4ull);
    if (e == hipSuccess) e = hipMemset(a, 0, R * s->h[l + 1] * 4ull);
    if (e != hipSuccess) {
      rc = hip_fail(e, "gbs_server_create: accumulators");
      break;
    }
    s->acc.push_back(a);
  }
  if (!rc) {
    hipError_t e = hipMalloc((void **)&s->d_counter, 256);
    if (e == hipSuccess) e = hipMemset(s->d_counter, 0, 256);
    s->in_half = align_up(R * s->h[0] + 16, 256);
    s->out_half = align_up(R * s->h[n_layers] + 4, 64);  // int32 elements
    if (e == hipSuccess) e = hipHostMalloc((void **)&s->h_in, 2 * s->in_half);
    if (e == hipSuccess) e = hipHostMalloc((void **)&s->h_out, 2 * s->out_half * 4);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&s->ev[0], hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&s->ev[1], hipEventDisableTiming);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) rc = hip_fail(e, "gbs_server_create: buffers");
  }
  if (rc) {
    server_free(s);
    return rc;
  }
  s->bytes_reserved = R * per_row;
  *out = s;
  return GBS_OK;
}

int gbs_server_destroy(gbs_server *s) {
  if (!s) return GBS_OK;
  int rc = ensure_device();
  if (rc) return rc;
  server_free(s);
  return GBS_OK;
}

int gbs_server_submit(gbs_server *s, uint64_t id, const uint8_t *x, uint32_t rows,
                      int64_t arrival_ns) {
  if (!s || !x) return fail(GBS_EINVAL, "gbs_server_submit: null argument");
  if (rows == 0 || rows > s->max_rows)
    return fail(GBS_ERANGE, "gbs_server_submit: %u rows outside [1, %u]", rows, s->max_rows);
  Req *r = new Req();
  r->id = id;
  r->rows = rows;
  r->off = 0;
  r->arrival = arrival_ns ? arrival_ns : now_ns();
  r->done = 0;
  r->x.assign(x, x + (size_t)rows * s->h[0]);
  s->waiting.push_back(r);
  return GBS_OK;
}

int gbs_server_step(gbs_server *s, uint32_t *finished_now) {
  if (!s) return fail(GBS_EINVAL, "gbs_server_step: null server");
  int rc = ensure_device();
  if (rc) return rc;
  hipStream_t st = thread_stream();
  const uint32_t L = s->L;
  const int cur = (int)(s->steps & 1), prev = cur ^ 1;
  uint8_t *h_in = s->h_in + (size_t)cur * s->in_half;
  int32_t *h_out = s->h_out + (size_t)cur * s->out_half;
  // this step's launch writes output half cur: results not yet taken from
  // it move to their own storage first
  for (Req *r : s->resident[cur]) {
    r->y.assign(r->rows_at, r->rows_at + (size_t)r->rows * s->h[L]);
    r->rows_at = nullptr;
    r->half = -1;
  }
  s->resident[cur].clear();
  // Between steps cohort[l] holds the requests whose next GEMM is layer l
  // (their activations sit in pool l); cohort[0] is empty.
  // 1. admission into layer 0: a cohort takes at most max_requests / L
  // requests, so that under load every layer holds one and the layers'
  // GEMMs share each launch (a cohort that took every slot would leave the
  // other layers idle until it finished)
  size_t in_flight = 0;
  for (uint32_t l = 1; l < L; l++) in_flight += s->cohort[l].size();
  const size_t per_cohort = std::max<size_t>(1, s->max_requests / L);
  uint32_t n0 = 0;
  while (!s->waiting.empty()) {
    Req *r = s->waiting.front();
    if (n0 + r->rows > s->R || in_flight + s->cohort[0].size() + 1 > s->max_requests ||
        s->cohort[0].size() >= per_cohort)
      break;
    s->waiting.pop_front();
    r->off = n0;
    std::copy(r->x.begin(), r->x.end(), h_in + (size_t)n0 * s->h[0]);
    n0 += r->rows;
    s->cohort[0].push_back(r);
  }
  s->cohort_rows[0] = n0;
  const int par = (int)(s->steps & 1);  // pools this step reads; it writes parity par ^ 1
  if (n0) {
    gbs_matrix *p = s->pool[par];
    const uint64_t units = (uint64_t)n0 * p->wpr();
    hipLaunchKernelGGL(input_pack_kernel, dim3((uint32_t)((units + 3) / 4)), dim3(256), 0, st,
                       (const uint8_t *)h_in, p->d_planes, p->d_vals, p->wpr(), p->wpp(),
                       p->vals_ld, s->h[0], units, s->act_bits, (1u << s->act_bits) - 1);
    GBS_HIP(hipGetLastError());
  }
  // 2. every layer's GEMM and requantization in one persistent launch
  // (weights re-imported since the last step get their int8 copies rebuilt
  // first)
  for (uint32_t l = 0; l < L; l++)
    if ((rc = ensure_vals(const_cast<gbs_matrix *>(s->W[l]), st, false))) return rc;
  MfmaOperands views[kMaxLayers];
  MfmaRequant rq[kMaxLayers];
  int nv = 0;
  for (uint32_t l = 0; l < L; l++) {
    const uint32_t n = s->cohort_rows[l];
    if (!n) continue;
    const gbs_matrix *in = s->pool[2 * l + par];
    MfmaOperands v{};
    v.R = in->d_vals;
    v.r_ld = in->vals_ld;
    v.L = s->W[l]->d_vals;
    v.l_ld = s->W[l]->vals_ld;
    v.res = s->acc[l];
    v.M = s->h[l + 1];
    v.N = n;
    v.kbytes = (uint32_t)align_up(s->h[l], kMfmaKS);
    MfmaRequant q{};
    if (l + 1 < L) {
      gbs_matrix *o = s->pool[2 * (l + 1) + (par ^ 1)];
      q.vals = o->d_vals;
      q.planes = o->d_planes;
      q.ld = o->vals_ld;
      q.wpr = o->wpr();
      q.wpp = o->wpp();
      q.shift = s->shift[l];
      q.nbits = s->act_bits;
      q.maxv = (1u << s->act_bits) - 1;
    } else {
      q.out = h_out;
    }
    views[nv] = v;
    rq[nv++] = q;
    s->ops_done += 2.0 * s->h[l + 1] * s->h[l] * n;
    // algorithmic bytes: weight planes, input codes, output codes (int32 rows last)
    s->hbm_bytes += (double)s->W[l]->nbits * s->h[l + 1] * s->h[l] / 8 +
                    (double)s->act_bits * n * s->h[l] / 8 +
                    (l + 1 < L ? (double)s->act_bits * n * s->h[l + 1] / 8
                               : 4.0 * n * s->h[l + 1]);
    s->layer_steps++;
  }
  if (nv) {
    unsigned int *cur_c = s->d_counter + (s->launches & 1), *next_c = s->d_counter + ((s->launches + 1) & 1);
    if ((rc = launch_views(views, rq, nv, cur_c, next_c, st))) return rc;
    s->launches++;
  }
  // 3. the last layer's rows are in h_out once this step's event has
  // passed (read in the next call)
  GBS_HIP(hipEventRecord(s->ev[cur], st));
  s->completing[cur].swap(s->cohort[L - 1]);
  s->cohort[L - 1].clear();
  s->pending[cur] = true;
  // every other cohort moves one layer on
  for (uint32_t l = L - 1; l > 0; l--) {
    s->cohort[l].swap(s->cohort[l - 1]);
    s->cohort_rows[l] = s->cohort_rows[l - 1];
  }
  s->cohort[0].clear();
  s->cohort_rows[0] = 0;
  s->steps++;
  // 4. complete the previous step
  uint32_t fin = 0;
  if (s->pending[prev]) {
    GBS_HIP(hipEventSynchronize(s->ev[prev]));
    const int64_t t = now_ns();
    const int32_t *out = s->h_out + (size_t)prev * s->out_half;
    for (Req *r : s->completing[prev]) {
      r->rows_at = out + (size_t)r->off * s->h[L];
      r->half = prev;
      s->resident[prev].push_back(r);
      r->done = t;
      s->finished[r->id] = r;
      s->done_requests++;
      s->done_rows += r->rows;
      fin++;
    }
    s->completing[prev].clear();
    s->pending[prev] = false;
  }
  if (finished_now) *finished_now = fin;
  return GBS_OK;
}

int gbs_server_take(gbs_server *s, uint64_t id, int32_t *out, uint64_t cap_elems,
                    int *ready, uint32_t *rows, int64_t *arrival_ns, int64_t *done_ns) {
  if (!s || !ready) return fail(GBS_EINVAL, "gbs_server_take: null argument");
  auto it = s->finished.find(id);
  if (it == s->finished.end()) {
    *ready = 0;
    return GBS_OK;
  }
  Req *r = it->second;
  const size_t n = (size_t)r->rows * s->h[s->L];
  if (out) {
    if (cap_elems < n)
      return fail(GBS_ERANGE, "gbs_server_take: output holds %llu of %zu values",
                  (unsigned long long)cap_elems, n);
    if (r->rows_at) std::copy(r->rows_at, r->rows_at + n, out);
    else std::copy(r->y.begin(), r->y.end(), out);
  }
  if (r->half >= 0) {
    std::vector<Req *> &res = s->resident[r->half];
    res.erase(std::find(res.begin(), res.end(), r));
  }
  *ready = 1;
  if (rows) *rows = r->rows;
  if (arrival_ns) *arrival_ns = r->arrival;
  if (done_ns) *done_ns = r->done;
  s->finished.erase(it);
  delete r;
  return GBS_OK;
}

int gbs_server_finished(const gbs_server *s, uint64_t *ids, uint32_t cap, uint32_t *n) {
  if (!s || !n || (cap && !ids)) return fail(GBS_EINVAL, "gbs_server_finished: null argument");
  uint32_t k = 0;
  for (const auto &kv : s->finished) {
    if (k >= cap) break;
    ids[k++] = kv.first;
  }
  *n = k;
  return GBS_OK;
}

int gbs_server_layer_input(gbs_server *s, uint32_t layer, const gbs_matrix **out,
                           uint32_t *rows) {
  if (!s || !out || !rows) return fail(GBS_EINVAL, "gbs_server_layer_input: null argument");
  if (layer >= s->L)
    return fail(GBS_ERANGE, "gbs_server_layer_input: layer %u of %u", layer, s->L);
  int rc = ensure_device();
  if (rc) return rc;
  GBS_HIP(hipStreamSynchronize(thread_stream()));
  // the pools the next step reads have the parity of the next step
  *out = s->pool[2 * layer + (s->steps & 1)];
  *rows = s->cohort_rows[layer];
  return GBS_OK;
}

int gbs_server_stats(const gbs_server *s, gbs_server_counters *st) {
  if (!s || !st) return fail(GBS_EINVAL, "gbs_server_stats: null argument");
  st->steps = s->steps;
  st->done_requests = s->done_requests;
  st->done_rows = s->done_rows;
  st->ops = s->ops_done;
  st->row_capacity = s->R;
  st->bytes_reserved = s->bytes_reserved;
  st->waiting = (uint32_t)s->waiting.size();
  uint32_t act = 0;
  for (const auto &c : s->cohort) act += (uint32_t)c.size();
  for (const auto &c : s->completing) act += (uint32_t)c.size();
  st->in_flight = act;
  st->layer_steps = s->layer_steps;
  st->hbm_bytes = s->hbm_bytes;
  return GBS_OK;
}

}  // extern "C"
