// test_dropin.cpp -- the reference's self-test suite (test/test.cpp:30-277)
// rewritten against the drop-in header include/gemmbitserial.hpp, with full
// comparisons (the reference compares only d bytes in two tests) and
// deterministic seeds.  Extra cases cover what the reference never tests:
// signed multi-bit operands, mixed bitwidths, column-major import, copy
// semantics and host-side edits after import.
//
//   test_dropin [mnist.bin]     exit status 0 = all passed
//   test_dropin --geometry      for each "M K N" on stdin, allocGEMMContext
//                               and print "M K N lhsBlock rhsBlock
//                               lhs.nrows_a lhs.ncols_a rhs.nrows_a
//                               rhs.ncols_a" as allocated
//
// mnist.bin (optional): int32 n_out, n_in, then n_out*n_in int8 weights,
// n_in uint8 inputs, n_out int32 golden outputs (written by the pytest
// wrapper from tests/golden/mnist.npz).
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "gemmbitserial.hpp"

using namespace gemmbitserial;

static uint64_t g_state = 0x67627300;
static uint32_t rnd() {  // splitmix64
  uint64_t z = (g_state += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return (uint32_t)((z ^ (z >> 31)) >> 32);
}

template <typename T>
static void random_values(size_t bits, size_t n, T *out, bool allowNeg) {
  if (bits == 1 && allowNeg) {
    for (size_t i = 0; i < n; i++) out[i] = (rnd() & 1) ? 1 : -1;
    return;
  }
  const int32_t range = 1 << bits;
  for (size_t i = 0; i < n; i++)
    out[i] = (T)((int32_t)(rnd() % range) - (allowNeg ? range / 2 : 0));
}

template <typename L, typename R>
static void naive_gemm(const L *lhs, const R *rhs, int32_t *res, int rows, int depth,
                       int cols) {
  for (int k = 0; k < cols; k++)
    for (int i = 0; i < rows; i++) {
      int64_t acc = 0;
      for (int j = 0; j < depth; j++) acc += (int64_t)lhs[i * depth + j] * rhs[k * depth + j];
      res[k * rows + i] = (int32_t)acc;
    }
}

static int g_fail = 0;
static void report(const char *name, int ok, int nok) {
  printf("%s: %d OK, %d NOK\n", name, ok, nok);
  if (nok) g_fail = 1;
}

static const size_t DIMS[] = {3, 5, 7, 16, 17, 18, 30, 31, 32, 100, 177, 256};

static void test_conversions() {
  const size_t bits[] = {1, 2, 3, 7};
  const size_t dims[] = {4, 16, 17, 32, 77, 100, 1023};
  int ok = 0, nok = 0;
  for (size_t b : bits)
    for (size_t d : dims)
      for (int sg = 0; sg < 2; sg++) {
        std::vector<int8_t> in(d * d), out(d * d);
        random_values(b, d * d, in.data(), sg);
        BitSerialMatrix bsm = BitSerialMatrix::alloc(b, d, d, sg);
        bsm.importRegular(in.data());
        bsm.exportRegular(out.data());
        (memcmp(in.data(), out.data(), d * d) == 0 ? ok : nok)++;
        BitSerialMatrix::dealloc(bsm);
      }
  report("Conversion tests", ok, nok);
}

static void test_rowwise_sum() {
  const size_t dims[] = {4, 16, 17, 32, 77, 100, 1023};
  int ok = 0, nok = 0;
  for (size_t b = 1; b <= 4; b++)
    for (size_t d : dims)
      for (int sg = 0; sg < 2; sg++) {
        std::vector<int8_t> m(d * d);
        std::vector<int32_t> got(d), want(d, 0);
        random_values(b, d * d, m.data(), sg);
        GEMMContext ctx = allocGEMMContext(d, d, 1, b, 1, sg, false);
        BitSerialMatrix bsm = ctx.lhs;  // import through a copy (test.cpp:47-48)
        bsm.importRegular(m.data());
        sumRows(bsm, got.data());
        for (size_t r = 0; r < d; r++)
          for (size_t c = 0; c < d; c++) want[r] += m[r * d + c];
        (memcmp(got.data(), want.data(), d * 4) == 0 ? ok : nok)++;
        deallocGEMMContext(ctx);
      }
  report("Row-wise sum tests", ok, nok);
}

static void test_mnist(const char *path) {
  FILE *f = fopen(path, "rb");
  if (!f) {
    printf("MNIST: cannot open %s\n", path);
    g_fail = 1;
    return;
  }
  int32_t dims[2];
  size_t rd = fread(dims, 4, 2, f);
  const int n_out = dims[0], n_in = dims[1];
  std::vector<int8_t> w((size_t)n_out * n_in);
  std::vector<uint8_t> x(n_in);
  std::vector<int32_t> golden(n_out);
  rd += fread(w.data(), 1, w.size(), f);
  rd += fread(x.data(), 1, x.size(), f);
  rd += fread(golden.data(), 4, golden.size(), f);
  fclose(f);
  // test.cpp:157-172: bipolar weights x 2-bit unsigned vector (GEMV)
  GEMMContext ctx = allocGEMMContext(n_out, n_in, 1, 1, 2, true, false);
  const int8_t *cw = w.data();
  const uint8_t *cx = x.data();
  ctx.lhs.importRegular(cw);
  ctx.rhs.importRegular(cx);
  gemmBitSerial(ctx);
  int res = memcmp(ctx.res, golden.data(), n_out * 4);
  printf("MNIST matrix-vector using bipolar times regular: %s\n", res == 0 ? "OK" : "NOK");
  if (res) g_fail = 1;
  deallocGEMMContext(ctx);
}

static void test_matrix_matrix() {
  int ok = 0, nok = 0;
  for (size_t b = 2; b <= 4; b++)
    for (size_t d : DIMS)
      for (int mv = 0; mv < 2; mv++) {
        const size_t cols = mv ? 1 : 3 * d;
        std::vector<uint8_t> a(d * 2 * d), bm(2 * d * cols);
        std::vector<int32_t> want(d * cols);
        random_values(b, a.size(), a.data(), false);
        random_values(b, bm.size(), bm.data(), false);
        naive_gemm(a.data(), bm.data(), want.data(), d, 2 * d, cols);
        GEMMContext ctx = allocGEMMContext(d, 2 * d, cols, b, b, false, false);
        ctx.lhs.importRegular(a.data());
        ctx.rhs.importRegular(bm.data());
        gemmBitSerial(ctx);
        (memcmp(ctx.res, want.data(), want.size() * 4) == 0 ? ok : nok)++;
        deallocGEMMContext(ctx);
      }
  report("Matrix matrix multiplication tests", ok, nok);
}

static void test_bipolar_times_regular() {
  int ok = 0, nok = 0;
  for (int sw = 0; sw < 2; sw++)
    for (size_t rb = 2; rb <= 4; rb++)
      for (size_t d : DIMS)
        for (int sg = 0; sg < 2; sg++) {
          std::vector<int8_t> bip(d * d), reg(d * d);
          std::vector<int32_t> want(d * d);
          random_values(1, bip.size(), bip.data(), true);
          random_values(rb, reg.size(), reg.data(), sg);
          GEMMContext ctx;
          if (sw) {
            ctx = allocGEMMContext(d, d, d, rb, 1, sg, true);
            ctx.rhs.importRegular(bip.data());
            ctx.lhs.importRegular(reg.data());
            naive_gemm(reg.data(), bip.data(), want.data(), d, d, d);
          } else {
            ctx = allocGEMMContext(d, d, d, 1, rb, true, sg);
            ctx.lhs.importRegular(bip.data());
            ctx.rhs.importRegular(reg.data());
            naive_gemm(bip.data(), reg.data(), want.data(), d, d, d);
          }
          gemmBitSerial(ctx);
          (memcmp(ctx.res, want.data(), want.size() * 4) == 0 ? ok : nok)++;
          deallocGEMMContext(ctx);
        }
  report("Bipolar times regular tests", ok, nok);
}

static void test_bipolar_times_bipolar() {
  int ok = 0, nok = 0;
  for (size_t d : DIMS)
    for (int mv = 0; mv < 2; mv++) {
      const size_t n = mv ? 1 : d;
      std::vector<int8_t> l(d * d), r(n * d);
      std::vector<int32_t> want(d * n);
      random_values(1, l.size(), l.data(), true);
      random_values(1, r.size(), r.data(), true);
      GEMMContext ctx = allocGEMMContext(d, d, n, 1, 1, true, true);
      ctx.lhs.importRegular(l.data());
      ctx.rhs.importRegular(r.data());
      gemmBitSerial(ctx);
      naive_gemm(l.data(), r.data(), want.data(), d, d, n);
      (memcmp(ctx.res, want.data(), want.size() * 4) == 0 ? ok : nok)++;
      deallocGEMMContext(ctx);
    }
  report("Bipolar times bipolar tests", ok, nok);
}

// Not in the reference's tests: signed multi-bit and mixed widths.
static void test_signed_mixed() {
  int ok = 0, nok = 0;
  const size_t dims[] = {3, 17, 64, 129};
  for (size_t lb = 1; lb <= 4; lb++)
    for (size_t rb = 1; rb <= 4; rb++)
      for (int ls = 0; ls < 2; ls++)
        for (int rs = 0; rs < 2; rs++)
          for (size_t d : dims) {
            const size_t M = d, K = 3 * d + 5, N = d + 2;
            std::vector<int8_t> a(M * K), b(N * K);
            std::vector<int32_t> want(M * N);
            random_values(lb, a.size(), a.data(), ls);
            random_values(rb, b.size(), b.data(), rs);
            naive_gemm(a.data(), b.data(), want.data(), M, K, N);
            GEMMContext ctx = allocateGEMMContext(M, K, N, lb, rb, ls, rs);
            ctx.lhs.importRegular(a.data());
            ctx.rhs.importRegular(b.data());
            gemmBitSerial(ctx);
            (memcmp(ctx.res, want.data(), want.size() * 4) == 0 ? ok : nok)++;
            deallocGEMMContext(ctx);
          }
  report("Signed / mixed bitwidth tests", ok, nok);
}

// Column-major import, host-side edits after import, copies sharing planes.
static void test_layout_and_coherence() {
  int ok = 0, nok = 0;
  const size_t M = 37, K = 150, N = 23;
  std::vector<int8_t> a(M * K), b(N * K), bt(N * K);
  std::vector<int32_t> want(M * N);
  random_values(3, a.size(), a.data(), true);
  random_values(2, b.size(), b.data(), false);
  for (size_t n = 0; n < N; n++)
    for (size_t k = 0; k < K; k++) bt[k * N + n] = b[n * K + k];
  GEMMContext ctx = allocGEMMContext(M, K, N, 3, 2, true, false);
  ctx.lhs.importRegular(a.data());
  BitSerialMatrix rcopy = ctx.rhs;
  rcopy.importRegular(bt.data(), true);  // column-major source through a copy
  gemmBitSerial(ctx);
  naive_gemm(a.data(), b.data(), want.data(), M, K, N);
  (memcmp(ctx.res, want.data(), want.size() * 4) == 0 ? ok : nok)++;
  // packed words are visible on the host: get() agrees with the values
  bool bits_ok = true;
  for (size_t n = 0; n < N && bits_ok; n++)
    for (size_t k = 0; k < K; k++)
      if (ctx.rhs.get(0, n, k) != (bool)(b[n * K + k] & 1) ||
          ctx.rhs.get(1, n, k) != (bool)(b[n * K + k] & 2)) {
        bits_ok = false;
        break;
      }
  (bits_ok ? ok : nok)++;
  // host-side set(): rhs element (0, 5) becomes 3; the next GEMM sees it
  ctx.rhs.set(0, 0, 5);
  ctx.rhs.set(1, 0, 5);
  b[5] = 3;
  gemmBitSerial(ctx);
  naive_gemm(a.data(), b.data(), want.data(), M, K, N);
  (memcmp(ctx.res, want.data(), want.size() * 4) == 0 ? ok : nok)++;
  // clearAll() then GEMM gives zeros
  ctx.rhs.clearAll();
  gemmBitSerial(ctx);
  bool zero = true;
  for (size_t i = 0; i < M * N; i++) zero &= ctx.res[i] == 0;
  (zero ? ok : nok)++;
  deallocGEMMContext(ctx);
  report("Layout and coherence tests", ok, nok);
}

// importRegularAndQuantize<T> (gemmbitserial.hpp:281-308): comparisons in
// T (uint32 values and thresholds above 2^31, floats), no thresholds (the
// element's own low bits), row- and column-major sources.
template <typename T>
static int threshold_case(size_t R, size_t C, int nthres, bool colmajor, T base, T step) {
  const uint64_t nbits = 3;
  std::vector<T> vals(R * C), thr((size_t)nthres * R + 1);
  for (size_t i = 0; i < R * C; i++) vals[i] = (T)(base + (T)(rnd() % 40) * step);
  for (size_t r = 0; r < R; r++) {
    T t0 = (T)(base + (T)(rnd() % 10) * step);
    for (int t = 0; t < nthres; t++) {
      thr[(size_t)t * R + r] = t0;
      t0 = (T)(t0 + (T)(1 + rnd() % 10) * step);
    }
  }
  std::vector<T> src(R * C);
  for (size_t r = 0; r < R; r++)
    for (size_t c = 0; c < C; c++) src[colmajor ? c * R + r : r * C + c] = vals[r * C + c];
  BitSerialMatrix m = BitSerialMatrix::alloc(nbits, R, C, false);
  m.importRegularAndQuantize(src.data(), thr.data(), nthres, colmajor);
  int bad = 0;
  for (size_t r = 0; r < R; r++)
    for (size_t c = 0; c < C; c++) {
      const T e = vals[r * C + c];
      uint32_t q;
      if (nthres == 0) {
        q = (uint8_t)e;
      } else {
        q = (uint32_t)nthres;
        for (int t = 0; t < nthres; t++)
          if (e <= thr[(size_t)t * R + r]) { q = (uint32_t)t; break; }
      }
      uint32_t got = 0;
      for (uint64_t b = 0; b < nbits; b++) got |= (uint32_t)m.get(b, r, c) << b;
      bad += got != (q & 7u);
    }
  BitSerialMatrix::dealloc(m);
  return bad;
}

static void test_threshold_import() {
  int ok = 0, nok = 0;
  for (int nthres : {0, 1, 3})
    for (int cm = 0; cm < 2; cm++) {
      (threshold_case<uint32_t>(19, 77, nthres, cm, 0x80000000u, 1000u) ? nok : ok)++;
      (threshold_case<uint32_t>(5, 130, nthres, cm, 0xfffff000u, 97u) ? nok : ok)++;
      (threshold_case<int16_t>(7, 65, nthres, cm, -200, 11) ? nok : ok)++;
      (threshold_case<float>(9, 100, nthres, cm, 0.25f, 0.5f) ? nok : ok)++;
      (threshold_case<double>(3, 64, nthres, cm, 1.0, 0.125) ? nok : ok)++;
      (threshold_case<uint8_t>(11, 70, nthres, cm, 3, 5) ? nok : ok)++;
    }
  report("Threshold import tests", ok, nok);
}

// The host mirror is filled lazily: after an import `data` is filled by
// syncToHost() or the first host accessor, and equals the device planes.
static void test_lazy_mirror() {
  int ok = 0, nok = 0;
  const size_t M = 70, K = 300;
  std::vector<int8_t> a(M * K);
  random_values(2, a.size(), a.data(), true);
  BitSerialMatrix m = BitSerialMatrix::alloc(2, M, K, true, 4, 128);
  m.importRegular(a.data());
  std::vector<uint64_t> dev(m.nbits * m.wordsPerBitplane());
  (gbs_planes_to_host(m.dev->dev, dev.data()) == GBS_OK ? ok : nok)++;
  BitSerialMatrix copy = m;  // copies share the mirror and its state
  copy.syncToHost();
  (memcmp(m.data, dev.data(), dev.size() * 8) == 0 ? ok : nok)++;
  // a fresh import marks the mirror stale again; word() fills it first
  for (auto &v : a) v = (int8_t)(-v);
  m.importRegular(a.data());
  (gbs_planes_to_host(m.dev->dev, dev.data()) == GBS_OK ? ok : nok)++;
  const uint64_t w = m.word(1, 3, 64);
  (w == dev[1 * m.wordsPerBitplane() + 3 * m.wordsPerRow() + 1] ? ok : nok)++;
  (memcmp(m.data, dev.data(), dev.size() * 8) == 0 ? ok : nok)++;
  BitSerialMatrix::dealloc(m);
  report("Lazy host mirror tests", ok, nok);
}

static void print_geometry() {
  unsigned long long M, K, N;
  while (scanf("%llu %llu %llu", &M, &K, &N) == 3) {
    GEMMContext ctx = allocGEMMContext(M, K, N, 1, 1, false, false);
    printf("%llu %llu %llu %llu %llu %llu %llu %llu %llu\n", M, K, N,
           (unsigned long long)ctx.lhsBlock, (unsigned long long)ctx.rhsBlock,
           (unsigned long long)ctx.lhs.nrows_a, (unsigned long long)ctx.lhs.ncols_a,
           (unsigned long long)ctx.rhs.nrows_a, (unsigned long long)ctx.rhs.ncols_a);
    deallocGEMMContext(ctx);
  }
}

static void test_errors() {
  int ok = 0, nok = 0;
  GEMMContext ctx = allocGEMMContext(4, 10, 4, 2, 2, false, false);
  BitSerialMatrix other = BitSerialMatrix::alloc(2, 4, 11, false);
  GEMMContext bad = ctx;
  bad.rhs = other;
  try {
    gemmBitSerial(bad);
    nok++;
  } catch (const Error &e) {
    (e.code == GBS_EINVAL ? ok : nok)++;
  }
  BitSerialMatrix::dealloc(other);
  deallocGEMMContext(ctx);
  report("Error handling tests", ok, nok);
}

int main(int argc, char **argv) {
  if (argc > 1 && strcmp(argv[1], "--geometry") == 0) {
    print_geometry();
    return 0;
  }
  test_conversions();
  test_rowwise_sum();
  if (argc > 1) test_mnist(argv[1]);
  test_matrix_matrix();
  test_bipolar_times_regular();
  test_bipolar_times_bipolar();
  test_signed_mixed();
  test_layout_and_coherence();
  test_threshold_import();
  test_lazy_mirror();
  test_errors();
  printf(g_fail ? "Some tests failed\n" : "All tests completed successfully\n");
  return g_fail;
}
