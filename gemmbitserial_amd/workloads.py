"""Seeded synthetic workloads of the benchmark configurations (shapes,
bitwidths, value distributions).  Values are int8 arrays in the regular
layout the importer takes: lhs M x K, rhs N x K."""
import numpy as np

SEED0 = 0x67627300


def values(rng, shape, nbits, issigned):
    """Uniform over the encoding's range: bipolar +-1 for 1-bit signed,
    [-2^(b-1), 2^(b-1)-1] signed, [0, 2^b - 1] unsigned."""
    if nbits == 1 and issigned:
        return np.where(rng.integers(0, 2, shape) == 1, 1, -1).astype(np.int8)
    if issigned:
        return rng.integers(-(1 << (nbits - 1)), 1 << (nbits - 1), shape).astype(np.int8)
    return rng.integers(0, 1 << nbits, shape).astype(np.int8)


def uniform_i8(rng, shape, lo, hi):
    """Uniform integers in [lo, hi] as int8."""
    return rng.integers(lo, hi + 1, shape).astype(np.int8)


def sparse_act(rng, shape, hi):
    """Activations: 0 with probability 0.5, else uniform in [1, hi]."""
    v = rng.integers(1, hi + 1, shape).astype(np.int8)
    v[rng.random(shape) < 0.5] = 0
    return v


# The single-product configurations.  Each returns a list of runs
# (encoding, lhs M x K, rhs N x K, lhs bits, rhs bits, lhs signed, rhs
# signed), drawn from np.random.default_rng(SEED0 + config number).
def plain_gemm(rng):
    """Plain GEMM 256 x 2048 x 256: lhs 2-bit unsigned in [0, 3], rhs 2-bit
    signed in [-2, 1]."""
    return [("2u x 2s", uniform_i8(rng, (256, 2048), 0, 3), uniform_i8(rng, (256, 2048), -2, 1),
             2, 2, False, True)]


def binary_layers(rng):
    """Binary layers 4096 x 8192 x 512, once bipolar (Bernoulli 0.5) and
    once unsigned {0, 1} with 30 % ones."""
    M, K, N = 4096, 8192, 512
    a_bip = np.where(rng.random((M, K)) < 0.5, 1, -1).astype(np.int8)
    b_bip = np.where(rng.random((N, K)) < 0.5, 1, -1).astype(np.int8)
    a_u = (rng.random((M, K)) < 0.3).astype(np.int8)
    b_u = (rng.random((N, K)) < 0.3).astype(np.int8)
    return [("bipolar", a_bip, b_bip, 1, 1, True, True),
            ("unsigned01", a_u, b_u, 1, 1, False, False)]


def conv_gemm(rng):
    """Conv-as-GEMM (im2col of a 3 x 3 conv over 256 channels, 16 x 56 x 56
    outputs): weights 256 x 2304 2-bit signed in [-2, 1], activations
    50176 x 2304 3-bit unsigned (50 % zeros, else [1, 7]); then at 4 x 4
    bits (weights [-8, 7], activations 50 % zeros, else [1, 15])."""
    w = uniform_i8(rng, (256, 2304), -2, 1)
    act = sparse_act(rng, (50176, 2304), 7)
    w4 = uniform_i8(rng, (256, 2304), -8, 7)
    act4 = sparse_act(rng, (50176, 2304), 15)
    return [("2s x 3u", w, act, 2, 3, True, False),
            ("4s x 4u", w4, act4, 4, 4, True, False)]


def skinny(rng):
    """Skinny products: lhs 32768 x 32768 2-bit signed times 1 and 8 rhs
    rows of 2-bit unsigned values."""
    a = rng.integers(-2, 2, (32768, 32768), dtype=np.int8)
    return [("gemv 1 row", a, rng.integers(0, 4, (1, 32768), dtype=np.int8), 2, 2, True, False),
            ("gemv 8 rows", a, rng.integers(0, 4, (8, 32768), dtype=np.int8), 2, 2, True, False)]


def log_uniform_int(rng, lo, hi):
    return int(round(float(np.exp(rng.uniform(np.log(lo), np.log(hi))))))


def batched_uneven(count=512, seed=SEED0 + 3, scale=1.0):
    """Batched uneven GEMMs: M in [64, 4096], K in [256, 8192] (multiple of
    64), N in [1, 256], each log-uniform; bitwidths uniform in {1,2,3,4};
    signedness random per side (1-bit signed = bipolar).  scale < 1 shrinks
    M and K for quick tests.  Returns a list of problem dicts."""
    rng = np.random.default_rng(seed)
    probs = []
    for _ in range(count):
        M = max(1, int(log_uniform_int(rng, 64, 4096) * scale))
        K = max(64, 64 * int(round(log_uniform_int(rng, 256, 8192) * scale / 64)))
        N = log_uniform_int(rng, 1, 256)
        lb, rb = int(rng.integers(1, 5)), int(rng.integers(1, 5))
        ls, rs = bool(rng.integers(0, 2)), bool(rng.integers(0, 2))
        probs.append(dict(M=M, K=K, N=N, lb=lb, rb=rb, ls=ls, rs=rs,
                          lhs=values(rng, (M, K), lb, ls), rhs=values(rng, (N, K), rb, rs)))
    return probs


# CaffeNet's GEMMs as (rows M, depth K, columns N), the list the reference
# benchmarks at 2-bit unsigned weights x 2-bit unsigned activations
# (benchmark/benchmark.cpp:173-198)
CAFFENET_GEMMS = [(96, 363, 3025), (256, 2400, 729), (384, 2304, 169), (384, 3456, 169),
                  (256, 3456, 169), (4096, 9216, 1), (4096, 4096, 1), (1000, 4096, 1)]


def caffenet(seed=SEED0 + 5, wbits=2, abits=2):
    """The CaffeNet GEMM list with uniform unsigned values (weights lhs
    M x K, activations rhs N x K).  Returns a list of problem dicts."""
    rng = np.random.default_rng(seed)
    probs = []
    for M, K, N in CAFFENET_GEMMS:
        probs.append(dict(M=M, K=K, N=N, lb=wbits, rb=abits, ls=False, rs=False,
                          lhs=values(rng, (M, K), wbits, False),
                          rhs=values(rng, (N, K), abits, False)))
    return probs


def mlp_shift(h_in, w_bits=4, a_bits=4):
    """Requant shift for a layer of depth h_in: acc >> s lands the typical
    accumulator (std of a sum of h_in products of uniform w_bits-bit signed
    weights and a_bits-bit unsigned activations) in the activation range."""
    ew2 = np.mean(np.arange(-(1 << (w_bits - 1)), 1 << (w_bits - 1)) ** 2.0)
    ex2 = np.mean(np.arange(0, 1 << a_bits) ** 2.0)
    std = np.sqrt(h_in * ew2 * ex2)
    return max(0, int(np.floor(np.log2(std / (1 << (a_bits - 1))))))


def serving_mlp(layers=8, lo=1024, hi=4096, seed=SEED0 + 4):
    """8-layer MLP: hidden sizes log-uniform in [lo, hi] (unrounded), 4-bit
    signed weights uniform in [-8, 7], requant shift per layer."""
    rng = np.random.default_rng(seed)
    h = [log_uniform_int(rng, lo, hi) for _ in range(layers + 1)]
    W = [values(rng, (h[l + 1], h[l]), 4, True) for l in range(layers)]
    shifts = [mlp_shift(h[l]) for l in range(layers - 1)]
    return h, W, shifts


def serving_requests(h0, count, rate_per_s, seed=SEED0 + 5, max_rows=64):
    """Poisson arrivals (exponential gaps at rate_per_s): each request a batch
    of 1..max_rows rows of 4-bit unsigned activations."""
    rng = np.random.default_rng(seed)
    gaps = rng.exponential(1.0 / rate_per_s, count)
    t = np.cumsum(gaps)
    out = []
    for i in range(count):
        rows = int(rng.integers(1, max_rows + 1))
        out.append(dict(id=i, t=float(t[i]), x=rng.integers(0, 16, (rows, h0)).astype(np.uint8)))
    return out


def mlp_reference(x, W, shifts):
    """Exact int64 forward pass of one request: acc = x W_l^T, then
    min(15, max(0, acc >> s_l)) between layers; returns the last int32
    accumulators."""
    a = x.astype(np.float64)
    for l, w in enumerate(W):
        acc = (a @ w.astype(np.float64).T).astype(np.int64)
        if l + 1 < len(W):
            a = np.clip(acc >> shifts[l], 0, 15).astype(np.float64)
    return acc.astype(np.int32)
