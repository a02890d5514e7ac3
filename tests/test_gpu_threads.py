"""The C ABI keeps a stream and a workspace per host thread: products run
from several threads at once stay exact."""
import threading

import numpy as np
import pytest

from conftest import rand_values
from oracle import orc as O

pytestmark = pytest.mark.gpu


def test_gemm_from_several_threads(gbs, orc):
    rng = np.random.default_rng(404)
    jobs = []
    for t in range(4):
        lb, rb = [(1, 1), (2, 3), (4, 4), (2, 2)][t]
        ls, rs = [(True, True), (True, False), (True, False), (False, True)][t]
        M, K, N = 200 + 50 * t, 1000 + 300 * t, 64 + 16 * t
        lhs, rhs = rand_values(rng, (M, K), lb, ls), rand_values(rng, (N, K), rb, rs)
        want = O.gemm(O.import_regular(lhs, lb, ls), O.import_regular(rhs, rb, rs))
        jobs.append((lhs, rhs, lb, rb, ls, rs, want))
    errors = []

    def work(job, algo):
        try:
            lhs, rhs, lb, rb, ls, rs, want = job
            gbs.init(0)
            l, r = gbs.from_values(lhs, lb, ls), gbs.from_values(rhs, rb, rs)
            for _ in range(5):
                got = gbs.gemm(l, r, algo)
                if not np.array_equal(got, want):
                    errors.append(("mismatch", lb, rb, algo))
                    return
        except Exception as e:  # reported below
            errors.append(repr(e))

    for algo in ("auto", "popcount"):
        threads = [threading.Thread(target=work, args=(j, algo)) for j in jobs]
        for th in threads:
            th.start()
        for th in threads:
            th.join()
    assert not errors, errors
