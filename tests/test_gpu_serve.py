"""Serving loop: every request's result equals the exact forward pass and is
the same alone as in the mix; admission respects the row capacity."""
import numpy as np
import pytest

from gemmbitserial_amd import workloads as W

pytestmark = pytest.mark.gpu


def make_server(gbs, h, Ws, shifts, **kw):
    mats = [gbs.from_values(w, 4, True) for w in Ws]
    return gbs.Server(mats, shifts, act_bits=4, **kw), mats


def drain(srv, ids, max_steps=10000):
    out = {}
    for _ in range(max_steps):
        srv.step()
        for i in list(ids):
            r = srv.take(i)
            if r is not None:
                out[i] = r
                ids.remove(i)
        if not ids:
            return out
    raise AssertionError("requests did not finish")


def small_mlp(seed, layers=3):
    rng = np.random.default_rng(seed)
    h = [int(x) for x in rng.integers(70, 300, layers + 1)]
    Ws = [W.values(rng, (h[l + 1], h[l]), 4, True) for l in range(layers)]
    shifts = [W.mlp_shift(h[l]) for l in range(layers - 1)]
    return h, Ws, shifts


def test_serve_mix_equals_alone_and_reference(gbs):
    h, Ws, shifts = small_mlp(1)
    reqs = W.serving_requests(h[0], 12, 1000.0, seed=2, max_rows=20)
    srv, mats = make_server(gbs, h, Ws, shifts, max_requests=5, max_rows=20)
    for r in reqs:
        srv.submit(r["id"], r["x"])
    mixed = drain(srv, [r["id"] for r in reqs])
    st = srv.stats()
    assert st["done_requests"] == len(reqs) and st["waiting"] == 0
    srv.close()
    for r in reqs:
        want = W.mlp_reference(r["x"], Ws, shifts)
        assert np.array_equal(mixed[r["id"]][0], want), r["id"]
    # alone: a fresh server per request
    for r in reqs[:4]:
        s1, _ = make_server(gbs, h, Ws, shifts, max_requests=1, max_rows=20)
        s1.submit(r["id"], r["x"])
        alone = drain(s1, [r["id"]])
        assert np.array_equal(alone[r["id"]][0], mixed[r["id"]][0])
        s1.close()


def test_serve_staggered_arrivals(gbs):
    """Requests submitted between steps join while others are in flight."""
    h, Ws, shifts = small_mlp(3, layers=4)
    reqs = W.serving_requests(h[0], 16, 1000.0, seed=4, max_rows=8)
    srv, mats = make_server(gbs, h, Ws, shifts, max_requests=8, max_rows=8)
    pending = [r["id"] for r in reqs]
    got = {}
    for r in reqs:
        srv.submit(r["id"], r["x"])
        srv.step()
        for i in list(pending):
            t = srv.take(i)
            if t is not None:
                got[i] = t
                pending.remove(i)
    got.update(drain(srv, pending))
    for r in reqs:
        assert np.array_equal(got[r["id"]][0], W.mlp_reference(r["x"], Ws, shifts))
        arr, done = got[r["id"]][1], got[r["id"]][2]
        assert done >= arr
    srv.close()


def test_serve_admission_by_memory(gbs):
    h, Ws, shifts = small_mlp(5)
    # a budget that holds only ~10 rows per cohort: requests queue, not fail
    # (per row: two pools per layer, each planes and int8 copy, every
    # layer's int32 accumulators, the input and the last layer's output
    # staging)
    per_row = sum(2 * (4 * ((x + 63) // 64 * 64) // 8 + (x + 255) // 256 * 256) for x in h[:-1]) \
        + sum(4 * x for x in h[1:]) + h[0] + 4 * h[-1]
    srv, mats = make_server(gbs, h, Ws, shifts, max_requests=64, max_rows=8,
                            mem_budget=per_row * 10)
    assert srv.stats()["row_capacity"] == 10
    reqs = W.serving_requests(h[0], 10, 1000.0, seed=6, max_rows=8)
    for r in reqs:
        srv.submit(r["id"], r["x"])
    srv.step()
    st = srv.stats()
    assert st["waiting"] > 0  # not everything fit into the first cohort
    got = drain(srv, [r["id"] for r in reqs])
    for r in reqs:
        assert np.array_equal(got[r["id"]][0], W.mlp_reference(r["x"], Ws, shifts))
    with pytest.raises(gbs.GbsError):
        srv.submit(99, np.zeros((9, h[0]), np.uint8))  # more rows than max_rows
    srv.close()
    with pytest.raises(gbs.GbsError, match="memory budget"):
        make_server(gbs, h, Ws, shifts, max_requests=4, max_rows=8, mem_budget=per_row * 4)


def test_serve_full_size_mlp(gbs):
    h, Ws, shifts = W.serving_mlp()
    reqs = W.serving_requests(h[0], 24, 1000.0, seed=7)
    srv, mats = make_server(gbs, h, Ws, shifts, max_requests=16, max_rows=64)
    for r in reqs:
        srv.submit(r["id"], r["x"])
    got = drain(srv, [r["id"] for r in reqs])
    for r in reqs:
        assert np.array_equal(got[r["id"]][0], W.mlp_reference(r["x"], Ws, shifts))
    srv.close()


def test_serve_uses_reimported_weights(gbs):
    """Weights re-imported between steps: requests served after the change
    get the new weights' exact forward pass."""
    h, Ws, shifts = small_mlp(7)
    reqs = W.serving_requests(h[0], 4, 1000.0, seed=8, max_rows=10)
    srv, mats = make_server(gbs, h, Ws, shifts, max_requests=4, max_rows=10)
    srv.submit(reqs[0]["id"], reqs[0]["x"])
    first = drain(srv, [reqs[0]["id"]])
    assert np.array_equal(first[reqs[0]["id"]][0], W.mlp_reference(reqs[0]["x"], Ws, shifts))
    rng = np.random.default_rng(9)
    Ws2 = [W.values(rng, w.shape, 4, True) for w in Ws]
    for m, w in zip(mats, Ws2):
        m.import_regular(w)
    for r in reqs[1:]:
        srv.submit(r["id"], r["x"])
    later = drain(srv, [r["id"] for r in reqs[1:]])
    for r in reqs[1:]:
        assert np.array_equal(later[r["id"]][0], W.mlp_reference(r["x"], Ws2, shifts)), r["id"]
    srv.close()


@pytest.mark.parametrize("layers", [1, 2])
def test_serve_shallow_stacks(gbs, layers):
    """One layer (its rows go straight out, no requantization) and two
    layers: every request equals the exact forward pass, over enough steps
    that accumulators are reused by later cohorts."""
    h, Ws, shifts = small_mlp(40 + layers, layers=layers)
    reqs = W.serving_requests(h[0], 9, 1000.0, seed=41, max_rows=12)
    srv, mats = make_server(gbs, h, Ws, shifts, max_requests=3, max_rows=12)
    for r in reqs:
        srv.submit(r["id"], r["x"])
    got = drain(srv, [r["id"] for r in reqs])
    for r in reqs:
        assert np.array_equal(got[r["id"]][0], W.mlp_reference(r["x"], Ws, shifts)), r["id"]
    srv.close()


def _layer_codes(x, Ws, shifts, upto):
    """Requantized activations entering layer `upto` (the exact forward pass)."""
    a = x.astype(np.int64)
    for l in range(upto):
        acc = a @ Ws[l].astype(np.int64).T
        a = np.clip(acc >> shifts[l], 0, 15)
    return a


@pytest.mark.parametrize("widths", [(150, 200, 90, 120), (3000, 1100, 4096, 1500)])
def test_serve_fused_requant_writes_planes(gbs, orc, widths):
    """The step launch requantizes each layer's sums into the next layer's
    input: narrow layers (whole-K tiles, converted from registers in the
    epilogue) and wide ones (split along K, converted after the pieces).  The bit-planes it leaves for every layer equal the oracle's
    import of the exact requantized activations, and the final rows equal
    the exact forward pass."""
    from oracle import orc as O
    rng = np.random.default_rng(sum(widths))
    h = list(widths)
    Ws = [W.values(rng, (h[l + 1], h[l]), 4, True) for l in range(len(h) - 1)]
    shifts = [W.mlp_shift(h[l]) for l in range(len(h) - 2)]
    srv, mats = make_server(gbs, h, Ws, shifts, max_requests=24, max_rows=40)
    xs = [rng.integers(0, 16, (n, h[0])).astype(np.uint8) for n in (7, 33, 1)]
    for i, x in enumerate(xs):
        srv.submit(i, x)
    stacked = np.concatenate(xs).astype(np.int64)
    got = {}
    for layer in range(1, len(h) - 1):
        srv.step()  # the cohort moves to `layer`
        planes, rows = srv.layer_input(layer)
        assert rows == stacked.shape[0]
        want = O.import_regular(_layer_codes(stacked, Ws, shifts, layer).astype(np.int8), 4,
                                False, nrows_a=rows, ncols_a=(h[layer] + 63) // 64 * 64)
        assert np.array_equal(planes[:, :rows, :].reshape(-1), want.planes.reshape(-1)), layer
        for i in range(len(xs)):
            t = srv.take(i)
            if t is not None:
                got[i] = t
    got.update(drain(srv, [i for i in range(len(xs)) if i not in got]))
    for i, x in enumerate(xs):
        assert np.array_equal(got[i][0], W.mlp_reference(x, Ws, shifts)), i
    st = srv.stats()
    assert st["layer_steps"] == len(Ws)
    assert st["hbm_bytes"] > 0
    srv.close()


def test_serve_finished_lists_ready_ids(gbs):
    """gbs_server_finished names exactly the requests that take() would
    return, and taking one removes it from the list."""
    h, Ws, shifts = small_mlp(11, layers=2)
    reqs = W.serving_requests(h[0], 6, 1000.0, seed=12, max_rows=5)
    srv, mats = make_server(gbs, h, Ws, shifts, max_requests=6, max_rows=5)
    for r in reqs:
        srv.submit(r["id"], r["x"])
    seen = {}
    for _ in range(50):
        srv.step()
        for i in srv.finished():
            assert i not in seen
            t = srv.take(i)
            assert t is not None
            seen[i] = t
        assert not srv.finished()
        if len(seen) == len(reqs):
            break
    assert sorted(seen) == [r["id"] for r in reqs]
    for r in reqs:
        assert np.array_equal(seen[r["id"]][0], W.mlp_reference(r["x"], Ws, shifts))
    srv.close()


def test_serve_cohorts_share_the_layers(gbs):
    """Under load a step admits at most max_requests / layers requests, so
    every layer holds a cohort: with 12 requests waiting and 4 layers x 2,
    the first cohort has 2 of them and the next step admits the next 2."""
    h, Ws, shifts = small_mlp(21, layers=4)
    reqs = W.serving_requests(h[0], 12, 1000.0, seed=22, max_rows=6)
    srv, mats = make_server(gbs, h, Ws, shifts, max_requests=8, max_rows=6)
    for r in reqs:
        srv.submit(r["id"], r["x"])
    srv.step()
    _, rows1 = srv.layer_input(1)
    assert rows1 == sum(r["x"].shape[0] for r in reqs[:2])
    srv.step()
    _, rows1 = srv.layer_input(1)
    _, rows2 = srv.layer_input(2)
    assert rows1 == sum(r["x"].shape[0] for r in reqs[2:4])
    assert rows2 == sum(r["x"].shape[0] for r in reqs[:2])
    got = drain(srv, [r["id"] for r in reqs])
    for r in reqs:
        assert np.array_equal(got[r["id"]][0], W.mlp_reference(r["x"], Ws, shifts)), r["id"]
    srv.close()


def test_serve_results_survive_late_take(gbs):
    """A finished request's rows stay valid however many steps pass before
    take(): rows still in the pinned output staging when it is reused are
    moved out first.  Half the requests are taken at once, half only after
    the server has run dry for a while."""
    h, Ws, shifts = small_mlp(31, layers=2)
    reqs = W.serving_requests(h[0], 10, 1000.0, seed=32, max_rows=9)
    srv, mats = make_server(gbs, h, Ws, shifts, max_requests=4, max_rows=9)
    for r in reqs:
        srv.submit(r["id"], r["x"])
    got = {}
    for _ in range(40):
        srv.step()
        for i in srv.finished():
            if i % 2 == 0:
                got[i] = srv.take(i)
    for _ in range(5):
        srv.step()
    for i in srv.finished():
        got[i] = srv.take(i)
    assert sorted(got) == [r["id"] for r in reqs]
    for r in reqs:
        assert np.array_equal(got[r["id"]][0], W.mlp_reference(r["x"], Ws, shifts)), r["id"]
    srv.close()
