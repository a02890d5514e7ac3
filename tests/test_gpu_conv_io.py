"""im2col packing straight from NCHW tensors, and packed-matrix files."""
import numpy as np
import pytest

from conftest import rand_values
from oracle import orc as O

pytestmark = pytest.mark.gpu


def out_hw(H, W, kh, kw, s, p, d):
    return ((H + 2 * p[0] - (d[0] * (kh - 1) + 1)) // s[0] + 1,
            (W + 2 * p[1] - (d[1] * (kw - 1) + 1)) // s[1] + 1)


def im2col(x, kh, kw, s, p, d):
    """Rows (n, oy, ox), columns (c, ky, kx); zeros outside the input."""
    N, Cc, H, W = x.shape
    OH, OW = out_hw(H, W, kh, kw, s, p, d)
    xp = np.zeros((N, Cc, H + 2 * p[0], W + 2 * p[1]), x.dtype)
    xp[:, :, p[0]:p[0] + H, p[1]:p[1] + W] = x
    cols = np.empty((N, OH, OW, Cc, kh, kw), x.dtype)
    for ky in range(kh):
        for kx in range(kw):
            win = xp[:, :, ky * d[0]:ky * d[0] + s[0] * (OH - 1) + 1:s[0],
                     kx * d[1]:kx * d[1] + s[1] * (OW - 1) + 1:s[1]]
            cols[:, :, :, :, ky, kx] = win.transpose(0, 2, 3, 1)
    return cols.reshape(N * OH * OW, Cc * kh * kw)


def conv_direct(w, x, s, p, d):
    """out[n, o, oy, ox] = sum w[o, c, ky, kx] * x[n, c, oy*s - p + ky*d, ...]
    with zeros outside x, in int64."""
    O_, Cc, kh, kw = w.shape
    N, _, H, W = x.shape
    OH, OW = out_hw(H, W, kh, kw, s, p, d)
    xp = np.zeros((N, Cc, H + 2 * p[0], W + 2 * p[1]), np.int64)
    xp[:, :, p[0]:p[0] + H, p[1]:p[1] + W] = x
    out = np.zeros((N, O_, OH, OW), np.int64)
    for ky in range(kh):
        for kx in range(kw):
            win = xp[:, :, ky * d[0]:ky * d[0] + s[0] * (OH - 1) + 1:s[0],
                     kx * d[1]:kx * d[1] + s[1] * (OW - 1) + 1:s[1]]
            out += np.tensordot(w[:, :, ky, kx].astype(np.int64), win, axes=([1], [1])
                                ).transpose(1, 0, 2, 3)
    return out


GEOMS = [(2, 5, 9, 7, 3, 3, (1, 1), (1, 1), (1, 1)),
         (1, 3, 17, 13, 5, 3, (2, 1), (2, 0), (1, 1)),
         (3, 4, 8, 8, 3, 3, (1, 2), (0, 1), (2, 1)),
         (1, 70, 6, 6, 1, 1, (1, 1), (0, 0), (1, 1)),
         (2, 9, 11, 11, 3, 3, (3, 3), (1, 1), (1, 1))]


@pytest.mark.parametrize("geo", GEOMS)
@pytest.mark.parametrize("enc", [(1, True), (2, False), (3, True), (4, False)])
def test_im2col_pack_equals_import_of_host_im2col(gbs, orc, geo, enc):
    N, Cc, H, W, kh, kw, s, p, d = geo
    nb, sg = enc
    rng = np.random.default_rng(hash((geo, enc)) & 0xffff)
    x = rand_values(rng, (N, Cc, H, W), nb, sg)
    g = gbs.conv2d(N, Cc, H, W, kh, kw, s, p, d)
    rows, cols, oh, ow = gbs.conv2d_shape(g)
    host = im2col(x, kh, kw, s, p, d)
    assert host.shape == (rows, cols)
    m = gbs.DeviceMatrix(nb, rows, cols, sg)
    got = m.import_im2col(x, g, want_planes=True)
    want = O.import_regular(host, nb, sg).planes
    assert np.array_equal(got, want.reshape(-1))
    # other element types of the same values pack the same words
    for dt in (np.int32, np.int16):
        m2 = gbs.DeviceMatrix(nb, rows, cols, sg)
        assert np.array_equal(m2.import_im2col(x.astype(dt), g, want_planes=True), got)


@pytest.mark.parametrize("geo", GEOMS[:3])
@pytest.mark.parametrize("algo", ["popcount", "auto"])
def test_gemm_on_im2col_planes_is_the_convolution(gbs, orc, geo, algo):
    N, Cc, H, W, kh, kw, s, p, d = geo
    rng = np.random.default_rng(7)
    x = rand_values(rng, (N, Cc, H, W), 3, False)
    w = rand_values(rng, (6, Cc, kh, kw), 2, True)
    g = gbs.conv2d(N, Cc, H, W, kh, kw, s, p, d)
    rows, cols, oh, ow = gbs.conv2d_shape(g)
    act = gbs.DeviceMatrix(3, rows, cols, False)
    act.import_im2col(x, g)
    wt = gbs.from_values(w.reshape(6, cols), 2, True)
    res = gbs.gemm(wt, act, algo)                    # rows x 6, res[pixel*6 + o]
    want = conv_direct(w, x, s, p, d).transpose(0, 2, 3, 1).reshape(rows, 6)
    assert np.array_equal(res, want.astype(np.int32))


def test_im2col_conv_layer_at_full_size(gbs, orc):
    """16 x 256 x 56 x 56, 3 x 3, pad 1: the conv shape of the benchmark
    (50176 x 2304 operand), packed from the 12.8 MB tensor."""
    rng = np.random.default_rng(11)
    x = rng.integers(0, 8, (16, 256, 56, 56)).astype(np.int8)
    g = gbs.conv2d(16, 256, 56, 56, 3, 3, 1, 1)
    rows, cols, oh, ow = gbs.conv2d_shape(g)
    assert (rows, cols) == (50176, 2304)
    m = gbs.DeviceMatrix(3, rows, cols, False)
    got = m.import_im2col(x, g, want_planes=True).reshape(3, rows, cols // 64)
    # spot rows against the host im2col of their images
    for n in (0, 15):
        host = im2col(x[n:n + 1], 3, 3, (1, 1), (1, 1), (1, 1))
        want = O.import_regular(host, 3, False).planes.reshape(3, oh * ow, cols // 64)
        assert np.array_equal(got[:, n * oh * ow:(n + 1) * oh * ow], want)


def test_matrix_file_round_trip(gbs, orc, tmp_path):
    rng = np.random.default_rng(3)
    for (nb, sg, M, K, rows_a, cols_a) in [(2, True, 70, 300, 72, 320), (1, True, 33, 64, 33, 64),
                                          (4, False, 5, 1000, 8, 1024)]:
        v = rand_values(rng, (M, K), nb, sg)
        m = gbs.DeviceMatrix(nb, M, K, sg, rows_a, cols_a)
        m.import_regular(v)
        path = tmp_path / f"m{nb}.gbsp"
        m.save(str(path))
        blob = path.read_bytes()
        assert blob[:8] == b"GBSPLN01"
        head = np.frombuffer(blob[8:64], "<u8")
        assert list(head) == [nb, M, K, rows_a, cols_a, int(sg), nb * rows_a * cols_a // 64]
        ref = O.import_regular(v, nb, sg, rows_a, cols_a)
        assert np.array_equal(np.frombuffer(blob[64:], "<u8"), ref.planes.reshape(-1))
        back = gbs.DeviceMatrix.load(str(path))
        assert (back.nbits, back.nrows, back.ncols, back.nrows_a, back.ncols_a, back.issigned) == \
            (nb, M, K, rows_a, cols_a, sg)
        assert np.array_equal(back.planes(), m.planes())
        other = gbs.from_values(rand_values(rng, (7, K), 2, False), 2, False)
        assert np.array_equal(gbs.gemm(back, other), gbs.gemm(m, other))
