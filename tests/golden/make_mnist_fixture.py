"""Write tests/golden/mnist.npz from the reference's MNIST known-answer test
data (test/mnistdata.h: weights, input vector and golden result of
test_mnist, test/test.cpp:157-172).

Reads the header as text (a data table) and stores the same numbers compactly:
    weights_bits  uint8, np.packbits of (weights == +1), 256 x 784 bipolar
    x             uint8, 784 values in [0, 3] (2-bit unsigned)
    golden        int32, 256 expected outputs of weights @ x

    python tests/golden/make_mnist_fixture.py [/path/to/reference]
"""
import os
import re
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def parse_array(text, name):
    m = re.search(r"%s\[\]\s*=\s*\{(.*?)\};" % re.escape(name), text, re.S)
    if not m:
        raise SystemExit(f"array {name} not found")
    return np.array([int(t) for t in m.group(1).replace("\n", " ").split(",") if t.strip()],
                    dtype=np.int64)


def main():
    ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    text = open(os.path.join(ref, "test", "mnistdata.h")).read()
    defs = dict(re.findall(r"#define\s+(MNIST_\w+)\s+(\w+)", text))
    n_in, n_out = int(defs["MNIST_IN"]), int(defs["MNIST_OUT"])
    assert (defs["MNIST_WBITS"], defs["MNIST_ABITS"]) == ("1", "2")
    assert (defs["MNIST_WSIGN"], defs["MNIST_ASIGN"]) == ("true", "false")
    x = parse_array(text, "mnist_in")
    golden = parse_array(text, "mnist_res_golden")
    w = parse_array(text, "mnist_weights")
    assert x.size == n_in and golden.size == n_out and w.size == n_in * n_out
    assert set(np.unique(w)) <= {-1, 1} and x.min() >= 0 and x.max() <= 3
    np.savez_compressed(os.path.join(HERE, "mnist.npz"),
                        weights_bits=np.packbits((w == 1).astype(np.uint8)),
                        x=x.astype(np.uint8), golden=golden.astype(np.int32),
                        shape=np.array([n_out, n_in], np.int64))
    print("wrote", os.path.join(HERE, "mnist.npz"))


if __name__ == "__main__":
    main()
