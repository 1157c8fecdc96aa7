"""Minimal `.qubo` loader for the tests (well-formed files only; the
grammar-faithful reader is the C++ one in include/onesolver/model/qubo.hpp,
covered by tests/cpp/io_test.cpp).  Returns the flattened symmetric matrix
of include/helpers/qubo_helpers.hpp:26-44 and N = max index + 1."""
import numpy as np


def load_qubo(path):
    lin, quad = {}, {}
    with open(path) as f:
        for line in f:
            t = line.split()
            if not t or t[0] == "c" or line.startswith("c") or t[0] == "p":
                continue
            i, j, v = int(t[0]), int(t[1]), float(t[2])
            if i == j:
                lin.setdefault(i, v)
            else:
                quad.setdefault((i, j), v)
    n = max([k for k in lin] + [max(k) for k in quad]) + 1
    q = np.zeros((n, n), dtype=np.float64)
    for i, v in lin.items():
        q[i, i] = v
    for (i, j), v in quad.items():
        q[i, j] += v
        q[j, i] += v
    return q
