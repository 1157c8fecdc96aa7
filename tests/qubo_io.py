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


def published_grid(n):
    """Rows of the reference's published annealer runs on the N=n Chimera
    instance (benchmarks/annealing/results/<n>/001.csv, committed as
    golden/data/chimera/<n>_results.csv): dicts with beta_min, num_iter,
    num_tries, schedule, the printed energy and the returned state."""
    import os

    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "data", "chimera",
                        f"{n}_results.csv")
    rows = []
    with open(path) as f:
        next(f)
        for line in f:
            bmin, it, k, sch, e, hx = line.strip().split(",")
            bits = format(int(hx, 16), f"0{n}b")
            rows.append({"beta_min": float(bmin), "num_iter": int(it), "num_tries": int(k),
                         "schedule": sch, "energy": float(e),
                         "state": np.frombuffer(bits.encode(), dtype=np.uint8) - ord("0")})
    return rows
