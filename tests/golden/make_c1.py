"""Config-1 fixture: dense N=28 QUBO with upper-triangle coefficients
i.i.d. U[-1,1) fp32 (Philox key (0, 0)), written with 9 significant digits,
and its exact ground state from the oracle's exhaustive search.

    python tests/golden/make_c1.py      (needs oracle/liboracle.so)
"""
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(HERE))

import numpy as np  # noqa: E402

import oracle  # noqa: E402
from onesolver_amd import instances  # noqa: E402
from qubo_io import load_qubo  # noqa: E402


def main():
    path = os.path.join(HERE, "data", "c1_dense28.qubo")
    q32 = instances.dense_uniform(28, seed=0, instance=0)
    instances.write_qubo(path, q32)
    q = load_qubo(path)
    state, e = oracle.exhaustive(q, 28, threads=0)
    with open(os.path.join(HERE, "data", "c1_ground.json"), "w") as f:
        json.dump({"n": 28, "energy": e, "state": [int(b) for b in state],
                   "energy_repr": repr(e)}, f, indent=1)
    print("ground energy", e)


if __name__ == "__main__":
    main()
