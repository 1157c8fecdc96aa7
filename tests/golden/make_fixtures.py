"""Copy the oneSolver data fixtures used by the parity tests into tests/golden/data.

Data only (MIT licence, (c) hyperQ - Ewa Hendzel): example .qubo instances,
the Chimera droplet benchmark instances in QUBO form, the published
annealer outputs and the first line of the tensor-network
ground-state files.  Run in a container where the oneSolver tree is mounted:

    python tests/golden/make_fixtures.py /root/reference

The files written are committed; nothing at test time reads the reference.
"""
import os
import shutil
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(HERE, "data")


def main(ref):
    os.makedirs(os.path.join(OUT, "examples"), exist_ok=True)
    os.makedirs(os.path.join(OUT, "chimera"), exist_ok=True)
    ex = os.path.join(ref, "benchmarks", "exhaustive_search", "examples")
    for name in sorted(os.listdir(ex)):
        if name.endswith(".qubo"):
            shutil.copyfile(os.path.join(ex, name), os.path.join(OUT, "examples", name))
    ann = os.path.join(ref, "benchmarks", "annealing")
    for n in (128, 512, 2048):
        shutil.copyfile(os.path.join(ann, "chimera_droplets_qbsolv", f"{n}power", "001.txt"),
                        os.path.join(OUT, "chimera", f"{n}_001.qubo"))
        with open(os.path.join(ann, "results", str(n), "groundstates_TN.txt")) as f:
            first = f.readline()
        with open(os.path.join(OUT, "chimera", f"{n}_groundstate_TN.txt"), "w") as f:
            f.write(first)
    # Published annealer outputs (results/<n>/001.csv): every row of the
    # grid, as beta_min,num_iter,num_tries,schedule,energy,state with the
    # state's N bits (variable 0 first) written as one hexadecimal number.
    for n in (128, 512):
        path = os.path.join(ann, "results", str(n), "001.csv")
        with open(path) as f:
            lines = f.read().splitlines()
        with open(os.path.join(OUT, "chimera", f"{n}_results.csv"), "w") as f:
            f.write("beta_min,num_iter,num_tries,schedule,energy,state\n")
            for r in lines[1:]:
                t = r.split(",")
                bits = "".join(t[:n])
                assert len(t) == n + 5 and set(bits) <= {"0", "1"}
                state = format(int(bits, 2), f"0{n // 4}x")
                f.write(",".join([t[n + 1], t[n + 2], t[n + 3], t[n + 4], t[n], state]) + "\n")


def kat_instances(ref):
    """The five (instance text, ground energy) pairs embedded in
    tests/exhaustive_test.cpp:12-142, written as kat/kat_<k>.qubo plus
    kat/expected.json."""
    import json
    import re

    with open(os.path.join(ref, "tests", "exhaustive_test.cpp")) as f:
        src = f.read()
    body = src[src.index("qubo_file_contents[]"):src.index("BOOST_AUTO_TEST_SUITE")]
    pairs = re.findall(r"make_pair<std::string, double>\(\s*((?:\"[^\"]*\"\s*)+),\s*([-0-9.]+)\)", body)
    os.makedirs(os.path.join(OUT, "kat"), exist_ok=True)
    expected = []
    for k, (lits, energy) in enumerate(pairs):
        text = "".join(bytes(s, "utf-8").decode("unicode_escape") for s in re.findall(r"\"([^\"]*)\"", lits))
        name = f"kat_{k}.qubo"
        with open(os.path.join(OUT, "kat", name), "w") as f:
            f.write(text)
        expected.append({"file": name, "energy": float(energy)})
    with open(os.path.join(OUT, "kat", "expected.json"), "w") as f:
        json.dump(expected, f, indent=1)


if __name__ == "__main__":
    kat_instances(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
