"""Texts for the `.qubo` reader tests: the reference's own test strings
(tests/io_test.cpp:12-124), the grammar cases of the product reader's tests,
and generated texts.  Shared by the oracle tests (CPU) and the GPU parity
tests."""
import numpy as np

# the reference's tests/io_test.cpp:12-19 (accepted, n = 3)
REF_VALID = ("c qubo Target MaxNodes NumNodes NumLinks\n"
             "p qubo 0 100 2 2\n"
             "c comment\n"
             "0 0 -0.5\n"
             "0 1 2.0\n"
             "1 2 4\n"
             "2 2 -0.7\n")

# tests/io_test.cpp:21-35 (grammar), :76-80 (lower triangle), :96-100 (count)
REF_REJECTED = [
    ("p qubo 0 1 100 3 492\n0 0 -0.5\n0 1 2.0\n", 1),
    ("p qubo 0 100 3 492\n0 0 1 -0.5\n0 1 2.0\n", 1),
    ("p qubo 0 100 3 492\n0 0 -0.5\n0 2.0\n", 1),
    ("p qubo 0 100 492\n0 0 -0.5\n0 1 2.0\n", 1),
    ("p qubo 0 100 3 492\n0 0 -0.5\nunexpected string\n0 1 2.0\n", 1),
    ("p qubo 0 100 2 2\n0 0 -0.5\n1 0 2.0\n1 2 4\n2 2 -0.7\n", 2),
    ("p qubo 0 100 3 3\n0 0 -0.5\n0 1 2.0\n1 2 4\n2 2 -0.7\n", 6),
]

H = "p qubo 0 4 1 0\n"
# (text, expected code) -- 0 accepted; codes as oracle.QP_MESSAGES
GRAMMAR = [
    ("p qubo   0\t4 1 0\n0 0 1\n", 0),
    ("p  qubo 0 4 1 0\n0 0 1\n", 1),
    ("p qubo0 4 1 0\n0 0 1\n", 0),
    (H + "0\t0 1\n", 1),
    (H + "0 0\t1\n", 1),
    (H + "  0   0 \t 1.5  \n", 0),
    (H + "\n0 0 1\n", 1),
    (H + "0 0 1\n\n", 1),
    (H + "0 0 1\n   ", 0),
    (H + "0 0 1\n \t ", 0),
    (H + "0 0 1", 0),
    ("p qubo 0 4 1 0\r\n0 0 1\r\n", 0),
    ("p qubo 0 4 1 0\r0 0 1\r", 0),
    ("p qubo 0 4 1 0\r\r0 0 1\r", 1),
    ("p qubo 0 4 1 0\n\r\n0 0 1\n", 1),
    ("p qubo 0 4 1 1\n0 0 1\n0 0 2\n0 1 3\n0 1 4\n", 0),
    ("p qubo 0 1 2 0\n0 0 1\n1 1 1\n", 5),
    ("p qubo 0 4 2 0\n0 0 1\n", 7),
    ("p qubo 0 4 0 0\n", 3),
    ("p qubo 0 4 0 0", 3),
    ("", 3),
    ("c only a comment\n", 3),
    ("0 0 1\n", 4),
    ("c x\n0 0 1\n", 4),
    ("0 0 1\np qubo 0 4 1 0\n", 1),
    ("p qubo 0 4 1 0\np qubo 0 4 1 0\n0 0 1\n", 1),
    ("c a\nc b\np qubo 0 4 1 0\nc c\n0 0 1\nc d", 0),
    ("c a\n  p qubo 0 4 1 0\n0 0 1\n", 0),
    ("p qubo 0 4 1\n0 0 1\n", 1),
    ("p qubo 0 4 1 0 7\n0 0 1\n", 1),
    (H + "c\n0 0 1\nc", 0),
    (H + "cfoo\t bar ~!\n0 0 1\n", 0),
    (H + "c caf\xc3\xa9\n0 0 1\n", 1),
    (H + "c bell\x07\n0 0 1\n", 1),
    (H + "0 0 1\x00\n", 1),
    (H + "0 0 -inf\n", 0),
    (H + "0 0 +nan\n", 0),
    (H + "0 0 NaN\n", 0),
    (H + "0 0 nan(abc)\n", 0),
    (H + "0 0 -nan()\n", 0),
    (H + "0 0 nan(\n)\n", 0),
    (H + "0 0 nan(\n0 1 2\n", 1),
    (H + "0 0 Infinity\n", 0),
    (H + "0 0 INF\n", 0),
    (H + "0 0 infin\n", 1),
    (H + "0 0 na\n", 1),
    (H + "0 0 1.\n", 0),
    (H + "0 0 .5\n", 0),
    (H + "0 0 .\n", 1),
    (H + "0 0 1e\n", 1),
    (H + "0 0 1e+\n", 1),
    (H + "0 0 1e5\n", 0),
    (H + "0 0 1E-5\n", 0),
    (H + "0 0 -\n", 1),
    (H + "0 0 +\n", 1),
    (H + "0 0 --1\n", 1),
    (H + "0 0 1 2\n", 1),
    (H + "-1 0 1\n", 1),
    (H + "+1 1 1\n", 1),
    (H + "0 0\n", 1),
    (H + "0 0 \n", 1),
    (H + "4294967296 4294967296 1\n", 1),
    ("p qubo 0 4294967296 1 0\n0 0 1\n", 1),
    (H + "1 0 2.0 junk\n", 2),
    (H + "0 0 1\njunk\n1 0 2.0\n", 1),
    (H + "1 0 2.0\njunk\n", 2),
    (H + "0 0 1 c\n", 1),
    (H + "00 0000 1\n", 0),
    ("p qubo 0 4 1 1\n0 0 1\n1 2 -0\n", 0),
    ("p qubo 0 4 1 1\n0 0 -0\n1 2 5\n", 0),
    ("p qubo 0 100 1 1\n7 7 1\n2 40 3\n", 0),
    ("p qubo 0 4 1 0\n0 0 1\n\t\n", 1),
]


def decimal_tokens(rng, count):
    """Decimal strings in the forms the grammar accepts: %.17g of doubles over
    the whole exponent range, short decimals, long mantissas, leading zeros,
    explicit signs and exponents, and hand-picked hard cases (ties, the
    smallest normal and subnormal, the overflow threshold)."""
    hard = ["9007199254740993", "9007199254740992.5", "9007199254740994.999",
            "1e23", "8.98846567431158e307", "2.2250738585072011e-308",
            "2.2250738585072014e-308", "4.9e-324", "2.4703282292062327e-324",
            "2.4703282292062328e-324", "1.7976931348623157e308", "1.7976931348623158e308",
            "1.7976931348623159e308", "1e309", "1e-400", "0.1", "0.3", "-0.0", "+0", "0e999",
            "123456789012345678901234567890", "0.000000000000000000000000000123",
            "1.00000000000000011102230246251565404236316680908203125",
            "1.00000000000000011102230246251565404236316680908203124",
            "1.00000000000000011102230246251565404236316680908203126",
            "7.2057594037927933e16", "5e-324", "1.5", "00000001.2500000", "18446744073709551615",
            "18446744073709551616", "9999999999999999999", "99999999999999999999",
            "3.0517578125e-05", "1." + "0" * 30 + "1", "4503599627370496.5",
            "4503599627370497.5", "1e22", "1e-22", "123e-2", "123.456e+2"]
    out = list(hard)
    k = max(0, count - len(out))
    bits = rng.integers(0, 2**63, k // 3, dtype=np.int64) | (
        rng.integers(0, 2, k // 3, dtype=np.int64) << 63)
    d = bits.view(np.float64)
    d = d[np.isfinite(d)]
    out += ["%.17g" % x for x in d]
    out += ["%.*g" % (int(p), x) for p, x in zip(rng.integers(1, 17, k // 3),
                                                rng.uniform(-1000, 1000, k // 3))]
    rest = count - len(out)
    for _ in range(max(0, rest)):
        nd = int(rng.integers(1, 24))
        digits = "".join(rng.choice(list("0123456789"), nd))
        dot = int(rng.integers(0, nd + 1))
        s = digits[:dot] + "." + digits[dot:] if rng.random() < 0.7 else digits
        if s in (".",):
            s = "0"
        if rng.random() < 0.5:
            s += "eE"[int(rng.integers(0, 2))] + ["", "+", "-"][int(rng.integers(0, 3))] + str(
                int(rng.integers(0, 340)))
        if rng.random() < 0.3:
            s = "+-"[int(rng.integers(0, 2))] + s
        out.append(s)
    return out[:max(count, len(hard))]


def linear_text(tokens):
    """One linear term per token: "k k token"."""
    m = len(tokens)
    return "p qubo 0 %d %d 0\n" % (m, m) + "".join("%d %d %s\n" % (k, k, t)
                                                   for k, t in enumerate(tokens))


def dense_text(q, digits=17, width=None):
    """`.qubo` text of the upper triangle of q (row-major, %.{digits}g);
    fixed-width index fields when `width` is given."""
    n = q.shape[0]
    iu, ju = np.triu_indices(n)
    nl = n
    nq = len(iu) - n
    lines = ["p qubo 0 %d %d %d\n" % (n, nl, nq)]
    if width:
        fmt = "%{0}d %{0}d %.{1}g\n".format(width, digits)
    else:
        fmt = "%d %d %.{0}g\n".format(digits)
    lines += [fmt % (i, j, q[i, j]) for i, j in zip(iu, ju)]
    return "".join(lines)
