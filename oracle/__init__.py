"""ctypes binding of the CPU oracle (``oracle/liboracle.so``).

TEST INFRASTRUCTURE ONLY.  Imported by ``tests/``, ``__graft_entry__.smoke()``
(as the checker) and ``bench.py``'s cpu_baseline leg; the product never imports
it.  Every function mirrors one in ``oracle/oracle.h``, which cites the
reference file:line each one restates.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

_u8p = ctypes.POINTER(ctypes.c_uint8)
_f64p = ctypes.POINTER(ctypes.c_double)
_f32p = ctypes.POINTER(ctypes.c_float)
_i32p = ctypes.POINTER(ctypes.c_int32)
_u32p = ctypes.POINTER(ctypes.c_uint32)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            raise RuntimeError(f"oracle library missing: {_LIB_PATH} (run `make -C oracle`)")
        L = ctypes.CDLL(_LIB_PATH)
        L.or_philox4x32_10.argtypes = [_u32p, _u32p, _u32p]
        L.or_neg_ln_u.argtypes = [ctypes.c_uint32]
        L.or_neg_ln_u.restype = ctypes.c_float
        L.or_neg_ln_u_many.argtypes = [_u32p, ctypes.c_int64, _f32p]
        L.or_flatten.argtypes = [ctypes.c_int, ctypes.c_int, _i32p, _f64p, ctypes.c_int,
                                 _i32p, _i32p, _f64p, _f64p]
        L.or_energy.argtypes = [_f64p, _u8p, ctypes.c_int]
        L.or_energy.restype = ctypes.c_double
        L.or_energy_many.argtypes = [_f64p, _u8p, ctypes.c_int, ctypes.c_int64, _f64p]
        L.or_linear_schedule.argtypes = [ctypes.c_double, ctypes.c_double, ctypes.c_uint32, _f64p]
        L.or_geometric_schedule.argtypes = [ctypes.c_double, ctypes.c_double, ctypes.c_uint32, _f64p]
        L.or_exhaustive_literal.argtypes = [ctypes.c_int, _f64p, _u8p, _f64p]
        L.or_exhaustive.argtypes = [ctypes.c_int, _f64p, _u8p, _f64p, ctypes.c_int]
        anneal_args = [ctypes.c_int, _f64p, _f64p, ctypes.c_int, ctypes.c_int, ctypes.c_uint32,
                       ctypes.c_uint64, ctypes.c_uint32, _u8p, _f64p, ctypes.c_int, _u8p]
        L.or_anneal_literal.argtypes = anneal_args
        L.or_anneal_fields.argtypes = anneal_args
        L.or_anneal_stream.argtypes = [ctypes.c_int, _f64p, _f64p, ctypes.c_int, ctypes.c_int,
                                       ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint64,
                                       ctypes.c_int, _u8p, _f64p, ctypes.c_int, _u8p]
        L.or_sweep_dense.argtypes = [ctypes.c_int, _f64p, _f32p, ctypes.c_int, ctypes.c_uint32,
                                     ctypes.c_uint64, ctypes.c_uint32, _u8p, _u8p, _u8p, _f64p,
                                     ctypes.c_int]
        L.or_sweep_csr.argtypes = [ctypes.c_int, _i32p, _i32p, _f32p, _f32p, _f32p, ctypes.c_int,
                                   ctypes.c_uint32, ctypes.c_uint64, ctypes.c_uint32, _u8p, _u8p,
                                   _u8p, _f64p, ctypes.c_int]
        L.or_sweep_csr_order.argtypes = [ctypes.c_int, _i32p, _i32p, _f32p, _f32p, _f32p,
                                         ctypes.c_int, ctypes.c_uint32, ctypes.c_uint64,
                                         ctypes.c_uint32, _i32p, _u8p, _u8p, _u8p, _f64p,
                                         ctypes.c_int]
        L.or_csr_colour_order.argtypes = [ctypes.c_int, _i32p, _i32p, _i32p, _i32p]
        L.or_resample.argtypes = [_f64p, ctypes.c_uint32, ctypes.c_double, ctypes.c_double,
                                  _u32p]
        L.or_resample_uniform.argtypes = [ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint32]
        L.or_resample_uniform.restype = ctypes.c_double
        L.or_population_dense.argtypes = [ctypes.c_int, _f64p, _f32p, ctypes.c_int, ctypes.c_int,
                                          ctypes.c_uint32, ctypes.c_uint64, ctypes.c_uint32, _u8p,
                                          _u8p, _f64p, _u32p, ctypes.c_int]
        L.or_population_csr.argtypes = [ctypes.c_int, _i32p, _i32p, _f32p, _f32p, _f32p,
                                        ctypes.c_int, ctypes.c_int, ctypes.c_uint32,
                                        ctypes.c_uint64, ctypes.c_uint32, _i32p, _u8p, _u8p,
                                        _f64p, _u32p, ctypes.c_int]
        L.or_tempering_csr.argtypes = [ctypes.c_int, _i32p, _i32p, _f32p, _f32p, _f32p,
                                       ctypes.c_int, ctypes.c_uint32, ctypes.c_int, ctypes.c_int,
                                       ctypes.c_uint64, ctypes.c_uint32, _i32p, _u8p, _u8p,
                                       _f64p, _u32p, ctypes.c_int]
        L.or_warm_init.argtypes = [ctypes.c_int, _f32p, ctypes.c_uint32, ctypes.c_uint64,
                                   ctypes.c_uint32, _u8p]
        L.or_sweep_init.argtypes = [ctypes.c_int, ctypes.c_uint32, ctypes.c_uint64,
                                    ctypes.c_uint32, _u8p]
        _lib = L
    return _lib


def _p(a, t):
    return a.ctypes.data_as(t) if a is not None else None


def philox(ctr, key):
    c = np.ascontiguousarray(ctr, dtype=np.uint32)
    k = np.ascontiguousarray(key, dtype=np.uint32)
    o = np.zeros(4, dtype=np.uint32)
    lib().or_philox4x32_10(_p(c, _u32p), _p(k, _u32p), _p(o, _u32p))
    return o


def neg_ln_u(w: int) -> float:
    return float(lib().or_neg_ln_u(ctypes.c_uint32(w)))


def neg_ln_u_many(w):
    ww = np.ascontiguousarray(w, dtype=np.uint32)
    out = np.zeros(len(ww), dtype=np.float32)
    lib().or_neg_ln_u_many(_p(ww, _u32p), len(ww), _p(out, _f32p))
    return out


def flatten(n, linear: dict, quadratic: dict):
    li = np.array(list(linear.keys()), dtype=np.int32)
    lv = np.array(list(linear.values()), dtype=np.float64)
    qi = np.array([k[0] for k in quadratic], dtype=np.int32)
    qj = np.array([k[1] for k in quadratic], dtype=np.int32)
    qv = np.array(list(quadratic.values()), dtype=np.float64)
    out = np.zeros(n * n, dtype=np.float64)
    lib().or_flatten(n, len(li), _p(li, _i32p), _p(lv, _f64p), len(qi), _p(qi, _i32p),
                     _p(qj, _i32p), _p(qv, _f64p), _p(out, _f64p))
    return out


def energy(flat, state):
    flat = np.ascontiguousarray(flat, dtype=np.float64).ravel()
    st = np.ascontiguousarray(state, dtype=np.uint8)
    n = st.shape[-1]
    if st.ndim == 1:
        return float(lib().or_energy(_p(flat, _f64p), _p(st, _u8p), n))
    st = st.reshape(-1, n)
    out = np.zeros(st.shape[0], dtype=np.float64)
    lib().or_energy_many(_p(flat, _f64p), _p(st, _u8p), n, st.shape[0], _p(out, _f64p))
    return out


def linear_schedule(beta_min, beta_max, n):
    out = np.zeros(n, dtype=np.float64)
    lib().or_linear_schedule(beta_min, beta_max, n, _p(out, _f64p))
    return out


def geometric_schedule(beta_min, beta_max, n):
    out = np.zeros(max(n, 1), dtype=np.float64)
    lib().or_geometric_schedule(beta_min, beta_max, n, _p(out, _f64p))
    return out[:n]


def exhaustive(flat, n, literal=False, threads=0):
    flat = np.ascontiguousarray(flat, dtype=np.float64).ravel()
    st = np.zeros(n, dtype=np.uint8)
    e = np.zeros(1, dtype=np.float64)
    if literal:
        rc = lib().or_exhaustive_literal(n, _p(flat, _f64p), _p(st, _u8p), _p(e, _f64p))
    else:
        rc = lib().or_exhaustive(n, _p(flat, _f64p), _p(st, _u8p), _p(e, _f64p), threads)
    if rc != 0:
        raise ValueError("exhaustive: unsupported size")
    return st, float(e[0])


def anneal(flat, n, beta, num_tries, sweeps_per_beta=1, seed=1234, first_replica=0,
           literal=False, threads=0, final=False):
    """Reference-mode chains: (best_states, best_energies), plus the final
    (current) states when final=True."""
    flat = np.ascontiguousarray(flat, dtype=np.float64).ravel()
    beta = np.ascontiguousarray(beta, dtype=np.float64)
    bs = np.zeros((num_tries, n), dtype=np.uint8)
    be = np.zeros(num_tries, dtype=np.float64)
    fs = np.zeros((num_tries, n), dtype=np.uint8) if final else None
    fn = lib().or_anneal_literal if literal else lib().or_anneal_fields
    fn(n, _p(flat, _f64p), _p(beta, _f64p), len(beta), sweeps_per_beta, num_tries, seed,
       first_replica, _p(bs, _u8p), _p(be, _f64p), threads, _p(fs, _u8p))
    return (bs, be, fs) if final else (bs, be)


def anneal_shared_stream(flat, n, beta, num_tries, sweeps_per_beta=1, seed=1234,
                         uniform_words=1, threads=0, final=False, block_tries=0):
    """Reference-mode chains whose replicas read one stream at offsets 0, 1,
    2, ... as the reference's engines do (or_anneal_stream); with
    block_tries, independent runs of that many replicas (run b: seed + b)."""
    flat = np.ascontiguousarray(flat, dtype=np.float64).ravel()
    beta = np.ascontiguousarray(beta, dtype=np.float64)
    bs = np.zeros((num_tries, n), dtype=np.uint8)
    be = np.zeros(num_tries, dtype=np.float64)
    fs = np.zeros((num_tries, n), dtype=np.uint8) if final else None
    lib().or_anneal_stream(n, _p(flat, _f64p), _p(beta, _f64p), len(beta), sweeps_per_beta,
                           num_tries, block_tries, seed, uniform_words, _p(bs, _u8p),
                           _p(be, _f64p), threads, _p(fs, _u8p))
    return (bs, be, fs) if final else (bs, be)


def sweep_dense(flat, n, tscale, num_replicas, seed=1, instance=0, init=None, threads=0):
    flat = np.ascontiguousarray(flat, dtype=np.float64).ravel()
    ts = np.ascontiguousarray(tscale, dtype=np.float32)
    fin = np.zeros((num_replicas, n), dtype=np.uint8)
    bs = np.zeros((num_replicas, n), dtype=np.uint8)
    be = np.zeros(num_replicas, dtype=np.float64)
    ini = None if init is None else np.ascontiguousarray(init, dtype=np.uint8)
    lib().or_sweep_dense(n, _p(flat, _f64p), _p(ts, _f32p), len(ts), num_replicas, seed, instance,
                         _p(ini, _u8p), _p(fin, _u8p), _p(bs, _u8p), _p(be, _f64p), threads)
    return fin, bs, be


def tempering_dense(flat, n, tscale, ladders, num_sweeps, interval, seed=1, instance=0,
                    threads=0):
    """(final states, best states, best energies, final slots) of
    or_tempering_dense; replicas = ladders * len(tscale)."""
    flat = np.ascontiguousarray(flat, dtype=np.float64).ravel()
    ts = np.ascontiguousarray(tscale, dtype=np.float32)
    R = ladders * len(ts)
    fin = np.zeros((R, n), dtype=np.uint8)
    bs = np.zeros((R, n), dtype=np.uint8)
    be = np.zeros(R, dtype=np.float64)
    sl = np.zeros(R, dtype=np.uint32)
    L = lib()
    L.or_tempering_dense.argtypes = [ctypes.c_int, _f64p, _f32p, ctypes.c_int, ctypes.c_uint32,
                                     ctypes.c_int, ctypes.c_int, ctypes.c_uint64, ctypes.c_uint32,
                                     _u8p, _u8p, _f64p, ctypes.POINTER(ctypes.c_uint32),
                                     ctypes.c_int]
    L.or_tempering_dense(n, _p(flat, _f64p), _p(ts, _f32p), len(ts), ladders, num_sweeps,
                         interval, seed, instance, _p(fin, _u8p), _p(bs, _u8p), _p(be, _f64p),
                         sl.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)), threads)
    return fin, bs, be, sl


def csr_colour_order(row_ptr, col):
    """Greedy colour classes (or_csr_colour_order): (order, class starts)."""
    rp = np.ascontiguousarray(row_ptr, dtype=np.int32)
    cl = np.ascontiguousarray(col, dtype=np.int32)
    n = len(rp) - 1
    order = np.zeros(max(n, 1), dtype=np.int32)
    maxdeg = int(np.diff(rp).max()) if n else 0
    cs = np.zeros(maxdeg + 2, dtype=np.int32)
    k = lib().or_csr_colour_order(n, _p(rp, _i32p), _p(cl, _i32p), _p(order, _i32p),
                                  _p(cs, _i32p))
    return order[:n], cs[:k + 1]


def resample(E, dbeta, u):
    """Systematic resampling (or_resample): ancestor of every new replica."""
    e = np.ascontiguousarray(E, dtype=np.float64)
    anc = np.zeros(len(e), dtype=np.uint32)
    lib().or_resample(_p(e, _f64p), len(e), dbeta, u, _p(anc, _u32p))
    return anc


def resample_uniform(seed, instance, k):
    return float(lib().or_resample_uniform(seed, instance, k))


def population_dense(flat, n, tscale, theta, num_replicas, seed=1, instance=0, threads=0):
    """Population annealing (or_population_dense): (final, best states, best
    energies, ancestors [(K-1), R])."""
    flat = np.ascontiguousarray(flat, dtype=np.float64).ravel()
    ts = np.ascontiguousarray(tscale, dtype=np.float32)
    K, R = len(ts), int(num_replicas)
    fin = np.zeros((R, n), dtype=np.uint8)
    bs = np.zeros((R, n), dtype=np.uint8)
    be = np.zeros(R, dtype=np.float64)
    anc = np.zeros((max(K - 1, 1), R), dtype=np.uint32)
    lib().or_population_dense(n, _p(flat, _f64p), _p(ts, _f32p), K, theta, R, seed, instance,
                              _p(fin, _u8p), _p(bs, _u8p), _p(be, _f64p), _p(anc, _u32p), threads)
    return fin, bs, be, anc[:K - 1]


def sweep_csr(row_ptr, col, val, diag, tscale, num_replicas, seed=1, instance=0, init=None,
              threads=0, order=None):
    """Sparse sweep-mode chains; `order`: the sites' visiting order per sweep
    (default 0..n-1)."""
    rp = np.ascontiguousarray(row_ptr, dtype=np.int32)
    cl = np.ascontiguousarray(col, dtype=np.int32)
    vl = np.ascontiguousarray(val, dtype=np.float32)
    dg = np.ascontiguousarray(diag, dtype=np.float32)
    ts = np.ascontiguousarray(tscale, dtype=np.float32)
    n = len(dg)
    fin = np.zeros((num_replicas, n), dtype=np.uint8)
    bs = np.zeros((num_replicas, n), dtype=np.uint8)
    be = np.zeros(num_replicas, dtype=np.float64)
    ini = None if init is None else np.ascontiguousarray(init, dtype=np.uint8)
    od = None if order is None else np.ascontiguousarray(order, dtype=np.int32)
    lib().or_sweep_csr_order(n, _p(rp, _i32p), _p(cl, _i32p), _p(vl, _f32p), _p(dg, _f32p),
                             _p(ts, _f32p), len(ts), num_replicas, seed, instance, _p(od, _i32p),
                             _p(ini, _u8p), _p(fin, _u8p), _p(bs, _u8p), _p(be, _f64p), threads)
    return fin, bs, be


def population_csr(row_ptr, col, val, diag, tscale, theta, num_replicas, seed=1, instance=0,
                   order=None, threads=0):
    """Sparse population annealing (or_population_csr): (final, best states,
    best energies, ancestors [(K-1), R])."""
    rp = np.ascontiguousarray(row_ptr, dtype=np.int32)
    cl = np.ascontiguousarray(col, dtype=np.int32)
    vl = np.ascontiguousarray(val, dtype=np.float32)
    dg = np.ascontiguousarray(diag, dtype=np.float32)
    ts = np.ascontiguousarray(tscale, dtype=np.float32)
    n, K, R = len(dg), len(ts), int(num_replicas)
    fin = np.zeros((R, n), dtype=np.uint8)
    bs = np.zeros((R, n), dtype=np.uint8)
    be = np.zeros(R, dtype=np.float64)
    anc = np.zeros((max(K - 1, 1), R), dtype=np.uint32)
    od = None if order is None else np.ascontiguousarray(order, dtype=np.int32)
    lib().or_population_csr(n, _p(rp, _i32p), _p(cl, _i32p), _p(vl, _f32p), _p(dg, _f32p),
                            _p(ts, _f32p), K, theta, R, seed, instance, _p(od, _i32p),
                            _p(fin, _u8p), _p(bs, _u8p), _p(be, _f64p), _p(anc, _u32p), threads)
    return fin, bs, be, anc[:K - 1]


def tempering_csr(row_ptr, col, val, diag, tscale, ladders, num_sweeps, interval, seed=1,
                  instance=0, order=None, threads=0):
    """Sparse replica exchange (or_tempering_csr): (final, best states, best
    energies, final slots); replicas = ladders * len(tscale)."""
    rp = np.ascontiguousarray(row_ptr, dtype=np.int32)
    cl = np.ascontiguousarray(col, dtype=np.int32)
    vl = np.ascontiguousarray(val, dtype=np.float32)
    dg = np.ascontiguousarray(diag, dtype=np.float32)
    ts = np.ascontiguousarray(tscale, dtype=np.float32)
    n, K = len(dg), len(ts)
    R = ladders * K
    fin = np.zeros((R, n), dtype=np.uint8)
    bs = np.zeros((R, n), dtype=np.uint8)
    be = np.zeros(R, dtype=np.float64)
    sl = np.zeros(R, dtype=np.uint32)
    od = None if order is None else np.ascontiguousarray(order, dtype=np.int32)
    lib().or_tempering_csr(n, _p(rp, _i32p), _p(cl, _i32p), _p(vl, _f32p), _p(dg, _f32p),
                           _p(ts, _f32p), K, ladders, num_sweeps, interval, seed, instance,
                           _p(od, _i32p), _p(fin, _u8p), _p(bs, _u8p), _p(be, _f64p),
                           _p(sl, _u32p), threads)
    return fin, bs, be, sl


def warm_init(p, num_replicas, seed=1, instance=0):
    pp = np.ascontiguousarray(p, dtype=np.float32)
    st = np.zeros((num_replicas, len(pp)), dtype=np.uint8)
    lib().or_warm_init(len(pp), _p(pp, _f32p), num_replicas, seed, instance, _p(st, _u8p))
    return st


def sweep_init(n, num_replicas, seed=1, instance=0):
    st = np.zeros((num_replicas, n), dtype=np.uint8)
    lib().or_sweep_init(n, num_replicas, seed, instance, _p(st, _u8p))
    return st


def csr_energy(row_ptr, col, val, diag, states):
    """FP64 energy of CSR instances: sum_j x_j diag_j + sum_{j<l} w_jl x_j x_l."""
    st = np.asarray(states, dtype=np.float64)
    n = len(diag)
    rows = np.repeat(np.arange(n), np.diff(row_ptr))
    upper = col > rows
    st2 = st.reshape(-1, n)
    e = st2 @ np.asarray(diag, dtype=np.float64)
    e += (st2[:, rows[upper]] * st2[:, col[upper]]) @ np.asarray(val, dtype=np.float64)[upper]
    return e if st.ndim > 1 else float(e[0])


QP_MESSAGES = {
    1: "Parsing failed. Incorrect file format.",
    2: "Incorrect file, encountered coefficient from lower triangle of QUBO matrix.",
    3: "An empty input, no coefficients defined.",
    4: "No header line or header line misformatted.",
    5: "Number of linear terms is greater than num nodes.",
    6: "Number of quadratic terms is not equal to the declared one.",
    7: "Number of linear terms is not equal to the declared one.",
}


def qubo_parse(text: bytes):
    """or_qubo_parse: returns (code, message, terms) where terms is None on
    error and otherwise a dict with the distinct terms in order of first
    appearance (i, j, v arrays), n, num_linear and num_quadratic."""
    L = lib()
    i64 = ctypes.c_int64
    L.or_qubo_parse.argtypes = [ctypes.c_char_p, i64, i64, _i32p, _i32p, _f64p,
                                ctypes.POINTER(i64), _i32p, ctypes.POINTER(i64),
                                ctypes.POINTER(i64)]
    text = bytes(text)
    cap = text.count(b"\n") + text.count(b"\r") + 2
    ti = np.zeros(cap, np.int32)
    tj = np.zeros(cap, np.int32)
    tv = np.zeros(cap, np.float64)
    nt, nl, nq = i64(), i64(), i64()
    n = ctypes.c_int32()
    code = L.or_qubo_parse(text, len(text), cap, _p(ti, _i32p), _p(tj, _i32p), _p(tv, _f64p),
                           ctypes.byref(nt), ctypes.byref(n), ctypes.byref(nl),
                           ctypes.byref(nq))
    if code:
        return code, QP_MESSAGES[code], None
    k = nt.value
    return 0, None, {"i": ti[:k].copy(), "j": tj[:k].copy(), "v": tv[:k].copy(), "n": n.value,
                     "num_linear": nl.value, "num_quadratic": nq.value}
